"""Seeded initial states for the benchmark workloads.

Patterns follow the reference drivers: a random thermal perturbation with a
sin(pi Z) envelope (nhqg/solver.py:1359-1361) and a barotropic (T_0
coefficient) vorticity annulus scaled to a target rms velocity
(scripts/run_polar.py:115-158).  numpy's default_rng makes them
reproducible.

thermal_noise_state: white noise -> rfft2 (Hermitian by construction) ->
amplitude filter k exp(-k^2 / (2 k_c^2)) (power spectrum proportional to
k^2 exp(-k^2/k_c^2)) -> 2/3 mask -> horizontal pattern of rms
``amplitude`` -> times sin(pi Z) at the CGL nodes -> Chebyshev coefficients
-> Dirichlet tau projection -> Galerkin coefficients.  w = q' = 0.
"""

from __future__ import annotations

import numpy as np

from nhqg_amd.grid import Grid
from nhqg_amd.solver import State, sanitize_state

K_ONSET = 1.3048  # onset wavenumber k_c of the stress-free problem


def _weights(grid: Grid) -> np.ndarray:
    w = np.full(grid.Nk, 2.0)
    w[0] = 1.0
    if grid.Nx % 2 == 0:
        w[-1] = 1.0
    return w[None, :]


def thermal_noise_state(grid: Grid, seed: int = 0, amplitude: float = 1e-3,
                        k_c: float = K_ONSET) -> State:
    Nx = grid.Nx
    rng = np.random.default_rng(seed)
    nh = np.fft.rfft2(rng.standard_normal((Nx, Nx)))
    filt = np.sqrt(grid.ksq) * np.exp(-grid.ksq / (2.0 * k_c ** 2))
    pat = np.fft.irfft2(nh * filt * grid.mask_23, s=(Nx, Nx))
    pat *= amplitude / pat.std()
    ph = np.fft.rfft2(pat) * grid.mask_23
    # Chebyshev coefficients of the sin(pi Z) envelope, with the ones far
    # below roundoff (they decay like (pi/2)^n / n! and reach the subnormal
    # range at Nz >= 128) set to zero
    env = grid.proj_dirichlet @ (grid.V_inv @ np.sin(np.pi * grid.Z))
    env[np.abs(env) < 1e-30 * np.abs(env).max()] = 0.0
    cheb = env[:, None, None] * ph[None]
    th = np.tensordot(grid.dirichlet_pinv, cheb, axes=1)
    Nq, Ng = grid.Nz + 1, grid.Nz - 1
    q = np.zeros((Nq, Nx, grid.Nk), np.complex128)
    w = np.zeros((Ng, Nx, grid.Nk), np.complex128)
    return sanitize_state(State(q, w, th, np.zeros(Nq)), grid)


def rms_velocity(q0: np.ndarray, grid: Grid) -> float:
    """rms of (u, v) = (-psi_y, psi_x) for one level's q spectrum."""
    psi = -q0 * grid.inv_denom
    e = np.sum(_weights(grid) * grid.ksq * np.abs(psi) ** 2) / float(grid.Nx) ** 4
    return float(np.sqrt(e))


def add_barotropic_flow(state: State, grid: Grid, seed: int = 1, u_rms: float = 15.0,
                        k_peak: float | None = None, k_width: float | None = None) -> State:
    """Add a barotropic vorticity annulus (T_0 coefficient of q') with rms
    velocity u_rms; k_peak defaults to k_c/4, k_width to k_peak/2."""
    Nx = grid.Nx
    k_peak = K_ONSET / 4.0 if k_peak is None else k_peak
    k_width = 0.5 * k_peak if k_width is None else k_width
    rng = np.random.default_rng(seed)
    zh = np.fft.rfft2(rng.standard_normal((Nx, Nx)))
    k = np.sqrt(grid.ksq)
    ann = np.exp(-0.5 * ((k - k_peak) / k_width) ** 2)
    ann[0, 0] = 0.0
    zh = zh * ann * grid.mask_23
    zh *= u_rms / rms_velocity(zh, grid)
    q = np.array(state.q_hat, copy=True)
    q[0] += zh
    return sanitize_state(State(q, state.w_hat, state.th_hat, state.th_bar), grid)


def barotropic_vorticity_state(grid: Grid, seed: int, k_peak: float, k_width: float,
                                amplitude: float, confine_radius: float | None = None) -> State:
    """The driver's annular barotropic vorticity initial state
    (scripts/run_polar.py:115-158): white noise filtered by a Gaussian
    annulus |k| ~ k_peak (width k_width, k = 0 removed) and the 2/3 band,
    optionally tapered to a polar cap of radius confine_radius around the
    domain centre (tanh ramp two initial wavelengths wide, band-limited
    again), scaled to standard deviation `amplitude` (over the cap when
    confined), placed in the T_0 (barotropic) coefficient of q; w, theta and
    th_bar zero."""
    Nx, Nz = grid.Nx, grid.Nz
    L = grid.cfg.L
    rng = np.random.default_rng(seed)
    zhat = np.fft.rfft2(rng.standard_normal((Nx, Nx)))
    kx = 2.0 * np.pi * np.fft.fftfreq(Nx, d=L / Nx)[:, None]
    ky = 2.0 * np.pi * np.arange(Nx // 2 + 1)[None, :] / L
    ann = np.exp(-0.5 * ((np.sqrt(kx ** 2 + ky ** 2) - k_peak) / k_width) ** 2)
    ann[0, 0] = 0.0
    kxi = np.rint(np.fft.fftfreq(Nx, d=1.0 / Nx)).astype(int)[:, None]
    kyi = np.arange(Nx // 2 + 1)[None, :]
    band = (np.abs(kxi) <= Nx // 3) & (kyi <= Nx // 3)
    zeta = np.fft.irfft2(zhat * ann * band, s=(Nx, Nx))
    if confine_radius is not None:
        xg = np.arange(Nx) * L / Nx
        X, Y = np.meshgrid(xg, xg, indexing="ij")
        r = np.hypot(X - L / 2.0, Y - L / 2.0)
        ramp = 2.0 * 2.0 * np.pi / k_peak
        zeta = zeta * 0.5 * (1.0 - np.tanh((r - confine_radius) / ramp))
        zeta = np.fft.irfft2(np.fft.rfft2(zeta) * band, s=(Nx, Nx))
        zeta *= amplitude / zeta[r < confine_radius].std()
    else:
        zeta *= amplitude / zeta.std()
    q = np.zeros((Nz + 1, Nx, Nx // 2 + 1), dtype=np.complex128)
    q[0] = np.fft.rfft2(zeta)
    zg = np.zeros((Nz - 1, Nx, Nx // 2 + 1), dtype=np.complex128)
    return State(q, zg, zg.copy(), np.zeros(Nz + 1))


def polar_config_256(**kw):
    """Plain FP64 production step on one MI355X: Nx=Ny=256, Nz=256,
    Ra=100, Pr=1, box 24 onset wavelengths, polar trap gamma=2.2e-3
    (r* = 0.4 L, A_d = 20), Rayleigh sponge (sigma_max = 50, r_s = 0.45 L,
    A_s = 30), flux advection, 2/3 rule, evolve_mean + balanced_sbp2_pc with
    4 corrector substeps, nu = 1, dt = 2.5e-5."""
    from nhqg_amd.config import NHQGConfig

    L = 24.0 * 2.0 * np.pi / K_ONSET
    d = dict(Nx=256, Nz=256, L=L, Ra_tilde=100.0, sigma=1.0, gamma=2.2e-3,
             trap_r_star=0.4 * L, trap_sharpness=20.0, sponge_rate=50.0,
             sponge_r_start=0.45 * L, sponge_sharpness=30.0, dt=2.5e-5, nu_q=1.0, nu_w=1.0,
             nu_theta=1.0, hyper_order=1, float_dtype="float64", nonlinear_advection="flux",
             horizontal_dealiasing="23_rule", thermal_closure="evolve_mean",
             mean_exchange_discretization="balanced_sbp2_pc", sbp_corrector_substeps=4,
             w_bc_top="dirichlet")
    d.update(kw)
    return NHQGConfig(**d)


def cpu_config_64(**kw):
    """CPU-runnable plain case: Nx=64, Nz=32, box 10 onset wavelengths,
    Ra=100, no trap, dt=5e-5, evolve_mean + balanced_sbp2_pc (4 substeps)."""
    from nhqg_amd.config import NHQGConfig

    L = 10.0 * 2.0 * np.pi / K_ONSET
    d = dict(Nx=64, Nz=32, L=L, Ra_tilde=100.0, sigma=1.0, dt=5e-5, nu_q=1.0, nu_w=1.0,
             nu_theta=1.0, hyper_order=1, float_dtype="float64", nonlinear_advection="flux",
             horizontal_dealiasing="23_rule", thermal_closure="evolve_mean",
             mean_exchange_discretization="balanced_sbp2_pc", sbp_corrector_substeps=4,
             w_bc_top="dirichlet")
    d.update(kw)
    return NHQGConfig(**d)


def polar_state_256(grid: Grid, seed: int = 0) -> State:
    """Seeded Hermitian thermal state of amplitude 1e-2 plus a low-k
    barotropic flow of rms velocity 15."""
    s = thermal_noise_state(grid, seed=seed, amplitude=1e-2)
    return add_barotropic_flow(s, grid, seed=seed + 1, u_rms=15.0)


def upsample_spectral(f_hat: np.ndarray, nx_new: int) -> np.ndarray:
    """Horizontal spectral (sinc) upsampling of a stack of rfft2 spectra
    (levels, Nx, Nx//2+1) to (levels, nx_new, nx_new//2+1), the method of
    scripts/upsample_checkpoint_horiz.py:27-59: back to the old grid, full
    FFT, the old Nyquist row and column split in half between +N/2 and
    -N/2, zero padding around the centred spectrum, and the factor
    (nx_new/Nx)^2 of the unnormalised forward FFT.  Grid values at the
    co-located points are unchanged."""
    f_hat = np.asarray(f_hat)
    nl, nx, nk = f_hat.shape
    if nk != nx // 2 + 1 or nx_new < nx or (nx_new - nx) % 2:
        raise ValueError("upsample_spectral: need an rfft2 stack and nx_new >= Nx, same parity")
    full = np.fft.fftshift(np.fft.fft2(np.fft.irfft2(f_hat, s=(nx, nx), axes=(1, 2)),
                                       axes=(1, 2)), axes=(1, 2))
    ext = np.zeros((nl, nx + 1, nx + 1), dtype=np.complex128)
    ext[:, :nx, :nx] = full
    # centred index 0 holds the -N/2 (Nyquist) row and column
    ext[:, 0, :nx] *= 0.5
    ext[:, nx, :nx] = ext[:, 0, :nx]
    ext[:, :, 0] *= 0.5
    ext[:, :, nx] = ext[:, :, 0]
    pad = (nx_new - nx) // 2
    big = np.zeros((nl, nx_new, nx_new), dtype=np.complex128)
    big[:, pad:pad + nx + 1, pad:pad + nx + 1] = ext
    out = np.fft.ifftshift(big, axes=(1, 2)) * (nx_new / nx) ** 2
    return np.ascontiguousarray(out[:, :, : nx_new // 2 + 1])


def upsample_state(state: State, grid_new: Grid) -> State:
    """A state on an Nx grid carried to grid_new (same Nz, larger Nx) by
    spectral upsampling of q', w, theta (th_bar is horizontal-mean only),
    then sanitised for the new grid's 2/3 band."""
    nx = grid_new.Nx
    q, w, th = (upsample_spectral(f, nx) for f in (state.q_hat, state.w_hat, state.th_hat))
    return sanitize_state(State(q, w, th, np.array(state.th_bar, dtype=np.float64)), grid_new)


def inject_cyclones(state: State, grid: Grid, amplitude: float, radius: float, count: int = 1,
                    rng: np.random.Generator | None = None) -> State:
    """Gaussian cyclones added to the barotropic PV (the T_0 coefficient of
    q') at random positions, 2/3-masked (scripts/run_polar.py:91-112)."""
    rng = np.random.default_rng() if rng is None else rng
    Nx, L = grid.Nx, grid.L
    x = np.arange(Nx) * L / Nx
    X, Y = np.meshgrid(x, x, indexing="ij")
    patch = np.zeros((Nx, Nx))
    for _ in range(count):
        x0, y0 = rng.uniform(0.0, L, size=2)
        dx = (X - x0 + L / 2.0) % L - L / 2.0
        dy = (Y - y0 + L / 2.0) % L - L / 2.0
        patch += amplitude * np.exp(-(dx ** 2 + dy ** 2) / (2.0 * radius ** 2))
    q = np.array(state.q_hat, dtype=np.complex128)
    q[0] += np.fft.rfft2(patch) * grid.mask_23
    return State(q, state.w_hat, state.th_hat, state.th_bar)


def turbulent_state(grid: Grid, seed: int = 0, u_rms: float = 20.0,
                    theta_rms: float = 0.1) -> State:
    """Turbulent-amplitude seeded state: thermal noise of rms theta_rms plus
    a barotropic flow of rms velocity u_rms."""
    s = thermal_noise_state(grid, seed=seed, amplitude=theta_rms)
    return add_barotropic_flow(s, grid, seed=seed + 1, u_rms=u_rms)
