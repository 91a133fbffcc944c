"""ORACLE -- CPU float64 restatement of the NHQG ARS(2,2,2) time step.

This file is test infrastructure.  It is imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker of
the HIP path; the product package (nhqg_amd) never imports it.

It restates, in numpy/scipy float64, the reference algorithm read as text:

  operators       nhqg/grid.py:148-342, 349-432, 439-698
  horizontal ops  nhqg/spectral.py:18-74, 187-358
  step            nhqg/solver.py:45-155, 204-245, 346-374, 389-424,
                  510-762, 873-1053, 1167-1242, 1249-1336
  diagnostics     nhqg/diagnostics.py:39-110, 190-212, 334-369, 583-710

Scope: ``vertical_dealiasing='none'``, ``q_boundary='none'``, both
``horizontal_dealiasing`` rules, both ``w_bc_top`` values, both advection
forms, trap and sponge, ``imex_scheme='ars222'`` with ``fixed_conduction``,
``evolve_mean`` + ``legacy`` exchange, or ``evolve_mean`` +
``balanced_sbp2_pc`` (the production branch), ``sbp_transfer_mode=
'interp'``; the RK4 validator for all three.

Pinning (tests/test_oracle_*.py): the reference's golden 20-step state
pins the whole step in the near-linear regime; the reference's own tests
restated here (mean-temperature eigenmode solve, analytic mean-flux RHS,
Nyquist dealiasing of the mean flux, the heat flux diagnostic vs the solver
flux path, the SBP exchange monitor, conservation of the nonlinear
transfers, the heat-flux response of the mean) pin the legacy and
diagnostic branches; analytic Jacobians, flux==Jacobian, numpy sponge
product, analytic trap eta, closed-form operator identities, closed-form
KE/max_speed of single Fourier modes, the SBP substep energy identity, the
agreement of the balanced_sbp2_pc exchange with the legacy Chebyshev
exchange as dt -> 0, Richardson ratios on q, theta and th_bar, dense
per-mode block solves and Parseval identities pin the rest.  Functions with
no such check: none on the step path or in the diagnostics.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np
import scipy.fft as sfft
import scipy.linalg as sla

_WORKERS = int(os.environ.get("NHQG_ORACLE_WORKERS", "1"))


def set_workers(n: int) -> None:
    global _WORKERS
    _WORKERS = max(1, int(n))


def _rfft2(a):
    return sfft.rfft2(a, axes=(-2, -1), workers=_WORKERS)


def _irfft2(a, n):
    return sfft.irfft2(a, s=(n, n), axes=(-2, -1), workers=_WORKERS)


# ---------------------------------------------------------------------------
# Operators (nhqg/grid.py)
# ---------------------------------------------------------------------------

def cheb_diff_coeff(N: int) -> np.ndarray:
    """d/dxi on Chebyshev coefficients by the backward recurrence
    b_{k} = b_{k+2} + 2(k+1) a_{k+1}, b_0 halved (grid.py:148-169)."""
    G = np.zeros((N + 1, N + 1))
    for n in range(N + 1):
        a = np.zeros(N + 1)
        a[n] = 1.0
        b = np.zeros(N + 1)
        if N >= 1:
            b[N - 1] = 2.0 * N * a[N]
            for k in range(N - 2, 0, -1):
                b[k] = b[k + 2] + 2.0 * (k + 1) * a[k + 1]
            b[0] = (0.5 * b[2] if N >= 2 else 0.0) + a[1]
        G[:, n] = b
    return G


def clenshaw_curtis(N: int) -> np.ndarray:
    """CC weights on [0, 1] at the N+1 CGL nodes (grid.py:176-189)."""
    th = np.pi * np.arange(N + 1) / N
    w = np.zeros(N + 1)
    for j in range(N + 1):
        s = 0.0
        for k in range(1, N // 2 + 1):
            b = 1.0 if k == N // 2 else 2.0
            s += b * np.cos(2.0 * k * th[j]) / (4.0 * k * k - 1.0)
        cj = 2.0 if j in (0, N) else 1.0
        w[j] = (1.0 - s) / (N * cj)
    return w


def dct1_pair(N: int):
    """V[j,n] = cos(pi j n / N) and its DCT-I inverse (grid.py:192-201)."""
    j = np.arange(N + 1, dtype=np.float64)
    V = np.cos(np.outer(j, j) * (np.pi / N))
    c = np.ones(N + 1)
    c[0] = c[N] = 2.0
    Vi = (2.0 / N) * np.outer(1.0 / c, 1.0 / c) * V.T
    return V, Vi


def linear_interp_matrix(src: np.ndarray, dst: np.ndarray) -> np.ndarray:
    """Row-stochastic piecewise-linear interpolation src -> dst; src may be
    ascending or descending (grid.py:229-258)."""
    desc = bool(src[0] > src[-1])
    s = src[::-1] if desc else src
    M = np.zeros((dst.size, s.size))
    for i, x in enumerate(dst):
        if x <= s[0]:
            M[i, 0] = 1.0
        elif x >= s[-1]:
            M[i, -1] = 1.0
        else:
            j = int(np.searchsorted(s, x, side="right") - 1)
            t = (x - s[j]) / (s[j + 1] - s[j])
            M[i, j] = 1.0 - t
            M[i, j + 1] = t
    return np.ascontiguousarray(M[:, ::-1]) if desc else M


def sbp2(N: int):
    """Second-order SBP norm H, first derivative D1, Laplacian L with zero
    boundary rows on the uniform grid (grid.py:261-286)."""
    dz = 1.0 / N
    h = np.ones(N + 1)
    h[0] = h[-1] = 0.5
    H = dz * np.diag(h)
    D1 = np.zeros((N + 1, N + 1))
    D1[0, 0], D1[0, 1] = -1.0 / dz, 1.0 / dz
    for i in range(1, N):
        D1[i, i - 1], D1[i, i + 1] = -0.5 / dz, 0.5 / dz
    D1[N, N - 1], D1[N, N] = -1.0 / dz, 1.0 / dz
    Lp = np.zeros((N + 1, N + 1))
    for i in range(1, N):
        Lp[i, i - 1] = Lp[i, i + 1] = 1.0 / dz ** 2
        Lp[i, i] = -2.0 / dz ** 2
    return H, D1, Lp


def tau_projection(rows: np.ndarray, N: int) -> np.ndarray:
    """Identity except rows N-1, N, which re-solve a_{N-1}, a_N so that the
    two constraint rows annihilate the result (grid.py:293-319)."""
    Minv = np.linalg.inv(rows[:, N - 1:N + 1])
    R = Minv @ rows[:, :N - 1]
    P = np.eye(N + 1)
    P[N - 1, :] = 0.0
    P[N, :] = 0.0
    P[N - 1, :N - 1] = -R[0]
    P[N, :N - 1] = -R[1]
    return P


def shen_mixed_stencil(N: int) -> np.ndarray:
    """phi_n = T_n + a_n T_{n+1} + b_n T_{n+2}, phi(-1)=0, phi'(+1)=0
    (grid.py:322-342)."""
    S = np.zeros((N + 1, N - 1))
    for n in range(N - 1):
        b = -(n ** 2 + (n + 1) ** 2) / ((n + 1) ** 2 + (n + 2) ** 2)
        S[n, n] = 1.0
        S[n + 1, n] = 1.0 + b
        S[n + 2, n] = b
    return S


def _left_inverse(S: np.ndarray) -> np.ndarray:
    n = S.shape[1]
    P = np.zeros((n, S.shape[0]))
    P[:, :n] = np.linalg.inv(S[:n, :])
    return P


def mask23(Nx: int) -> np.ndarray:
    K = Nx // 3
    i = np.arange(Nx)
    kxi = np.where(i <= Nx // 2, i, i - Nx)
    kyi = np.arange(Nx // 2 + 1)
    return ((np.abs(kxi)[:, None] <= K) & (kyi[None, :] <= K)).astype(np.float64)


def shell_inverses(ops, cfg):
    """Per-|k|^2 shell inverses of the buoyancy-eliminated w operator
    (grid.py:349-432), q_boundary='none'."""
    g, dt = ops.gamma_imex, cfg.dt
    Ng = ops.Nz - 1
    ksq_flat = ops.ksq.ravel()
    uniq, inv_idx = np.unique(np.round(ksq_flat, decimals=8), return_inverse=True)
    p = cfg.hyper_order
    Ra_s = cfg.Ra_tilde / cfg.sigma
    B0 = ops.w_pinv @ ops.G_Z @ ops.G_Z @ ops.w_stencil
    Kb = ops.buoyancy_K
    mats = np.zeros((uniq.size, Ng, Ng))
    I = np.eye(Ng)
    for s, k2 in enumerate(uniq):
        kp = k2 ** p
        a_q = 1.0 + g * dt * (cfg.nu_q * kp + cfg.drag)
        a_w = 1.0 + g * dt * cfg.nu_w * kp
        a_t = 1.0 + g * dt * (cfg.nu_theta / cfg.sigma) * kp
        if Kb is None:
            couple = (a_w - (g * dt) ** 2 * Ra_s / a_t) * I
        else:
            couple = a_w * I - (g * dt) ** 2 * (Ra_s / a_t) * Kb
        den = k2 + ops.Ld_inv_sq
        if den == 0.0:
            mats[s] = np.linalg.inv(couple)
        else:
            # q_solve = I / alpha_q for q_boundary='none'
            Bk = B0 / a_q
            mats[s] = np.linalg.inv(couple - (g * dt) ** 2 * (1.0 / den) * Bk)
    return mats, inv_idx.reshape(ops.ksq.shape).astype(np.int64)


def build_operators(cfg, with_shells: bool = True):
    """Restatement of make_grid (grid.py:439-698) for the supported branch."""
    if cfg.horizontal_dealiasing not in ("23_rule", "32_rule"):
        raise ValueError(f"Unsupported horizontal_dealiasing={cfg.horizontal_dealiasing!r}")
    if cfg.vertical_dealiasing != "none":
        raise NotImplementedError("oracle covers vertical_dealiasing='none'")
    if cfg.q_boundary != "none":
        raise NotImplementedError("oracle covers q_boundary='none'")
    if cfg.sbp_transfer_mode != "interp":
        raise NotImplementedError("oracle covers sbp_transfer_mode='interp'")
    Nx, N, L = cfg.Nx, cfg.Nz, cfg.L
    Nk = Nx // 2 + 1
    # grid.Npad is 3Nx/2 under both rules (grid.py:444); the 2/3 rule uses it
    # only for the diagnostics' and the legacy exchange's horizontal means
    o = SimpleNamespace(Nx=Nx, Nk=Nk, Nz=N, L=L, Npad=3 * Nx // 2,
                        rule=cfg.horizontal_dealiasing)
    xi = np.cos(np.pi * np.arange(N + 1) / N)
    o.xi, o.Z = xi, 0.5 * (1.0 + xi)
    o.cc = clenshaw_curtis(N)
    o.G_Z = 2.0 * cheb_diff_coeff(N)
    o.G_Z2 = o.G_Z @ o.G_Z
    o.V, o.V_inv = dct1_pair(N)
    Sd = np.zeros((N + 1, N - 1))
    for j in range(N - 1):
        Sd[j, j], Sd[j + 2, j] = -1.0, 1.0
    o.S_d, o.d_pinv = Sd, _left_inverse(Sd)
    ep = np.ones(N + 1)
    em = (-1.0) ** np.arange(N + 1)
    tau_dir = np.stack([ep, em])
    tau_neu = np.stack([ep @ o.G_Z, em @ o.G_Z])
    o.proj_d = tau_projection(tau_dir, N)
    if cfg.w_bc_top == "dirichlet":
        o.w_stencil, o.w_pinv, o.proj_w = Sd, o.d_pinv, o.proj_d
        o.map_w_to_th = o.map_th_to_w = o.buoyancy_K = None
    elif cfg.w_bc_top == "neumann":
        o.w_stencil = shen_mixed_stencil(N)
        o.w_pinv = _left_inverse(o.w_stencil)
        o.proj_w = tau_projection(np.stack([tau_neu[0], tau_dir[1]]), N)
        o.map_w_to_th = o.d_pinv @ o.w_stencil
        o.map_th_to_w = o.w_pinv @ Sd
        o.buoyancy_K = o.map_th_to_w @ o.map_w_to_th
    else:
        raise ValueError(f"Unsupported w_bc_top={cfg.w_bc_top!r}")
    o.Z_sbp = np.linspace(0.0, 1.0, N + 1)
    o.sbp_H, o.sbp_D1, o.sbp_L = sbp2(N)
    o.c2s = linear_interp_matrix(o.Z, o.Z_sbp)
    o.s2c = linear_interp_matrix(o.Z_sbp, o.Z)

    kx = 2.0 * np.pi * np.fft.fftfreq(Nx, d=L / Nx)
    ky = 2.0 * np.pi * np.arange(Nk) / L
    o.kx, o.ky = kx[:, None], ky[None, :]
    o.ksq = o.kx ** 2 + o.ky ** 2
    o.Ld_inv_sq = 0.0 if cfg.Ld == float("inf") else 1.0 / cfg.Ld ** 2
    den = o.ksq + o.Ld_inv_sq
    o.inv_denom = np.zeros_like(den)
    np.divide(1.0, den, out=o.inv_denom, where=den > 0)
    o.mask = mask23(Nx)

    if cfg.gamma != 0.0 and cfg.beta != 0.0:
        raise ValueError("gamma (polar trap) and beta (flat beta-plane) are mutually exclusive")
    xy = np.arange(Nx) * (L / Nx)
    X, Y = np.meshgrid(xy, xy, indexing="ij")
    r = np.hypot(X - L / 2.0, Y - L / 2.0)
    o.eta_hat = None
    r_star = None
    if cfg.gamma != 0.0:
        r_star = cfg.trap_r_star if cfg.trap_r_star is not None else 0.45 * (L / 2.0)
        trap = 0.5 * (1.0 - np.tanh(cfg.trap_sharpness * (r - r_star) / r_star))
        eh = np.fft.rfft2(-0.5 * cfg.gamma * r ** 2 * trap)
        if np.sum(np.abs(eh * (1 - o.mask)) ** 2) > 1e-10 * np.sum(np.abs(eh) ** 2):
            raise ValueError("Trap background PV is under-resolved")
        o.eta_hat = eh * o.mask if o.rule == "23_rule" else eh
    o.sponge_phys = o.sponge_phys_pad = None
    if cfg.sponge_rate < 0.0:
        raise ValueError("sponge_rate must be >= 0")
    if cfg.sponge_rate > 0.0:
        if cfg.sponge_r_start is not None:
            r_s = cfg.sponge_r_start
        elif cfg.gamma != 0.0:
            r_s = 1.18 * r_star
        else:
            raise ValueError("sponge_r_start is required when gamma == 0")
        sig = cfg.sponge_rate * 0.5 * (1.0 + np.tanh(cfg.sponge_sharpness * (r - r_s) / r_s))
        sh = np.fft.rfft2(sig)
        if np.sum(np.abs(sh * (1 - o.mask)) ** 2) > 1e-10 * np.sum(np.abs(sh) ** 2):
            raise ValueError("Sponge profile is under-resolved")
        if o.rule == "23_rule":
            sh = sh * o.mask
        o.sponge_phys = np.fft.irfft2(sh, s=(Nx, Nx))
        # sigma on the padded grid, zero-padded spectrum, no rescale (grid.py:663-667)
        o.sponge_phys_pad = np.fft.irfft2(zero_pad(sh[None], Nx, o.Npad)[0], s=(o.Npad, o.Npad))

    if cfg.imex_scheme != "ars222":
        raise NotImplementedError("oracle covers imex_scheme='ars222'")
    o.gamma_imex = 1.0 - 1.0 / np.sqrt(2.0)
    p = cfg.hyper_order
    o.diss_q = cfg.nu_q * o.ksq ** p + cfg.drag
    o.diss_w = cfg.nu_w * o.ksq ** p
    o.diss_th = (cfg.nu_theta / cfg.sigma) * o.ksq ** p
    o.inv_alpha_q = 1.0 / (1.0 + o.gamma_imex * cfg.dt * o.diss_q)
    o.inv_alpha_th = 1.0 / (1.0 + o.gamma_imex * cfg.dt * o.diss_th)
    o.Ra_sigma = cfg.Ra_tilde / cfg.sigma
    o.sigma = cfg.sigma
    o.beta = cfg.beta
    o.dt = cfg.dt
    o.mu = cfg.mean_temp_eps_sq
    o.thermal_closure = cfg.thermal_closure
    o.exchange = cfg.mean_exchange_discretization
    if o.exchange not in ("legacy", "balanced_sbp2_pc"):
        raise NotImplementedError("oracle covers the legacy and balanced_sbp2_pc exchanges")
    if o.thermal_closure not in ("fixed_conduction", "evolve_mean"):
        raise ValueError(f"Unsupported thermal_closure={o.thermal_closure!r}")
    o.sigma_mean = cfg.sigma
    # evolve_mean tau solve of (I - alpha G_Z2) with the two Dirichlet rows
    # (solver.py:746-758); kept as the matrix, applied with a dense solve
    a_m = o.gamma_imex * cfg.dt * cfg.mean_temp_eps_sq / cfg.sigma
    Am = np.eye(N + 1) - a_m * o.G_Z2
    Am[N - 1, :] = ep
    Am[N, :] = em
    o.mean_solve_A = Am
    o.substeps = int(cfg.sbp_corrector_substeps)
    o.advection = cfg.nonlinear_advection
    o.cutoff = cfg.vertical_cutoff_n
    if with_shells:
        o.imex_inv, o.ksq_idx = shell_inverses(o, cfg)
    return o


# ---------------------------------------------------------------------------
# Step pieces (nhqg/solver.py)
# ---------------------------------------------------------------------------

@dataclass
class OState:
    q: np.ndarray
    w: np.ndarray
    th: np.ndarray
    thb: np.ndarray

    def copy(self):
        return OState(self.q.copy(), self.w.copy(), self.th.copy(), self.thb.copy())


def vmat(M: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Apply a real vertical matrix to a level-major field (einsum 'ij,j...')."""
    M = np.ascontiguousarray(M)
    if np.iscomplexobj(f):
        fr = np.ascontiguousarray(f).view(np.float64)
        out = (M @ fr.reshape(fr.shape[0], -1)).reshape((M.shape[0],) + fr.shape[1:])
        return out.view(np.complex128)
    return (M @ f.reshape(f.shape[0], -1)).reshape((M.shape[0],) + f.shape[1:])


def zero_pad(f, Nx, Npad):
    """Zero-pad (..., Nx, Nk) rfft2 spectra to (..., Npad, Npad//2+1): rows
    0..Nx/2-1 keep their place, rows Nx/2..Nx-1 move to the top of the padded
    kx range (spectral.py:18-34)."""
    Nk = Nx // 2 + 1
    out = np.zeros(f.shape[:-2] + (Npad, Npad // 2 + 1), dtype=np.complex128)
    out[..., :Nx // 2, :Nk] = f[..., :Nx // 2, :]
    out[..., Npad - Nx // 2:, :Nk] = f[..., Nx // 2:, :]
    return out


def truncate(fp, Nx, Npad):
    """Inverse of zero_pad with the (Npad/Nx)^2 factor (spectral.py:37-50)."""
    Nk = Nx // 2 + 1
    keep = np.concatenate([fp[..., :Nx // 2, :Nk], fp[..., Npad - Nx // 2:, :Nk]], axis=-2)
    return keep * (Npad / Nx) ** 2


def padded_wavenumbers(o):
    """Padded-grid wavenumbers from the base fundamental (spectral.py:80-88)."""
    k0 = o.ky[0, 1] if o.Nk >= 2 else 0.0
    kxp = k0 * np.fft.fftfreq(o.Npad, d=1.0 / o.Npad)
    kyp = k0 * np.arange(o.Npad // 2 + 1)
    return kxp[:, None], kyp[None, :]


def flux23(psi, fields, o):
    """2/3-rule div(u f, v f) with u=-psi_y, v=psi_x on the Nx grid, output
    masked (spectral.py:307-328)."""
    Nx = o.Nx
    u = _irfft2(-1j * o.ky * psi, Nx)
    v = _irfft2(1j * o.kx * psi, Nx)
    out = []
    for f in fields:
        fp = _irfft2(f, Nx)
        d = 1j * o.kx * _rfft2(u * fp) + 1j * o.ky * _rfft2(v * fp)
        out.append(d * o.mask)
    return out


def jacobian23(psi, fields, o):
    """2/3-rule J[psi, f] (spectral.py:281-304)."""
    Nx = o.Nx
    px = _irfft2(1j * o.kx * psi, Nx)
    py = _irfft2(1j * o.ky * psi, Nx)
    out = []
    for f in fields:
        fx = _irfft2(1j * o.kx * f, Nx)
        fy = _irfft2(1j * o.ky * f, Nx)
        out.append(_rfft2(px * fy - py * fx) * o.mask)
    return out


def flux32(psi, fields, o):
    """3/2-rule div(u f, v f): padded products, derivatives with the padded
    wavenumbers, truncation (spectral.py:187-212)."""
    Nx, Np = o.Nx, o.Npad
    u = _irfft2(zero_pad(-1j * o.ky * psi, Nx, Np), Np)
    v = _irfft2(zero_pad(1j * o.kx * psi, Nx, Np), Np)
    kxp, kyp = padded_wavenumbers(o)
    out = []
    for f in fields:
        fp = _irfft2(zero_pad(f, Nx, Np), Np)
        d = 1j * kxp * _rfft2(u * fp) + 1j * kyp * _rfft2(v * fp)
        out.append(truncate(d, Nx, Np))
    return out


def jacobian32(psi, fields, o):
    """3/2-rule J[psi, f] with shared psi_x, psi_y (spectral.py:155-184)."""
    Nx, Np = o.Nx, o.Npad
    px = _irfft2(zero_pad(1j * o.kx * psi, Nx, Np), Np)
    py = _irfft2(zero_pad(1j * o.ky * psi, Nx, Np), Np)
    out = []
    for f in fields:
        fx = _irfft2(zero_pad(1j * o.kx * f, Nx, Np), Np)
        fy = _irfft2(zero_pad(1j * o.ky * f, Nx, Np), Np)
        out.append(truncate(_rfft2(px * fy - py * fx), Nx, Np))
    return out


def advection(psi, fields, o):
    """_triple_horizontal_advection (solver.py:389-418)."""
    if o.advection not in ("flux", "jacobian"):
        raise ValueError(f"Unsupported nonlinear_advection={o.advection!r}")
    if o.rule == "23_rule":
        fn = flux23 if o.advection == "flux" else jacobian23
    else:
        fn = flux32 if o.advection == "flux" else jacobian32
    return fn(psi, fields, o)


def sponge23(f, o):
    """rfft2(sigma * irfft2(f)) * mask per level (spectral.py:53-60)."""
    return _rfft2(o.sponge_phys[None] * _irfft2(f, o.Nx)) * o.mask


def sponge32(f, o):
    """Padded sigma*f, truncated (spectral.py:62-74)."""
    fp = _irfft2(zero_pad(f, o.Nx, o.Npad), o.Npad)
    return truncate(_rfft2(o.sponge_phys_pad[None] * fp), o.Nx, o.Npad)


def sponge_tendency(f, o):
    """_sponge_tendency (solver.py:421-424)."""
    return sponge23(f, o) if o.rule == "23_rule" else sponge32(f, o)


def hmean(a, b, o, npad=None):
    """<a b>_xy per level (horizontal_mean_from_nodal_spectral,
    solver.py:231-245): irfft2 on the Nx grid when npad is None or Nx, else
    on the zero-padded grid with both factors scaled by (Npad/Nx)^2."""
    if npad is None or npad == o.Nx:
        return np.mean(_irfft2(a, o.Nx) * _irfft2(b, o.Nx), axis=(1, 2))
    sc = (npad / o.Nx) ** 2
    ap = sc * _irfft2(zero_pad(a, o.Nx, npad), npad)
    bp = sc * _irfft2(zero_pad(b, o.Nx, npad), npad)
    return np.mean(ap * bp, axis=(1, 2))


def _closure(o, base):
    # base=True is the reference's base_grid inside balanced_sbp2_pc:
    # fixed_conduction with the legacy exchange (solver.py:1173-1176)
    return "fixed_conduction" if base else o.thermal_closure


def explicit_rhs(s: OState, o, base: bool = False) -> OState:
    """explicit_rhs (solver.py:510-599).  Under evolve_mean the legacy
    exchange adds -dTheta_bar/dZ * w to E_theta and -eps^2 d<w theta>/dZ as
    E_th_bar (solver.py:575-590); split exchanges (balanced_sbp2_pc) leave
    both to their corrector (:552-556)."""
    psi = -s.q * o.inv_denom[None]
    psi_n = vmat(o.V, psi)
    q_n = vmat(o.V, s.q)
    w_n = vmat(o.V, vmat(o.w_stencil, s.w))
    th_n = vmat(o.V, vmat(o.S_d, s.th))
    q_adv = q_n if o.eta_hat is None else q_n + o.eta_hat[None]
    Aq_n, Aw_n, Ath_n = advection(psi_n, [q_adv, w_n, th_n], o)
    Eq = -vmat(o.V_inv, Aq_n) - 1j * o.beta * o.kx[None] * psi
    Ew = vmat(o.w_pinv, vmat(o.proj_w, -vmat(o.V_inv, Aw_n)))
    Etb = np.zeros_like(s.thb)
    if _closure(o, base) == "evolve_mean" and o.exchange == "legacy":
        g = o.V @ (o.G_Z @ s.thb)
        prod = vmat(o.V_inv, g[:, None, None] * w_n)
        Eth = vmat(o.d_pinv, vmat(o.proj_d, -vmat(o.V_inv, Ath_n) - prod))
        flux = hmean(w_n, th_n, o, o.Npad)          # horizontal_mean_wtheta (solver.py:204-228)
        Etb = o.mu * (-(o.G_Z @ (o.V_inv @ flux)))  # mean_flux_exchange_rhs_coeffs (:324-329)
    else:
        Eth = vmat(o.d_pinv, vmat(o.proj_d, -vmat(o.V_inv, Ath_n)))
    if o.sponge_phys is not None:
        Eq = Eq - sponge_tendency(s.q, o)
        Ew = Ew - sponge_tendency(s.w, o)
        Eth = Eth - sponge_tendency(s.th, o)
    return OState(Eq, Ew, Eth, Etb)


def implicit_tendency(s: OState, o, base: bool = False) -> OState:
    """solver.py:615-641; I_th_bar = (eps^2/sigma) G_Z2 th_bar under evolve_mean."""
    dw = vmat(o.G_Z, vmat(o.w_stencil, s.w))
    Iw_c = o.inv_denom[None] * vmat(o.G_Z, s.q) + o.Ra_sigma * vmat(o.S_d, s.th)
    Iw = vmat(o.w_pinv, vmat(o.proj_w, Iw_c))
    Ith = s.w if o.map_w_to_th is None else vmat(o.map_w_to_th, s.w)
    if _closure(o, base) == "evolve_mean":
        Itb = (o.mu / o.sigma_mean) * (o.G_Z2 @ s.thb)
    else:
        Itb = np.zeros_like(s.thb)
    return OState(dw, Iw, Ith, Itb)


def shell_apply(mats, idx, f):
    """Per-mode application of the per-shell matrices (solver.py:648-675),
    evaluated shell by shell instead of by a full gather."""
    Nl, Nx, Nk = f.shape
    ff = f.reshape(Nl, Nx * Nk)
    out = np.zeros_like(ff)
    flat = idx.ravel()
    order = np.argsort(flat, kind="stable")
    bounds = np.searchsorted(flat[order], np.arange(mats.shape[0] + 1))
    for s in range(mats.shape[0]):
        cols = order[bounds[s]:bounds[s + 1]]
        if cols.size:
            out[:, cols] = mats[s] @ ff[:, cols]
    return out.reshape(Nl, Nx, Nk)


def implicit_solve(Rq, Rw, Rth, o):
    """imex_implicit_solve with the scalar q stage (solver.py:678-738)."""
    gdt = o.gamma_imex * o.dt
    rth = Rth * o.inv_alpha_th[None]
    if o.map_th_to_w is not None:
        rth = vmat(o.map_th_to_w, rth)
    Rw_eff = Rw + gdt * o.Ra_sigma * rth
    temp = Rq * o.inv_alpha_q[None]
    dtemp = vmat(o.G_Z, temp)
    dtg = vmat(o.w_pinv, vmat(o.proj_w, o.inv_denom[None] * dtemp))
    w_new = shell_apply(o.imex_inv, o.ksq_idx, Rw_eff + gdt * dtg)
    dw = vmat(o.G_Z, vmat(o.w_stencil, w_new))
    q_new = (Rq + gdt * dw) * o.inv_alpha_q[None]
    wth = w_new if o.map_w_to_th is None else vmat(o.map_w_to_th, w_new)
    th_new = (Rth + gdt * wth) * o.inv_alpha_th[None]
    return q_new, w_new, th_new


def mean_temp_solve(R, o, base: bool = False):
    """imex_mean_temp_solve (solver.py:741-758): the Dirichlet projection
    without an evolving mean, else the tau solve of (I - alpha G_Z2)."""
    if _closure(o, base) != "evolve_mean":
        return o.proj_d @ R
    rhs = R.copy()
    rhs[o.Nz - 1] = 0.0
    rhs[o.Nz] = 0.0
    return np.linalg.solve(o.mean_solve_A, rhs)


def mean_temp_solve_base(R, o):
    return mean_temp_solve(R, o, base=True)


def sbp_matrices(w2, sdt, mu, kappa, D1, Lp):
    """A = I - half, B = I + half with half = mu kappa sdt L / 2 +
    mu sdt^2 D1 diag<w^2> D1 / 4 and A's boundary rows set to the identity
    (solver.py:928-942)."""
    half = 0.5 * mu * kappa * sdt * Lp + 0.25 * mu * sdt ** 2 * (D1 @ np.diag(w2) @ D1)
    I = np.eye(D1.shape[0])
    A = I - half
    B = I + half
    A[0, :] = 0.0
    A[-1, :] = 0.0
    A[0, 0] = A[-1, -1] = 1.0
    return A, B


def sbp_substep(th_sbp, thb_sbp, w_sbp, A, B, sdt, mu, D1, mean):
    """One SBP-grid substep before the transfer back (solver.py:885-898):
    implicit mean update from the flux <w theta>, then the fluctuation
    update from the midpoint gradient of the mean, boundary rows zeroed.
    ``mean(a, b)`` is the per-level horizontal mean."""
    flux = mean(w_sbp, th_sbp)
    rhs = B @ thb_sbp - mu * sdt * (D1 @ flux)
    rhs[0] = rhs[-1] = 0.0
    thb_new = np.linalg.solve(A, rhs)
    gh = 0.5 * (D1 @ (thb_sbp + thb_new))
    th_new = th_sbp - sdt * w_sbp * gh[:, None, None]
    th_new[0] = 0.0
    th_new[-1] = 0.0
    thb_new[0] = thb_new[-1] = 0.0
    return th_new, thb_new


def sbp_corrector(s: OState, o, total_dt: float) -> OState:
    """SBP2 thermal corrector with hoisted invariants and n substeps
    (solver.py:909-950, 873-906, 346-358); identity without an evolving
    mean (:917-918)."""
    if o.thermal_closure != "evolve_mean":
        return s
    n = max(1, o.substeps)
    sdt = total_dt / n
    npad = o.Nx if o.rule == "23_rule" else o.Npad
    mean = lambda a, b: hmean(a, b, o, npad)
    w_sbp = vmat(o.c2s, vmat(o.V, vmat(o.w_stencil, s.w)))
    A, B = sbp_matrices(mean(w_sbp, w_sbp), sdt, o.mu, 1.0 / o.sigma, o.sbp_D1, o.sbp_L)
    th, thb = s.th, s.thb
    for _ in range(n):
        th_sbp = vmat(o.c2s, vmat(o.V, vmat(o.S_d, th)))
        thb_sbp = o.c2s @ (o.V @ thb)
        th_new_sbp, thb_new = sbp_substep(th_sbp, thb_sbp, w_sbp, A, B, sdt, o.mu,
                                          o.sbp_D1, mean)
        th = vmat(o.d_pinv, vmat(o.proj_d, vmat(o.V_inv, vmat(o.s2c, th_new_sbp))))
        thb = o.proj_d @ (o.V_inv @ (o.s2c @ thb_new))
    return OState(s.q, s.w, th, thb)


def thermal_correction_tendency(p: OState, c: OState, alpha: float) -> OState:
    """(corrected - predicted) / alpha for theta and th_bar (solver.py:953-962)."""
    return OState(np.zeros_like(p.q), np.zeros_like(p.w), (c.th - p.th) / alpha,
                  (c.thb - p.thb) / alpha)


def hermitian_project(f, Nx):
    neg = (-np.arange(Nx)) % Nx
    f = f.copy()
    for c in ([0, -1] if Nx % 2 == 0 else [0]):
        col = f[:, :, c]
        f[:, :, c] = 0.5 * (col + np.conj(col[:, neg]))
    return f


def sanitize(s: OState, o) -> OState:
    """solver.py:107-146: Hermitian columns, and the 2/3 band under 23_rule."""
    q, w, th = (hermitian_project(f, o.Nx) for f in (s.q, s.w, s.th))
    if o.rule == "23_rule":
        m = o.mask[None]
        q, w, th = q * m, w * m, th * m
    return OState(q, w, th, s.thb)


def vertical_cutoff(s: OState, o) -> OState:
    """solver.py:177-191."""
    n = o.cutoff
    if n is None or n >= o.Nz:
        return s
    keep = (np.arange(o.Nz + 1) <= n).astype(np.float64)[:, None, None]
    wc = vmat(o.proj_w, vmat(o.w_stencil, s.w) * keep)
    tc = vmat(o.proj_d, vmat(o.S_d, s.th) * keep)
    return OState(s.q, vmat(o.w_pinv, wc), vmat(o.d_pinv, tc), s.thb)


def finalize(s: OState, o) -> OState:
    """zero_mean on q, w, theta then the optional cutoff (solver.py:993-999)."""
    q, w, th = s.q.copy(), s.w.copy(), s.th.copy()
    q[:, 0, 0] = 0.0
    w[:, 0, 0] = 0.0
    th[:, 0, 0] = 0.0
    return vertical_cutoff(OState(q, w, th, s.thb), o)


def step_balanced_sbp2_pc(s: OState, o) -> OState:
    """imex_step_balanced_sbp2_pc (solver.py:1167-1225): predictors on the
    base grid, a corrector after each stage, and stage 2 carrying omega*C1."""
    g = o.gamma_imex
    dt = o.dt
    delta = -np.sqrt(2.0) / 2.0
    alpha = g * dt
    omg = dt * (1.0 - g)
    E1 = explicit_rhs(s, o, base=True)
    q1p, w1p, th1p = implicit_solve(s.q + alpha * E1.q, s.w + alpha * E1.w,
                                    s.th + alpha * E1.th, o)
    p1 = OState(q1p, w1p, th1p, mean_temp_solve(s.thb + alpha * E1.thb, o, base=True))
    s1 = sbp_corrector(p1, o, alpha)
    C1 = thermal_correction_tendency(p1, s1, alpha)
    E2 = explicit_rhs(s1, o, base=True)
    I1 = implicit_tendency(s1, o, base=True)
    Rq = s.q + dt * (delta * E1.q + (1 - delta) * E2.q) + omg * I1.q - omg * o.diss_q[None] * s1.q
    Rw = s.w + dt * (delta * E1.w + (1 - delta) * E2.w) + omg * I1.w - omg * o.diss_w[None] * s1.w
    Rth = (s.th + dt * (delta * E1.th + (1 - delta) * E2.th) + omg * I1.th
           - omg * o.diss_th[None] * s1.th + omg * C1.th)
    Rtb = s.thb + dt * (delta * E1.thb + (1 - delta) * E2.thb) + omg * I1.thb + omg * C1.thb
    q2, w2, th2 = implicit_solve(Rq, Rw, Rth, o)
    p2 = OState(q2, w2, th2, mean_temp_solve(Rtb, o, base=True))
    return finalize(sbp_corrector(p2, o, alpha), o)


def step_ars222(s: OState, o) -> OState:
    """imex_step_ars222 (solver.py:1012-1053), fixed_conduction or
    evolve_mean with the legacy exchange."""
    g = o.gamma_imex
    dt = o.dt
    delta = -np.sqrt(2.0) / 2.0
    E1 = explicit_rhs(s, o)
    q1, w1, th1 = implicit_solve(s.q + g * dt * E1.q, s.w + g * dt * E1.w,
                                 s.th + g * dt * E1.th, o)
    s1 = OState(q1, w1, th1, mean_temp_solve(s.thb + g * dt * E1.thb, o))
    E2 = explicit_rhs(s1, o)
    I1 = implicit_tendency(s1, o)
    omg = dt * (1 - g)
    Rq = s.q + dt * (delta * E1.q + (1 - delta) * E2.q) + omg * I1.q - omg * o.diss_q[None] * q1
    Rw = s.w + dt * (delta * E1.w + (1 - delta) * E2.w) + omg * I1.w - omg * o.diss_w[None] * w1
    Rth = s.th + dt * (delta * E1.th + (1 - delta) * E2.th) + omg * I1.th - omg * o.diss_th[None] * th1
    Rtb = s.thb + dt * (delta * E1.thb + (1 - delta) * E2.thb) + omg * I1.thb
    q2, w2, th2 = implicit_solve(Rq, Rw, Rth, o)
    return finalize(OState(q2, w2, th2, mean_temp_solve(Rtb, o)), o)


step_ars222_fixed = step_ars222


def imex_step(s: OState, o) -> OState:
    """imex_step dispatch + sanitize (solver.py:1228-1242)."""
    if o.exchange == "balanced_sbp2_pc":
        new = step_balanced_sbp2_pc(s, o)
    else:
        new = step_ars222(s, o)
    return sanitize(new, o)


def full_rhs(s: OState, o) -> OState:
    """explicit + implicit tendency (solver.py:1249-1258)."""
    E = explicit_rhs(s, o)
    I = implicit_tendency(s, o)
    return OState(E.q + I.q, E.w + I.w, E.th + I.th, E.thb + I.thb)


def apply_bcs(s: OState, o) -> OState:
    """_apply_bcs (solver.py:1261-1273), q_boundary='none'."""
    w = vmat(o.w_pinv, vmat(o.proj_w, vmat(o.w_stencil, s.w)))
    th = vmat(o.d_pinv, vmat(o.proj_d, vmat(o.S_d, s.th)))
    return vertical_cutoff(OState(s.q, w, th, o.proj_d @ s.thb), o)


def rk4_step(s: OState, o) -> OState:
    """rk4_step, the explicit validator (solver.py:1276-1336)."""
    dt = o.dt

    def axpy(a, x, k):
        return OState(x.q + a * k.q, x.w + a * k.w, x.th + a * k.th, x.thb + a * k.thb)

    k1 = full_rhs(s, o)
    k2 = full_rhs(apply_bcs(axpy(0.5 * dt, s, k1), o), o)
    k3 = full_rhs(apply_bcs(axpy(0.5 * dt, s, k2), o), o)
    k4 = full_rhs(apply_bcs(axpy(dt, s, k3), o), o)
    comb = [getattr(s, f) + (dt / 6) * (getattr(k1, f) + 2 * getattr(k2, f) + 2 * getattr(k3, f)
                                         + getattr(k4, f)) for f in ("q", "w", "th", "thb")]
    # post-step dissipation exp(-rate dt), RK4 only (grid.py:682-684)
    q = comb[0] * np.exp(-o.diss_q * dt)[None]
    w = comb[1] * np.exp(-o.diss_w * dt)[None]
    th = comb[2] * np.exp(-o.diss_th * dt)[None]
    for f in (q, w, th):
        f[:, 0, 0] = 0.0                                    # zero_mean
    new = apply_bcs(OState(q, w, th, comb[3]), o)
    return sanitize(new, o)


# ---------------------------------------------------------------------------
# Initial states
# ---------------------------------------------------------------------------

def jax_initial_state(o, seed=0, amplitude=1e-3, partitionable=True) -> OState:
    """make_initial_state (solver.py:1343-1378) with the JAX PRNG restated."""
    from oracle import jax_prng as jp
    Nz1 = o.Nz + 1
    k1, k2 = jp.split(jp.prng_key(seed), 2, partitionable)
    shape = (Nz1, o.Nx, o.Nk)
    qr = jp.normal64(k1, shape, partitionable)
    qi = jp.normal64(k2, shape, partitionable)
    qn = amplitude * (qr + 1j * qi) * np.sin(np.pi * o.Z)[:, None, None]
    q = vmat(o.V_inv, qn)
    q[:, 0, 0] = 0.0
    gal = (o.Nz - 1, o.Nx, o.Nk)
    s = OState(q, np.zeros(gal, complex), np.zeros(gal, complex), np.zeros(Nz1))
    return sanitize(s, o)


# ---------------------------------------------------------------------------
# Diagnostics (nhqg/diagnostics.py)
# ---------------------------------------------------------------------------

def rfft_weight(Nk):
    w = np.ones(Nk)
    if Nk > 2:
        w[1:Nk - 1] = 2.0
    return w[None, :]


def shell_bins(o):
    """_shell_bins (diagnostics.py:56-66): |k| shells of width 2 pi / L."""
    k_mag = np.sqrt(o.ksq)
    dk = 2.0 * np.pi / o.L
    n_bins = int(float(np.sqrt(np.max(o.ksq)) / dk)) + 1
    return k_mag, dk, np.arange(n_bins) * dk + dk / 2


def shell_bin_sum(mode_quantity, o):
    """_shell_bin_sum (diagnostics.py:69-77)."""
    k_mag, dk, k_bins = shell_bins(o)
    spec = np.array([np.sum(np.where((k_mag >= i * dk) & (k_mag < (i + 1) * dk),
                                     mode_quantity, 0.0)) for i in range(k_bins.size)])
    return k_bins, spec


def high_mode_fraction(spec, frac=0.25):
    """diagnostics.py:590-596."""
    n_tail = max(1, int(spec.shape[0] * frac))
    total = np.sum(spec)
    return float(np.sum(spec[-n_tail:]) / total) if total > 0 else 0.0


def sbp_exchange_monitor(s: OState, o) -> dict:
    """compute_sbp_internal_exchange_budget (diagnostics.py:334-369): the
    fluctuation-side and mean-side exchange terms under the SBP norm."""
    keys = ("th_mean_feedback_sum_sbp", "mean_flux_exchange_tendency_sbp",
            "mean_theta_exchange_boundary_sbp", "mean_theta_exchange_residual_sbp",
            "mean_theta_exchange_residual_sbp_rel")
    if o.thermal_closure != "evolve_mean" or o.exchange != "balanced_sbp2_pc":
        return {k: float("nan") for k in keys}
    w_sbp = vmat(o.c2s, vmat(o.V, vmat(o.w_stencil, s.w)))
    th_sbp = vmat(o.c2s, vmat(o.V, vmat(o.S_d, s.th)))
    thb_sbp = o.c2s @ (o.V @ s.thb)
    dthb = o.sbp_D1 @ thb_sbp
    flux = hmean(w_sbp, th_sbp, o, o.Npad)
    fb = -(dthb @ (o.sbp_H @ flux))
    ex = -(thb_sbp @ (o.sbp_H @ (o.sbp_D1 @ flux)))
    bnd = thb_sbp[-1] * flux[-1] - thb_sbp[0] * flux[0]
    res = fb + ex
    return dict(zip(keys, (float(fb), float(ex), float(bnd), float(res),
                           float(res / max(abs(fb) + abs(ex), 1e-300)))))


def diagnostics(s: OState, o) -> dict:
    """compute_diagnostics' scalar set, vertical and horizontal spectra and
    the SBP exchange monitor (diagnostics.py:597-710).  Nusselt_dealiased
    uses the solver-consistent flux profile whole (diagnostics.py:206-212);
    the reference re-sums it from its per-shell parts, equal to roundoff."""
    Nx = o.Nx
    norm = float(Nx) ** 4
    wr = rfft_weight(o.Nk)
    psi_n = vmat(o.V, -s.q * o.inv_denom[None])
    q_n = vmat(o.V, s.q)
    w_c = vmat(o.w_stencil, s.w)
    th_c = vmat(o.S_d, s.th)
    w_n, th_n = vmat(o.V, w_c), vmat(o.V, th_c)
    cc = o.cc
    psi_bt = np.tensordot(cc, psi_n, axes=1)
    KE_bt = 0.5 * np.sum(o.ksq * np.abs(psi_bt) ** 2 * wr) / norm
    psi2 = np.tensordot(cc, np.abs(psi_n) ** 2, axes=1)
    KE_tot = 0.5 * np.sum(o.ksq * psi2 * wr) / norm
    u = _irfft2(-1j * o.ky * psi_n, Nx)
    v = _irfft2(1j * o.kx * psi_n, Nx)
    w_p = _irfft2(w_n, Nx)
    th_p = _irfft2(th_n, Nx)
    max_speed = float(np.max(np.sqrt(u * u + v * v)))
    q2 = np.tensordot(cc, np.abs(q_n) ** 2, axes=1)
    w2 = np.tensordot(cc, np.abs(w_n) ** 2, axes=1)
    t2 = np.tensordot(cc, np.abs(th_n) ** 2, axes=1)
    enstrophy = 0.5 * np.sum(q2 * wr) / norm
    wth = np.tensordot(cc, np.real(w_n * np.conj(th_n)), axes=1)
    vol_avg_tw = float(np.sum(wth * wr) / norm)
    if o.exchange == "balanced_sbp2_pc":          # sbp2_flux_profile_nodal (solver.py:370-374)
        prof = o.s2c @ hmean(vmat(o.c2s, w_n), vmat(o.c2s, th_n), o, o.Npad)
    else:                                          # horizontal_mean_wtheta (solver.py:204-228)
        prof = hmean(w_n, th_n, o, o.Npad)
    vol_dealiased = float(np.sum(cc * prof))
    q_spec = np.sum(np.abs(s.q) ** 2, axis=(1, 2))
    w_spec = np.sum(np.abs(w_c) ** 2, axis=(1, 2))
    th_spec = np.sum(np.abs(th_c) ** 2, axis=(1, 2))
    k_bins, q_h = shell_bin_sum(q2 * wr, o)
    _, w_h = shell_bin_sum(w2 * wr, o)
    _, th_h = shell_bin_sum(t2 * wr, o)
    out = {
        "KE_tot": float(KE_tot), "KE_bt": float(KE_bt), "KE_bc": float(KE_tot - KE_bt),
        "max_speed": max_speed, "max_w": float(np.max(np.abs(w_p))),
        "max_theta": float(np.max(np.abs(th_p))), "max_tw": float(np.max(np.abs(w_p * th_p))),
        "enstrophy": float(enstrophy),
        "Nusselt": 1.0 + vol_avg_tw, "Nusselt_dealiased": 1.0 + vol_dealiased,
        "vol_avg_tw": vol_avg_tw, "vol_avg_tw_dealiased": vol_dealiased,
        "heat_flux_mismatch": vol_dealiased - vol_avg_tw,
        "flux_profile_dealiased": prof,
        "q_rms": float(np.sqrt(np.sum(q2 * wr) / norm)),
        "w_rms": float(np.sqrt(np.sum(w2 * wr) / norm)),
        "th_rms": float(np.sqrt(np.sum(t2 * wr) / norm)),
        "q_vert_spec": q_spec, "w_vert_spec": w_spec, "th_vert_spec": th_spec,
        "q_high_frac": high_mode_fraction(q_spec), "w_high_frac": high_mode_fraction(w_spec),
        "th_high_frac": high_mode_fraction(th_spec),
        "k_bins": k_bins, "q_horiz_spec": q_h, "w_horiz_spec": w_h, "th_horiz_spec": th_h,
        "th_bar_max": float(np.max(np.abs(s.thb))),
    }
    out.update(sbp_exchange_monitor(s, o))
    return out
