"""Host-side operator construction: ``make_grid(cfg) -> Grid``.

Builds, in float64 numpy, the vertical Chebyshev/Galerkin/SBP operators, the
horizontal wavenumber tables, the trap and sponge fields and the per-|k|^2
shell inverses of the reference's ``make_grid`` (nhqg/grid.py:439-789), then
the device tables of the MI355X step: fused vertical operators, the
band-compact mode layout, shell groupings and FFT twiddles.  The device
context is created lazily on first use.

Supported branch (anything else raises): float64, ``horizontal_dealiasing``
either ``'23_rule'`` (band-compact layout) or ``'32_rule'`` (full layout, products
on the zero-padded 3Nx/2 grid), ``vertical_dealiasing='none'``,
``q_boundary='none'``, ``nonlinear_advection`` either ``'flux'`` or
``'jacobian'``, ``imex_scheme='ars222'``, ``sbp_transfer_mode='interp'``,
``thermal_closure`` either ``'fixed_conduction'`` or ``'evolve_mean'``,
``mean_exchange_discretization`` either ``'balanced_sbp2_pc'`` or ``'legacy'``,
either ``w_bc_top``, any ``vertical_cutoff_n``, power-of-two ``Nx`` in [16, 1024],
``Nz`` <= 256.
"""

from __future__ import annotations

import os
import weakref
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from nhqg_amd.config import NHQGConfig

# ---------------------------------------------------------------------------
# vertical operators (nhqg/grid.py:148-342)
# ---------------------------------------------------------------------------


def cheb_derivative(N: int) -> np.ndarray:
    """d/dxi on Chebyshev coefficients in closed form:
    (T_n)' = 2n sum_{k<n, n-k odd} T_k / c_k,  c_0 = 2, c_k = 1 otherwise.
    Same operator as the recurrence of nhqg/grid.py:148-169."""
    n = np.arange(N + 1)
    k = n[:, None]
    G = np.where((n[None, :] > k) & ((n[None, :] - k) % 2 == 1), 2.0 * n[None, :], 0.0)
    G[0, :] *= 0.5
    return G.astype(np.float64)


def cc_weights(N: int) -> np.ndarray:
    """Clenshaw-Curtis weights on [0, 1] (nhqg/grid.py:176-189)."""
    th = np.pi * np.arange(N + 1) / N
    kk = np.arange(1, N // 2 + 1)
    b = np.where(kk == N // 2, 1.0, 2.0)
    s = (b[None, :] * np.cos(2.0 * kk[None, :] * th[:, None]) / (4.0 * kk[None, :] ** 2 - 1.0)).sum(1)
    c = np.ones(N + 1)
    c[0] = c[N] = 2.0
    return (1.0 - s) / (N * c)


def cgl_transforms(N: int):
    """V[j, n] = T_n(xi_j) = cos(pi j n / N) and its DCT-I inverse
    (nhqg/grid.py:192-201)."""
    idx = np.arange(N + 1, dtype=np.float64)
    V = np.cos(np.pi / N * np.outer(idx, idx))
    c = np.ones(N + 1)
    c[0] = c[N] = 2.0
    Vinv = (2.0 / N) * (V.T / c[:, None]) / c[None, :]
    return V, Vinv


def interp_matrix(src: np.ndarray, dst: np.ndarray) -> np.ndarray:
    """Piecewise-linear interpolation from nodes src (any monotone order) to
    points dst; rows sum to one (nhqg/grid.py:229-258)."""
    order = np.argsort(src)
    s = src[order]
    M = np.zeros((dst.size, src.size))
    j = np.clip(np.searchsorted(s, dst, side="right") - 1, 0, s.size - 2)
    t = (dst - s[j]) / (s[j + 1] - s[j])
    lo = dst <= s[0]
    hi = dst >= s[-1]
    for i in range(dst.size):
        if lo[i]:
            M[i, order[0]] = 1.0
        elif hi[i]:
            M[i, order[-1]] = 1.0
        else:
            M[i, order[j[i]]] = 1.0 - t[i]
            M[i, order[j[i] + 1]] = t[i]
    return M


def sbp2_operators(N: int):
    """Uniform-grid second-order SBP norm, D1 and Laplacian with zero boundary
    rows (nhqg/grid.py:261-286)."""
    dz = 1.0 / N
    n = N + 1
    H = dz * np.diag(np.r_[0.5, np.ones(n - 2), 0.5])
    D1 = (np.diag(np.full(n - 1, 0.5), 1) - np.diag(np.full(n - 1, 0.5), -1)) / dz
    D1[0, :2] = [-1.0 / dz, 1.0 / dz]
    D1[N, N - 1:] = [-1.0 / dz, 1.0 / dz]
    Lp = (np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1) - 2.0 * np.eye(n)) / dz ** 2
    Lp[0, :] = 0.0
    Lp[N, :] = 0.0
    return H, D1, Lp


def tau_projection(rows: np.ndarray, N: int) -> np.ndarray:
    """Projection that re-solves a_{N-1}, a_N from the other coefficients so
    both constraint rows vanish (nhqg/grid.py:293-319)."""
    C = -np.linalg.solve(rows[:, N - 1:], rows[:, :N - 1])
    P = np.eye(N + 1)
    P[N - 1:, :] = 0.0
    P[N - 1:, :N - 1] = C
    return P


def dirichlet_stencil(N: int) -> np.ndarray:
    """Basis -T_j + T_{j+2} (nhqg/grid.py:497-502)."""
    S = np.zeros((N + 1, N - 1))
    j = np.arange(N - 1)
    S[j, j] = -1.0
    S[j + 2, j] = 1.0
    return S


def mixed_stencil(N: int) -> np.ndarray:
    """Bottom-Dirichlet / top-Neumann Shen basis (nhqg/grid.py:322-342)."""
    S = np.zeros((N + 1, N - 1))
    n = np.arange(N - 1, dtype=np.float64)
    b = -(n ** 2 + (n + 1) ** 2) / ((n + 1) ** 2 + (n + 2) ** 2)
    j = np.arange(N - 1)
    S[j, j] = 1.0
    S[j + 1, j] = 1.0 + b
    S[j + 2, j] = b
    return S


def unique_left_inverse(S: np.ndarray) -> np.ndarray:
    """Exact left inverse from the leading square block (nhqg/grid.py:503-506)."""
    g = S.shape[1]
    P = np.zeros((g, S.shape[0]))
    P[:, :g] = np.linalg.inv(S[:g, :])
    return P


def mask_23(Nx: int) -> np.ndarray:
    K = Nx // 3
    kxi = np.fft.fftfreq(Nx, d=1.0 / Nx).round().astype(int)
    kyi = np.arange(Nx // 2 + 1)
    return ((np.abs(kxi)[:, None] <= K) & (kyi[None, :] <= K)).astype(np.float64)


def zero_pad(f: np.ndarray, Nx: int, Npad: int) -> np.ndarray:
    """An (..., Nx, Nx/2+1) rfft2 spectrum on the (..., Npad, Npad/2+1) grid:
    rows kx = 0..Nx/2-1 keep their index, rows kx = -Nx/2..-1 sit at
    Npad-Nx/2..Npad-1, every other entry zero (nhqg/spectral.py:18-34)."""
    h = Nx // 2
    out = np.zeros(f.shape[:-2] + (Npad, Npad // 2 + 1), dtype=np.complex128)
    out[..., :h, :h + 1] = f[..., :h, :]
    out[..., Npad - h:, :h + 1] = f[..., h:, :]
    return out


# ---------------------------------------------------------------------------
# shell inverses (nhqg/grid.py:349-432)
# ---------------------------------------------------------------------------


def _invert_stack(A: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    if n == 0:
        return A.copy()
    workers = max(1, min(16, os.cpu_count() or 1))
    chunk = max(1, -(-n // (4 * workers)))
    parts = [(i, min(n, i + chunk)) for i in range(0, n, chunk)]
    out = np.empty_like(A)

    def run(se):
        s, e = se
        out[s:e] = np.linalg.inv(A[s:e])

    # one BLAS thread per worker: nested BLAS threading oversubscribes the
    # cores by orders of magnitude here
    try:
        from threadpoolctl import threadpool_limits
        limit = threadpool_limits(1)
    except Exception:  # threadpoolctl absent: serial BLAS calls in one thread
        limit = None
        workers = 1
    try:
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(run, parts))
    finally:
        if limit is not None:
            limit.restore_original_limits()
    return out


def shell_coefficients(ksq_vals, cfg, gamma_imex, Ld_inv_sq, neumann: bool):
    """Per-shell scalars of A'(k) = alpha I - beta K - c B0 (w's Galerkin
    basis, buoyancy eliminated, scalar q stage for q_boundary='none';
    nhqg/grid.py:395-429).  For the Dirichlet top K = I is folded into
    alpha and beta = 0."""
    g = gamma_imex * cfg.dt
    kp = ksq_vals ** cfg.hyper_order
    a_q = 1.0 + g * (cfg.nu_q * kp + cfg.drag)
    a_w = 1.0 + g * cfg.nu_w * kp
    a_t = 1.0 + g * (cfg.nu_theta / cfg.sigma) * kp
    Ra_s = cfg.Ra_tilde / cfg.sigma
    den = ksq_vals + Ld_inv_sq
    ck = np.where(den == 0.0, 0.0, 1.0 / np.where(den == 0.0, 1.0, den))
    c = g * g * ck / a_q
    if neumann:
        return a_w, g * g * Ra_s / a_t, c
    return a_w - g * g * Ra_s / a_t, np.zeros_like(a_w), c


def shell_inverses(ksq_vals, cfg, gamma_imex, B0, Kb, Ld_inv_sq) -> np.ndarray:
    """inv(A'(k)) for each listed shell value on the host (numpy/LAPACK)."""
    alpha, beta, c = shell_coefficients(ksq_vals, cfg, gamma_imex, Ld_inv_sq, Kb is not None)
    Ng = B0.shape[0]
    I = np.eye(Ng)
    out = np.empty((ksq_vals.size, Ng, Ng))
    step = 512
    for s in range(0, ksq_vals.size, step):
        e = min(ksq_vals.size, s + step)
        sl = slice(s, e)
        if Kb is None:
            couple = alpha[sl][:, None, None] * I[None]
        else:
            couple = alpha[sl][:, None, None] * I[None] - beta[sl][:, None, None] * Kb[None]
        A = couple - c[sl][:, None, None] * B0[None]
        out[sl] = _invert_stack(A)
    return out


# ---------------------------------------------------------------------------
# Grid
# ---------------------------------------------------------------------------


@dataclass
class Grid:
    """Precomputed operators; attribute names follow the reference's Grid
    (nhqg/grid.py:29-141).  The per-shell implicit factors (the reference's
    ``imex_inv``) are not held here: a device context builds and inverts
    the shells its mode layout touches, on the GPU in FP64
    (``shell_build='device'``) or from host LAPACK inverses ('host')."""

    cfg: NHQGConfig
    Nx: int
    Nk: int
    Nz: int
    Nz_gal: int
    Z: np.ndarray
    xi: np.ndarray
    cc_weights: np.ndarray
    G_Z: np.ndarray
    G_Z2: np.ndarray
    V: np.ndarray
    V_inv: np.ndarray
    dirichlet_stencil: np.ndarray
    dirichlet_pinv: np.ndarray
    w_stencil: np.ndarray
    w_pinv: np.ndarray
    map_w_to_th: np.ndarray | None
    map_th_to_w: np.ndarray | None
    proj_dirichlet: np.ndarray
    proj_neumann: np.ndarray
    proj_w: np.ndarray
    Z_sbp: np.ndarray
    sbp_H: np.ndarray
    sbp_D1: np.ndarray
    sbp_L: np.ndarray
    cgl_to_sbp: np.ndarray
    sbp_to_cgl: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    ksq: np.ndarray
    inv_denom: np.ndarray
    mask_23: np.ndarray
    diss_rate_q: np.ndarray
    diss_rate_w: np.ndarray
    diss_rate_th: np.ndarray
    inv_alpha_q: np.ndarray
    inv_alpha_th: np.ndarray
    ksq_idx: np.ndarray
    eta_hat: np.ndarray | None
    sponge_phys: np.ndarray | None
    sponge_phys_pad: np.ndarray | None
    Npad: int
    beta: float
    Ra_sigma: float
    sigma: float
    L: float
    Ld_inv_sq: float
    dt: float
    gamma_imex: float
    mean_temp_eps_sq: float
    thermal_closure: str
    q_boundary: str
    w_bc_top: str
    nonlinear_advection: str
    vertical_cutoff_n: int | None
    imex_scheme: str
    vertical_dealiasing: str
    horizontal_dealiasing: str
    mean_exchange_discretization: str
    sbp_transfer_mode: str
    sbp_corrector_substeps: int
    buoyancy_K: np.ndarray | None = None
    # 'device': shell factors built and inverted on the GPU (FP64);
    # 'host': numpy/LAPACK inverses uploaded (used to cross-check)
    shell_build: str = "device"
    # Dirichlet-top w: 'eigen' solves every shell in the eigenbasis of B0
    # (two dense contractions, no per-shell matrices); 'matrices' streams
    # the per-shell inverses like the reference
    shell_solver: str = "eigen"
    # 0: plain FP64 step; 4..7: mixed-precision step with this many int8
    # slices per operand in the Chebyshev transforms and implicit-solve maps
    # (set before the first context is created)
    mixed_slices: int = 0
    # True: apply every vertical operator as a dense product, not as the two
    # half-size products its reflection symmetry allows (set before the
    # first context is created)
    dense_contractions: bool = False
    _layouts: dict = field(default_factory=dict, repr=False)
    _contexts: dict = field(default_factory=dict, repr=False)
    _live: "weakref.WeakSet" = field(default_factory=weakref.WeakSet, repr=False)
    _closure: tuple | None = field(default=None, repr=False)

    @property
    def evolve_mean(self) -> bool:
        return self.thermal_closure == "evolve_mean"

    @property
    def rule32(self) -> bool:
        return self.horizontal_dealiasing == "32_rule"

    def layout(self, full: bool = False) -> "ModeLayout":
        # the 3/2 rule keeps every mode: its only layout is the full one
        full = full or self.rule32
        key = "full" if full else "band"
        if key not in self._layouts:
            self._layouts[key] = ModeLayout.build(self, full)
        return self._layouts[key]

    def new_context(self, full: bool = False):
        """A device context owned by the caller (each Stepper holds its own,
        so no other Stepper or per-call function writes its state).  An
        attached closure is applied to it."""
        from nhqg_amd import _native

        full = full or self.rule32
        ctx = _native.Context(self, self.layout(full))
        if self._closure is not None and not full:
            ctx.set_closure(*self._closure)
        self._live.add(ctx)
        return ctx

    def context(self, full: bool = False):
        """The grid's scratch context of one layout, shared by the per-call
        functions (imex_step, rk4_step, compute_diagnostics), which set the
        whole state on every call; created on first use."""
        full = full or self.rule32
        key = "full" if full else "band"
        if key not in self._contexts:
            self._contexts[key] = self.new_context(full)
        return self._contexts[key]

    def attach_closure(self, mlp, n0: int, in_scale: float, out_scale: float) -> None:
        """Attach a trained closure to every band-layout context of this
        grid, present and future (mlp=None detaches)."""
        if mlp is not None and self.rule32:
            raise NotImplementedError("the learned closure is built for the 2/3 rule's "
                                      "band layout")
        self._closure = None if mlp is None else (mlp, int(n0), float(in_scale),
                                                  float(out_scale))
        for ctx in list(self._live):
            if not ctx.layout.full and getattr(ctx, "h", None):
                ctx.set_closure(mlp, int(n0), float(in_scale), float(out_scale))

    def close(self) -> None:
        for ctx in self._contexts.values():
            ctx.close()
        self._contexts.clear()


def _check_config(cfg: NHQGConfig) -> None:
    # Unknown switch values raise ValueError as in the reference
    # (nhqg/grid.py:488, 548, 583, 676; nhqg/solver.py:406, 418, 608, 1241).
    choices = {
        "vertical_dealiasing": ("none", "cheb_3o2", "cheb_2x"),
        "w_bc_top": ("dirichlet", "neumann"),
        "sbp_transfer_mode": ("interp", "mass_adjoint", "weighted_polar"),
        "imex_scheme": ("ars222", "rk443"),
        "nonlinear_advection": ("jacobian", "flux"),
        "horizontal_dealiasing": ("32_rule", "23_rule"),
        "q_boundary": ("none", "neumann"),
        "thermal_closure": ("fixed_conduction", "evolve_mean"),
    }
    for name, allowed in choices.items():
        v = getattr(cfg, name)
        if v not in allowed:
            raise ValueError(f"Unsupported {name}={v!r}")
    built = {
        "float_dtype": "float64",
        "vertical_dealiasing": "none",
        "q_boundary": "none",
        "imex_scheme": "ars222",
        "sbp_transfer_mode": "interp",
    }
    for name, want in built.items():
        if getattr(cfg, name) != want:
            raise NotImplementedError(
                f"the MI355X step is built for {name}={want!r} (got {getattr(cfg, name)!r})")
    if cfg.mean_exchange_discretization not in ("legacy", "balanced_sbp2_pc"):
        raise NotImplementedError(
            "the MI355X step is built for mean_exchange_discretization='balanced_sbp2_pc' or "
            f"'legacy' (got {cfg.mean_exchange_discretization!r})")
    if cfg.Nx < 16 or cfg.Nx > 1024 or cfg.Nx & (cfg.Nx - 1):
        raise NotImplementedError("Nx must be a power of two in [16, 1024]")
    if cfg.Nz < 3 or cfg.Nz > 256:
        raise NotImplementedError("Nz must be in [3, 256]")


def make_grid(cfg: NHQGConfig, shells: str = "all") -> Grid:
    """Build all host operators for ``cfg`` (nhqg/grid.py:439-789).

    shells: accepted for call compatibility ('all', 'band', 'none'); the
    distinct |k|^2 shell values are always recorded and the factors of the
    shells a context touches are built when the context is created."""
    _check_config(cfg)
    Nx, N, L = cfg.Nx, cfg.Nz, cfg.L
    Nk = Nx // 2 + 1
    xi = np.cos(np.pi * np.arange(N + 1) / N)
    Z = 0.5 * (1.0 + xi)
    G_Z = 2.0 * cheb_derivative(N)
    V, Vinv = cgl_transforms(N)
    Sd = dirichlet_stencil(N)
    Pd = unique_left_inverse(Sd)
    ep = np.ones(N + 1)
    em = (-1.0) ** np.arange(N + 1)
    tau_dir = np.vstack([ep, em])
    tau_neu = np.vstack([ep @ G_Z, em @ G_Z])
    proj_d = tau_projection(tau_dir, N)
    proj_n = tau_projection(tau_neu, N)
    if cfg.w_bc_top == "dirichlet":
        Sw, Pw_, proj_w = Sd, Pd, proj_d
        m_w2t = m_t2w = Kb = None
    else:
        Sw = mixed_stencil(N)
        Pw_ = unique_left_inverse(Sw)
        proj_w = tau_projection(np.vstack([tau_neu[0], tau_dir[1]]), N)
        m_w2t = Pd @ Sw
        m_t2w = Pw_ @ Sd
        Kb = m_t2w @ m_w2t
    Z_sbp = np.linspace(0.0, 1.0, N + 1)
    H, D1, Lp = sbp2_operators(N)
    c2s = interp_matrix(Z, Z_sbp)
    s2c = interp_matrix(Z_sbp, Z)

    kx1 = 2.0 * np.pi * np.fft.fftfreq(Nx, d=L / Nx)
    ky1 = 2.0 * np.pi * np.arange(Nk) / L
    ksq = kx1[:, None] ** 2 + ky1[None, :] ** 2
    den = ksq + cfg.Ld_inv_sq
    inv_denom = np.divide(1.0, den, out=np.zeros_like(den), where=den > 0)
    m23 = mask_23(Nx)

    if cfg.gamma != 0.0 and cfg.beta != 0.0:
        raise ValueError("gamma (polar trap) and beta (flat beta-plane) are mutually exclusive")
    xy = np.arange(Nx) * (L / Nx)
    X, Y = np.meshgrid(xy, xy, indexing="ij")
    r = np.hypot(X - L / 2.0, Y - L / 2.0)
    eta_hat = None
    r_star = None
    if cfg.gamma != 0.0:
        r_star = cfg.trap_r_star if cfg.trap_r_star is not None else 0.45 * (L / 2.0)
        trap = 0.5 * (1.0 - np.tanh(cfg.trap_sharpness * (r - r_star) / r_star))
        eta_hat = np.fft.rfft2(-0.5 * cfg.gamma * r ** 2 * trap)
        e_tot = np.sum(np.abs(eta_hat) ** 2)
        e_out = np.sum(np.abs(eta_hat * (1.0 - m23)) ** 2)
        if e_out > 1e-10 * e_tot:
            raise ValueError(
                f"Trap background PV is under-resolved: {e_out / e_tot:.2e} of its spectral "
                f"energy lies outside the 2/3 band (tolerance 1e-10). Decrease trap_sharpness "
                f"(tanh width r_star/A_d = {r_star / cfg.trap_sharpness:.3g} vs dx = "
                f"{L / Nx:.3g}) or increase Nx.")
        if cfg.horizontal_dealiasing == "23_rule":
            eta_hat = eta_hat * m23
    sponge = sponge_pad = None
    Npad = 3 * Nx // 2
    if cfg.sponge_rate < 0.0:
        raise ValueError("sponge_rate must be >= 0")
    if cfg.sponge_rate > 0.0:
        if cfg.sponge_r_start is not None:
            r_s = cfg.sponge_r_start
        elif cfg.gamma != 0.0:
            r_s = 1.18 * r_star
        else:
            raise ValueError("sponge_r_start is required when gamma == 0 "
                             "(no trap radius to derive it from)")
        sig = cfg.sponge_rate * 0.5 * (1.0 + np.tanh(cfg.sponge_sharpness * (r - r_s) / r_s))
        sh = np.fft.rfft2(sig)
        s_tot = np.sum(np.abs(sh) ** 2)
        s_out = np.sum(np.abs(sh * (1.0 - m23)) ** 2)
        if s_out > 1e-10 * s_tot:
            raise ValueError(
                f"Sponge profile is under-resolved: {s_out / s_tot:.2e} of its spectral energy "
                f"lies outside the 2/3 band (tolerance 1e-10). Decrease sponge_sharpness "
                f"(tanh width r_s/A_s = {r_s / cfg.sponge_sharpness:.3g} vs dx = "
                f"{L / Nx:.3g}) or increase Nx.")
        if cfg.horizontal_dealiasing == "23_rule":
            sh = sh * m23
        sponge = np.fft.irfft2(sh, s=(Nx, Nx))
        # sigma on the padded grid from its zero-padded spectrum, without the
        # (Npad/Nx)^2 rescale (the 3/2 rule's sponge product, grid.py:663-667)
        sponge_pad = np.fft.irfft2(zero_pad(sh, Nx, Npad), s=(Npad, Npad))

    gamma_imex = 1.0 - 1.0 / np.sqrt(2.0)
    p = cfg.hyper_order
    dq = cfg.nu_q * ksq ** p + cfg.drag
    dw = cfg.nu_w * ksq ** p
    dth = (cfg.nu_theta / cfg.sigma) * ksq ** p
    iaq = 1.0 / (1.0 + gamma_imex * cfg.dt * dq)
    iath = 1.0 / (1.0 + gamma_imex * cfg.dt * dth)
    _, ksq_idx = np.unique(np.round(ksq.ravel(), decimals=8), return_inverse=True)
    ksq_idx = ksq_idx.reshape(Nx, Nk).astype(np.int32)

    grid = Grid(
        cfg=cfg, Nx=Nx, Nk=Nk, Nz=N, Nz_gal=N - 1, Z=Z, xi=xi, cc_weights=cc_weights(N),
        G_Z=G_Z, G_Z2=G_Z @ G_Z, V=V, V_inv=Vinv, dirichlet_stencil=Sd, dirichlet_pinv=Pd,
        w_stencil=Sw, w_pinv=Pw_, map_w_to_th=m_w2t, map_th_to_w=m_t2w, proj_dirichlet=proj_d,
        proj_neumann=proj_n, proj_w=proj_w, Z_sbp=Z_sbp, sbp_H=H, sbp_D1=D1, sbp_L=Lp,
        cgl_to_sbp=c2s, sbp_to_cgl=s2c, kx=kx1[:, None], ky=ky1[None, :], ksq=ksq,
        inv_denom=inv_denom, mask_23=m23, diss_rate_q=dq, diss_rate_w=dw, diss_rate_th=dth,
        inv_alpha_q=iaq, inv_alpha_th=iath, ksq_idx=ksq_idx, eta_hat=eta_hat,
        sponge_phys=sponge, sponge_phys_pad=sponge_pad, Npad=Npad, beta=float(cfg.beta), Ra_sigma=cfg.Ra_tilde / cfg.sigma,
        sigma=float(cfg.sigma), L=float(L), Ld_inv_sq=cfg.Ld_inv_sq, dt=float(cfg.dt),
        gamma_imex=float(gamma_imex), mean_temp_eps_sq=float(cfg.mean_temp_eps_sq),
        thermal_closure=cfg.thermal_closure, q_boundary=cfg.q_boundary, w_bc_top=cfg.w_bc_top,
        nonlinear_advection=cfg.nonlinear_advection, vertical_cutoff_n=cfg.vertical_cutoff_n,
        imex_scheme=cfg.imex_scheme, vertical_dealiasing=cfg.vertical_dealiasing,
        horizontal_dealiasing=cfg.horizontal_dealiasing,
        mean_exchange_discretization=cfg.mean_exchange_discretization,
        sbp_transfer_mode=cfg.sbp_transfer_mode,
        sbp_corrector_substeps=int(cfg.sbp_corrector_substeps), buoyancy_K=Kb)
    grid._shell_values = np.unique(np.round(ksq.ravel(), decimals=8))
    grid._shell_mode = shells
    return grid


def shell_matrix_B0(grid: Grid) -> np.ndarray:
    """Fixed part of the per-shell operator: w_pinv G_Z G_Z w_stencil."""
    return grid.w_pinv @ grid.G_Z @ grid.G_Z @ grid.w_stencil


# ---------------------------------------------------------------------------
# device mode layout
# ---------------------------------------------------------------------------


@dataclass
class ModeLayout:
    """Band-compact (or full) ky-major mode layout and all device tables."""

    full: bool
    rows: np.ndarray
    nky: int
    tables: dict

    @property
    def nkx(self) -> int:
        return int(self.rows.size)

    @property
    def nm(self) -> int:
        return self.nkx * self.nky

    @staticmethod
    def build(g: Grid, full: bool) -> "ModeLayout":
        Nx, Nk = g.Nx, g.Nk
        K = Nx // 3
        if full:
            rows = np.arange(Nx, dtype=np.int32)
            nky = Nk
        else:
            rows = np.r_[np.arange(K + 1), np.arange(Nx - K, Nx)].astype(np.int32)
            nky = K + 1
        nkx = rows.size
        ky_i = np.repeat(np.arange(nky), nkx)
        kx_i = np.tile(np.arange(nkx), nky)
        row_m = rows[kx_i]
        full_idx = (row_m * Nk + ky_i).astype(np.int32)
        kidx_of_x = np.full(Nx, -1, dtype=np.int32)
        kidx_of_x[rows] = np.arange(nkx, dtype=np.int32)
        sel = (row_m, ky_i)
        self_conj = (ky_i == 0) | ((Nx % 2 == 0) & (ky_i == Nx // 2))
        neg_row = (-row_m) % Nx
        partner = np.where(self_conj, ky_i * nkx + kidx_of_x[neg_row], -1).astype(np.int32)
        m_all = np.arange(nkx * nky)
        pm = partner >= 0
        pairs = np.stack([m_all[pm], partner[pm]], 1)
        pairs = pairs[pairs[:, 0] <= pairs[:, 1]].astype(np.int32)
        wgt = np.where(self_conj, 1.0, 2.0) / float(Nx) ** 4
        # the same means on the zero-padded 3/2 grid (the reference's
        # horizontal means with Npad): there ky = Nx/2 is an ordinary column
        # (weight 2, no partner) and the kx = Nx/2 row of ky = 0 sits at
        # -Nx/2 with no partner (Hermitian part of the column: weight 1/2)
        nyq_row = (Nx % 2 == 0) & (row_m == Nx // 2)
        wgt_pad = np.where(ky_i == 0, np.where(nyq_row, 0.5, 1.0), 2.0) / float(Nx) ** 4
        partner_pad = np.where((ky_i == 0) & ~nyq_row, partner, -1).astype(np.int32)
        # shells touched by the layout
        shell_of_m = g.ksq_idx[sel].astype(np.int64)
        shell_ids, inv = np.unique(shell_of_m, return_inverse=True)
        order = np.argsort(inv, kind="stable")
        counts = np.bincount(inv, minlength=shell_ids.size)
        shell_ptr = np.r_[0, np.cumsum(counts)].astype(np.int32)
        t = dict(
            rows=rows, kidx_of_x=kidx_of_x, full_idx=full_idx,
            kx_row=g.kx[rows, 0].copy(), ky_col=g.ky[0, :nky].copy(),
            kx_mode=g.kx[row_m, 0].copy(), ksq=g.ksq[sel].copy(),
            inv_denom=g.inv_denom[sel].copy(), inv_alpha_q=g.inv_alpha_q[sel].copy(),
            inv_alpha_th=g.inv_alpha_th[sel].copy(), diss_q=g.diss_rate_q[sel].copy(),
            diss_w=g.diss_rate_w[sel].copy(), diss_th=g.diss_rate_th[sel].copy(),
            mask=g.mask_23[sel].copy(), pweight=wgt, partner=partner, pweight_pad=wgt_pad,
            partner_pad=partner_pad,
            herm_pairs=np.ascontiguousarray(pairs.ravel()),
            shell_ids=shell_ids, shell_ptr=shell_ptr,
            shell_modes=order.astype(np.int32),
        )
        if g.rule32:
            # rows of the zero-padded grid (nhqg/spectral.py:18-34): kx row r
            # at r for r < Nx/2, at Npad - Nx + r above
            Np, h = g.Npad, Nx // 2
            rows_pad = np.where(rows < h, rows, rows - Nx + Np).astype(np.int32)
            kidx_pad = np.full(Np, -1, dtype=np.int32)
            kidx_pad[rows_pad] = np.arange(nkx, dtype=np.int32)
            t.update(rows_pad=rows_pad, kidx_of_x_pad=kidx_pad)
        return ModeLayout(full=full, rows=rows, nky=nky, tables=t)


def _eigen_by_parity(B0: np.ndarray):
    """Real eigendecomposition B0 = Q diag(lam) Q^-1 for the eigen-basis shell
    solve.  B0 couples only Galerkin indices of equal parity (the reflection
    z -> -z), so when that holds exactly the decomposition is taken per
    parity block: eigenvectors of pure parity, the even block's first, and
    Q, Q^-1 exact parity blocks.  A full eig() can mix eigenvectors of the two
    blocks whose eigenvalues nearly coincide (at 1e-10 for Nz = 256), which
    hides that structure.  Otherwise the full matrix is decomposed."""
    n = B0.shape[0]
    par = np.arange(n) % 2
    blocks = [np.flatnonzero(par == 0), np.flatnonzero(par == 1)]
    if np.all(B0[par[:, None] != par[None, :]] == 0.0) and blocks[1].size > 0:
        lam = np.zeros(n)
        Q = np.zeros((n, n))
        Qinv = np.zeros((n, n))
        c0 = 0
        for ix in blocks:
            lb, qb = np.linalg.eig(B0[np.ix_(ix, ix)])
            _check_real_spectrum(lb, qb)
            cs = slice(c0, c0 + ix.size)
            lam[cs] = lb.real
            Q[ix, cs] = np.real(qb)
            Qinv[cs, ix] = np.linalg.inv(np.real(qb))
            c0 += ix.size
        return lam, Q, Qinv
    lam, Q = np.linalg.eig(B0)
    _check_real_spectrum(lam, Q)
    Q = np.real(Q).copy()
    return lam.real.copy(), Q, np.linalg.inv(Q)


def _check_real_spectrum(lam, Q):
    if np.abs(lam.imag).max() > 1e-12 * np.abs(lam.real).max() or np.iscomplexobj(Q) and \
            np.abs(Q.imag).max() > 1e-12:
        raise NotImplementedError("B0 has a complex spectrum; use shell_solver='matrices'")


def device_operators(g: Grid, lay: ModeLayout) -> dict:
    """Fused vertical operators and physical fields for the device."""
    V, Vi, GZ = g.V, g.V_inv, g.G_Z
    Sw, Sd = g.w_stencil, g.dirichlet_stencil
    c2s, s2c = g.cgl_to_sbp, g.sbp_to_cgl
    Pw = g.w_pinv @ g.proj_w
    Pd = g.dirichlet_pinv @ g.proj_dirichlet
    d1 = g.sbp_D1
    n = g.Nz + 1

    def band3(M):
        b = np.zeros((n, 3))
        for o in (-1, 0, 1):
            i = np.arange(max(0, -o), min(n, n - o))
            b[i, o + 1] = M[i, i + o]
        resid = M.copy()
        for o in (-1, 0, 1):
            i = np.arange(max(0, -o), min(n, n - o))
            resid[i, i + o] = 0.0
        if np.any(resid != 0.0):
            raise ValueError("SBP operator is not tridiagonal")
        return b

    ops = dict(
        V=V, VSw=V @ Sw, VSd=V @ Sd, Vinv=Vi, Pw=Pw @ Vi, Pd=Pd @ Vi, Qw=Pw @ GZ, GSw=GZ @ Sw,
        W2S=c2s @ V @ Sw, T2S=c2s @ V @ Sd, S2T=Pd @ Vi @ s2c, thb2s=c2s @ V,
        s2thb=g.proj_dirichlet @ Vi @ s2c, proj_d=g.proj_dirichlet, d1_band=band3(d1),
        lap_band=band3(g.sbp_L), cc=g.cc_weights, row_cc=g.cc_weights[None, :].copy(),
        cc_sbp=g.cc_weights @ s2c, Sw=Sw, Sd=Sd,
    )
    if g.w_bc_top == "neumann":
        ops.update(Mtw=Pw @ Sd, Mwt=g.map_w_to_th, Mth2w=g.map_th_to_w)
    Nx = g.Nx
    # the trap and the sponge on the grid of the spectral products: the Nx
    # grid (2/3 rule) or the zero-padded Npad grid (3/2 rule), in irfft2's
    # normalisation on that grid, like the products' other factors
    if g.rule32:
        Np = g.Npad

        def phys(f):
            return np.fft.irfft2(zero_pad(f, Nx, Np), s=(Np, Np))

        sfx = "_pad"
        ops["sponge_phys_pad"] = g.sponge_phys_pad
        k = np.arange(Np)
        ops["twiddle_pad"] = np.stack([np.cos(2.0 * np.pi * k / Np),
                                       -np.sin(2.0 * np.pi * k / Np)], 1)
    else:
        def phys(f):
            return np.fft.irfft2(f, s=(Nx, Nx))

        sfx = ""
        ops["sponge_phys"] = g.sponge_phys
    ops["eta_phys" + sfx] = None if g.eta_hat is None else phys(g.eta_hat)
    if g.eta_hat is not None:
        if g.nonlinear_advection == "jacobian":
            # grad(eta) for J(psi, q + eta) (nhqg/solver.py:534-535 adds eta to
            # q before the Jacobian, nhqg/spectral.py:155-184, 293-298
            # differentiate)
            ops["eta_x_phys" + sfx] = phys(1j * g.kx * g.eta_hat)
            ops["eta_y_phys" + sfx] = phys(1j * g.ky * g.eta_hat)
    # mean temperature (nhqg/solver.py:575-590, 638, 746-758)
    N = g.Nz
    a_m = g.gamma_imex * g.dt * g.mean_temp_eps_sq / g.sigma
    A = np.eye(N + 1) - a_m * g.G_Z2
    A[N - 1, :] = 1.0
    A[N, :] = (-1.0) ** np.arange(N + 1)
    ms = np.linalg.inv(A)
    ms[:, N - 1:] = 0.0          # the tau rows of the right-hand side are dropped
    ops.update(VGZ=V @ GZ, mean_solve=ms,
               mean_diff=(g.mean_temp_eps_sq / g.sigma) * g.G_Z2,
               mean_flux=-g.mean_temp_eps_sq * (GZ @ Vi))
    k = np.arange(Nx)
    ops["twiddle"] = np.stack([np.cos(2.0 * np.pi * k / Nx), -np.sin(2.0 * np.pi * k / Nx)], 1)
    # vertical_cutoff_n (nhqg/solver.py:177-191): keep Chebyshev degrees
    # n <= cutoff of w and theta, re-project onto the boundary conditions
    cut = g.vertical_cutoff_n
    if cut is not None and cut < g.Nz:
        keep = (np.arange(n) <= cut).astype(np.float64)[:, None]
        ops["cutoff_w"] = g.w_pinv @ g.proj_w @ (keep * Sw)
        ops["cutoff_th"] = g.dirichlet_pinv @ g.proj_dirichlet @ (keep * Sd)
    # shell factors of the touched shells: built and inverted on the device
    # in FP64 from B0, K and three scalars per shell (default), or inverted
    # here with LAPACK and uploaded transposed
    vals = g._shell_values[lay.tables["shell_ids"]]
    B0 = shell_matrix_B0(g)
    if g.shell_solver == "eigen" and g.buoyancy_K is None:
        # Dirichlet top: A_k = alpha_k I - c_k B0 for every shell, and B0 has
        # a real spectrum with well-conditioned eigenvectors, so the solve
        # is Q diag(1 / (alpha_k - c_k lam)) Q^-1 with no per-shell matrix
        lam, Q, Qinv = _eigen_by_parity(B0)
        alpha, _, c = shell_coefficients(vals, g.cfg, g.gamma_imex, g.Ld_inv_sq, False)
        ptr, modes = lay.tables["shell_ptr"], lay.tables["shell_modes"]
        ma = np.ones(lay.nm)
        mc = np.zeros(lay.nm)
        for s in range(vals.size):
            ms = modes[ptr[s]:ptr[s + 1]]
            ma[ms] = alpha[s]
            mc[ms] = c[s]
        ops.update(shell_inv=None, eig_Q=Q, eig_Qinv=Qinv, eig_lam=lam,
                   mode_alpha=ma, mode_c=mc)
        return ops
    if g.shell_build == "host":
        inv = shell_inverses(vals, g.cfg, g.gamma_imex, B0, g.buoyancy_K, g.Ld_inv_sq)
        ops["shell_inv"] = np.ascontiguousarray(inv.transpose(0, 2, 1))
    else:
        alpha, beta, c = shell_coefficients(vals, g.cfg, g.gamma_imex, g.Ld_inv_sq,
                                            g.buoyancy_K is not None)
        ops["shell_inv"] = None
        ops["shell_B0"] = B0
        ops["shell_K"] = g.buoyancy_K
        ops["shell_coef"] = np.stack([alpha, beta, c], axis=1)
    return ops
