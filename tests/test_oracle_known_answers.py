"""Closed-form and brute-force checks of the oracle's pieces.

Each check is chosen so that a plausible mistake (dropped term, wrong sign
or index, transposed operand) fails it.  Several restate the reference's
own tests (tests/test_grid.py, test_spectral.py, test_sponge.py,
test_polar.py, test_solver.py of the reference).
"""

import numpy as np
import pytest

from nhqg_amd import NHQGConfig
from oracle import nhqg_oracle as O
from tests import cases


def _cfg(**kw):
    base = dict(Nx=32, Nz=8, L=20.0, Ra_tilde=100.0, dt=1e-3, float_dtype="float64",
                horizontal_dealiasing="23_rule", nonlinear_advection="flux")
    base.update(kw)
    return NHQGConfig(**base)


@pytest.fixture(scope="module")
def ops8():
    return O.build_operators(_cfg(), with_shells=False)


# ---- vertical operators (reference tests/test_grid.py:34-271) ----

def test_V_evaluates_chebyshev(ops8):
    for n in range(ops8.Nz + 1):
        np.testing.assert_allclose(ops8.V[:, n], np.cos(n * np.arccos(ops8.xi)), atol=1e-14)


def test_V_Vinv_roundtrip(ops8):
    I = np.eye(ops8.Nz + 1)
    np.testing.assert_allclose(ops8.V @ ops8.V_inv, I, atol=1e-13)
    np.testing.assert_allclose(ops8.V_inv @ ops8.V, I, atol=1e-13)


@pytest.mark.parametrize("deg", [1, 2, 3, 5, 8])
def test_GZ_differentiates_polynomials(ops8, deg):
    Z = ops8.Z
    a = ops8.V_inv @ Z ** deg
    np.testing.assert_allclose(ops8.V @ (ops8.G_Z @ a), deg * Z ** (deg - 1), atol=1e-10)


def test_GZ_sin():
    o = O.build_operators(_cfg(Nz=32), with_shells=False)
    a = o.V_inv @ np.sin(np.pi * o.Z)
    np.testing.assert_allclose(o.V @ (o.G_Z @ a), np.pi * np.cos(np.pi * o.Z), atol=1e-9)
    np.testing.assert_allclose(o.V @ (o.G_Z @ (o.G_Z @ a)), -np.pi ** 2 * np.sin(np.pi * o.Z),
                               atol=1e-6)


def test_dirichlet_projection_and_stencil(ops8):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(ops8.Nz + 1)
    p = ops8.proj_d @ a
    assert abs((ops8.V @ p)[0]) < 1e-13 and abs((ops8.V @ p)[-1]) < 1e-13
    np.testing.assert_allclose(ops8.proj_d @ p, p, atol=1e-13)  # idempotent
    np.testing.assert_allclose(ops8.d_pinv @ ops8.S_d, np.eye(ops8.Nz - 1), atol=1e-13)
    g = rng.standard_normal(ops8.Nz - 1)
    f = ops8.V @ (ops8.S_d @ g)
    assert abs(f[0]) < 1e-13 and abs(f[-1]) < 1e-13


def test_mixed_stencil_boundary_conditions():
    o = O.build_operators(_cfg(w_bc_top="neumann"), with_shells=False)
    S = o.w_stencil
    bottom = (o.V @ S)[-1]             # Z = 0 (xi = -1): Dirichlet
    top_slope = (o.V @ (o.G_Z @ S))[0]  # Z = 1: Neumann
    assert np.abs(bottom).max() < 1e-12 and np.abs(top_slope).max() < 1e-10
    np.testing.assert_allclose(o.w_pinv @ S, np.eye(o.Nz - 1), atol=1e-12)


def test_cc_quadrature_exact_for_polynomials(ops8):
    for d in range(0, ops8.Nz + 1):
        assert abs(np.sum(ops8.cc * ops8.Z ** d) - 1.0 / (d + 1)) < 1e-13


def test_sbp_operators():
    o = O.build_operators(_cfg(Nz=16), with_shells=False)
    z = o.Z_sbp
    # D1 exact on linear functions; SBP property H D1 + (H D1)^T = diag(-1, 0.., 1)
    np.testing.assert_allclose(o.sbp_D1 @ (3 * z + 1), 3.0, atol=1e-12)
    Q = o.sbp_H @ o.sbp_D1
    Bnd = np.zeros_like(Q)
    Bnd[0, 0], Bnd[-1, -1] = -1.0, 1.0
    np.testing.assert_allclose(Q + Q.T, Bnd, atol=1e-12)
    np.testing.assert_allclose((o.sbp_L @ z ** 2)[1:-1], 2.0, atol=1e-9)
    # transfers are row-stochastic interpolations, exact on linear functions
    np.testing.assert_allclose(o.c2s.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(o.c2s @ (2 * o.Z - 1), 2 * z - 1, atol=1e-13)
    np.testing.assert_allclose(o.s2c @ (2 * z - 1), 2 * o.Z - 1, atol=1e-13)


# ---- horizontal products (reference tests/test_spectral.py, test_sponge.py) ----

def _xy(N, L):
    x = np.arange(N) * (L / N)
    return np.meshgrid(x, x, indexing="ij")


@pytest.fixture(scope="module")
def ops64():
    return O.build_operators(_cfg(Nx=64, L=2 * np.pi), with_shells=False)


def test_flux_analytic_jacobian(ops64):
    X, Y = _xy(64, 2 * np.pi)
    psi = np.fft.rfft2(np.sin(X))[None]
    f = np.fft.rfft2(np.sin(Y))[None]
    (A,) = O.flux23(psi, [f], ops64)
    np.testing.assert_allclose(np.fft.irfft2(A[0], s=(64, 64)), np.cos(X) * np.cos(Y), atol=1e-12)
    f2 = np.fft.rfft2(np.sin(3 * Y))[None]
    psi2 = np.fft.rfft2(np.sin(2 * X))[None]
    (A2,) = O.flux23(psi2, [f2], ops64)
    np.testing.assert_allclose(np.fft.irfft2(A2[0], s=(64, 64)),
                               6 * np.cos(2 * X) * np.cos(3 * Y), atol=1e-11)


def test_flux_equals_jacobian_and_antisymmetry(ops64):
    X, Y = _xy(64, 2 * np.pi)
    a = np.fft.rfft2(np.sin(X) + 0.3 * np.cos(Y))[None]
    b = np.fft.rfft2(np.cos(2 * X - Y))[None]
    (F,) = O.flux23(a, [b], ops64)
    (J,) = O.jacobian23(a, [b], ops64)
    (Jba,) = O.jacobian23(b, [a], ops64)
    (Jaa,) = O.jacobian23(a, [a], ops64)
    np.testing.assert_allclose(F, J, atol=1e-9)
    np.testing.assert_allclose(J + Jba, 0.0, atol=1e-9)
    assert np.abs(Jaa).max() < 1e-9


def test_flux_brute_force_dft():
    """Direct O(N^4) DFT sums of div(u f, v f) at N=8, no FFT library."""
    N, L = 8, 5.0
    o = O.build_operators(_cfg(Nx=N, Nz=4, L=L), with_shells=False)
    rng = np.random.default_rng(1)
    x = np.arange(N) * (L / N)
    kx = 2 * np.pi * np.fft.fftfreq(N, d=L / N)
    ky = 2 * np.pi * np.arange(N // 2 + 1) / L
    psi_p = rng.standard_normal((N, N))
    f_p = rng.standard_normal((N, N))
    E = np.exp(-1j * np.outer(kx, x))             # (k, x) forward kernel along x
    Ey = np.exp(-1j * np.outer(ky, x))

    def dft2(a):  # rfft2 by direct sums
        return np.einsum("kx,ly,xy->kl", E, Ey, a)

    def idft_full(ah):  # inverse of the full spectrum built from the half plane
        full = np.zeros((N, N), complex)
        kyf = 2 * np.pi * np.fft.fftfreq(N, d=L / N)
        for i in range(N):
            for j in range(N):
                jj = j if j <= N // 2 else N - j
                ii = i if j <= N // 2 else (-i) % N
                v = ah[ii, jj]
                full[i, j] = v if j <= N // 2 else np.conj(v)
        Ex = np.exp(1j * np.outer(x, kx))
        Eyf = np.exp(1j * np.outer(x, kyf))
        return np.real(np.einsum("xk,yl,kl->xy", Ex, Eyf, full)) / N ** 2

    ph, fh = dft2(psi_p) * o.mask, dft2(f_p) * o.mask
    ph[:, -1] = ph[:, -1].real  # self-conjugate columns of a real field
    u = idft_full(-1j * ky[None] * ph)
    v = idft_full(1j * kx[:, None] * ph)
    ff = idft_full(fh)
    want = (1j * kx[:, None] * dft2(u * ff) + 1j * ky[None] * dft2(v * ff)) * o.mask
    (got,) = O.flux23(ph[None], [fh[None]], o)
    np.testing.assert_allclose(got[0], want, atol=1e-10 * np.abs(want).max())


def test_sponge_matches_numpy_product():
    cfg = _cfg(Nx=128, Nz=8, Ra_tilde=50.0, dt=1e-4, sponge_rate=3.0, sponge_r_start=6.0,
               sponge_sharpness=8.0)
    o = O.build_operators(cfg, with_shells=False)
    q = np.fft.rfft2(np.random.default_rng(0).standard_normal((cfg.Nz + 1, 128, 128)))
    X, Y = _xy(128, cfg.L)
    r = np.hypot(X - cfg.L / 2, Y - cfg.L / 2)
    sig = 3.0 * 0.5 * (1 + np.tanh(8.0 * (r - 6.0) / 6.0))
    sig = np.fft.irfft2(np.fft.rfft2(sig) * o.mask, s=(128, 128))
    want = np.stack([np.fft.rfft2(sig * np.fft.irfft2(l, s=(128, 128))) * o.mask for l in q])
    np.testing.assert_allclose(O.sponge23(q, o), want, atol=1e-12, rtol=1e-12)


def test_trap_eta_analytic_and_guard():
    cfg = _cfg(Nx=128, Nz=8, Ra_tilde=50.0, gamma=2e-3, trap_sharpness=6.0)
    o = O.build_operators(cfg, with_shells=False)
    X, Y = _xy(128, cfg.L)
    r = np.hypot(X - cfg.L / 2, Y - cfg.L / 2)
    rs = 0.45 * cfg.L / 2
    want = -0.5 * 2e-3 * r ** 2 * 0.5 * (1 - np.tanh(6.0 * (r - rs) / rs))
    # under the 2/3 rule the stored eta is the band-limited analytic field
    want_hat = np.fft.rfft2(want) * o.mask
    assert np.abs(o.eta_hat - want_hat).max() < 1e-12 * np.abs(want_hat).max()
    got = np.fft.irfft2(o.eta_hat, s=(128, 128))
    assert np.abs(got - want).max() < 1e-4 * np.abs(want).max()
    with pytest.raises(ValueError, match="under-resolved"):
        O.build_operators(_cfg(Nx=16, gamma=1e-3, trap_sharpness=50.0), with_shells=False)
    with pytest.raises(ValueError, match="mutually exclusive"):
        O.build_operators(_cfg(gamma=1e-3, beta=0.1), with_shells=False)


# ---- horizontal means: Parseval == physical mean (solver.py:231-245) ----

def test_parseval_mean_equals_physical_mean():
    o = O.build_operators(_cfg(Nx=16, Nz=4), with_shells=False)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((5, 16, 9)) + 1j * rng.standard_normal((5, 16, 9))
    b = rng.standard_normal((5, 16, 9)) + 1j * rng.standard_normal((5, 16, 9))
    phys = O.hmean(a, b, o)
    neg = (-np.arange(16)) % 16

    def herm(f):
        f = f.copy()
        for c in (0, -1):
            f[:, :, c] = 0.5 * (f[:, :, c] + np.conj(f[:, neg, c]))
        return f

    A, B = herm(a), herm(b)
    w = O.rfft_weight(9)
    pars = np.sum(w * (A.real * B.real + A.imag * B.imag), axis=(1, 2)) / 16.0 ** 4
    np.testing.assert_allclose(pars, phys, rtol=1e-12, atol=1e-15)


# ---- implicit stage: block elimination == dense per-mode solve ----

@pytest.mark.parametrize("w_bc_top", ["dirichlet", "neumann"])
def test_implicit_solve_equals_dense_block_solve(w_bc_top):
    cfg = _cfg(Nx=16, Nz=8, nu_q=0.7, nu_w=0.5, nu_theta=0.3, hyper_order=1, w_bc_top=w_bc_top)
    o = O.build_operators(cfg)
    rng = np.random.default_rng(4)
    Nq, Ng = o.Nz + 1, o.Nz - 1
    shp = lambda n: (n, 16, 9)
    Rq = rng.standard_normal(shp(Nq)) + 1j * rng.standard_normal(shp(Nq))
    Rw = rng.standard_normal(shp(Ng)) + 1j * rng.standard_normal(shp(Ng))
    Rth = rng.standard_normal(shp(Ng)) + 1j * rng.standard_normal(shp(Ng))
    q, w, th = O.implicit_solve(Rq, Rw, Rth, o)
    al = o.gamma_imex * cfg.dt
    Pw = o.w_pinv @ o.proj_w
    wth = np.eye(Ng) if o.map_w_to_th is None else o.map_w_to_th
    for (i, j) in [(1, 0), (3, 2), (0, 5), (15, 8)]:
        idn = o.inv_denom[i, j]
        # y - al * (I(y) - diss y) = R for y = (q, w, theta)
        M = np.zeros((Nq + 2 * Ng, Nq + 2 * Ng))
        M[:Nq, :Nq] = (1 + al * o.diss_q[i, j]) * np.eye(Nq)
        M[:Nq, Nq:Nq + Ng] = -al * (o.G_Z @ o.w_stencil)
        M[Nq:Nq + Ng, :Nq] = -al * idn * (Pw @ o.G_Z)
        M[Nq:Nq + Ng, Nq:Nq + Ng] = (1 + al * o.diss_w[i, j]) * np.eye(Ng)
        M[Nq:Nq + Ng, Nq + Ng:] = -al * o.Ra_sigma * (Pw @ o.S_d)
        M[Nq + Ng:, Nq:Nq + Ng] = -al * wth
        M[Nq + Ng:, Nq + Ng:] = (1 + al * o.diss_th[i, j]) * np.eye(Ng)
        rhs = np.concatenate([Rq[:, i, j], Rw[:, i, j], Rth[:, i, j]])
        y = np.linalg.solve(M, rhs)
        got = np.concatenate([q[:, i, j], w[:, i, j], th[:, i, j]])
        np.testing.assert_allclose(got, y, rtol=1e-10, atol=1e-10 * np.abs(y).max())


# ---- stepping: boundary conditions, band limit, ghost, time order ----

def test_step_keeps_bcs_zero_mean_and_band():
    cfg = cases.fixture_cfg()
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    for _ in range(5):
        s = O.imex_step(s, o)
    wn = O.vmat(o.V, O.vmat(o.w_stencil, s.w))
    tn = O.vmat(o.V, O.vmat(o.S_d, s.th))
    for f in (wn, tn):
        assert np.abs(f[0]).max() < 1e-12 and np.abs(f[-1]).max() < 1e-12
    assert abs((o.V @ s.thb)[0]) < 1e-14 and abs((o.V @ s.thb)[-1]) < 1e-14
    for f in (s.q, s.w, s.th):
        assert np.abs(f[:, 0, 0]).max() < 1e-14
        assert np.abs(f * (1 - o.mask)[None]).max() == 0.0


def test_ghost_is_invisible():
    cfg = cases.fixture_cfg()
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    clean = O.imex_step(s, o)
    rng = np.random.default_rng(99)
    neg = (-np.arange(cfg.Nx)) % cfg.Nx

    def poison(f):
        f = f.copy()
        r = rng.normal(size=f[:, :, 0].shape) + 1j * rng.normal(size=f[:, :, 0].shape)
        g = 0.5 * (r - np.conj(r[:, neg]))
        g[:, 0] = 0.0
        # the kx-Nyquist row is excluded as in the reference's test: there the
        # anti-Hermitian part is imaginary and i*kx makes it a visible real value
        g[:, cfg.Nx // 2] = 0.0
        f[:, :, 0] += 1e-3 * g
        return f

    dirty = O.imex_step(O.OState(poison(s.q), poison(s.w), poison(s.th), s.thb), o)
    for a, b in ((clean.q, dirty.q), (clean.w, dirty.w), (clean.th, dirty.th)):
        assert np.abs(a - b).max() < 1e-12 * max(np.abs(a).max(), 1e-30) + 1e-18
    assert np.abs(clean.thb - dirty.thb).max() < 1e-16


def test_second_order_in_time():
    """Richardson ratio of ARS222 errors (reference tests/test_solver.py:85-134)."""
    base = dict(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, float_dtype="float64",
                horizontal_dealiasing="23_rule", nonlinear_advection="flux",
                thermal_closure="evolve_mean", mean_exchange_discretization="balanced_sbp2_pc",
                sbp_corrector_substeps=2)
    n, dtr = 8, 1e-5
    runs = {}
    for m in (1, 2, 4):
        cfg = NHQGConfig(**base, dt=dtr * m)
        o = O.build_operators(cfg)
        s = O.jax_initial_state(o, 42, 1e-2, True)
        for _ in range(n * 4 // m):
            s = O.imex_step(s, o)
        runs[m] = s.q
    e1 = np.abs(runs[4] - runs[1]).max()
    e2 = np.abs(runs[2] - runs[1]).max()
    assert 2.0 < e1 / e2 < 8.0
