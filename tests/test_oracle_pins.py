"""Pins for the oracle's mean-temperature, exchange, 3/2-rule and diagnostic
branches.

Three kinds of check, each chosen so that a plausible mistake fails it:

* restatements of the reference's own tests (tests/test_solver.py and
  tests/test_spectral.py of the reference, cited per test) against the
  oracle, with the reference's configurations and tolerances;
* identities the mathematics fixes: the SBP substep's energy balance, exact
  (alias-free) 3/2-rule products by direct convolution, closed-form
  diagnostics of single Fourier modes, second-order convergence of the
  SBP-grid heat flux, and the agreement of the balanced_sbp2_pc step with
  the legacy Chebyshev exchange once the grid transfers are exact;
* planted mistakes (monkeypatched into the oracle, on the CPU) that each of
  those checks must reject: a sign flip of the fluctuation update, a dropped
  omega*C1 stage-2 term, uniform instead of Clenshaw-Curtis weights.
"""

from __future__ import annotations

import contextlib

import numpy as np
import pytest

from nhqg_amd import NHQGConfig
from oracle import nhqg_oracle as O


def _cfg(**kw):
    base = dict(float_dtype="float64")
    base.update(kw)
    return NHQGConfig(**base)


@contextlib.contextmanager
def planted(name, replacement):
    """Temporarily replace an oracle function (a mistake planted on the CPU)."""
    orig = getattr(O, name)
    setattr(O, name, replacement(orig))
    try:
        yield
    finally:
        setattr(O, name, orig)


def flip_theta_update(orig):
    """Sign flip of theta_sbp - dt * w_sbp * g in the SBP substep."""
    def sub(th, thb, w, A, B, sdt, mu, D1, mean):
        t, b = orig(th, thb, w, A, B, sdt, mu, D1, mean)
        t = 2.0 * th - t
        t[0] = 0.0
        t[-1] = 0.0
        return t, b
    return sub


def drop_c1(orig):
    """Stage 2 without the omega * C1 correction tendency."""
    return lambda p, c, a: O.OState(0 * p.q, 0 * p.w, 0 * p.th, 0 * p.thb)


def uniform_weights(orig):
    """Uniform quadrature weights in place of Clenshaw-Curtis."""
    return lambda N: np.full(N + 1, 1.0 / (N + 1))


# ---------------------------------------------------------------------------
# Reference tests restated (tests/test_solver.py of the reference)
# ---------------------------------------------------------------------------

def test_mean_temperature_bcs_preserved():
    """Reference tests/test_solver.py:284-297 (legacy exchange, 3/2 rule)."""
    cfg = _cfg(Nx=32, Nz=8, L=20.0, Ra_tilde=100.0, dt=1e-3, thermal_closure="evolve_mean")
    o = O.build_operators(cfg)
    s = O.jax_initial_state(o, 0, 1e-2)
    for _ in range(10):
        s = O.imex_step(s, o)
    assert abs(s.thb[0]) < 1e-14 and abs(s.thb[-1]) < 1e-14


def test_mean_temperature_responds_to_heat_flux():
    """Reference tests/test_solver.py:299-320: a correlated w, theta seed makes
    th_bar evolve through the legacy explicit E_th_bar."""
    cfg = _cfg(Nx=32, Nz=8, L=20.0, Ra_tilde=100.0, dt=1e-3, thermal_closure="evolve_mean",
               mean_temp_eps_sq=1.0)
    o = O.build_operators(cfg)
    s = O.jax_initial_state(o, 0, 1e-2)
    env = np.sin(np.pi * o.Z)[:, None, None]
    seed = O.vmat(o.d_pinv, O.vmat(o.V_inv, np.ones((cfg.Nz + 1, cfg.Nx, o.Nk)) * env))
    s = O.OState(s.q, seed.astype(complex), 0.5 * seed.astype(complex), s.thb)
    s = O.imex_step(s, o)
    assert np.max(np.abs(s.thb[1:-1])) > 0.0
    # stronger: with th_bar = 0 the first step's th_bar comes from
    # -eps^2 d<w theta>/dZ alone, whose interior is O(dt) and of one sign
    # pattern set by sin^2; its size must be far above roundoff
    assert np.max(np.abs(s.thb[1:-1])) > 1e-6


def test_nonlinear_ke_transfer_sums_to_zero():
    """Reference tests/test_solver.py:396-414 (Jacobian, 3/2 rule): the KE
    shell transfer of the nonlinear q tendency sums to zero.

    Also checked relative to the summed magnitudes of the per-mode terms,
    on the same state with its Nyquist row and column removed: on the
    padded grid the ky = Nx/2 column is not self-conjugate, so with Nyquist
    content the Nx-grid Parseval weights of the diagnostic and the padded
    product's own energy identity differ (by 4e-4 of the terms here)."""
    cfg = _cfg(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, dt=1e-4, nu_q=0.0,
               nu_w=0.0, nu_theta=0.0, drag=0.0, hyper_order=1,
               thermal_closure="fixed_conduction")
    o = O.build_operators(cfg)
    s = O.jax_initial_state(o, 9, 1e-3)
    for _ in range(3):
        s = O.imex_step(s, o)

    def transfer(st):
        E = O.explicit_rhs(st, o)
        psi_n = O.vmat(o.V, -st.q * o.inv_denom[None])
        qnl_n = O.vmat(o.V, E.q)                           # beta = 0: E_q is all nonlinear
        dens = np.tensordot(o.cc, np.real(np.conj(psi_n) * qnl_n), axes=1)
        mode = -(o.ksq * o.inv_denom) * dens * O.rfft_weight(o.Nk) / cfg.Nx ** 4
        return mode, O.shell_bin_sum(mode, o)[1]

    mode, shells = transfer(s)
    assert abs(shells.sum()) < 1e-10
    q = s.q.copy()
    q[:, cfg.Nx // 2, :] = 0.0
    q[:, :, -1] = 0.0
    mode, shells = transfer(O.OState(q, s.w, s.th, s.thb))
    assert abs(shells.sum()) < 1e-12 * np.abs(mode).sum()


def test_w_theta_nonlinear_transfers_sum_to_zero():
    """Reference tests/test_solver.py:468-485: the nonlinear w and theta
    shell transfers sum to zero (fixed conduction: E_theta is advection)."""
    cfg = _cfg(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, dt=1e-4, nu_q=0.0,
               nu_w=0.0, nu_theta=0.0, drag=0.0, hyper_order=1,
               thermal_closure="fixed_conduction")
    o = O.build_operators(cfg)
    s = O.jax_initial_state(o, 10, 1e-3)
    for _ in range(3):
        s = O.imex_step(s, o)
    E = O.explicit_rhs(s, o)
    wr = O.rfft_weight(o.Nk) / cfg.Nx ** 4
    for f, ef, S in ((s.w, E.w, o.w_stencil), (s.th, E.th, o.S_d)):
        fn = O.vmat(o.V, O.vmat(S, f))
        en = O.vmat(o.V, O.vmat(S, ef))
        _, shells = O.shell_bin_sum(np.tensordot(o.cc, np.real(np.conj(fn) * en), axes=1) * wr, o)
        assert abs(shells.sum()) < 1e-10


def test_mean_temp_solve_matches_dirichlet_eigenmode():
    """Reference tests/test_solver.py:491-508: the evolve_mean tau solve of
    (I - alpha d2/dZ2) on sin(pi Z) returns sin(pi Z)/(1 + alpha pi^2)."""
    cfg = _cfg(Nx=16, Nz=16, L=10.0, Ra_tilde=10.0, dt=2e-2, thermal_closure="evolve_mean",
               mean_temp_eps_sq=1.0)
    o = O.build_operators(cfg, with_shells=False)
    nod = np.sin(np.pi * o.Z)
    alpha = o.gamma_imex * cfg.dt * cfg.mean_temp_eps_sq / cfg.sigma
    exact = o.V_inv @ (nod / (1.0 + alpha * np.pi ** 2))
    assert np.abs(O.mean_temp_solve(o.V_inv @ nod, o) - exact).max() < 1e-10
    # the base grid (no evolving mean) is the plain Dirichlet projection
    assert np.abs(O.mean_temp_solve(o.V_inv @ nod, o, base=True)
                  - o.proj_d @ (o.V_inv @ nod)).max() == 0.0


def _mean_flux_cfg(**kw):
    return _cfg(Nx=16, Nz=16, L=10.0, Ra_tilde=10.0, dt=1e-3, thermal_closure="evolve_mean",
                mean_temp_eps_sq=1.0, **kw)


def test_mean_temp_flux_rhs_matches_analytic_derivative():
    """Reference tests/test_solver.py:510-543: horizontally uniform
    w = sin(pi Z), theta = w/2 give <w theta> = sin^2(pi Z)/2 and
    E_th_bar = -d/dZ of it = -(pi/2) sin(2 pi Z)."""
    cfg = _mean_flux_cfg()
    o = O.build_operators(cfg, with_shells=False)
    env_c = o.V_inv @ np.sin(np.pi * o.Z)
    w_c = np.zeros((cfg.Nz + 1, cfg.Nx, o.Nk), complex)
    w_c[:, 0, 0] = cfg.Nx * cfg.Nx * env_c
    s = O.OState(np.zeros_like(w_c), O.vmat(o.d_pinv, w_c), O.vmat(o.d_pinv, 0.5 * w_c),
                 np.zeros(cfg.Nz + 1))
    w_n = O.vmat(o.V, O.vmat(o.w_stencil, s.w))
    th_n = O.vmat(o.V, O.vmat(o.S_d, s.th))
    flux = O.hmean(w_n, th_n, o, o.Npad)
    assert np.abs(flux - 0.5 * np.sin(np.pi * o.Z) ** 2).max() < 1e-12
    E = O.explicit_rhs(s, o)
    assert np.abs(o.V @ E.thb + 0.5 * np.pi * np.sin(2 * np.pi * o.Z)).max() < 1e-10
    # the split (balanced_sbp2_pc) exchange leaves E_th_bar to its corrector
    o2 = O.build_operators(_mean_flux_cfg(mean_exchange_discretization="balanced_sbp2_pc"),
                           with_shells=False)
    assert np.abs(O.explicit_rhs(s, o2).thb).max() == 0.0


def _nyquist_state(o, cfg):
    env = np.sin(np.pi * o.Z)
    x = 2.0 * np.pi * np.arange(cfg.Nx) / cfg.Nx
    phys = env[:, None, None] * np.cos((cfg.Nx // 2) * x)[None, :, None] * np.ones((1, 1, cfg.Nx))
    f = O.vmat(o.d_pinv, O.vmat(o.V_inv, np.fft.rfft2(phys, axes=(1, 2))))
    return O.OState(np.zeros((cfg.Nz + 1, cfg.Nx, o.Nk), complex), f, f.copy(),
                    np.zeros(cfg.Nz + 1))


def test_mean_temp_flux_is_horizontally_dealiased_at_nyquist():
    """Reference tests/test_solver.py:545-579: a kx-Nyquist field squared
    aliases to a constant on the Nx grid; the padded mean gives sin^2/2."""
    cfg = _mean_flux_cfg()
    o = O.build_operators(cfg, with_shells=False)
    s = _nyquist_state(o, cfg)
    w_n = O.vmat(o.V, O.vmat(o.w_stencil, s.w))
    th_n = O.vmat(o.V, O.vmat(o.S_d, s.th))
    expected = 0.5 * np.sin(np.pi * o.Z) ** 2
    assert np.abs(O.hmean(w_n, th_n, o) - expected).max() > 1e-1
    assert np.abs(O.hmean(w_n, th_n, o, o.Npad) - expected).max() < 1e-10


def test_dealiased_heat_flux_matches_solver_flux_path():
    """Reference tests/test_solver.py:606-634, plus its closed form: the
    dealiased profile is sin^2(pi Z)/2, whose Clenshaw-Curtis integral is
    1/4 to spectral accuracy.  Uniform weights must fail the closed form."""
    cfg = _mean_flux_cfg()
    o = O.build_operators(cfg, with_shells=False)
    s = _nyquist_state(o, cfg)
    w_n = O.vmat(o.V, O.vmat(o.w_stencil, s.w))
    th_n = O.vmat(o.V, O.vmat(o.S_d, s.th))
    expected = float(np.sum(o.cc * O.hmean(w_n, th_n, o, o.Npad)))
    d = O.diagnostics(s, o)
    assert abs(d["vol_avg_tw_dealiased"] - expected) < 1e-12
    assert abs(d["Nusselt_dealiased"] - (1.0 + expected)) < 1e-12
    assert abs(d["vol_avg_tw_dealiased"] - 0.25) < 1e-9
    with planted("clenshaw_curtis", uniform_weights):
        bad = O.diagnostics(s, O.build_operators(cfg, with_shells=False))
    assert abs(bad["vol_avg_tw_dealiased"] - 0.25) > 1e-3


def _exchange_monitor_state(o, cfg):
    rng = np.random.default_rng(321)
    env = np.sin(np.pi * o.Z)
    wp = env[:, None, None] * rng.standard_normal((cfg.Nz + 1, cfg.Nx, cfg.Nx))
    tp = env[:, None, None] * rng.standard_normal((cfg.Nz + 1, cfg.Nx, cfg.Nx))
    w = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, np.fft.rfft2(wp, axes=(1, 2)))))
    th = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, np.fft.rfft2(tp, axes=(1, 2)))))
    tb = rng.standard_normal(cfg.Nz + 1)
    tb[0] = tb[-1] = 0.0
    tb = o.proj_d @ (o.V_inv @ tb)
    return O.OState(np.zeros((cfg.Nz + 1, cfg.Nx, o.Nk), complex), w, th, tb)


def test_balanced_sbp2_internal_exchange_monitor_closes():
    """Reference tests/test_solver.py:637-678 (3/2 rule, full-band random
    w, theta): under the SBP norm the theta-side and mean-side exchange
    terms cancel up to the boundary term."""
    cfg = _cfg(Nx=16, Nz=12, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, dt=1e-4,
               thermal_closure="evolve_mean", mean_exchange_discretization="balanced_sbp2_pc",
               imex_scheme="ars222")
    o = O.build_operators(cfg, with_shells=False)
    d = O.diagnostics(_exchange_monitor_state(o, cfg), o)
    assert abs(d["mean_theta_exchange_boundary_sbp"]) < 1e-12
    assert abs(d["mean_theta_exchange_residual_sbp"]) < 1e-10
    assert abs(d["mean_theta_exchange_residual_sbp_rel"]) < 1e-10
    # both sides are O(1), so the closure is not a cancellation of zeros
    assert abs(d["th_mean_feedback_sum_sbp"]) > 1e-2
    # a derivative without the SBP property (forward differences inside)
    # does not pair with the norm, and the terms no longer cancel
    dz = 1.0 / cfg.Nz
    D1 = o.sbp_D1.copy()
    for i in range(1, cfg.Nz):
        D1[i] = 0.0
        D1[i, i], D1[i, i + 1] = -1.0 / dz, 1.0 / dz
    o.sbp_D1 = D1
    assert abs(O.diagnostics(_exchange_monitor_state(o, cfg), o)
               ["mean_theta_exchange_residual_sbp_rel"]) > 1e-3


# ---------------------------------------------------------------------------
# 3/2 rule: reference tests/test_spectral.py and exact convolutions
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ops32():
    return O.build_operators(_cfg(Nx=64, Nz=4, L=2 * np.pi, horizontal_dealiasing="32_rule",
                                  nonlinear_advection="jacobian"), with_shells=False)


def _xy(N, L):
    x = np.arange(N) * (L / N)
    return np.meshgrid(x, x, indexing="ij")


def test_jacobian32_analytic(ops32):
    """Reference tests/test_spectral.py:40-60."""
    X, Y = _xy(64, 2 * np.pi)
    (J,) = O.jacobian32(np.fft.rfft2(np.sin(X))[None], [np.fft.rfft2(np.sin(Y))[None]], ops32)
    np.testing.assert_allclose(np.fft.irfft2(J[0], s=(64, 64)), np.cos(X) * np.cos(Y), atol=1e-12)
    (J2,) = O.jacobian32(np.fft.rfft2(np.sin(2 * X))[None], [np.fft.rfft2(np.sin(3 * Y))[None]],
                         ops32)
    np.testing.assert_allclose(np.fft.irfft2(J2[0], s=(64, 64)),
                               6 * np.cos(2 * X) * np.cos(3 * Y), atol=1e-11)


def test_flux32_matches_jacobian32(ops32):
    """Reference tests/test_spectral.py:110-120 and :148-167."""
    X, Y = _xy(64, 2 * np.pi)
    psi = np.fft.rfft2(np.sin(X) + 0.3 * np.cos(Y))[None].repeat(4, 0)
    fs = [np.fft.rfft2(g)[None].repeat(4, 0) for g in
          (np.cos(2 * X - Y), np.sin(X + 2 * Y), np.cos(3 * X) * np.sin(Y))]
    for F, J in zip(O.flux32(psi, fs, ops32), O.jacobian32(psi, fs, ops32)):
        np.testing.assert_allclose(F, J, atol=1e-11)


def _full_spectrum(h, N):
    """Hermitian completion of an (N, N/2+1) half spectrum, indexed by signed
    integer wavenumbers in [-N/2, N/2)."""
    full = np.zeros((N, N), complex)
    for i in range(N):
        for j in range(N):
            kx = i if i < N // 2 else i - N
            ky = j if j < N // 2 else j - N
            if ky >= 0:
                full[i, j] = h[i, ky]
            else:
                full[i, j] = np.conj(h[(-kx) % N, -ky])
    return full


def _convolve(a, b, N):
    """Exact (unaliased) convolution of two full spectra over signed
    wavenumbers; returns a dict (kx, ky) -> coefficient."""
    out = {}
    ks = [k if k < N // 2 else k - N for k in range(N)]
    for i1, kx1 in enumerate(ks):
        for j1, ky1 in enumerate(ks):
            if a[i1, j1] == 0:
                continue
            for i2, kx2 in enumerate(ks):
                for j2, ky2 in enumerate(ks):
                    k = (kx1 + kx2, ky1 + ky2)
                    out[k] = out.get(k, 0.0) + a[i1, j1] * b[i2, j2]
    return out


def test_flux32_and_sponge32_are_exact_convolutions():
    """The 3/2-rule products equal the unaliased convolution of the input
    spectra on every kept mode (direct sums at N = 8, no FFT library); the
    kx row N/2 of the output is the wavenumber -N/2 of the padded grid."""
    N, L = 8, 5.0
    o = O.build_operators(_cfg(Nx=N, Nz=2, L=L, horizontal_dealiasing="32_rule",
                               nonlinear_advection="flux"), with_shells=False)
    rng = np.random.default_rng(7)
    k0 = 2 * np.pi / L

    def real_spec():
        h = np.fft.rfft2(rng.standard_normal((N, N)))
        h[N // 2, :] = 0.0                      # no Nyquist content in the inputs
        h[:, N // 2] = 0.0
        return h

    ph, fh = real_spec(), real_spec()
    P, F = _full_spectrum(ph, N), _full_spectrum(fh, N)
    ks = np.array([k if k < N // 2 else k - N for k in range(N)])
    KX, KY = np.meshgrid(ks * k0, ks * k0, indexing="ij")
    U, Vv = -1j * KY * P, 1j * KX * P
    cu, cv = _convolve(U, F, N), _convolve(Vv, F, N)
    (got,) = O.flux32(ph[None], [fh[None]], o)
    sig = real_spec()        # a sponge profile too sharp for the guard at N = 8, set directly
    o.sponge_phys_pad = np.fft.irfft2(O.zero_pad(sig[None], N, o.Npad)[0], s=(o.Npad, o.Npad))
    cs = _convolve(_full_spectrum(sig, N), F, N)
    got_s = O.sponge32(fh[None], o)[0]
    for i in range(N):
        kx = i if i < N // 2 else i - N
        for ky in range(N // 2 + 1):
            want = (1j * kx * k0 * cu.get((kx, ky), 0) + 1j * ky * k0 * cv.get((kx, ky), 0)) / N ** 2
            assert abs(got[0, i, ky] - want) < 1e-12 * max(1.0, abs(want)), (i, ky)
            ws = cs.get((kx, ky), 0) / N ** 2
            assert abs(got_s[i, ky] - ws) < 1e-12 * max(1.0, abs(ws)), (i, ky)


def _padded_components(h, N):
    """Fourier components (signed wavenumbers) of M^2 irfft2(zero_pad(h),
    (M, M)) from numpy's documented irfft2 semantics: the inverse along x,
    then along y a half spectrum whose ky = 0 entry counts by its real part
    only, every other stored column ky also by its conjugate at (-kx, -ky).
    Row N/2 of h is kx = -N/2, column N/2 an ordinary column of the M grid."""
    comp = {}

    def add(k, v):
        comp[k] = comp.get(k, 0.0) + v

    for i in range(N):
        kx = i if i < N // 2 else i - N
        for ky in range(N // 2 + 1):
            v = h[i, ky]
            if v == 0:
                continue
            if ky == 0:
                add((kx, 0), v / 2)
                add((-kx, 0), np.conj(v) / 2)
            else:
                add((kx, ky), v)
                add((-kx, -ky), np.conj(v))
    return comp


def _aliased_product(a, b, M):
    """Components of the product of two fields on the M grid: sums over
    k1 + k2, wavenumbers folded into [-M/2, M/2) (the aliasing of the grid)."""
    out = {}
    for (ax, ay), va in a.items():
        for (bx, by), vb in b.items():
            k = ((ax + bx + M // 2) % M - M // 2, (ay + by + M // 2) % M - M // 2)
            out[k] = out.get(k, 0.0) + va * vb
    return out


@pytest.mark.parametrize("modulus", ["3N/2", "2N"])
def test_rule32_products_alias_nyquist_content_on_the_3n2_grid(modulus):
    """With content on the kx = -N/2 row and the ky = N/2 column, the 3/2
    rule's products (flux, Jacobian, sponge) keep the alias of the products
    of two such components that the 3N/2 grid folds onto kept modes: the
    oracle equals the modular convolution with M = 3N/2 to roundoff, and
    differs from the one of a 2N grid (which has no such alias)."""
    N, L = 8, 5.0
    o = O.build_operators(_cfg(Nx=N, Nz=2, L=L, horizontal_dealiasing="32_rule",
                               nonlinear_advection="flux"), with_shells=False)
    rng = np.random.default_rng(11)
    k0 = 2 * np.pi / L
    ph = np.fft.rfft2(rng.standard_normal((N, N))) + 0.5 * (
        rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1)))
    fh = np.fft.rfft2(rng.standard_normal((N, N)))
    fh[N // 2, 1:] += 1.0 - 2.0j                     # Nyquist row and column content
    fh[2:5, N // 2] += 0.7 + 0.3j
    M = 3 * N // 2 if modulus == "3N/2" else 2 * N
    kxr = np.array([i if i < N // 2 else i - N for i in range(N)])[:, None] * k0
    kyr = np.arange(N // 2 + 1)[None, :] * k0
    C = lambda h: _padded_components(h, N)
    cu = _aliased_product(C(-1j * kyr * ph), C(fh), M)
    cv = _aliased_product(C(1j * kxr * ph), C(fh), M)
    cj = _aliased_product(C(1j * kxr * ph), C(1j * kyr * fh), M)
    cj2 = _aliased_product(C(1j * kyr * ph), C(1j * kxr * fh), M)
    sig = np.fft.rfft2(rng.standard_normal((N, N)))
    sig[N // 2, 2] += 0.4
    o.sponge_phys_pad = np.fft.irfft2(O.zero_pad(sig[None], N, o.Npad)[0], s=(o.Npad, o.Npad))
    cs = _aliased_product(C(sig), C(fh), M)
    (F,) = O.flux32(ph[None], [fh[None]], o)
    (J,) = O.jacobian32(ph[None], [fh[None]], o)
    S = O.sponge32(fh[None], o)[0]
    worst = 0.0
    for i in range(N):
        kx = i if i < N // 2 else i - N
        for ky in range(N // 2 + 1):
            k = (kx, ky)
            wf = (1j * kx * k0 * cu.get(k, 0) + 1j * ky * k0 * cv.get(k, 0)) / N ** 2
            wj = (cj.get(k, 0) - cj2.get(k, 0)) / N ** 2
            ws = cs.get(k, 0) / N ** 2
            worst = max(worst, abs(F[0, i, ky] - wf) / max(1.0, abs(wf)),
                        abs(J[0, i, ky] - wj) / max(1.0, abs(wj)),
                        abs(S[i, ky] - ws) / max(1.0, abs(ws)))
    if modulus == "3N/2":
        assert worst < 1e-12
    else:
        assert worst > 1e-3


# ---------------------------------------------------------------------------
# SBP substep: energy balance
# ---------------------------------------------------------------------------

def _sbp_fields(o, Nx, seed=5):
    rng = np.random.default_rng(seed)
    N = o.Nz
    env = np.sin(np.pi * o.Z_sbp)[:, None, None]
    w = np.fft.rfft2(env * rng.standard_normal((N + 1, Nx, Nx)))
    th = np.fft.rfft2(env * rng.standard_normal((N + 1, Nx, Nx)))
    thb = rng.standard_normal(N + 1)
    thb[0] = thb[-1] = 0.0
    return w, th, thb


@pytest.mark.parametrize("kappa", [0.0, 1.0])
def test_sbp_substep_energy_balance(kappa):
    """E = th_bar^T H th_bar / (2 mu) + 1/2 sum_l H_l <theta_l^2>.  One SBP
    substep changes it by exactly kappa dt m^T H L m (m the midpoint mean):
    the flux and gradient terms cancel through H D1 + (H D1)^T = boundary
    (zero ends of w, theta, th_bar); with kappa = 0 E is conserved.  A sign
    flip of the fluctuation update must break the balance."""
    cfg = _cfg(Nx=16, Nz=24, L=10.0, dt=1e-3, thermal_closure="evolve_mean",
               mean_exchange_discretization="balanced_sbp2_pc", mean_temp_eps_sq=0.7)
    o = O.build_operators(cfg, with_shells=False)
    w, th, thb = _sbp_fields(o, 16)
    mean = lambda a, b: O.hmean(a, b, o)
    sdt, mu = 3e-3, o.mu
    A, B = O.sbp_matrices(mean(w, w), sdt, mu, kappa, o.sbp_D1, o.sbp_L)
    H = np.diag(o.sbp_H)
    Em = lambda b: b @ (H * b) / (2 * mu)
    Ef = lambda t: 0.5 * np.sum(H * mean(t, t))
    t1, b1 = O.sbp_substep(th, thb, w, A, B, sdt, mu, o.sbp_D1, mean)
    m = 0.5 * (b1 + thb)
    e0 = Em(thb) + Ef(th)
    assert abs(Em(b1) - Em(thb)) > 1e-4            # both sides move ...
    got = Em(b1) + Ef(t1) - e0
    want = kappa * sdt * (m @ (H * (o.sbp_L @ m)))
    assert abs(got - want) < 1e-13 * e0            # ... and balance to roundoff
    bad = flip_theta_update(O.sbp_substep)(th, thb, w, A, B, sdt, mu, o.sbp_D1, mean)
    assert abs(Em(bad[1]) + Ef(bad[0]) - e0 - want) > 1e-4


# ---------------------------------------------------------------------------
# Stage assembly: balanced_sbp2_pc against the legacy Chebyshev exchange
# ---------------------------------------------------------------------------

def _exact_transfers(o):
    """Polynomial interpolation between the CGL and SBP grids (Chebyshev
    series evaluated at the uniform nodes, and its inverse), so that the
    CGL -> SBP -> CGL round trip is the identity."""
    n = np.arange(o.Nz + 1)
    U = np.cos(np.outer(np.arccos(np.clip(2 * o.Z_sbp - 1, -1, 1)), n))
    o.c2s = U @ o.V_inv
    o.s2c = np.linalg.inv(o.c2s)


def _exchange_state(o):
    Nx, L = o.Nx, o.L
    X, Y = np.meshgrid(np.arange(Nx) * (L / Nx), np.arange(Nx) * (L / Nx), indexing="ij")
    k0 = 2 * np.pi / L

    def gal(phys, env):
        c = O.vmat(o.V_inv, env[:, None, None] * np.fft.rfft2(phys)[None])
        return O.vmat(o.d_pinv, O.vmat(o.proj_d, c))

    env = np.sin(np.pi * o.Z)
    w = gal(np.cos(k0 * X) * np.cos(2 * k0 * Y) + 0.5 * np.sin(3 * k0 * X + k0 * Y), env)
    th = gal(0.7 * np.cos(k0 * X + 0.3) * np.cos(2 * k0 * Y) + 0.2 * np.cos(2 * k0 * X),
             env * (1 + 0.3 * o.Z))
    q = O.vmat(o.V_inv, np.cos(np.pi * o.Z)[:, None, None]
               * np.fft.rfft2(0.5 * np.sin(k0 * X) * np.cos(k0 * Y))[None])
    thb = o.proj_d @ (o.V_inv @ (2.0 * np.sin(np.pi * o.Z) + 0.5 * np.sin(2 * np.pi * o.Z)))
    return O.sanitize(O.OState(q, w, th, thb), o)


def _exchange_increments(exchange, Nz, exact=True, dt=2e-4, T=4e-3):
    cfg = _cfg(Nx=16, Nz=Nz, L=10.0, Ra_tilde=20.0, sigma=1.0, dt=dt,
               horizontal_dealiasing="23_rule", nonlinear_advection="flux", nu_q=0.2,
               nu_w=0.2, nu_theta=0.2, hyper_order=1, thermal_closure="evolve_mean",
               mean_exchange_discretization=exchange, sbp_corrector_substeps=4,
               mean_temp_eps_sq=1.0)
    o = O.build_operators(cfg)
    if exact:
        _exact_transfers(o)
    s0 = _exchange_state(o)
    s = s0
    for _ in range(int(round(T / dt))):
        s = O.imex_step(s, o)
    return s.th - s0.th, s.thb - s0.thb


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(a).max()


def test_pc_exchange_converges_to_legacy_exchange():
    """With exact grid transfers, balanced_sbp2_pc (SBP-grid exchange in two
    stage correctors, omega*C1 carried into stage 2) and the legacy step
    (Chebyshev exchange as explicit/implicit RHS terms) discretise the same
    equations: theta and th_bar increments over 20 steps agree to the
    SBP grid's second-order truncation, shrinking from Nz = 16 to 24.
    Dropping omega*C1 leaves 1 - gamma = 71% of the exchange missing;
    flipping the fluctuation update reverses theta's exchange term."""
    e = {}
    for Nz in (16, 24):
        lt, lb = _exchange_increments("legacy", Nz)
        pt, pb = _exchange_increments("balanced_sbp2_pc", Nz)
        e[Nz] = (_rel(lt, pt), _rel(lb, pb))
    assert e[24][0] < 1e-2 and e[24][1] < 1e-2
    assert e[24][1] < 0.6 * e[16][1]
    lt, lb = _exchange_increments("legacy", 24)
    with planted("thermal_correction_tendency", drop_c1):
        pt, pb = _exchange_increments("balanced_sbp2_pc", 24)
    assert _rel(lt, pt) > 0.5 and _rel(lb, pb) > 0.5
    with planted("sbp_substep", flip_theta_update):
        pt, pb = _exchange_increments("balanced_sbp2_pc", 24)
    assert _rel(lt, pt) > 1.0


def test_interp_transfer_round_trip_is_second_order():
    """The reference's 'interp' transfers (piecewise-linear both ways) do not
    invert each other: one CGL -> SBP -> CGL round trip changes a smooth
    th_bar at second order in 1/Nz.  The balanced_sbp2_pc step applies it
    once per substep, so at fixed Nz this change does not vanish with dt
    (the reason the stage-assembly check above uses exact transfers)."""
    ch = {}
    for Nz in (32, 64, 128):
        o = O.build_operators(_cfg(Nx=16, Nz=Nz, horizontal_dealiasing="23_rule"),
                              with_shells=False)
        tb = o.proj_d @ (o.V_inv @ np.sin(np.pi * o.Z))
        P = o.proj_d @ o.V_inv @ o.s2c @ o.c2s @ o.V
        ch[Nz] = np.abs(P @ tb - tb).max() / np.abs(tb).max()
    assert 3.0 < ch[32] / ch[64] < 5.0 and 3.0 < ch[64] / ch[128] < 5.0
    assert ch[64] > 1e-4


def _pc_scalar_step(y, dt, a, c):
    """The stage formulas of imex_step_balanced_sbp2_pc (solver.py:1190-1225)
    on y' = a y + c y with a the implicit part and c the corrector's
    part, the corrector over alpha being the Crank-Nicolson map of c (as
    the SBP substep is for its linear part)."""
    g = 1 - 1 / np.sqrt(2)
    al, om = g * dt, (1 - g) * dt
    R = (1 + al * c / 2) / (1 - al * c / 2)
    P1 = y / (1 - al * a)
    Y1 = R * P1
    C1 = (Y1 - P1) / al
    return R * (y + om * a * Y1 + om * C1) / (1 - al * a)


def test_time_order_of_the_exchange():
    """Richardson ratios (errors at 4dt and 2dt against dt: 5 for second
    order, 3 for first) from a state with active w, theta and th_bar, exact
    transfers.  The legacy ARS222 step is second order in q, theta, th_bar.
    The balanced_sbp2_pc step is first order in theta and th_bar: its
    stage-2 exchange term omega*C1 is the corrector's rate at the stage-1
    predictor, not at the corrected stage value, which leaves an O(dt^2)
    local error.  The scalar model of the same stage formulas shows the
    same first order, so this is the reference's scheme, not the oracle."""
    errs = []
    for dt in (1e-2, 5e-3, 2.5e-3):
        y = 1.0
        for _ in range(int(round(1.0 / dt))):
            y = _pc_scalar_step(y, dt, -1.0, -2.0)
        errs.append(abs(y - np.exp(-3.0)))
    assert 1.9 < errs[0] / errs[1] < 2.1 and 1.9 < errs[1] / errs[2] < 2.1
    for exchange, lo, hi in (("legacy", 4.7, 5.3), ("balanced_sbp2_pc", 2.8, 3.2)):
        runs = {m: _exchange_increments(exchange, 16, dt=1e-4 * m, T=3.2e-3) for m in (1, 2, 4)}
        for k in (0, 1):
            e1 = np.abs(runs[4][k] - runs[1][k]).max()
            e2 = np.abs(runs[2][k] - runs[1][k]).max()
            assert lo < e1 / e2 < hi, (exchange, k, e1 / e2)


# ---------------------------------------------------------------------------
# Diagnostics: closed forms of single Fourier modes
# ---------------------------------------------------------------------------

def test_diagnostics_single_mode_closed_forms():
    """psi = a cos(k.x) P(Z), w = b cos(k.x) R(Z), theta = c cos(k.x) S(Z)
    with P = Z(1-Z), R = S = sin(pi Z) (spectrally resolved) give
    KE_tot = a^2 |k|^2 / 120, KE_bt = a^2 |k|^2 / 144, enstrophy =
    a^2 |k|^4 / 120, max_speed = |a| |k| / 4 (attained at Z = 1/2 on the
    grid), vol_avg_tw = b c / 4, max_w = |b|, max_theta = |c|,
    max_tw = |b c|.  Uniform weights in place of Clenshaw-Curtis must fail."""
    def check(uniform):
        cfg = _cfg(Nx=16, Nz=24, L=10.0, horizontal_dealiasing="23_rule",
                   thermal_closure="fixed_conduction")
        o = O.build_operators(cfg, with_shells=False)
        X, Y = _xy(16, cfg.L)
        k0 = 2 * np.pi / cfg.L
        mx, my = 1, 2
        k2 = (k0 * mx) ** 2 + (k0 * my) ** 2
        a, b, c = 0.8, -1.3, 0.6
        mode = np.fft.rfft2(np.cos(k0 * (mx * X + my * Y)))
        P = o.Z * (1 - o.Z)
        R = np.sin(np.pi * o.Z)
        psi_n = a * P[:, None, None] * mode[None]
        q = O.vmat(o.V_inv, -k2 * psi_n)
        w = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, b * R[:, None, None] * mode[None])))
        th = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, c * R[:, None, None] * mode[None])))
        return O.diagnostics(O.OState(q, w, th, np.zeros(cfg.Nz + 1)), o), k2, a, b, c

    d, k2, a, b, c = check(False)
    assert abs(d["KE_tot"] - a * a * k2 / 120) < 1e-13 * d["KE_tot"]
    assert abs(d["KE_bt"] - a * a * k2 / 144) < 1e-13 * d["KE_bt"]
    assert abs(d["enstrophy"] - a * a * k2 * k2 / 120) < 1e-12 * d["enstrophy"]
    assert abs(d["max_speed"] - abs(a) * np.sqrt(k2) / 4) < 1e-13
    assert abs(d["vol_avg_tw"] - b * c / 4) < 1e-12
    assert abs(d["vol_avg_tw_dealiased"] - b * c / 4) < 1e-3   # SBP-grid profile: O(dz^2)
    assert abs(d["max_w"] - abs(b)) < 1e-12 and abs(d["max_theta"] - abs(c)) < 1e-12
    assert abs(d["max_tw"] - abs(b * c)) < 1e-12
    with planted("clenshaw_curtis", uniform_weights):
        bad = check(True)[0]
    assert abs(bad["KE_tot"] - a * a * k2 / 120) > 1e-2 * d["KE_tot"]


def test_sbp_heat_flux_profile_converges_at_second_order():
    """Nu_d - 1 on the balanced_sbp2_pc branch integrates the SBP-grid flux
    <w_sbp theta_sbp> carried back to the CGL nodes (solver.py:370-374);
    for w = theta = cos(k.x) sin(pi Z) it tends to 1/4 at second order."""
    err = {}
    for Nz in (16, 32, 64):
        cfg = _cfg(Nx=16, Nz=Nz, L=10.0, horizontal_dealiasing="23_rule",
                   thermal_closure="evolve_mean", mean_exchange_discretization="balanced_sbp2_pc")
        o = O.build_operators(cfg, with_shells=False)
        X, Y = _xy(16, cfg.L)
        mode = np.fft.rfft2(np.cos(2 * np.pi / cfg.L * (X + Y)))
        f = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(
            o.V_inv, np.sin(np.pi * o.Z)[:, None, None] * mode[None])))
        s = O.OState(np.zeros((Nz + 1, 16, o.Nk), complex), f, f.copy(), np.zeros(Nz + 1))
        err[Nz] = abs(O.diagnostics(s, o)["vol_avg_tw_dealiased"] - 0.25)
    assert 3.0 < err[16] / err[32] < 5.0 and 3.0 < err[32] / err[64] < 5.0
    assert err[64] < 4e-4


# ---------------------------------------------------------------------------
# vertical_cutoff_n (solver.py:177-191)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_vertical_cutoff_is_a_boundary_preserving_projection(top):
    """The cutoff keeps Chebyshev degrees n <= cutoff and re-imposes the
    boundary conditions: it is idempotent, the identity on fields with no
    content above the cutoff (which an off-by-one in the degree test
    breaks), and its result satisfies the boundary conditions of w and
    theta."""
    N, cut = 16, 9
    o = O.build_operators(_cfg(Nx=16, Nz=N, L=20.0, vertical_cutoff_n=cut, w_bc_top=top,
                               horizontal_dealiasing="23_rule"), with_shells=False)
    rng = np.random.default_rng(4)
    shape = (N - 1, 16, 9)
    w = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    th = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    s = O.OState(np.zeros((N + 1, 16, 9), complex), w, th, np.zeros(N + 1))
    a = O.vertical_cutoff(s, o)
    b = O.vertical_cutoff(a, o)
    assert np.abs(b.w - a.w).max() <= 1e-13 * np.abs(a.w).max()
    assert np.abs(b.th - a.th).max() <= 1e-13 * np.abs(a.th).max()
    assert np.abs(a.w - w).max() > 1e-2          # it did remove something
    # Galerkin index j spans degrees j and j+2: indices <= cut-2 stay below the cutoff
    low = np.zeros(shape, complex)
    low[:cut - 1] = w[:cut - 1]
    c = O.vertical_cutoff(O.OState(s.q, low, low, s.thb), o)
    assert np.abs(c.w - low).max() <= 1e-13 * np.abs(low).max()
    assert np.abs(c.th - low).max() <= 1e-13 * np.abs(low).max()
    # boundary conditions of the result on the CGL grid (Z = 1 first, Z = 0 last)
    wn = O.vmat(o.V, O.vmat(o.w_stencil, a.w))
    tn = O.vmat(o.V, O.vmat(o.S_d, a.th))
    sc = np.abs(wn).max()
    assert np.abs(wn[-1]).max() <= 1e-12 * sc                     # w = 0 at the bottom
    assert np.abs(tn[0]).max() <= 1e-12 * np.abs(tn).max()        # theta = 0 at both ends
    assert np.abs(tn[-1]).max() <= 1e-12 * np.abs(tn).max()
    if top == "dirichlet":
        assert np.abs(wn[0]).max() <= 1e-12 * sc
    else:
        dw = O.vmat(o.V, O.vmat(o.G_Z, O.vmat(o.w_stencil, a.w)))
        assert np.abs(dw[0]).max() <= 1e-11 * np.abs(dw).max()   # dw/dZ = 0 at the top
