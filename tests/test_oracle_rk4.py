"""The oracle's RK4 validator (solver.py:1276-1336) pinned without the
reference: fourth-order convergence of the full-RHS integration (a wrong
stage weight drops the error ratio under dt halving from 16 to 2-4), and
the reference's own IMEX-vs-RK4 agreement test (tests/test_solver.py:
141-167) on its config and initial draw."""

import numpy as np

from nhqg_amd import NHQGConfig
from oracle import nhqg_oracle as O


def _run(cfg, s0, n, step):
    o = O.build_operators(cfg)
    s = s0
    for _ in range(n):
        s = step(s, o)
    return s


def test_rk4_is_fourth_order():
    base = dict(Nx=16, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, Ld=float("inf"),
                float_dtype="float64", thermal_closure="fixed_conduction",
                horizontal_dealiasing="23_rule")
    dts = [4e-3, 2e-3, 1e-3]
    T = 8e-3
    o0 = O.build_operators(NHQGConfig(dt=dts[0], **base))
    s0 = O.jax_initial_state(o0, 3, 5e-2, True)
    ends = [_run(NHQGConfig(dt=dt, **base), s0, int(round(T / dt)), O.rk4_step) for dt in dts]
    e1 = np.abs(ends[0].q - ends[1].q).max()
    e2 = np.abs(ends[1].q - ends[2].q).max()
    ratio = e1 / e2
    assert e2 > 1e-13 * np.abs(ends[2].q).max()     # above roundoff
    assert 12.0 < ratio < 20.0, ratio


def test_imex_and_rk4_agree_at_small_dt():
    """tests/test_solver.py:141-167: 20 steps at dt = 1e-4 from the seed-7
    draw, relative q difference < 1e-3 (IMEX second order, RK4 fourth).
    Run under the 2/3 rule: the config's default 3/2 padding is not built."""
    cfg = NHQGConfig(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, Ld=float("inf"),
                     dt=1e-4, float_dtype="float64", horizontal_dealiasing="23_rule")
    o = O.build_operators(cfg)
    s0 = O.jax_initial_state(o, 7, 1e-3, True)
    a, b = s0, s0
    for _ in range(20):
        a = O.imex_step(a, o)
        b = O.rk4_step(b, o)
    rel = np.abs(a.q - b.q).max() / np.abs(b.q).max()
    assert rel < 1e-3, rel
    assert rel > 0.0


def test_imex_and_rk4_agree_reference_config():
    """The same test on the reference's own config: its defaults select
    the 3/2 rule and the Jacobian form."""
    cfg = NHQGConfig(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, Ld=float("inf"),
                     dt=1e-4, float_dtype="float64")
    assert cfg.horizontal_dealiasing == "32_rule" and cfg.nonlinear_advection == "jacobian"
    o = O.build_operators(cfg)
    s0 = O.jax_initial_state(o, 7, 1e-3, True)
    a, b = s0, s0
    for _ in range(20):
        a = O.imex_step(a, o)
        b = O.rk4_step(b, o)
    rel = np.abs(a.q - b.q).max() / np.abs(b.q).max()
    assert 0.0 < rel < 1e-3, rel


def _mean_state(o, seed=2):
    """Active w, theta and th_bar (boundary conditions satisfied)."""
    rng = np.random.default_rng(seed)
    s = O.jax_initial_state(o, seed, 5e-2, True)
    env = np.sin(np.pi * o.Z)[:, None, None]
    w = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, env * O.vmat(o.V, s.q) * 0.8j)))
    th = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, env * O.vmat(o.V, s.q) * 0.5)))
    tb = o.proj_d @ (o.V_inv @ (0.3 * np.sin(np.pi * o.Z) + 0.1 * rng.standard_normal() *
                                np.sin(2 * np.pi * o.Z)))
    return O.sanitize(O.OState(s.q, w, th, tb), o)


def test_rk4_evolve_mean_legacy_fourth_order_and_imex_agreement():
    """RK4 with the legacy mean exchange (E_theta's -dTheta/dZ w term,
    E_th_bar = -eps^2 d<w theta>/dZ, I_th_bar = eps^2/sigma G_Z2 th_bar):
    fourth order under dt halving in theta and th_bar, and within the IMEX
    step's second-order error of the legacy ARS222 step."""
    base = dict(Nx=16, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, float_dtype="float64",
                thermal_closure="evolve_mean", horizontal_dealiasing="23_rule",
                nonlinear_advection="flux", mean_temp_eps_sq=0.5)
    o0 = O.build_operators(NHQGConfig(dt=1e-4, **base))
    s0 = _mean_state(o0)
    T = 1.6e-3
    dts = [4e-4, 2e-4, 1e-4]
    ends = [_run(NHQGConfig(dt=dt, **base), s0, int(round(T / dt)), O.rk4_step) for dt in dts]
    for f in ("th", "thb"):
        e1 = np.abs(getattr(ends[0], f) - getattr(ends[1], f)).max()
        e2 = np.abs(getattr(ends[1], f) - getattr(ends[2], f)).max()
        assert e2 > 1e-13 * np.abs(getattr(ends[2], f)).max()
        assert 12.0 < e1 / e2 < 24.0, (f, e1 / e2)
    imex = _run(NHQGConfig(dt=1e-4, **base), s0, int(round(T / 1e-4)), O.imex_step)
    for f in ("th", "thb"):
        d_rk = getattr(ends[2], f) - getattr(s0, f)
        assert np.abs(getattr(imex, f) - getattr(ends[2], f)).max() < 1e-3 * np.abs(d_rk).max()


def test_rk4_split_exchange_mean_diffusion_closed_form():
    """Under a split exchange (balanced_sbp2_pc) the full RHS carries no
    exchange (E_th_bar = 0, solver.py:552-556) but keeps the mean diffusion
    I_th_bar = eps^2/sigma G_Z2 th_bar: with q = w = theta = 0 a
    th_bar = sin(pi Z) decays as exp(-eps^2 pi^2 t / sigma)."""
    cfg = NHQGConfig(Nx=16, Nz=16, L=20.0, sigma=2.0, mean_temp_eps_sq=0.5, dt=2e-5,
                     float_dtype="float64", thermal_closure="evolve_mean",
                     mean_exchange_discretization="balanced_sbp2_pc",
                     horizontal_dealiasing="23_rule")
    o = O.build_operators(cfg)
    z = np.zeros((cfg.Nz - 1, 16, 9), complex)
    s = O.OState(np.zeros((cfg.Nz + 1, 16, 9), complex), z, z.copy(),
                 o.proj_d @ (o.V_inv @ np.sin(np.pi * o.Z)))
    n = 50
    for _ in range(n):
        s = O.rk4_step(s, o)
    want = np.exp(-0.5 * np.pi ** 2 * n * cfg.dt / 2.0) * np.sin(np.pi * o.Z)
    assert np.abs(o.V @ s.thb - want).max() < 1e-10
