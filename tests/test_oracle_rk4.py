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
