"""RK4 validator on the GPU (nhqg_amd.rk4_step, solver.py:1276-1336)
against the oracle's restatement, and the reference's IMEX-vs-RK4 test run
on the device: fixed_conduction configs, and evolve_mean with the split
(balanced_sbp2_pc: mean diffusion only) and the legacy exchange."""

import numpy as np
import pytest

from nhqg_amd import NHQGConfig, State, imex_step, make_grid, rk4_step
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


FIXED = dict(thermal_closure="fixed_conduction", mean_exchange_discretization="legacy")
CONFIGS = {
    "fixed": lambda: cases.fixture_cfg(**FIXED),
    "fixed_neumann": lambda: cases.neumann_cfg(**FIXED),
    "fixed_trap_sponge": lambda: cases.trap_sponge_cfg(**FIXED),
    "fixed_jacobian": lambda: cases.fixture_cfg(nonlinear_advection="jacobian", **FIXED),
    # explicit mean diffusion: dt inside RK4's stability region at Nz = 12-16
    "evolve_pc": lambda: cases.fixture_cfg(dt=2e-5),
    "evolve_legacy": lambda: cases.fixture_cfg(mean_exchange_discretization="legacy", dt=2e-5),
    "evolve_legacy_trap_sponge": lambda: cases.trap_sponge_cfg(
        mean_exchange_discretization="legacy", dt=2e-5),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_rk4_matches_oracle(name):
    cfg = CONFIGS[name]()
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    g = make_grid(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    ref = s
    say("rk4", name)
    for n in range(3):
        st = rk4_step(st, g)
        ref = O.rk4_step(ref, o)
        errs = [rel(st.q_hat, ref.q), rel(st.w_hat, ref.w), rel(st.th_hat, ref.th),
                rel(st.th_bar, ref.thb)]
        say("  step", n + 1, errs)
        assert max(errs) < TOL
    g.close()


def test_imex_and_rk4_agree_on_device():
    """tests/test_solver.py:141-167 on the GPU paths (under the 2/3 rule)."""
    cfg = NHQGConfig(Nx=32, Nz=8, L=20.0, Ra_tilde=50.0, sigma=1.0, beta=0.0, Ld=float("inf"),
                     dt=1e-4, float_dtype="float64", horizontal_dealiasing="23_rule")
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s0 = O.jax_initial_state(o, 7, 1e-3, True)
    a = b = State(s0.q, s0.w, s0.th, s0.thb)
    for _ in range(20):
        a = imex_step(a, g)
        b = rk4_step(b, g)
    r = rel(a.q_hat, b.q_hat)
    say("imex vs rk4", r)
    assert 0.0 < r < 1e-3
    g.close()
