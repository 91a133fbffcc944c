"""vertical_cutoff_n on the GPU against the oracle: the cutoff of w and
theta at the end of every step (_finalize_state, nhqg/solver.py:993-999,
177-191) and on every RK4 stage state (_apply_bcs, :1261-1273).

Bar: 1e-12 relative per field for one step and an RK4 step, 1e-11 after
ten steps.
"""

import dataclasses

import numpy as np
import pytest

from nhqg_amd import State, Stepper, imex_step, make_grid, rk4_step
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


CONFIGS = {
    "fixture": lambda: cases.fixture_cfg(vertical_cutoff_n=10),
    "trap_sponge": lambda: cases.trap_sponge_cfg(vertical_cutoff_n=8),
    "neumann": lambda: cases.neumann_cfg(vertical_cutoff_n=11),
    "legacy_rule32": lambda: cases.fixture_cfg(vertical_cutoff_n=12,
                                               mean_exchange_discretization="legacy",
                                               horizontal_dealiasing="32_rule"),
}


def _errs(got, ref):
    return [rel(got[0], ref.q), rel(got[1], ref.w), rel(got[2], ref.th), rel(got[3], ref.thb)]


@pytest.mark.parametrize("name", list(CONFIGS))
def test_cutoff_steps_match_oracle(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    say("cutoff", name)
    got = imex_step(State(s.q, s.w, s.th, s.thb), g)
    ref = O.imex_step(s, o)
    e = _errs(got, ref)
    say("  one step", e)
    assert max(e) < 1e-12
    # the cutoff acted: the step differs from one without it
    o2 = O.build_operators(cfg)
    o2.cutoff = None
    assert rel(O.imex_step(s, o2).w, ref.w) > 1e-6
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    sp.run(10)
    for _ in range(9):
        ref = O.imex_step(ref, o)
    e = _errs(sp.state(), ref)
    say("  ten steps", e)
    assert max(e) < 1e-11
    sp.close()
    g.close()


@pytest.mark.parametrize("name", ["fixture", "neumann"])
def test_cutoff_rk4_step_matches_oracle(name):
    cfg = CONFIGS[name]()
    cfg = dataclasses.replace(cfg, thermal_closure="fixed_conduction",
                              mean_exchange_discretization="legacy")
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    say("cutoff rk4", name)
    got = rk4_step(State(s.q, s.w, s.th, s.thb), g)
    e = _errs(got, O.rk4_step(s, o))
    say("  ", e)
    assert max(e[:3]) < 1e-12
    g.close()
