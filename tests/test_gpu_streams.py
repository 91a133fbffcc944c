"""The step graph forks the SBP corrector's preamble (W2S, <w_sbp^2>, the
factorisation and the substep maps) onto a side stream as soon as w is
final and joins it before the first substep; the profiled step runs the
same step on one stream without the fork.  Every kernel of the step is
deterministic (fixed reduction trees, no atomics) and a contraction's
result does not depend on the batch it is launched in, so both must give
bit-identical states, also after several graph replays with eager work on
the context between them."""

import numpy as np
import pytest

from nhqg_amd import make_grid
from tests import cases

pytestmark = pytest.mark.gpu

CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
}


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_forked_graph_step_equals_single_stream_step(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    assert g.evolve_mean, "the corrector (and its fork) runs with evolve_mean"
    s = cases.active_state(cfg)
    ctx = g.context(full=False)
    say("streams", name, "graph replays")
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.step(3)
    forked = [np.array(a, copy=True) for a in ctx.get_state()]
    say("streams", name, "profiled single-stream steps")
    ctx.set_state(s.q, s.w, s.th, s.thb)
    for _ in range(3):
        ctx.profile_step()
    serial = [np.array(a, copy=True) for a in ctx.get_state()]
    # graph replays again after the eager steps, from the same state
    say("streams", name, "graph replays after eager work")
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.step(3)
    again = [np.array(a, copy=True) for a in ctx.get_state()]
    g.close()
    for label, x, y, z in zip(("q", "w", "theta", "th_bar"), forked, serial, again):
        d = float(np.abs(x - y).max())
        say("streams", name, label, "max |forked - serial|", d)
        assert np.array_equal(x, y), (name, label, d)
        assert np.array_equal(x, z), (name, label, "replay after eager work")
