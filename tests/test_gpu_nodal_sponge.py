"""The nodal w/theta sponge route against the coefficient route (GPU).

In the flux form the row passes multiply the nodal w and theta rows they
already hold by sigma and add F_y(sigma f) into the H slot of the same
field; q' keeps the reference's coefficient-level sponge.  With
NHQG_NODAL_SPONGE=0 at context creation all three sponges run on
coefficient levels through their own X5 / Y6 slots, as the reference
computes them.  The routes are equal in exact arithmetic (the sponge acts
on every level alike and commutes with the vertical maps); five steps agree
to 1e-12 relative per field on the 256-point row pass, the radix-4 row pass
and the 3/2 rule's padded row pass.
"""

import os

import numpy as np
import pytest

from nhqg_amd import make_grid
from nhqg_amd.init import polar_config_256, polar_state_256
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _run(cfg, state, nodal, steps):
    old = os.environ.get("NHQG_NODAL_SPONGE")
    os.environ["NHQG_NODAL_SPONGE"] = "1" if nodal else "0"
    try:
        g = make_grid(cfg, shells="band")
        ctx = g.context(full=False)
    finally:
        if old is None:
            del os.environ["NHQG_NODAL_SPONGE"]
        else:
            os.environ["NHQG_NODAL_SPONGE"] = old
    ctx.set_state(*state)
    ctx.step(steps)
    out = [np.array(a, copy=True) for a in ctx.get_state()]
    g.close()
    return out


CONFIGS = {
    # radix-4 row pass (Nx = 128)
    "trap_sponge_128": lambda: cases.trap_sponge_cfg(),
    # the 3/2 rule's padded passes
    "trap_sponge_128_rule32": lambda: cases.trap_sponge_cfg(horizontal_dealiasing="32_rule"),
    # the 256-point row pass
    "polar_256x32": lambda: polar_config_256(Nx=256, Nz=32, dt=2.5e-5),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_nodal_sponge_matches_the_coefficient_route(name):
    cfg = CONFIGS[name]()
    if name.startswith("polar"):
        g = make_grid(cfg, shells="band")
        s = polar_state_256(g, seed=0)
        g.close()
        st = (s.q_hat, s.w_hat, s.th_hat, s.th_bar)
    else:
        s = cases.active_state(cfg)
        st = (s.q, s.w, s.th, s.thb)
    say("nodal sponge", name, "5 steps each route")
    a = _run(cfg, st, True, 5)
    b = _run(cfg, st, False, 5)
    for label, x, y in zip(("q", "w", "theta", "th_bar"), a, b):
        e = float(np.abs(x - y).max() / max(float(np.abs(y).max()), 1e-300))
        say("nodal sponge", name, label, "rel diff", e)
        assert e <= 1e-12, (name, label, e)
