"""The banded SBP substep transfer against the split contraction (GPU).

Each corrector substep but the last maps the updated SBP-grid theta through
Psbp = T2S S2T.  By default the device applies it as the five-diagonal
stencil of Psbp Z (sbp_band_kernel, with the next substep's horizontal mean
summed in the same pass and the self-conjugate columns' cross terms added
by sbp_cross_kernel); NHQG_SBP_BAND=0 at context creation selects the
reflection-split contraction of the full Psbp and the separate Parseval
pass.  The two differ by the construction roundoff of Psbp's off-band
entries (<= 1.25e-14 of its largest at Nz = 256): five steps agree to 1e-12
relative per field, and the stencil path is the one compared with the
oracle by every other GPU test.
"""

import os

import numpy as np
import pytest

from nhqg_amd import make_grid
from nhqg_amd.init import polar_config_256, polar_state_256
from tests import cases

pytestmark = pytest.mark.gpu

CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "trap_sponge_rule32": lambda: cases.trap_sponge_cfg(horizontal_dealiasing="32_rule"),
}


def say(*a):
    print(*a, flush=True)


def _run(cfg, state, band, steps):
    old = os.environ.get("NHQG_SBP_BAND")
    os.environ["NHQG_SBP_BAND"] = "1" if band else "0"
    try:
        g = make_grid(cfg, shells="band")
        ctx = g.context(full=False)
    finally:
        if old is None:
            del os.environ["NHQG_SBP_BAND"]
        else:
            os.environ["NHQG_SBP_BAND"] = old
    ctx.set_state(*state)
    ctx.step(steps)
    out = [np.array(a, copy=True) for a in ctx.get_state()]
    g.close()
    return out


def _compare(name, a, b, tol):
    for label, x, y in zip(("q", "w", "theta", "th_bar"), a, b):
        e = float(np.abs(x - y).max() / max(float(np.abs(y).max()), 1e-300))
        say("sbp band", name, label, "rel diff", e)
        assert e <= tol, (name, label, e)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_band_substeps_match_the_split_contraction(name):
    cfg = CONFIGS[name]()
    s = cases.active_state(cfg)
    st = (s.q, s.w, s.th, s.thb)
    say("sbp band", name, "5 steps each path")
    _compare(name, _run(cfg, st, True, 5), _run(cfg, st, False, 5), 1e-12)


def test_band_substeps_match_at_256_cubed():
    cfg = polar_config_256(Nx=256, Nz=256, dt=2.5e-5)
    g = make_grid(cfg, shells="band")
    s = polar_state_256(g, seed=0)
    g.close()
    st = (s.q_hat, s.w_hat, s.th_hat, s.th_bar)
    say("sbp band", "polar 256^3", "3 steps each path")
    _compare("polar256", _run(cfg, st, True, 3), _run(cfg, st, False, 3), 1e-12)
