"""Reflection-split contractions (gemm_split.hip) against the dense ones.

The vertical operators commute with z -> -z up to the roundoff of their
construction, and each split product differs from the dense product by that
roundoff (<= 5e-14 of the field in the CPU emulation).  Bar: every component
of the step and whole steps within 1e-12 relative of the dense contractions,
on configs where every operator splits (Dirichlet top) and where some do not
(Neumann top, which falls back to dense for its asymmetric maps)."""

import numpy as np
import pytest

from nhqg_amd import State, imex_step, make_grid
from nhqg_amd import _native
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "neumann": cases.neumann_cfg,
    "nz256": lambda: cases.fixture_cfg(Nx=32, Nz=256, dt=1e-4),
    "nz48": lambda: cases.fixture_cfg(Nx=32, Nz=48),
}


def _grid(cfg, dense):
    g = make_grid(cfg)
    g.dense_contractions = dense
    return g


@pytest.mark.parametrize("name", list(CONFIGS))
def test_split_components_match_dense(name):
    cfg = CONFIGS[name]()
    s = cases.active_state(cfg)
    out = {}
    for dense in (True, False):
        g = _grid(cfg, dense)
        ctx = g.context(full=False)
        res = []
        ops = [_native.OP_EXPLICIT_RHS, _native.OP_IMPLICIT_TEND, _native.OP_IMPLICIT_SOLVE]
        if g.evolve_mean:
            ops.append(_native.OP_CORRECTOR)
        for op in ops:
            ctx.set_state(s.q, s.w, s.th, s.thb)
            ctx.apply(op, g.gamma_imex * g.dt)
            res.append(ctx.get_aux() if op in (_native.OP_EXPLICIT_RHS, _native.OP_IMPLICIT_TEND)
                       else ctx.get_state())
        out[dense] = res
        g.close()
    errs = [[rel(x, y) for x, y in zip(a, b)] for a, b in zip(out[False], out[True])]
    say("split components", name, errs)
    assert max(max(e) for e in errs) < TOL


@pytest.mark.parametrize("name", list(CONFIGS))
def test_split_steps_match_dense(name):
    cfg = CONFIGS[name]()
    st = cases.active_state(cfg)
    res = {}
    for dense in (True, False):
        g = _grid(cfg, dense)
        s = State(st.q, st.w, st.th, st.thb)
        for _ in range(5):
            s = imex_step(s, g)
        res[dense] = s
        g.close()
    errs = [rel(x, y) for x, y in zip(res[False], res[True])]
    say("split 5 steps", name, errs)
    assert max(errs) < TOL


def test_split_step_256():
    """The 256^2 x 256 polar trap workload: one step, split vs dense."""
    from nhqg_amd.init import polar_config_256, polar_state_256

    cfg = polar_config_256()
    res = {}
    for dense in (True, False):
        g = _grid(cfg, dense)
        s0 = polar_state_256(g)
        res[dense] = imex_step(s0, g)
        g.close()
    errs = [rel(x, y) for x, y in zip(res[False], res[True])]
    say("split step 256^3", errs)
    assert max(errs) < TOL
