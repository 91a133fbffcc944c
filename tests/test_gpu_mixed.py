"""Mixed-precision step (int8 slice products on the matrix cores, FP64
reconstruction) against the plain FP64 step and the oracle.

Bars: components and one step within 1e-10 relative of FP64; after 1000
steps every field within 1e-8 relative of the plain GPU run and KE, Nu_d,
max_speed within 1e-9.  Slices per operand: 6 (~42 bits) up to Nz = 128,
7 at Nz = 256, where d/dZ-type operators have entries ~Nz^2 whose sums
cancel (the slice remainder is relative to a column's largest entry)."""

import numpy as np
import pytest

from nhqg_amd import State, Stepper, imex_step, make_grid
from nhqg_amd import _native
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "neumann": cases.neumann_cfg,
    # K = 257: the last k of every Nq-long contraction runs in FP64
    "nz256": lambda: cases.fixture_cfg(Nx=32, Nz=256, dt=1e-4),
}


def slices_for(cfg):
    return 7 if cfg.Nz > 128 else 6


def _grid(cfg, slices):
    g = make_grid(cfg)
    g.mixed_slices = slices
    return g


@pytest.mark.parametrize("name", list(CONFIGS))
def test_mixed_components_match_fp64(name):
    cfg = CONFIGS[name]()
    s = cases.active_state(cfg)
    out = {}
    S = slices_for(cfg)
    for sl in (0, S):
        g = _grid(cfg, sl)
        ctx = g.context(full=False)
        res = []
        for op in (_native.OP_EXPLICIT_RHS, _native.OP_IMPLICIT_TEND, _native.OP_IMPLICIT_SOLVE):
            ctx.set_state(s.q, s.w, s.th, s.thb)
            ctx.apply(op)
            res.append(ctx.get_aux() if op != _native.OP_IMPLICIT_SOLVE else ctx.get_state()[:3])
        out[sl] = res
        g.close()
    errs = [[rel(x, y) for x, y in zip(a, b)] for a, b in zip(out[S], out[0])]
    say("mixed components", name, "slices", S, errs)
    assert max(max(e) for e in errs) < TOL


@pytest.mark.parametrize("name", list(CONFIGS))
def test_mixed_one_step_vs_oracle(name):
    cfg = CONFIGS[name]()
    s = cases.active_state(cfg)
    ref = O.imex_step(s, O.build_operators(cfg))
    g = _grid(cfg, slices_for(cfg))
    say("mixed one step", name)
    got = imex_step(State(s.q, s.w, s.th, s.thb), g)
    errs = [rel(got.q_hat, ref.q), rel(got.w_hat, ref.w), rel(got.th_hat, ref.th)]
    if np.abs(ref.thb).max() > 0:
        errs.append(rel(got.th_bar, ref.thb))
    say("  errors", errs)
    assert max(errs) < TOL
    g.close()


def test_mixed_slices_rejected_out_of_range():
    g = _grid(cases.fixture_cfg(), 3)
    with pytest.raises(_native.NativeError):
        g.context(full=False)


def test_mixed_1000_steps_track_fp64():
    """256^2 x 128 polar-trap workload: 1000 graph-replayed mixed steps vs
    1000 plain FP64 steps from the same seeded state."""
    from nhqg_amd.init import polar_config_256, polar_state_256

    cfg = polar_config_256(Nz=128)
    res = {}
    for sl in (0, 6):
        g = make_grid(cfg, shells="band")
        g.mixed_slices = sl
        st = polar_state_256(g, seed=0)
        sp = Stepper(g)
        sp.set_state(st)
        say("1000 steps, slices", sl)
        sp.run(1000)
        res[sl] = (sp.state(), sp.diagnostics())
        g.close()
    (a, da), (b, db) = res[6], res[0]
    errs = {n: rel(x, y) for n, x, y in zip(("q", "w", "th", "thb"), a, b)}
    derr = {k: abs(da[k] - db[k]) / abs(db[k]) for k in ("KE_tot", "Nusselt_dealiased",
                                                          "max_speed")}
    say("  field errors", errs, "diagnostic errors", derr)
    assert max(errs.values()) < 1e-8
    assert max(derr.values()) < 1e-9
