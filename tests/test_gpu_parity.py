"""GPU parity of the HIP step against the oracle, through the C ABI.

Bar: per field max|gpu - oracle| / max|oracle| <= 1e-12 (the FP64 per-step
tolerance the north star states), unless noted.
"""

import numpy as np
import pytest

from nhqg_amd import make_grid, imex_step, State, Stepper
from nhqg_amd import _native
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


def _ctx(cfg):
    g = make_grid(cfg)
    return g, g.context(full=False)


CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "neumann": cases.neumann_cfg,
    "fixed": lambda: cases.fixture_cfg(thermal_closure="fixed_conduction",
                                       mean_exchange_discretization="legacy"),
    "jacobian": lambda: cases.fixture_cfg(nonlinear_advection="jacobian"),
    "jacobian_trap_sponge": lambda: cases.trap_sponge_cfg(nonlinear_advection="jacobian"),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_state_roundtrip(name):
    cfg = CONFIGS[name]()
    g, ctx = _ctx(cfg)
    s = cases.active_state(cfg)
    say("roundtrip", name)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    q, w, th, thb = ctx.get_state()
    assert np.array_equal(q, s.q) and np.array_equal(w, s.w)
    assert np.array_equal(th, s.th) and np.array_equal(thb, s.thb)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_explicit_rhs(name):
    cfg = CONFIGS[name]()
    g, ctx = _ctx(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg)
    say("explicit_rhs", name)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    Eq, Ew, Eth = ctx.get_aux()
    ref = O.explicit_rhs(s, o)
    assert rel(Eq, ref.q) < TOL
    assert rel(Ew, ref.w) < TOL
    assert rel(Eth, ref.th) < TOL


@pytest.mark.parametrize("name", list(CONFIGS))
def test_implicit_tendency(name):
    cfg = CONFIGS[name]()
    g, ctx = _ctx(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg)
    say("implicit_tendency", name)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_IMPLICIT_TEND)
    Iq, Iw, Ith = ctx.get_aux()
    ref = O.implicit_tendency(s, o)
    assert rel(Iq, ref.q) < TOL
    assert rel(Iw, ref.w) < TOL
    assert rel(Ith, ref.th) < TOL


@pytest.mark.parametrize("name", list(CONFIGS))
def test_implicit_solve(name):
    cfg = CONFIGS[name]()
    g, ctx = _ctx(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    say("implicit_solve", name)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_IMPLICIT_SOLVE)
    q, w, th, thb = ctx.get_state()
    rq, rw, rth = O.implicit_solve(s.q, s.w, s.th, o)
    assert rel(q, rq) < TOL
    assert rel(w, rw) < TOL
    assert rel(th, rth) < TOL
    assert rel(thb, o.proj_d @ s.thb) < TOL


@pytest.mark.parametrize("name", ["fixture", "trap_sponge", "neumann"])
def test_sbp_corrector(name):
    cfg = CONFIGS[name]()
    g, ctx = _ctx(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg)
    total = g.gamma_imex * g.dt
    say("corrector", name)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_CORRECTOR, total)
    q, w, th, thb = ctx.get_state()
    ref = O.sbp_corrector(s, o, total)
    assert rel(th, ref.th) < TOL
    assert rel(thb, ref.thb) < TOL
    assert np.array_equal(q, s.q) and np.array_equal(w, s.w)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_one_step(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    say("one step", name)
    got = imex_step(State(s.q, s.w, s.th, s.thb), g)
    ref = O.imex_step(s, o)
    assert rel(got.q_hat, ref.q) < TOL
    assert rel(got.w_hat, ref.w) < TOL
    assert rel(got.th_hat, ref.th) < TOL
    if np.abs(ref.thb).max() > 0:
        assert rel(got.th_bar, ref.thb) < TOL


WIDE = {
    # Nx = 256: the 16-thread row pass, with and without the trap and sponge
    "trap_sponge_256": lambda: cases.trap_sponge_cfg(Nx=256, Nz=10),
    "trap_256": lambda: cases.trap_sponge_cfg(Nx=256, Nz=8, sponge_rate=0.0),
    "sponge_256": lambda: cases.trap_sponge_cfg(Nx=256, Nz=8, gamma=0.0, sponge_r_start=7.0),
    "fixture_256": lambda: cases.fixture_cfg(Nx=256, Nz=8),
    "trap_sponge_512": lambda: cases.trap_sponge_cfg(Nx=512, Nz=10),
    "fixture_1024": lambda: cases.fixture_cfg(Nx=1024, Nz=8),
}


@pytest.mark.parametrize("name", list(WIDE))
def test_one_step_wide(name):
    """The Nx = 256, 512 and 1024 instantiations of the FFT passes, at small
    Nz so the oracle step stays at seconds."""
    cfg = WIDE[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    say("one step", name)
    got = imex_step(State(s.q, s.w, s.th, s.thb), g)
    ref = O.imex_step(s, o)
    assert rel(got.q_hat, ref.q) < TOL
    assert rel(got.w_hat, ref.w) < TOL
    assert rel(got.th_hat, ref.th) < TOL
    assert rel(got.th_bar, ref.thb) < TOL
    g.close()


def test_golden_fixture_20_steps():
    """The reference's own 20-step regression state
    (tests/data/m2b_dirichlet_evolvemean_ref.npz of the reference), from
    make_initial_state(seed=0, amplitude=1e-3) with the JAX draw restated."""
    import os

    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s0 = O.jax_initial_state(o, 0, 1e-3, True)
    path = os.path.join(os.path.dirname(__file__), "golden", "m2b_dirichlet_evolvemean_ref.npz")
    ref = np.load(path)
    say("golden 20 steps")
    sp = Stepper(g)
    sp.set_state(State(s0.q, s0.w, s0.th, s0.thb))
    sp.run(20)
    st = sp.state()
    assert rel(st.q_hat, ref["q_hat"]) < TOL
    assert rel(st.w_hat, ref["w_hat"]) < TOL
    assert rel(st.th_hat, ref["th_hat"]) < TOL
    # th_bar is ~1e-15 in this near-linear fixture: compare absolutely
    assert np.abs(st.th_bar - ref["th_bar"]).max() < 1e-20


def test_multi_step_tracks_oracle():
    cfg = cases.trap_sponge_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg, amp=3e-2)
    say("10 steps trap_sponge")
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    sp.run(10)
    got = sp.state()
    ref = s
    for _ in range(10):
        ref = O.imex_step(ref, o)
    assert rel(got.q_hat, ref.q) < 1e-11
    assert rel(got.w_hat, ref.w) < 1e-11
    assert rel(got.th_hat, ref.th) < 1e-11
    assert rel(got.th_bar, ref.thb) < 1e-11


def test_unsanitized_input_full_layout():
    """Out-of-band and anti-Hermitian input content is handled like the
    reference: the first step runs in the full layout."""
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    rng = np.random.default_rng(3)
    q = s.q + 1e-4 * (1.0 - o.mask)[None] * (rng.standard_normal(s.q.shape)
                                              + 1j * rng.standard_normal(s.q.shape))
    q[:, 1:5, 0] += 1e-4j  # anti-Hermitian ghost on the ky=0 column
    s2 = O.OState(q, s.w, s.th, s.thb)
    say("unsanitized full layout")
    got = imex_step(State(q, s.w, s.th, s.thb), g)
    ref = O.imex_step(s2, o)
    assert rel(got.q_hat, ref.q) < TOL
    assert rel(got.w_hat, ref.w) < TOL
    assert rel(got.th_hat, ref.th) < TOL
    out = np.abs(got.q_hat) * (1.0 - o.mask)[None]
    assert float(out.max()) == 0.0


def test_graph_replay_deterministic():
    cfg = cases.trap_sponge_cfg()
    g = make_grid(cfg)
    s = cases.active_state(cfg)
    say("determinism")
    outs = []
    for _ in range(2):
        sp = Stepper(g)
        sp.set_state(State(s.q, s.w, s.th, s.thb))
        sp.run(3)
        outs.append(sp.state())
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["fixture", "trap_sponge", "fixed"])
def test_diagnostics(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg)
    say("diagnostics", name)
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    d = sp.diagnostics()
    r = O.diagnostics(s, o)
    for k in ("KE_tot", "KE_bt", "max_speed", "enstrophy", "Nusselt", "Nusselt_dealiased"):
        assert abs(d[k] - r[k]) <= 1e-12 * max(abs(r[k]), 1e-300), (k, d[k], r[k])
    for k in ("q_vert_spec", "w_vert_spec", "th_vert_spec"):
        assert rel(d[k], r[k]) < 1e-12


def test_scale_separated_fields():
    """A strong barotropic q' next to weak w, theta: each field must keep
    its own relative accuracy (no roundoff leaking between fields)."""
    cfg = cases.trap_sponge_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    q = s.q.copy()
    q[0] += 1e6 * s.q[0] / np.abs(s.q[0]).max()
    w0 = np.zeros_like(s.w)
    st = O.OState(q, w0, 1e-6 * s.th, s.thb)
    say("scale separated")
    ctx = g.context(full=False)
    ctx.set_state(st.q, st.w, st.th, st.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    Eq, Ew, Eth = ctx.get_aux()
    ref = O.explicit_rhs(st, o)
    assert np.abs(Ew).max() == 0.0  # advecting and damping w = 0 gives exactly 0
    assert rel(Eq, ref.q) < TOL
    assert rel(Eth, ref.th) < TOL
    got = imex_step(State(st.q, st.w, st.th, st.thb), g)
    r1 = O.imex_step(st, o)
    r2 = cases.oracle_reordered(O.imex_step, st, o)
    # with 12 decades between q' and theta the weak fields carry amplified
    # roundoff of q' in any implementation; the bar is the oracle's own
    # spread under reordered sums (x10), or 1e-12 if that is smaller
    for a, b, c in ((got.q_hat, r1.q, r2.q), (got.w_hat, r1.w, r2.w), (got.th_hat, r1.th, r2.th)):
        floor = rel(c, b)
        assert rel(a, b) < max(TOL, 10.0 * floor), (rel(a, b), floor)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_device_shell_factors(name):
    """Shell factors built and inverted on the GPU (FP64 Gauss-Jordan) match
    the host LAPACK inverses, and the step is the same with either."""
    from nhqg_amd.grid import shell_inverses, shell_matrix_B0

    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    g.shell_solver = "matrices"
    ctx = g.context(full=False)
    lay = ctx.layout
    vals = g._shell_values[lay.tables["shell_ids"]]
    ref = shell_inverses(vals, g.cfg, g.gamma_imex, shell_matrix_B0(g), g.buoyancy_K, g.Ld_inv_sq)
    say("device shell factors", name, vals.size)
    for s in range(vals.size):
        assert rel(ctx.shell_inverse(s), ref[s]) < 1e-13, s
    st = cases.active_state(cfg)
    gh = make_grid(cfg)
    gh.shell_solver = "matrices"
    gh.shell_build = "host"
    a = imex_step(State(st.q, st.w, st.th, st.thb), g)
    b = imex_step(State(st.q, st.w, st.th, st.thb), gh)
    for x, y in zip(a, b):
        assert rel(x, y) < 1e-13


def test_device_shell_factors_256():
    """At 256^2 x 256 (2,755 touched shells) the device inverses match LAPACK."""
    from nhqg_amd.grid import shell_inverses, shell_matrix_B0
    from nhqg_amd.init import polar_config_256

    cfg = polar_config_256()
    g = make_grid(cfg, shells="band")
    g.shell_solver = "matrices"
    ctx = g.context(full=False)
    vals = g._shell_values[ctx.layout.tables["shell_ids"]]
    pick = np.unique(np.linspace(0, vals.size - 1, 24).astype(int))
    ref = shell_inverses(vals[pick], g.cfg, g.gamma_imex, shell_matrix_B0(g), g.buoyancy_K,
                         g.Ld_inv_sq)
    say("device shell factors 256^3", vals.size)
    for k, s in enumerate(pick):
        assert rel(ctx.shell_inverse(int(s)), ref[k]) < 1e-13, s


@pytest.mark.parametrize("name", ["fixture", "trap_sponge", "fixed"])
def test_eigen_shell_solve_matches_matrices(name):
    """Dirichlet top: the eigenbasis solve and the per-shell matrices give
    the same step."""
    cfg = CONFIGS[name]()
    st = cases.active_state(cfg)
    ge = make_grid(cfg)
    gm = make_grid(cfg)
    gm.shell_solver = "matrices"
    say("eigen vs matrices", name)
    a = imex_step(State(st.q, st.w, st.th, st.thb), ge)
    b = imex_step(State(st.q, st.w, st.th, st.thb), gm)
    for x, y in zip(a, b):
        assert rel(x, y) < 1e-13
