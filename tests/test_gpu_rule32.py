"""The 3/2 dealiasing rule on the GPU, against the oracle.

horizontal_dealiasing='32_rule' is the reference's default
(nhqg/config.py:62): the products run on the zero-padded 3Nx/2 grid and are
truncated with the factor (Npad/Nx)^2 (nhqg/spectral.py:18-50, 62-74,
155-253), the state is full-band and sanitized by the Hermitian projection
alone (solver.py:128-146), and the SBP corrector's horizontal means use the
padded grid (solver.py:873-906).  The states here carry content on every
mode, the kx = -Nx/2 row and ky = Nx/2 column included, so the alias of the
products of two such components that the 3Nx/2 grid folds back onto kept
modes is in every comparison.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12 per field for one step or one
explicit RHS (FP64), 1e-11 after ten steps.
"""

import numpy as np
import pytest

from nhqg_amd import State, Stepper, imex_step, make_grid, rk4_step
from nhqg_amd import _native
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12

R32 = dict(horizontal_dealiasing="32_rule")


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


def full_band_state(cfg, seed=0, amp=2e-3):
    """cases.active_state plus seeded content on every mode (decaying with
    the Chebyshev index), Nyquist row and column included, Hermitian on the
    self-conjugate columns (the oracle's sanitize under the 3/2 rule)."""
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg, seed)
    rng = np.random.default_rng(1000 + seed)

    def noise(nl):
        f = rng.standard_normal((nl, cfg.Nx, cfg.Nx // 2 + 1)) + \
            1j * rng.standard_normal((nl, cfg.Nx, cfg.Nx // 2 + 1))
        f *= amp * np.exp(-np.arange(nl) / 3.0)[:, None, None]
        f[:, 0, 0] = 0.0
        return f

    scale = np.abs(s.q).max()
    st = O.OState(s.q + scale * noise(cfg.Nz + 1), s.w + scale * noise(cfg.Nz - 1),
                  s.th + scale * noise(cfg.Nz - 1), s.thb)
    st = O.sanitize(st, o)
    assert np.abs(st.q[:, cfg.Nx // 2, 1:]).max() > 0 and np.abs(st.q[:, :, -1]).max() > 0
    return st


def _check(got, ref, tol, what):
    errs = [rel(got[0], ref.q), rel(got[1], ref.w), rel(got[2], ref.th)]
    if len(got) > 3 and np.abs(ref.thb).max() > 0:
        errs.append(rel(got[3], ref.thb))
    say("  ", what, ["%.2e" % e for e in errs])
    assert max(errs) < tol


# Every transform length the padded passes instantiate: Npad = 3Nx/2 for
# Nx = 16 .. 1024 (P = Nx/2 = 8 .. 512), both advection forms; small Nz so
# the oracle's products on the padded grid stay at seconds.
SIZES = [16, 32, 64, 128, 256, 512, 1024]


@pytest.mark.parametrize("adv", ["flux", "jacobian"])
@pytest.mark.parametrize("nx", SIZES)
def test_rule32_explicit_rhs_every_size(nx, adv):
    cfg = cases.fixture_cfg(Nx=nx, Nz=4, nonlinear_advection=adv, **R32)
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = full_band_state(cfg)
    say("rule32 explicit_rhs", nx, adv)
    ctx = g.context()
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    _check(ctx.get_aux(), O.explicit_rhs(s, o), TOL, "E")
    g.close()


CONFIGS = {
    "fixture": lambda: cases.fixture_cfg(**R32),
    "trap_sponge": lambda: cases.trap_sponge_cfg(**R32),
    "neumann": lambda: cases.neumann_cfg(**R32),
    "legacy": lambda: cases.fixture_cfg(mean_exchange_discretization="legacy", **R32),
    "fixed": lambda: cases.fixture_cfg(thermal_closure="fixed_conduction",
                                       mean_exchange_discretization="legacy", **R32),
    "jacobian": lambda: cases.fixture_cfg(nonlinear_advection="jacobian", **R32),
    "jacobian_trap_sponge": lambda: cases.trap_sponge_cfg(nonlinear_advection="jacobian",
                                                          **R32),
    # the reference's bare defaults (3/2 rule, Jacobian form, legacy
    # exchange, evolve_mean) in FP64
    "defaults": lambda: cases.fixture_cfg(nonlinear_advection="jacobian",
                                          mean_exchange_discretization="legacy", **R32),
}


@pytest.mark.parametrize("name", ["trap_sponge", "jacobian_trap_sponge"])
def test_rule32_explicit_rhs_trap_sponge(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = full_band_state(cfg)
    say("rule32 explicit_rhs", name)
    ctx = g.context()
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    _check(ctx.get_aux(), O.explicit_rhs(s, o), TOL, "E")
    g.close()


@pytest.mark.parametrize("name", list(CONFIGS))
def test_rule32_one_step(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = full_band_state(cfg)
    say("rule32 one step", name)
    got = imex_step(State(s.q, s.w, s.th, s.thb), g)
    _check(got, O.imex_step(s, o), TOL, "step")
    g.close()


@pytest.mark.parametrize("name", ["fixture", "trap_sponge", "defaults"])
def test_rule32_ten_steps_resident(name):
    """Ten steps of the captured step graph on a resident state (the
    Stepper), each tracked against the oracle."""
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = full_band_state(cfg, seed=2)
    say("rule32 ten steps", name)
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    ref = s
    for n in range(10):
        sp.run(1)
        ref = O.imex_step(ref, o)
    _check(sp.state(), ref, 1e-11, "10 steps")
    sp.close()
    g.close()


def test_rule32_unsanitized_input():
    """Anti-Hermitian content on the self-conjugate columns enters the
    products like the reference's and leaves with the final sanitation."""
    cfg = CONFIGS["fixture"]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = full_band_state(cfg, seed=4)
    q = s.q.copy()
    q[:, 1:5, 0] += 1e-4j
    q[:, 3, -1] += 2e-4
    s2 = O.OState(q, s.w, s.th, s.thb)
    say("rule32 unsanitized input")
    got = imex_step(State(q, s.w, s.th, s.thb), g)
    _check(got, O.imex_step(s2, o), TOL, "step")
    g.close()


RK4 = {
    "fixed": CONFIGS["fixed"],
    # explicit mean diffusion: dt inside RK4's stability region at Nz = 16
    # (at dt = 1e-3 the RK4 polynomial amplifies th_bar's roundoff)
    "evolve_pc": lambda: cases.fixture_cfg(dt=2e-5, **R32),
    "evolve_legacy": lambda: cases.fixture_cfg(mean_exchange_discretization="legacy", dt=2e-5,
                                               **R32),
}


@pytest.mark.parametrize("name", list(RK4))
def test_rule32_rk4_step(name):
    cfg = RK4[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = full_band_state(cfg, seed=6)
    say("rule32 rk4", name)
    got = rk4_step(State(s.q, s.w, s.th, s.thb), g)
    _check(got, O.rk4_step(s, o), TOL, "rk4")
    g.close()


def test_rule32_diagnostics():
    cfg = CONFIGS["trap_sponge"]()
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = full_band_state(cfg, seed=8)
    say("rule32 diagnostics")
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    d, r = sp.diagnostics(), O.diagnostics(s, o)
    for k in ("KE_tot", "KE_bt", "max_speed", "max_w", "enstrophy", "Nusselt",
              "Nusselt_dealiased", "vol_avg_tw_dealiased", "q_rms", "th_rms"):
        assert abs(d[k] - r[k]) <= 1e-12 * max(abs(r[k]), 1e-300), (k, d[k], r[k])
    sp.close()
    g.close()
