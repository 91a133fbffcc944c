"""The legacy mean-temperature exchange and the full diagnostic set on the
GPU, against the oracle.

* evolve_mean with mean_exchange_discretization='legacy' (the reference's
  default): imex_step_ars222 with -dTheta/dZ w in E_theta, E_th_bar =
  -eps^2 d<w theta>/dZ and the mean's tau solve (solver.py:575-590,
  746-758, 1012-1053), one and ten steps.
* compute_diagnostics' scalars (max_w, max_theta, max_tw, heat-flux
  mismatch, high-mode fractions, the SBP exchange monitor), vertical and
  horizontal shell spectra (diagnostics.py:334-369, 583-710), including a
  full-band state with Nyquist content, where the heat-flux profile and the
  monitor use the padded-grid means of the reference.
"""

import numpy as np
import pytest

from nhqg_amd import State, Stepper, imex_step, make_grid
from nhqg_amd.diagnostics import compute_diagnostics
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


LEGACY = dict(mean_exchange_discretization="legacy")
CONFIGS = {
    "legacy": lambda: cases.fixture_cfg(**LEGACY),
    "legacy_jacobian": lambda: cases.fixture_cfg(nonlinear_advection="jacobian", **LEGACY),
    "legacy_trap_sponge": lambda: cases.trap_sponge_cfg(**LEGACY),
    "legacy_neumann": lambda: cases.neumann_cfg(**LEGACY),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_legacy_exchange_steps_match_oracle(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    st, ref = State(s.q, s.w, s.th, s.thb), s
    say("legacy exchange", name)
    for n in range(10):
        st = imex_step(st, g)
        ref = O.imex_step(ref, o)
        errs = [rel(st.q_hat, ref.q), rel(st.w_hat, ref.w), rel(st.th_hat, ref.th),
                rel(st.th_bar, ref.thb)]
        if n in (0, 9):
            say("  step", n + 1, errs)
        assert max(errs) < (TOL if n == 0 else 1e-11)
    # the mean moves through the exchange (not a trivially zero th_bar)
    assert np.abs(ref.thb - s.thb).max() > 1e-3 * np.abs(s.thb).max()
    g.close()


SCALARS = ("KE_tot", "KE_bt", "KE_bc", "max_speed", "max_w", "max_theta", "max_tw",
           "enstrophy", "Nusselt", "Nusselt_dealiased", "vol_avg_tw", "vol_avg_tw_dealiased",
           "q_rms", "w_rms", "th_rms", "th_bar_max", "q_high_frac", "w_high_frac",
           "th_high_frac")
MONITOR = ("th_mean_feedback_sum_sbp", "mean_flux_exchange_tendency_sbp",
           "mean_theta_exchange_boundary_sbp", "mean_theta_exchange_residual_sbp")

DIAG_CONFIGS = {
    "pc": cases.fixture_cfg,
    "pc_trap_sponge": cases.trap_sponge_cfg,
    "legacy": lambda: cases.fixture_cfg(**LEGACY),
    "fixed_pc": lambda: cases.fixture_cfg(thermal_closure="fixed_conduction"),
    "fixed_legacy": lambda: cases.fixture_cfg(thermal_closure="fixed_conduction", **LEGACY),
}


def _check_diagnostics(d, r, scale_ref=None):
    for k in SCALARS:
        assert abs(d[k] - r[k]) <= 1e-12 * max(abs(r[k]), 1e-300), (k, d[k], r[k])
    # the mismatch is a difference of two near-equal sums: absolute bar on
    # the scale of the sums
    assert abs(d["heat_flux_mismatch"] - r["heat_flux_mismatch"]) <= \
        1e-12 * max(abs(r["vol_avg_tw"]), 1e-300)
    for k in ("q_vert_spec", "w_vert_spec", "th_vert_spec", "q_horiz_spec", "w_horiz_spec",
              "th_horiz_spec"):
        assert d[k].shape == r[k].shape, k
        assert rel(d[k], r[k]) < 1e-12, k
    assert np.array_equal(d["k_bins"], r["k_bins"])
    if np.isnan(r["th_mean_feedback_sum_sbp"]):
        assert all(np.isnan(d[k]) for k in MONITOR)
    else:
        s = abs(r["th_mean_feedback_sum_sbp"]) + abs(r["mean_flux_exchange_tendency_sbp"])
        for k in MONITOR:
            assert abs(d[k] - r[k]) <= 1e-12 * s, (k, d[k], r[k])


@pytest.mark.parametrize("name", list(DIAG_CONFIGS))
def test_full_diagnostics_match_oracle(name):
    cfg = DIAG_CONFIGS[name]()
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    s = cases.active_state(cfg)
    say("diagnostics", name)
    sp = Stepper(g)
    sp.set_state(State(s.q, s.w, s.th, s.thb))
    _check_diagnostics(sp.diagnostics(), O.diagnostics(s, o))


def test_diagnostics_of_a_full_band_state_use_padded_means():
    """Random full-band w, theta and th_bar (the reference's exchange-monitor
    test, tests/test_solver.py:637-678, under the 2/3 rule): Nyquist
    content makes the padded-grid means differ from the Nx-grid ones, and
    the monitor still closes."""
    cfg = cases.fixture_cfg(Nx=16, Nz=12, L=20.0, Ra_tilde=50.0, dt=1e-4)
    g = make_grid(cfg)
    o = O.build_operators(cfg, with_shells=False)
    rng = np.random.default_rng(321)
    env = np.sin(np.pi * o.Z)
    wp = env[:, None, None] * rng.standard_normal((cfg.Nz + 1, 16, 16))
    tp = env[:, None, None] * rng.standard_normal((cfg.Nz + 1, 16, 16))
    w = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, np.fft.rfft2(wp, axes=(1, 2)))))
    th = O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, np.fft.rfft2(tp, axes=(1, 2)))))
    tb = rng.standard_normal(cfg.Nz + 1)
    tb[0] = tb[-1] = 0.0
    tb = o.proj_d @ (o.V_inv @ tb)
    q = np.zeros((cfg.Nz + 1, 16, 9), complex)
    s = O.OState(q, w, th, tb)
    say("full-band diagnostics")
    d = compute_diagnostics(State(q, w, th, tb), g)
    r = O.diagnostics(s, o)
    w_n = O.vmat(o.V, O.vmat(o.S_d, w))
    th_n = O.vmat(o.V, O.vmat(o.S_d, th))
    assert np.abs(O.hmean(w_n, th_n, o) - O.hmean(w_n, th_n, o, o.Npad)).max() > 1e-3
    _check_diagnostics(d, r)
    assert abs(d["mean_theta_exchange_residual_sbp_rel"]) < 1e-10
