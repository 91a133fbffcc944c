"""The reference-side binding (nhqg_amd/reference_binding.py) driven with a
reference-shaped Grid NamedTuple (tests/reference_grid.py).

CPU: the binding rebuilds this package's grid from the reference field
arrays exactly (same device operators, same shell coefficients recovered
without a config) and rejects what the device does not build.  GPU: its
step is bit-identical to nhqg_amd.imex_step on the same arrays, and within
the per-step bar of the oracle when the tuple holds the oracle's own
operators (also for input that is not band-limited)."""

import numpy as np
import pytest

from nhqg_amd import NHQGConfig, State, imex_step, make_grid, rk4_step
from nhqg_amd import reference_binding as RB
from nhqg_amd.grid import device_operators, shell_coefficients
from oracle import nhqg_oracle as O
from tests import cases
from tests.reference_grid import FIELDS, RefGrid, from_oracle, from_package_grid

CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "neumann": cases.neumann_cfg,
    "fixed_hyper": lambda: cases.fixture_cfg(thermal_closure="fixed_conduction",
                                             mean_exchange_discretization="legacy",
                                             hyper_order=2, nu_q=1e-3, nu_w=2e-3,
                                             nu_theta=3e-3, drag=0.05),
    # the reference's default dealiasing: products on the 3Nx/2 grid
    "rule32_trap_sponge": lambda: cases.trap_sponge_cfg(horizontal_dealiasing="32_rule",
                                                        nonlinear_advection="jacobian"),
    "cutoff": lambda: cases.fixture_cfg(vertical_cutoff_n=10),
}


def test_reference_grid_stand_in_has_the_reference_fields():
    assert len(FIELDS) == 79 and RefGrid._fields == FIELDS


@pytest.mark.parametrize("name", list(CONFIGS))
def test_binding_rebuilds_the_device_operators(name):
    cfg = CONFIGS[name]()
    g = make_grid(cfg)
    gb = RB.grid_from_reference(from_package_grid(g, cfg))
    a = device_operators(g, g.layout(False))
    b = device_operators(gb, gb.layout(False))
    assert set(a) == set(b)
    for k in a:
        if a[k] is None:
            assert b[k] is None, k
        elif k in ("mode_alpha", "mode_c", "shell_coef"):
            # from rates recovered off the dissipation arrays: equal to roundoff
            x, y = np.asarray(a[k]), np.asarray(b[k])
            assert np.abs(x - y).max() <= 1e-15 * np.abs(x).max(), k
        else:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    vals = g._shell_values
    neu = g.buoyancy_K is not None
    for x, y in zip(shell_coefficients(vals, cfg, g.gamma_imex, g.Ld_inv_sq, neu),
                    shell_coefficients(vals, gb.cfg, gb.gamma_imex, gb.Ld_inv_sq, neu)):
        assert np.abs(x - y).max() <= 1e-15 * max(np.abs(x).max(), 1.0)


def test_binding_recovers_rates_from_oracle_arrays():
    cfg = CONFIGS["fixed_hyper"]()
    r = RB._rates(from_oracle(cfg), O.build_operators(cfg, with_shells=False).ksq)
    assert r.hyper_order == 2
    for k in ("nu_q", "nu_w", "nu_theta", "drag", "Ra_tilde", "sigma", "dt"):
        assert abs(getattr(r, k) - getattr(cfg, k)) <= 1e-14 * max(abs(getattr(cfg, k)), 1.0), k


def test_binding_rejects_branches_not_built():
    # the reference defaults (3/2 rule, Jacobian form, legacy exchange) in FP64
    ref = from_oracle(NHQGConfig(Nx=32, Nz=8, L=20.0, dt=1e-3, float_dtype="float64"))
    assert RB.supported(ref)
    ref = ref._replace(sbp_transfer_mode="mass_adjoint")
    assert not RB.supported(ref)
    assert "sbp_transfer_mode" in RB.unsupported_reason(ref)
    with pytest.raises(NotImplementedError):
        RB.grid_from_reference(ref)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CONFIGS))
def test_binding_step_is_bit_identical_to_the_package_step(name):
    """Bit-identical where the recovered rates are exact (nu = 1, p = 1, no
    drag); within 1e-14 with hyperviscosity and drag."""
    cfg = CONFIGS[name]()
    print("binding bit-identity", name, flush=True)
    g = make_grid(cfg)
    ref = from_package_grid(g, cfg)
    s = cases.active_state(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    got = RB.step(st, ref)
    want = imex_step(st, g)
    assert type(got) is State
    for a, b in zip(got, want):
        if name == "fixed_hyper":
            assert _rel(a, b) < 1e-14
        else:
            assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fixture", "trap_sponge", "neumann", "rule32_trap_sponge",
                                  "cutoff"])
def test_binding_on_oracle_operators_matches_oracle(name):
    cfg = CONFIGS[name]()
    print("binding vs oracle", name, flush=True)
    ref = from_oracle(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    got = RB.step(State(s.q, s.w, s.th, s.thb), ref)
    want = O.imex_step(s, o)
    for a, b in zip(got, (want.q, want.w, want.th, want.thb)):
        assert _rel(a, b) < 1e-12
    # input that is neither band-limited nor Hermitian: full device layout
    rng = np.random.default_rng(3)
    noisy = O.OState(s.q + 1e-4 * (rng.standard_normal(s.q.shape)
                                   + 1j * rng.standard_normal(s.q.shape)), s.w, s.th, s.thb)
    got = RB.step(State(noisy.q, noisy.w, noisy.th, noisy.thb), ref)
    want = O.imex_step(noisy, o)
    for a, b in zip(got, (want.q, want.w, want.th, want.thb)):
        assert _rel(a, b) < 1e-12


@pytest.mark.gpu
def test_binding_rk4_and_run():
    cfg = cases.fixture_cfg(thermal_closure="fixed_conduction",
                            mean_exchange_discretization="legacy")
    print("binding rk4/run", flush=True)
    g = make_grid(cfg)
    ref = from_package_grid(g, cfg)
    s = cases.active_state(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    a = RB.rk4_step(st, ref)
    b = rk4_step(st, g)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    seen = []
    final, snaps = RB.run(ref, st, 5, 2, True, lambda s_, k, t: seen.append(k))
    assert seen == [2, 4] and len(snaps) == 2
    x = st
    for _ in range(5):
        x = imex_step(x, g)
    for u, v in zip(final, x):
        assert np.array_equal(u, v)
