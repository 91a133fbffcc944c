"""The driver's init generators on the device (nhqg_init_barotropic,
nhqg_inject_cyclones) against their host restatements in nhqg_amd.init
(scripts/run_polar.py:91-158), from the same numpy draws.

Bar: the T_0 level of q within 1e-12 of the host state (relative to its
largest mode) on the modes the context holds; every other field exactly
zero after the init.
"""

import numpy as np
import pytest

from nhqg_amd import Stepper, make_grid
from nhqg_amd.init import barotropic_vorticity_state, inject_cyclones, thermal_noise_state
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


def _held(g, f):
    """f on the modes a context of g holds (the 2/3 band, or all under the
    3/2 rule)."""
    return f if g.rule32 else f * g.mask_23


@pytest.mark.parametrize("rule", ["23_rule", "32_rule"])
@pytest.mark.parametrize("nx", [16, 64, 256, 1024])
@pytest.mark.parametrize("confine", [False, True])
def test_device_barotropic_init_matches_host(nx, confine, rule):
    cfg = cases.fixture_cfg(Nx=nx, Nz=4, horizontal_dealiasing=rule)
    g = make_grid(cfg, shells="band")
    L = cfg.L
    k0 = 2.0 * np.pi / L
    kp, kw = k0 * max(2.0, nx / 16.0), k0 * max(1.0, nx / 64.0)
    R = 0.3 * L if confine else None
    say("barotropic init", nx, rule, "confine" if confine else "")
    host = barotropic_vorticity_state(g, seed=7, k_peak=kp, k_width=kw, amplitude=0.5,
                                      confine_radius=R)
    sp = Stepper(g)
    sp.init_barotropic(seed=7, k_peak=kp, k_width=kw, amplitude=0.5, confine_radius=R)
    dev = sp.state()
    assert rel(dev.q_hat[0], _held(g, host.q_hat[0])) < 1e-12
    assert np.abs(dev.q_hat[1:]).max() == 0.0
    assert np.abs(dev.w_hat).max() == 0.0 and np.abs(dev.th_hat).max() == 0.0
    assert np.abs(dev.th_bar).max() == 0.0
    sp.close()
    g.close()


@pytest.mark.parametrize("rule", ["23_rule", "32_rule"])
@pytest.mark.parametrize("nx", [32, 512])
def test_device_cyclone_injection_matches_host(nx, rule):
    cfg = cases.fixture_cfg(Nx=nx, Nz=4, horizontal_dealiasing=rule)
    g = make_grid(cfg, shells="band")
    s = thermal_noise_state(g, seed=3, amplitude=1e-3)
    say("cyclones", nx, rule)
    host = inject_cyclones(s, g, amplitude=2.0, radius=0.05 * cfg.L, count=3,
                           rng=np.random.default_rng(11))
    sp = Stepper(g)
    sp.set_state(s)
    sp.inject_cyclones(amplitude=2.0, radius=0.05 * cfg.L, count=3,
                       rng=np.random.default_rng(11))
    dev = sp.state()
    assert rel(dev.q_hat[0], _held(g, host.q_hat[0])) < 1e-12
    # the other levels and fields are untouched
    assert np.array_equal(dev.q_hat[1:], _held(g, s.q_hat[1:]))
    assert np.array_equal(dev.th_hat, _held(g, s.th_hat))
    sp.close()
    g.close()


def test_device_init_then_step_matches_host_init_then_step():
    """A run started from the device init steps like one started from the
    host init (first-step difference at roundoff)."""
    cfg = cases.fixture_cfg(Nx=64, Nz=8)
    g = make_grid(cfg, shells="band")
    k0 = 2.0 * np.pi / cfg.L
    host = barotropic_vorticity_state(g, seed=2, k_peak=4 * k0, k_width=k0, amplitude=0.3)
    a, b = Stepper(g), Stepper(g)
    a.set_state(host)
    b.init_barotropic(seed=2, k_peak=4 * k0, k_width=k0, amplitude=0.3)
    a.run(5)
    b.run(5)
    sa, sb = a.state(), b.state()
    # w, theta and th_bar of a purely barotropic start stay at roundoff
    # (~1e-18): every field is held to the state's scale, max |q|
    scale = np.abs(sa.q_hat).max()
    for x, y in zip(sb, sa):
        assert np.abs(x - y).max() <= 1e-12 * scale
    a.close()
    b.close()
    g.close()
