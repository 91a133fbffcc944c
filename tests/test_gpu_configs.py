"""The workload configurations beyond the 256^2 x 256 headline, on the GPU.

* Device spectral upsampling (nhqg_upsample) against the host restatement
  of scripts/upsample_checkpoint_horiz.py:27-59.
* 512^2 x 128 polar trap, mixed-precision step from a turbulent state
  (velocity rms 20, theta' rms 0.1): 1000 mixed steps against 1000 plain
  FP64 steps, fields within 1e-8 and KE / Nu_d / max_speed within 1e-9.
* 256^3 polar trap, mixed step with 6 slices: 1000 steps against FP64.
* 1024^2 x 128, dt 1e-5, mixed step with the learned closure, from a
  turbulent 256^2 state upsampled on the device: finite, band-limited and
  Hermitian after 100 steps, and within 1e-8 of the plain step with the
  same closure.
* 256^2 x 64 plus the learned closure (committed weights) against the
  256^3 plain run over 1000 steps: the closure lowers the top-16 psi-KE
  error and the psi-KE spectrum and Nu_d stay within 5%.
"""

import os

import numpy as np
import pytest

from nhqg_amd import Stepper, make_grid
from nhqg_amd.init import polar_config_256, turbulent_state, upsample_state
from nhqg_amd.solver import is_band_limited, is_hermitian

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "nhqg_amd", "data")


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def say(*a):
    print(*a, flush=True)


def _diag_errs(da, db):
    return {k: abs(da[k] - db[k]) / abs(db[k]) for k in ("KE_tot", "Nusselt_dealiased",
                                                         "max_speed")}


@pytest.mark.parametrize("nx_new", [128, 256])
def test_device_upsampling_matches_host(nx_new):
    from tests import cases

    g0 = make_grid(cases.fixture_cfg(Nx=64, L=40.0), shells="band")
    s0 = turbulent_state(g0, seed=3)
    # also a source with content on the old Nyquist row and column (full layout)
    q = np.array(s0.q_hat)
    q[:, 32, 5] += 1e-3 * (1 + 2j)
    q[:, 7, 32] += 2e-3 * (1 - 1j)
    noisy = type(s0)(q, s0.w_hat, s0.th_hat, s0.th_bar)
    g1 = make_grid(cases.fixture_cfg(Nx=nx_new, L=40.0), shells="band")
    for name, src in (("band", s0), ("nyquist", noisy)):
        say("upsample", name, 64, "->", nx_new)
        a = Stepper(g0)
        a.set_state(src)
        b = Stepper(g1)
        b.upsample_from(a)
        got = b.state()
        want = upsample_state(src, g1)
        for x, y in zip(got, want):
            assert rel(x, y) < 1e-13
        assert np.array_equal(got.th_bar, want.th_bar)


def _mixed_vs_plain(cfg, state_fn, steps, slices, closure=None):
    res = {}
    for sl in (0, slices):
        g = make_grid(cfg, shells="band")
        g.mixed_slices = sl
        if closure is not None:
            closure.attach(g)
        sp = Stepper(g)
        state_fn(g, sp)
        say("  slices", sl, "steps", steps)
        sp.run(steps)
        res[sl] = (sp.state(), sp.diagnostics())
        sp.close()
        g.close()
    (a, da), (b, db) = res[slices], res[0]
    errs = {n: rel(x, y) for n, x, y in zip(("q", "w", "th", "thb"), a, b)}
    return a, errs, _diag_errs(da, db)


def test_c3_mixed_512x512x128_turbulent_1000_steps():
    cfg = polar_config_256(Nx=512, Nz=128)
    say("C3: 512^2 x 128 mixed vs FP64, turbulent state")
    _, errs, derr = _mixed_vs_plain(
        cfg, lambda g, sp: sp.set_state(turbulent_state(g, seed=0)), 1000, 6)
    say("  field errors", errs, "diagnostic errors", derr)
    assert max(errs.values()) < 1e-8
    assert max(derr.values()) < 1e-9


def test_mixed_256_cubed_six_slices_1000_steps():
    from nhqg_amd.init import polar_state_256

    cfg = polar_config_256()
    say("256^3 mixed (6 slices) vs FP64")
    _, errs, derr = _mixed_vs_plain(
        cfg, lambda g, sp: sp.set_state(polar_state_256(g, seed=0)), 1000, 6)
    say("  field errors", errs, "diagnostic errors", derr)
    assert max(errs.values()) < 1e-8
    assert max(derr.values()) < 1e-9


def test_c5_1024x1024x128_mixed_closure_from_upsampled_state():
    from nhqg_amd.closure import Closure

    path = os.path.join(DATA, "closure_256x256x128.npz")
    cfg = polar_config_256(Nx=1024, Nz=128, dt=1e-5)
    g0 = make_grid(polar_config_256(Nx=256, Nz=128, dt=1e-5), shells="band")
    src = Stepper(g0)
    src.set_state(turbulent_state(g0, seed=0))

    def start(g, sp):
        sp.upsample_from(src)

    say("C5: 1024^2 x 128 mixed + closure, 100 steps")
    c = Closure.load(path, nz=128)
    st, errs, derr = _mixed_vs_plain(cfg, start, 100, 6, closure=c)
    say("  field errors", errs, "diagnostic errors", derr)
    g = make_grid(cfg, shells="band")
    for f in st:
        assert np.all(np.isfinite(f))
    assert is_band_limited(st, g) and is_hermitian(st, g)
    assert max(errs.values()) < 1e-8
    assert max(derr.values()) < 1e-9


def test_c4_coarse_64_with_closure_against_256_cubed():
    from nhqg_amd.closure import Closure, judge
    from nhqg_amd.init import add_barotropic_flow, thermal_noise_state

    path = os.path.join(DATA, "closure_256x256x64.npz")
    cfg_f = polar_config_256()
    cfg_c = polar_config_256(Nz=64)
    c = Closure.load(path, nz=64)
    gj = make_grid(cfg_f, shells="band")
    s = add_barotropic_flow(thermal_noise_state(gj, seed=10, amplitude=0.1), gj, seed=11,
                            u_rms=20.0)
    gj.close()
    say("C4: 256^2 x 64 + closure vs 256^3, 1000 steps")
    res = judge(cfg_f, cfg_c, s, 1000, 500, closure=c)["summary"]
    say("  ", res)
    plain, clos = res["coarse"], res["coarse+closure"]
    # the committed weights carry their validated gain (4): over these 1000
    # steps the top-16 error falls to about a third of the plain run's
    assert c.gain > 1.0
    assert clos["psi_ke_top16_err_mean"] < 0.6 * plain["psi_ke_top16_err_mean"]
    assert clos["psi_ke_spectrum_err_final"] < 0.05
    assert clos["Nu_d_rel_err_final"] < 0.05
