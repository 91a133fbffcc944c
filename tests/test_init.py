"""Initial-state tools: spectral upsampling (the reference's
upsample_checkpoint_horiz.py method) and cyclone injection, pinned by closed
forms: co-located grid values are preserved, the band of an upsampled
spectrum is the old band times (N_new/N_old)^2, and an injected cyclone
appears only in the barotropic level."""

import numpy as np
import pytest

from nhqg_amd import make_grid
from nhqg_amd.init import (barotropic_vorticity_state, inject_cyclones, thermal_noise_state,
                           turbulent_state, upsample_spectral, upsample_state)
from tests import cases


@pytest.mark.parametrize("nx,nx_new", [(16, 32), (8, 16), (16, 64), (32, 96)])
def test_upsample_keeps_colocated_values(nx, nx_new):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, nx, nx))
    X = np.fft.rfft2(x, axes=(1, 2))
    Y = upsample_spectral(X, nx_new)
    assert Y.shape == (3, nx_new, nx_new // 2 + 1)
    y = np.fft.irfft2(Y, s=(nx_new, nx_new), axes=(1, 2))
    step = nx_new // nx
    assert np.abs(y[:, ::step, ::step] - x).max() < 1e-12 * np.abs(x).max()


def test_upsample_band_is_the_old_band_scaled():
    nx, nx_new = 32, 128
    rng = np.random.default_rng(1)
    kx = np.fft.fftfreq(nx, d=1.0 / nx).round().astype(int)
    m = ((np.abs(kx)[:, None] <= nx // 3) & (np.arange(nx // 2 + 1)[None, :] <= nx // 3))
    x = rng.standard_normal((2, nx, nx))
    X = np.fft.rfft2(x, axes=(1, 2)) * m
    Y = upsample_spectral(X, nx_new)
    rows = np.where(kx >= 0, kx, nx_new + kx)
    for i in range(nx):
        if np.abs(kx[i]) > nx // 3:
            continue
        np.testing.assert_allclose(Y[:, rows[i], : nx // 3 + 1],
                                   X[:, i, : nx // 3 + 1] * (nx_new / nx) ** 2, rtol=0,
                                   atol=1e-12 * np.abs(X).max() * (nx_new / nx) ** 2)
    # nothing outside the old band
    keep = np.zeros(Y.shape[1:], bool)
    keep[rows[np.abs(kx) <= nx // 3][:, None], np.arange(nx // 3 + 1)[None, :]] = True
    assert np.abs(Y[:, ~keep]).max() < 1e-12 * np.abs(Y).max()


def test_upsample_state_to_a_finer_grid():
    g0 = make_grid(cases.fixture_cfg(Nx=32))
    g1 = make_grid(cases.fixture_cfg(Nx=64))
    s0 = turbulent_state(g0, seed=2)
    s1 = upsample_state(s0, g1)
    assert s1.q_hat.shape == (g1.Nz + 1, 64, 33)
    for a, b in ((s0.q_hat, s1.q_hat), (s0.th_hat, s1.th_hat)):
        pa = np.fft.irfft2(a, s=(32, 32), axes=(1, 2))
        pb = np.fft.irfft2(b, s=(64, 64), axes=(1, 2))
        assert np.abs(pb[:, ::2, ::2] - pa).max() < 1e-12 * np.abs(pa).max()
    assert np.array_equal(s1.th_bar, s0.th_bar)


def test_inject_cyclones_touches_only_the_barotropic_level():
    g = make_grid(cases.fixture_cfg(Nx=64))
    s = thermal_noise_state(g, seed=3, amplitude=1e-2)
    out = inject_cyclones(s, g, amplitude=5.0, radius=2.0, count=2,
                          rng=np.random.default_rng(4))
    assert np.array_equal(out.q_hat[1:], s.q_hat[1:])
    assert np.array_equal(out.th_hat, s.th_hat)
    d = np.fft.irfft2(out.q_hat[0] - s.q_hat[0], s=(64, 64))
    # two Gaussians of peak 5 (slightly smoothed by the 2/3 band)
    assert 4.5 < d.max() < 10.5
    assert np.abs((out.q_hat[0] - s.q_hat[0]) * (1.0 - g.mask_23)).max() == 0.0


@pytest.mark.parametrize("confine", [None, 6.0])
def test_barotropic_vorticity_state(confine):
    """scripts/run_polar.py:115-158 by its properties: vorticity standard
    deviation = amplitude (over the cap when confined), only the barotropic
    T_0 level set, band-limited, no mean and the annulus's spectrum peaked
    near k_peak (unconfined), deterministic in the seed."""
    g = make_grid(cases.fixture_cfg(Nx=64, L=20.0))
    kp, kw, amp = 2.0, 0.5, 3.0
    s = barotropic_vorticity_state(g, 5, kp, kw, amp, confine)
    zeta = np.fft.irfft2(s.q_hat[0], s=(64, 64))
    if confine is None:
        assert abs(zeta.std() - amp) < 1e-12 * amp
        spec = np.abs(s.q_hat[0]) ** 2
        k = np.sqrt(g.ksq)
        kbar = (spec * k).sum() / spec.sum()
        assert abs(kbar - kp) < 2 * kw
    else:
        xg = np.arange(64) * 20.0 / 64
        X, Y = np.meshgrid(xg, xg, indexing="ij")
        inside = np.hypot(X - 10.0, Y - 10.0) < confine
        assert abs(zeta[inside].std() - amp) < 1e-12 * amp
        assert np.abs(zeta[~inside]).max() < np.abs(zeta[inside]).max()
    assert np.all(s.q_hat[1:] == 0) and np.all(s.w_hat == 0) and np.all(s.th_hat == 0)
    assert np.all(s.th_bar == 0)
    assert np.abs(s.q_hat[0] * (1.0 - g.mask_23)).max() < 1e-12 * np.abs(s.q_hat[0]).max()
    if confine is None:
        # (the cap taper leaves a mean; the solver's k = 0 projection removes it)
        assert abs(s.q_hat[0, 0, 0]) < 1e-12 * np.abs(s.q_hat[0]).max()
    s2 = barotropic_vorticity_state(g, 5, kp, kw, amp, confine)
    assert np.array_equal(s2.q_hat, s.q_hat)
