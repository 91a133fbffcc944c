"""Why the mixed step keeps its FFT intermediates in FP64.

The device's flux products run as three passes: a column pass (the inverse
transform over kx), a row pass (the inverse over ky, the products, the
forward over ky) and a column pass (the forward over kx).  The two
intermediates between them are the largest HBM streams of the step.  Storing
them in FP32 would halve that traffic; this test measures what it would cost
in accuracy with the oracle: each 2-D transform of the flux products and the
sponge is split into the same two 1-D passes, the intermediate is rounded to
complex64 between them, and the products stay FP64.  From a turbulent state
(velocity rms 20, theta' rms 0.1) w leaves the mixed step's bar (1e-8
relative field error) within ten steps, through the rounded velocities that
every product uses, while q and theta stay well below it.
"""

import numpy as np

from nhqg_amd import make_grid
from nhqg_amd.init import polar_config_256, turbulent_state
from oracle import nhqg_oracle as O


def _r32(z):
    return z.astype(np.complex64).astype(np.complex128)


def _inv2(S, n):
    """irfft2 with the x-pass result stored in FP32."""
    return np.fft.irfft(_r32(np.fft.ifft(S, axis=-2)), n=n, axis=-1)


def _fwd2(p):
    """rfft2 with the y-pass result stored in FP32."""
    return np.fft.fft(_r32(np.fft.rfft(p, axis=-1)), axis=-2)


def _flux23_fp32(psi, fields, o):
    n = o.Nx
    u = _inv2(-1j * o.ky * psi, n)
    v = _inv2(1j * o.kx * psi, n)
    out = []
    for f in fields:
        fp = _inv2(f, n)
        out.append((1j * o.kx * _fwd2(u * fp) + 1j * o.ky * _fwd2(v * fp)) * o.mask)
    return out


def _sponge23_fp32(f, o):
    return _fwd2(o.sponge_phys[None] * _inv2(f, o.Nx)) * o.mask


def test_fp32_fft_intermediates_miss_the_mixed_bar():
    # the polar trap and sponge are under-resolved at Nx = 128; the turbulent
    # flow is what the rounding acts on
    cfg = polar_config_256(Nx=128, Nz=32, gamma=0.0, sponge_rate=0.0)
    s0 = turbulent_state(make_grid(cfg), seed=0)
    o = O.build_operators(cfg)
    a = O.sanitize(O.OState(s0.q_hat, s0.w_hat, s0.th_hat, s0.th_bar), o)
    b = a.copy()
    orig = (O.flux23, O.sponge23)
    try:
        for _ in range(10):
            O.flux23, O.sponge23 = orig
            a = O.imex_step(a, o)
            O.flux23, O.sponge23 = _flux23_fp32, _sponge23_fp32
            b = O.imex_step(b, o)
    finally:
        O.flux23, O.sponge23 = orig
    err = {k: float(np.abs(getattr(b, k) - getattr(a, k)).max() / np.abs(getattr(a, k)).max())
           for k in ("q", "w", "th", "thb")}
    # measured: q 5.0e-11, w 4.8e-8, theta 7.1e-11, th_bar 6.7e-12 after 10 steps
    assert err["w"] > 1e-8
    assert max(err["q"], err["th"], err["thb"]) < 1e-9
