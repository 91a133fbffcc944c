"""Learned vertical-tail closure: oracle pinned by finite differences, the
fine -> coarse filter, GPU forward / gradient / training against the oracle,
the closure term inside the step, and the gain over the coarse solver
without it."""

import numpy as np
import pytest

from nhqg_amd import State, make_grid
from nhqg_amd import closure as CL
from oracle import mlp_oracle as MO
from tests import cases


def test_oracle_gradient_matches_finite_differences():
    n_in, width, n_out = 4, 8, 3
    rng = np.random.default_rng(0)
    w = MO.init_params(n_in, width, n_out, seed=1) + 0.05 * rng.standard_normal(
        width * n_in + width + width * width + width + n_out * width + n_out)
    x = rng.standard_normal((7, n_in))
    t = rng.standard_normal((7, n_out))
    loss, g = MO.loss_and_grad(w, x, t, n_in, width, n_out)
    h = 1e-6
    for k in rng.choice(w.size, 25, replace=False):
        e = np.zeros_like(w)
        e[k] = h
        fd = (MO.loss_and_grad(w + e, x, t, n_in, width, n_out)[0]
              - MO.loss_and_grad(w - e, x, t, n_in, width, n_out)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_filter_is_exact_on_coarse_representable_states():
    """A fine state built from a coarse one by zero-padding its Chebyshev
    series filters back to the coarse state."""
    cfg_c = cases.fixture_cfg()
    gc = make_grid(cfg_c)
    gf = make_grid(cases.fixture_cfg(Nz=2 * cfg_c.Nz))
    s = cases.active_state(cfg_c)
    nqc, nqf = gc.Nz + 1, gf.Nz + 1
    pad = lambda a: np.concatenate([a, np.zeros((nqf - nqc,) + a.shape[1:], a.dtype)])
    q = pad(s.q)
    wcheb = pad(np.tensordot(gc.w_stencil, s.w, axes=1))
    tcheb = pad(np.tensordot(gc.dirichlet_stencil, s.th, axes=1))
    w = np.tensordot(gf.w_pinv @ gf.proj_w, wcheb, axes=1)
    th = np.tensordot(gf.dirichlet_pinv @ gf.proj_dirichlet, tcheb, axes=1)
    thb = np.zeros(nqf)
    back = CL.filter_state(State(q, w, th, thb), gf, gc)
    for a, b in ((back.q_hat, s.q), (back.w_hat, s.w), (back.th_hat, s.th)):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def _perturbed_params(seed):
    n = 16 * 128 + 128 + 128 * 128 + 128 + 16 * 128 + 16
    w = MO.init_params(16, 128, 16, seed=seed) + 0.02 * np.random.default_rng(seed + 1).standard_normal(n)
    return w.astype(np.float32)


def _bias_slices():
    o1 = 16 * 128
    o2 = o1 + 128 + 128 * 128
    o3 = o2 + 128 + 16 * 128
    return [slice(o1, o1 + 128), slice(o2, o2 + 128), slice(o3, o3 + 16)]


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    from nhqg_amd import _native
    m = _native.MLP(16, 128, 16, max_batch=1024)
    w = _perturbed_params(3)
    m.set_params(w)
    w = w.astype(np.float64)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((300, 16)).astype(np.float32)
    t = rng.standard_normal((300, 16)).astype(np.float32)
    y = m.forward(x)
    yr, _ = MO.forward(w, x.astype(np.float64), 16, 128, 16)
    assert np.abs(y - yr).max() <= 1e-5 * np.abs(yr).max()
    g, loss = m.gradient(x, t)
    lr, gr = MO.loss_and_grad(w, x.astype(np.float64), t.astype(np.float64), 16, 128, 16)
    assert abs(loss - lr) <= 1e-5 * lr
    assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max()
    m.close()


@pytest.mark.gpu
def test_gpu_odd_network_holds_biases_at_zero():
    from nhqg_amd import _native
    m = _native.MLP(16, 128, 16, max_batch=1024, odd=True)
    m.set_params(_perturbed_params(7))
    w = m.get_params()
    for sl in _bias_slices():
        assert not w[sl].any()
    rng = np.random.default_rng(8)
    x = rng.standard_normal((256, 16)).astype(np.float32)
    m.set_data(x, np.tanh(x[:, ::-1]).astype(np.float32))
    m.train(20, 128, 1e-3)
    w2 = m.get_params()
    for sl in _bias_slices():
        assert not w2[sl].any()
    assert not np.array_equal(w, w2)
    y = m.forward(np.concatenate([x, -x]))
    assert np.array_equal(y[:256], -y[256:])
    m.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_learnable_map():
    from nhqg_amd import _native
    rng = np.random.default_rng(6)
    A = rng.standard_normal((16, 16)) / 4.0
    x = rng.standard_normal((8192, 16))
    t = np.tanh(x @ A.T) * 0.8
    m = _native.MLP(16, 128, 16, max_batch=1024)
    m.set_params(CL.init_params(0))
    m.set_data(x.astype(np.float32), t.astype(np.float32))
    loss = m.train(1500, 512, 2e-3)
    assert np.isfinite(loss).all()
    assert loss[-50:].mean() < 0.05 * loss[:10].mean(), (loss[:10].mean(), loss[-50:].mean())


@pytest.mark.gpu
def test_closure_term_in_the_step():
    """E_q gains out_scale * MLP_bf16(in_scale * q_top) on the top levels
    and nothing elsewhere; a network with biases is refused."""
    from nhqg_amd import _native
    m = _native.MLP(16, 128, 16, max_batch=1024, odd=True)
    m.set_params(_perturbed_params(3))
    w = m.get_params().astype(np.float64)
    cfg = cases.fixture_cfg(Nz=24)
    g = make_grid(cfg)
    ctx = g.context(full=False)
    s = cases.active_state(cfg)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    e0 = ctx.get_aux()[0]
    n0, si, so = g.Nz + 1 - 16, 3.0e3, 2.0e-4
    ctx.set_closure(m, n0, si, so)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    e1 = ctx.get_aux()[0]
    ctx.set_closure(None, 0, 1.0, 1.0)
    # numpy evaluation with the weights rounded to BF16 (nearest even)
    u = w.astype(np.float32).view(np.uint32)
    wb = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)
    x = CL.band_columns(s.q[n0:], g.mask_23) * si
    y, _ = MO.forward(wb.astype(np.float64), x, 16, 128, 16)
    d = CL.band_columns(e1[n0:] - e0[n0:], g.mask_23)
    assert np.abs(d - so * y).max() <= 1e-4 * np.abs(so * y).max()
    assert np.array_equal(e1[:n0], e0[:n0])
    plain = _native.MLP(16, 128, 16, max_batch=64)
    plain.set_params(_perturbed_params(3))
    with pytest.raises(_native.NativeError):
        ctx.set_closure(plain, n0, si, so)
    plain.close()
    m.close()


@pytest.mark.gpu
def test_closure_gains_over_the_coarse_solver():
    """Trained on filtered fine/coarse pairs (fine Nz = 2 x coarse), the
    closure reduces the coarse step's missing top-16 q tendency on held-out
    snapshots."""
    from nhqg_amd.init import add_barotropic_flow, cpu_config_64, thermal_noise_state
    cfg_f = cpu_config_64(Nz=64)
    cfg_c = cpu_config_64(Nz=32)
    gf, gc = make_grid(cfg_f), make_grid(cfg_c)
    s0 = add_barotropic_flow(thermal_noise_state(gf, seed=0, amplitude=1e-2), gf, seed=1,
                             u_rms=15.0)
    train_set = CL.make_pairs(gf, gc, s0, spinup=20, snapshots=6, every=20)
    test_set = CL.make_pairs(gf, gc, s0, spinup=200, snapshots=2, every=20)
    c = CL.train(train_set, steps=2000, batch=512, lr=1e-3)
    res = CL.tendency_gain(c, test_set)
    print("closure held-out", res, flush=True)
    assert np.isfinite(c.loss).all()
    assert res["gain"] > 0.3, res   # 0.52 measured; a closure that learned nothing scores ~0


@pytest.mark.gpu
def test_closure_gain_scales_the_step_term_and_survives_save_load(tmp_path):
    """Closure.gain multiplies the tendency the attached network adds to E_q
    (a gain of 2 doubles the increment, a power of two: exactly up to the
    rounding of the sums), is saved with the weights, and loads as 1 from a
    file without it; predict stays the fitted, unscaled map in batches
    beyond the device network's max_batch."""
    from nhqg_amd import _native
    m = _native.MLP(16, 128, 16, max_batch=4096, odd=True)
    m.set_params(_perturbed_params(5))
    c = CL.Closure(m, 0, 3.0e3, 2.0e-4, np.zeros(1), gain=1.0)
    cfg = cases.fixture_cfg(Nz=24)
    g = make_grid(cfg)
    c.n0 = g.Nz + 1 - 16
    s = cases.active_state(cfg)
    ctx = g.context(full=False)
    ctx.set_state(s.q, s.w, s.th, s.thb)
    ctx.apply(_native.OP_EXPLICIT_RHS)
    e0 = ctx.get_aux()[0]
    inc = {}
    for gain in (1.0, 2.0):
        c.gain = gain
        c.attach(g)
        ctx.apply(_native.OP_EXPLICIT_RHS)
        inc[gain] = ctx.get_aux()[0] - e0
    g.attach_closure(None, 0, 1.0, 1.0)
    big = np.abs(inc[2.0]).max()
    assert big > 0.0
    # the sums E + inc round at the scale of the increments (half an ulp
    # of them measured)
    assert np.abs(inc[2.0] - 2.0 * inc[1.0]).max() <= 1e-15 * big
    # save / load keeps the gain; a file written without one loads gain 1
    c.gain = 4.0
    path = str(tmp_path / "c.npz")
    c.save(path, note="test")
    d = CL.Closure.load(path)
    assert d.gain == 4.0 and d.n0 == c.n0 and d.in_scale == c.in_scale
    assert np.array_equal(d.mlp.get_params(), c.mlp.get_params())
    with np.load(path) as z:
        np.savez(str(tmp_path / "old.npz"), **{k: z[k] for k in z.files if k != "gain"})
    assert CL.Closure.load(str(tmp_path / "old.npz")).gain == 1.0
    x = np.random.default_rng(2).standard_normal((9000, 16)) * 1e-3
    y = c.predict(x)
    assert y.shape == (9000, 16)
    assert np.array_equal(y[:100], c.predict(x[:100]))
    d.mlp.close()
    m.close()
