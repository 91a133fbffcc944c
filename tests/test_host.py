"""CPU checks of the product's host side: operators vs the oracle, the
device tables (through the numpy model of the device algorithm), config
guards, init generators, and the C ABI surface (no GPU calls)."""

import os
import re

import numpy as np
import pytest

from nhqg_amd import NHQGConfig, make_grid
from nhqg_amd import _native
from nhqg_amd.grid import device_operators
from nhqg_amd.init import cpu_config_64, thermal_noise_state, add_barotropic_flow, rms_velocity
from oracle import nhqg_oracle as O
from tests import cases
from tests.device_model import DeviceModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


@pytest.mark.parametrize("cfgf", [cases.fixture_cfg, cases.neumann_cfg, cases.trap_sponge_cfg])
def test_operators_match_oracle(cfgf):
    cfg = cfgf()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    pairs = [(g.V, o.V), (g.V_inv, o.V_inv), (g.G_Z, o.G_Z), (g.cc_weights, o.cc),
             (g.dirichlet_stencil, o.S_d), (g.dirichlet_pinv, o.d_pinv), (g.w_stencil, o.w_stencil),
             (g.w_pinv, o.w_pinv), (g.proj_dirichlet, o.proj_d), (g.proj_w, o.proj_w),
             (g.sbp_D1, o.sbp_D1), (g.sbp_L, o.sbp_L), (g.cgl_to_sbp, o.c2s), (g.sbp_to_cgl, o.s2c),
             (g.ksq, o.ksq), (g.inv_denom, o.inv_denom), (g.mask_23, o.mask),
             (g.inv_alpha_q, o.inv_alpha_q), (g.inv_alpha_th, o.inv_alpha_th)]
    for a, b in pairs:
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1.0)
    if o.eta_hat is not None:
        assert rel(g.eta_hat, o.eta_hat) < 1e-14
    if o.sponge_phys is not None:
        assert rel(g.sponge_phys, o.sponge_phys) < 1e-14
    lay = g.layout(False)
    ids = lay.tables["shell_ids"]
    if g.buoyancy_K is None:
        # Dirichlet top: eigenbasis solve Q diag(1 / (alpha - c lam)) Q^-1
        ops = device_operators(g, lay)
        assert ops["shell_inv"] is None
        Q, Qi, lam = ops["eig_Q"], ops["eig_Qinv"], ops["eig_lam"]
        ptr, modes = lay.tables["shell_ptr"], lay.tables["shell_modes"]
        for k in range(0, ids.size, max(1, ids.size // 7)):
            m = modes[ptr[k]]
            inv = Q @ np.diag(1.0 / (ops["mode_alpha"][m] - ops["mode_c"][m] * lam)) @ Qi
            assert rel(inv, o.imex_inv[ids[k]]) < 1e-13
    g.shell_solver = "matrices"
    g.shell_build = "host"
    inv = device_operators(g, lay)["shell_inv"].transpose(0, 2, 1)
    for k in range(0, ids.size, max(1, ids.size // 7)):
        assert rel(inv[k], o.imex_inv[ids[k]]) < 1e-13
    # the device recipe: A_s = alpha I - beta K - c B0 from three scalars
    g.shell_build = "device"
    ops = device_operators(g, lay)
    assert ops["shell_inv"] is None
    B0, K, coef = ops["shell_B0"], ops["shell_K"], ops["shell_coef"]
    I = np.eye(B0.shape[0])
    for k in range(0, ids.size, max(1, ids.size // 7)):
        A = coef[k, 0] * I - coef[k, 2] * B0
        if K is not None:
            A = coef[k, 0] * I - coef[k, 1] * K - coef[k, 2] * B0
        assert rel(np.linalg.inv(A), o.imex_inv[ids[k]]) < 1e-13


def test_layout_tables():
    g = make_grid(cases.fixture_cfg())
    lay = g.layout(False)
    t = lay.tables
    K = g.Nx // 3
    assert lay.nkx == 2 * K + 1 and lay.nky == K + 1
    # every kept mode is in the 2/3 band and the band is covered exactly
    flat = g.mask_23.ravel()
    assert np.all(flat[t["full_idx"]] == 1.0)
    assert t["full_idx"].size == int(flat.sum())
    # Hermitian partners on ky = 0: (-kx) mod Nx
    rows = t["rows"]
    for m in range(lay.nkx):
        p = t["partner"][m]
        assert rows[p] == (-rows[m]) % g.Nx
    # shells group equal |k|^2
    ks = t["ksq"]
    ptr, modes = t["shell_ptr"], t["shell_modes"]
    for s in range(ptr.size - 1):
        v = ks[modes[ptr[s]:ptr[s + 1]]]
        assert np.ptp(v) < 1e-8 * max(1.0, v.max())


@pytest.mark.parametrize("cfgf", [cases.fixture_cfg, cases.neumann_cfg, cases.trap_sponge_cfg])
def test_device_algorithm_model_matches_oracle(cfgf):
    """The fused/restructured device algorithm, run in numpy on the exact
    tables handed to nhqg_create, agrees with the oracle."""
    cfg = cfgf()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    dm = DeviceModel(g, g.layout(False))
    got = dm.run_full((s.q, s.w, s.th, s.thb), 2)
    ref = O.imex_step(O.imex_step(s, o), o)
    for a, b in zip(got, (ref.q, ref.w, ref.th, ref.thb)):
        assert rel(a, b) < 1e-12


def test_full_layout_model_handles_out_of_band_input():
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    rng = np.random.default_rng(5)
    q = s.q + 1e-4 * (1 - o.mask)[None] * rng.standard_normal(s.q.shape)
    st = O.OState(q, s.w, s.th, s.thb)
    dm = DeviceModel(g, g.layout(True))
    got = dm.run_full((st.q, st.w, st.th, st.thb), 1)
    ref = O.imex_step(st, o)
    for a, b in zip(got, (ref.q, ref.w, ref.th, ref.thb)):
        assert rel(a, b) < 1e-12


def test_config_guards():
    with pytest.raises(ValueError, match="Unsupported w_bc_top"):
        make_grid(cases.fixture_cfg(w_bc_top="open"))
    with pytest.raises(NotImplementedError):
        make_grid(cases.fixture_cfg(vertical_dealiasing="cheb_3o2"))
    with pytest.raises(NotImplementedError):
        make_grid(cases.fixture_cfg(float_dtype="float32"))
    with pytest.raises(ValueError, match="mutually exclusive"):
        make_grid(cases.fixture_cfg(gamma=1e-3, beta=0.1))
    with pytest.raises(ValueError, match="under-resolved"):
        make_grid(cases.fixture_cfg(Nx=16, gamma=1e-3, trap_sharpness=50.0))
    with pytest.raises(ValueError, match="sponge_r_start is required"):
        make_grid(cases.fixture_cfg(sponge_rate=1.0))


def test_init_generators():
    cfg = cpu_config_64()
    g = make_grid(cfg)
    s = thermal_noise_state(g, seed=0, amplitude=1e-3)
    assert np.abs(s.q_hat).max() == 0.0 and np.abs(s.w_hat).max() == 0.0
    th = np.tensordot(g.V, np.tensordot(g.dirichlet_stencil, s.th_hat, axes=1), axes=1)
    mid = np.argmin(np.abs(g.Z - 0.5))
    rms = np.fft.irfft2(th[mid], s=(g.Nx, g.Nx)).std()
    assert abs(rms - 1e-3) < 1e-6  # rms 1e-3 of the pattern at mid depth (sin = 1)
    assert np.abs(th[0]).max() < 1e-15 and np.abs(th[-1]).max() < 1e-15
    s2 = add_barotropic_flow(s, g, u_rms=15.0)
    assert abs(rms_velocity(s2.q_hat[0], g) - 15.0) < 1e-9
    assert np.abs(s2.q_hat * (1 - g.mask_23)[None]).max() == 0.0
    s3 = thermal_noise_state(g, seed=0, amplitude=1e-3)
    assert np.array_equal(s.th_hat, s3.th_hat)  # seeded


def _header():
    with open(os.path.join(ROOT, "include", "nhqg.h")) as f:
        return f.read()


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = set(re.findall(r"\b(nhqg_\w+)\s*\(", _header()))
    assert declared, "no declarations parsed"
    assert declared == set(_native.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_ctypes_structs_match_header():
    h = _header()
    ops = re.search(r"typedef struct \{(.*?)\} nhqg_operators;", h, re.S).group(1)
    names = []
    for decl in re.findall(r"const\s+(?:double|int32_t)\s*\*([^;]+);", ops):
        names += [n.strip().lstrip("*") for n in decl.split(",")]
    assert names == [n for n, _ in _native._OPS_FIELDS]
    par = re.search(r"typedef struct \{(.*?)\} nhqg_params;", h, re.S).group(1)
    pn = []
    for decl in re.findall(r"(?:int32_t|double)\s+([^;]+);", par):
        pn += [n.strip() for n in decl.split(",")]
    assert pn == [n for n, _ in _native.Params._fields_]


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one MI355X" in r.stderr


def test_rule32_grid_tables_match_the_padding():
    """3/2 rule: the padded-grid row tables place every kept row where the
    oracle's zero_pad (nhqg/spectral.py:18-34) puts it, the trap is not
    masked, sigma on the padded grid is the oracle's, and the device's
    physical trap / gradient fields are irfft2 of the zero-padded spectra."""
    from nhqg_amd.grid import device_operators, zero_pad

    for adv in ("flux", "jacobian"):
        cfg = cases.trap_sponge_cfg(horizontal_dealiasing="32_rule", nonlinear_advection=adv)
        g = make_grid(cfg)
        o = O.build_operators(cfg, with_shells=False)
        lay = g.layout(False)
        assert lay.full and g.Npad == 3 * cfg.Nx // 2
        t = lay.tables
        Nx, Np = cfg.Nx, g.Npad
        # a field that is 1 + its row index on every kept mode lands on the
        # padded rows exactly as the oracle's zero_pad places it
        f = (1.0 + np.arange(Nx))[:, None] * np.ones((1, Nx // 2 + 1))
        zp = O.zero_pad(f[None].astype(complex), Nx, Np)[0][:, :Nx // 2 + 1].real
        for x in range(Np):
            i = t["kidx_of_x_pad"][x]
            assert (zp[x] == 0).all() if i < 0 else (zp[x] == 1.0 + i).all()
        assert np.array_equal(t["kidx_of_x_pad"][t["rows_pad"]], np.arange(Nx))
        assert np.array_equal(g.eta_hat, o.eta_hat)
        assert np.abs(g.eta_hat * (1 - g.mask_23)).max() > 0     # unmasked
        assert np.array_equal(g.sponge_phys_pad, o.sponge_phys_pad)
        ops = device_operators(g, lay)
        assert ops["eta_phys_pad"].shape == (Np, Np) and ops.get("eta_phys") is None
        want = np.fft.irfft2(zero_pad(o.eta_hat, Nx, Np), s=(Np, Np))
        assert np.abs(ops["eta_phys_pad"] - want).max() <= 1e-15 * np.abs(want).max()
        assert np.array_equal(zero_pad(o.eta_hat, Nx, Np), O.zero_pad(o.eta_hat, Nx, Np))
        if adv == "jacobian":
            gx = np.fft.irfft2(O.zero_pad(1j * o.kx * o.eta_hat, Nx, Np), s=(Np, Np))
            assert np.abs(ops["eta_x_phys_pad"] - gx).max() <= 1e-15 * np.abs(gx).max()
        k = np.arange(Np)
        tw = ops["twiddle_pad"]
        assert np.abs(tw[:, 0] + 1j * tw[:, 1] - np.exp(-2j * np.pi * k / Np)).max() < 1e-15

