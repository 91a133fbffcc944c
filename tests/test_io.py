"""Checkpoint format (nhqg/io.py:107-158) and driver-loop semantics
(scripts/run_polar.py:332-378)."""

import os

import numpy as np
import pytest

from nhqg_amd import State, make_grid
from nhqg_amd.io import load_checkpoint, run_loop, save_checkpoint
from tests import cases

REF_KEYS = {"q_hat_real", "q_hat_imag", "w_hat_real", "w_hat_imag", "th_hat_real",
            "th_hat_imag", "th_bar", "step", "t"}


def test_checkpoint_keys_and_roundtrip(tmp_path):
    cfg = cases.fixture_cfg()
    s = cases.active_state(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    f = save_checkpoint(st, 120, cfg, tmp_path)
    assert os.path.basename(f) == "checkpoint_00000120.npz"
    with np.load(f) as d:
        assert set(d.files) == REF_KEYS
        assert float(d["t"]) == 120 * cfg.dt
    got, step, t = load_checkpoint(f)
    assert step == 120 and t == 120 * cfg.dt
    for a, b in zip(got, st):   # a sanitized state is a fixed point of the projection
        assert np.array_equal(a, b)


def test_load_projects_hermitian_ghost(tmp_path):
    cfg = cases.fixture_cfg()
    s = cases.active_state(cfg)
    q = s.q.copy()
    q[:, 1:4, 0] += 1e-3j          # anti-Hermitian ghost on ky = 0
    q[:, 2, -1] += 5e-4            # and on the Nyquist column
    f = save_checkpoint(State(q, s.w, s.th, s.thb), 0, cfg, tmp_path)
    got, _, _ = load_checkpoint(f)
    Nx = q.shape[1]
    neg = (-np.arange(Nx)) % Nx
    for c in (0, -1):
        col = got.q_hat[:, :, c]
        assert np.array_equal(col, np.conj(col[:, neg]))
        # equals the reference's projection 0.5 (f + conj f[-kx])
        ref = 0.5 * (q[:, :, c] + np.conj(q[:, neg, c]))
        assert np.abs(col - ref).max() == 0.0
    assert np.array_equal(got.q_hat[:, :, 1:-1], q[:, :, 1:-1])


def test_load_without_th_bar(tmp_path):
    cfg = cases.fixture_cfg()
    s = cases.active_state(cfg)
    f = os.path.join(tmp_path, "old.npz")
    np.savez(f, q_hat_real=s.q.real, q_hat_imag=s.q.imag, w_hat_real=s.w.real,
             w_hat_imag=s.w.imag, th_hat_real=s.th.real, th_hat_imag=s.th.imag, step=3, t=0.1)
    got, step, t = load_checkpoint(f)
    assert np.array_equal(got.th_bar, np.zeros(s.q.shape[0])) and step == 3


@pytest.mark.gpu
def test_run_loop_cadence_and_checkpoints(tmp_path):
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    s = cases.active_state(cfg)
    status, final, diag = run_loop(g, State(s.q, s.w, s.th, s.thb), 30, 10, tmp_path,
                                   checkpoint_steps=20)
    assert status == 0
    rows = open(os.path.join(tmp_path, "diagnostics.csv")).read().strip().splitlines()
    assert len(rows) == 4 and rows[0].startswith("step,t,")
    assert [int(r.split(",")[0]) for r in rows[1:]] == [10, 20, 30]
    assert os.path.exists(os.path.join(tmp_path, "checkpoint_00000020.npz"))
    back, step, _ = load_checkpoint(os.path.join(tmp_path, "checkpoint_00000030.npz"))
    assert step == 30
    for a, b in zip(back, final):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_run_loop_stops_on_nonfinite(tmp_path):
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    s = cases.active_state(cfg)
    th = s.th.copy()
    th[3, 2, 1] = np.nan
    status, final, diag = run_loop(g, State(s.q, s.w, th, s.thb), 30, 10, tmp_path)
    assert status == 2
    assert os.path.exists(os.path.join(tmp_path, "checkpoint_00000010.npz"))
    rows = open(os.path.join(tmp_path, "diagnostics.csv")).read().strip().splitlines()
    assert len(rows) == 2


def test_run_config_roundtrip_and_no_clobber(tmp_path):
    from nhqg_amd.io import read_run_config, write_run_config
    from nhqg_amd.init import polar_config_256

    cfg = polar_config_256(Nz=64)
    p1 = write_run_config(tmp_path, cfg, {"steps": 10})
    p2 = write_run_config(tmp_path, cfg)
    assert p1.name == "run_config.json" and p2.name == "run_config.1.json"
    assert read_run_config(p1) == cfg
    assert read_run_config(p2) == cfg
