"""Host-side semantics of the device step, on the GPU.

* A state that needs the full layout (the driver's barotropic init carries
  roundoff outside the 2/3 band) is counted exactly: run_loop for N steps
  equals N imex_step calls, and its rows and checkpoints carry the right
  step numbers (scripts/run_polar.py:348-378).
* Steppers own their device state: the per-call imex_step and a second
  Stepper on the same grid leave a running Stepper untouched.
* A band-limited state whose ky = 0 column is not Hermitian takes the full
  layout and matches the oracle.
* run(..., use_imex=False) is the RK4 validator stepped n times.
"""

import os

import numpy as np
import pytest

from nhqg_amd import State, Stepper, imex_step, make_grid, rk4_step, run
from nhqg_amd.init import barotropic_vorticity_state
from nhqg_amd.io import load_checkpoint, run_loop
from nhqg_amd.solver import is_band_limited, is_hermitian, needs_full_layout
from oracle import nhqg_oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def test_run_loop_from_the_barotropic_init_counts_steps(tmp_path):
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    s0 = barotropic_vorticity_state(g, seed=0, k_peak=1.3, k_width=0.4, amplitude=0.5)
    assert not is_band_limited(s0, g)          # roundoff outside the band
    print("run_loop barotropic init", flush=True)
    status, final, _ = run_loop(g, s0, 12, 4, tmp_path, checkpoint_steps=8)
    assert status == 0
    rows = open(os.path.join(tmp_path, "diagnostics.csv")).read().strip().splitlines()
    assert [int(r.split(",")[0]) for r in rows[1:]] == [4, 8, 12]
    x = s0
    for _ in range(12):
        x = imex_step(x, g)
    for a, b in zip(final, x):
        assert np.array_equal(a, b)
    back, step, _ = load_checkpoint(os.path.join(tmp_path, "checkpoint_00000008.npz"))
    y = s0
    for _ in range(8):
        y = imex_step(y, g)
    assert step == 8
    for a, b in zip(back, y):
        assert np.array_equal(a, b)


def test_stepper_state_is_private():
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    s = cases.active_state(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    other = State(2 * s.q, s.w, 0.5 * s.th, s.thb)
    print("stepper isolation", flush=True)
    a = Stepper(g)
    a.set_state(st)
    a.run(3)
    imex_step(other, g)                 # per-call step on the grid's scratch context
    b = Stepper(g)                      # a second resident run on the same grid
    b.set_state(other)
    b.run(2)
    a.run(3)
    c = Stepper(g)
    c.set_state(st)
    c.run(6)
    assert a.steps == 6
    for x, y in zip(a.state(), c.state()):
        assert np.array_equal(x, y)


def test_band_limited_anti_hermitian_ghost_takes_the_full_layout():
    cfg = cases.fixture_cfg()
    g = make_grid(cfg)
    o = O.build_operators(cfg)
    s = cases.active_state(cfg)
    q = s.q.copy()
    q[:, 2, 0] += 1e-3j                 # inside the band, ky = 0, no conjugate partner
    q[:, cfg.Nx - 2, 0] += 1e-3j
    st = State(q, s.w, s.th, s.thb)
    assert is_band_limited(st, g) and not is_hermitian(st, g) and needs_full_layout(st, g)
    print("anti-hermitian ghost", flush=True)
    got = imex_step(st, g)
    want = O.imex_step(O.OState(q, s.w, s.th, s.thb), o)
    for a, b in zip(got, (want.q, want.w, want.th, want.thb)):
        assert rel(a, b) < 1e-12


def test_run_without_imex_is_rk4():
    cfg = cases.fixture_cfg(thermal_closure="fixed_conduction",
                            mean_exchange_discretization="legacy")
    g = make_grid(cfg)
    s = cases.active_state(cfg)
    st = State(s.q, s.w, s.th, s.thb)
    print("run use_imex=False", flush=True)
    seen = []
    final, snaps = run(g, st, 5, 2, False, lambda x, k, t: seen.append((k, t)))
    assert [k for k, _ in seen] == [2, 4] and len(snaps) == 2
    x = st
    for _ in range(5):
        x = rk4_step(x, g)
    for a, b in zip(final, x):
        assert np.array_equal(a, b)
