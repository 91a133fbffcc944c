"""Oracle trajectories for the long-horizon GPU parity tests.

For each case, runs the numpy/scipy FP64 oracle (oracle/nhqg_oracle.py)
for up to 1000 steps from a seeded initial state and writes
tests/golden/<case>_trajectory.json with, at steps 0, 1, 10, 100, 200, ...,
1000: the judged diagnostics (KE_tot, KE_bt, max_speed, Nusselt_dealiased,
enstrophy), per-level L2 norms of q, w, theta, the L2 norm of th_bar and a
few probe coefficients.

Cases:
  c1       64^2 x 32, 10 onset wavelengths, no trap, dt 5e-5; theta' seeded
           noise of amplitude 1e-3, w = q' = 0 (nhqg_amd.init.cpu_config_64,
           thermal_noise_state seed 0)
  c1_flow  c1 plus a low-k barotropic flow of rms velocity 15 (seed 1) and
           theta' amplitude 1e-2: the nonlinear, evolving-mean regime
  c2       256^2 x 256 polar trap + sponge, dt 2.5e-5
           (nhqg_amd.init.polar_state_256, seed 0); hours of CPU time

    python tests/golden/make_trajectory.py --case c1 --workers 4

--probes1 writes tests/golden/<case>_step1_probes.npz instead: after one
oracle step, 4096 seeded random in-band coefficients of each of q, w,
theta, each field's largest magnitude, and the whole th_bar (full-field
per-step parity at the case's size without storing the fields).

    python tests/golden/make_trajectory.py --case c2 --probes1 --workers 8
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import nhqg_oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
CHECKPOINTS = [0, 1, 10] + list(range(100, 1001, 100))
PROBES = [(0, 1, 1), (5, 3, 2), (12, 10, 7), (20, 40, 9)]


def case_state(case: str):
    """(cfg, initial State) of a case, built with the product's host-side
    generators (numpy only; no device code)."""
    from nhqg_amd import make_grid
    from nhqg_amd.init import (add_barotropic_flow, cpu_config_64, polar_config_256,
                               polar_state_256, thermal_noise_state)
    if case == "c1":
        cfg = cpu_config_64()
        g = make_grid(cfg, shells="band")
        return cfg, thermal_noise_state(g, seed=0, amplitude=1e-3)
    if case == "c1_flow":
        cfg = cpu_config_64()
        g = make_grid(cfg, shells="band")
        return cfg, add_barotropic_flow(thermal_noise_state(g, seed=0, amplitude=1e-2), g,
                                        seed=1, u_rms=15.0)
    if case == "c2":
        cfg = polar_config_256()
        g = make_grid(cfg, shells="band")
        return cfg, polar_state_256(g, seed=0)
    raise ValueError(case)


def summarize(q, w, th, thb, diag: dict, step: int) -> dict:
    rec = {"step": step}
    for k in ("KE_tot", "KE_bt", "max_speed", "Nusselt_dealiased", "enstrophy"):
        rec[k] = float(diag[k])
    for name, f in (("q_hat", q), ("w_hat", w), ("th_hat", th)):
        rec[name + "_level_l2"] = [float(v) for v in np.sqrt(np.sum(np.abs(f) ** 2, axis=(1, 2)))]
        rec[name + "_probe"] = [[float(f[p].real), float(f[p].imag)]
                                for p in PROBES if p[0] < f.shape[0]]
    rec["th_bar_l2"] = float(np.sqrt(np.sum(np.asarray(thb) ** 2)))
    return rec


N_PROBES = 4096


def step1_probes(case: str, s, o) -> int:
    t0 = time.time()
    s1 = O.imex_step(s, o)
    rng = np.random.default_rng(2024)
    kept = np.argwhere(o.mask > 0)
    out = {"th_bar": s1.thb}
    for name, f in (("q_hat", s1.q), ("w_hat", s1.w), ("th_hat", s1.th)):
        lev = rng.integers(0, f.shape[0], N_PROBES)
        pick = kept[rng.integers(0, kept.shape[0], N_PROBES)]
        idx = np.stack([lev, pick[:, 0], pick[:, 1]], 1).astype(np.int16)
        out[name + "_idx"] = idx
        out[name + "_val"] = f[idx[:, 0], idx[:, 1], idx[:, 2]]
        out[name + "_max"] = np.abs(f).max()
    path = os.path.join(HERE, f"{case}_step1_probes.npz")
    np.savez_compressed(path, **out)
    print(f"{case}: one oracle step and probes in {time.time() - t0:.0f} s -> {path}")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", required=True, choices=["c1", "c1_flow", "c2"])
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--probes1", action="store_true")
    args = ap.parse_args()
    O.set_workers(args.workers)
    cfg, st = case_state(args.case)
    o = O.build_operators(cfg)
    s = O.OState(st.q_hat.copy(), st.w_hat.copy(), st.th_hat.copy(), st.th_bar.copy())
    if args.probes1:
        return step1_probes(args.case, s, o)
    out_path = os.path.join(HERE, f"{args.case}_trajectory.json")
    recs = [summarize(s.q, s.w, s.th, s.thb, O.diagnostics(s, o), 0)]
    t0 = time.time()
    for n in range(1, args.steps + 1):
        s = O.imex_step(s, o)
        if n in CHECKPOINTS:
            recs.append(summarize(s.q, s.w, s.th, s.thb, O.diagnostics(s, o), n))
            out = {"case": args.case, "oracle": "oracle/nhqg_oracle.py imex_step, FP64",
                   "probes": PROBES, "records": recs}
            with open(out_path + ".tmp", "w") as f:
                json.dump(out, f)
            os.replace(out_path + ".tmp", out_path)
            print(f"{args.case}: step {n}, {time.time() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
