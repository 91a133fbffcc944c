"""Oracle trajectory of the 256^2 x 256 polar-trap workload (1000 steps).

Writes tests/golden/c2_trajectory.json: the judged diagnostics (KE_tot,
KE_bt, max_speed, Nusselt_dealiased, enstrophy) and per-field L2 norms and
a few probe coefficients at steps 0, 1, 10, 100, 200, ..., 1000, computed
by the numpy/scipy FP64 oracle from the seeded initial state
(nhqg_amd.init.polar_state_256, seed 0).  The GPU test replays the same
1000 steps and compares.  Run time: hours on a few CPU cores, e.g.

    OMP_NUM_THREADS=6 python tests/golden/make_c2_trajectory.py --workers 6
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

from nhqg_amd import make_grid  # noqa: E402
from nhqg_amd.init import polar_config_256, polar_state_256  # noqa: E402
from oracle import nhqg_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_trajectory.json")
CHECKPOINTS = [0, 1, 10] + list(range(100, 1001, 100))
# probe coefficients (level, kx row, ky column) of each field
PROBES = [(0, 1, 1), (5, 3, 2), (128, 10, 7), (200, 250, 40)]


def record(s, o, step):
    d = O.diagnostics(s, o)
    rec = {"step": step}
    for k in ("KE_tot", "KE_bt", "max_speed", "Nusselt_dealiased", "enstrophy"):
        rec[k] = float(d[k])
    for name, f in (("q_hat", s.q), ("w_hat", s.w), ("th_hat", s.th)):
        rec[name + "_l2"] = float(np.sqrt(np.sum(np.abs(f) ** 2)))
        rec[name + "_probe"] = [[float(f[p].real), float(f[p].imag)]
                                for p in PROBES if p[0] < f.shape[0]]
    rec["th_bar_l2"] = float(np.sqrt(np.sum(s.thb ** 2)))
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--steps", type=int, default=1000)
    args = ap.parse_args()
    O.set_workers(args.workers)
    cfg = polar_config_256()
    g = make_grid(cfg, shells="none")
    st = polar_state_256(g, seed=0)
    o = O.build_operators(cfg)
    s = O.OState(st.q_hat.copy(), st.w_hat.copy(), st.th_hat.copy(), st.th_bar.copy())
    recs = [record(s, o, 0)]
    t0 = time.time()
    for n in range(1, args.steps + 1):
        s = O.imex_step(s, o)
        if n in CHECKPOINTS:
            recs.append(record(s, o, n))
            out = {"workload": "polar trap 256^2 x 256, seed 0 (nhqg_amd.init.polar_state_256)",
                   "oracle": "oracle/nhqg_oracle.py imex_step, FP64",
                   "probes": PROBES, "records": recs}
            with open(OUT + ".tmp", "w") as f:
                json.dump(out, f, indent=1)
            os.replace(OUT + ".tmp", OUT)
            print(f"step {n} done, {time.time() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
