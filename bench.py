#!/usr/bin/env python3
"""Benchmark: ARS(2,2,2) steps/s of the plain FP64 step at 256^2 x 256.

Workload: the production step (evolve_mean + balanced_sbp2_pc with 4 SBP
substeps, flux advection, 2/3 rule) at Nx=Ny=256, Nz=256, Ra=100, Pr=1,
box 24 onset wavelengths, polar trap gamma=2.2e-3 with a Rayleigh sponge,
dt=2.5e-5, from a seeded Hermitian thermal state (amplitude 1e-2) plus a
low-k barotropic flow of rms velocity 15 (synthetic, see nhqg_amd/init.py).

One JSON line on stdout.  ``value`` is steps/s of K graph replays with the
state resident in HBM, bracketed by device synchronisation.  ``roofline``
is measured live: one eager step with HIP events around every launch on
the context stream gives each kernel class's time and algorithmic work.
``cpu_baseline`` times the oracle (numpy/scipy FP64 restatement of the
reference algorithm) on this host for a bounded sample of the same
workload; the same sample gives the full-size one-step parity error.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ARS222 steps/s at 256²×256: FP64 plain vs mixed/closure at stated error"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def roofline(prof: dict, g, lay) -> dict:
    """Dominant kernel class of one step and its achieved rate."""
    name, (ms, n, work) = max(prof.items(), key=lambda kv: kv[1][0])
    out = {"kernel": name, "launches_per_step": n, "ms_per_step": ms}
    nq, nm = g.Nz + 1, lay.nm
    if name == "gemm_f64_mfma":
        ach = work / (ms * 1e-3) / 1e12
        out.update(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                   frac=ach / PEAK_FP64_TFLOPS,
                   work="4*M*K flops per complex mode (real operator x complex field), "
                        f"{nm} modes of the 2/3 band per launch")
    elif name.startswith("flux"):
        # algorithmic bytes of the whole FFT-product pipeline: read q, w, theta
        # nodal + write three tendencies, per RHS (band-compact modes)
        t = sum(v[0] for k, v in prof.items() if k.startswith("flux"))
        nrhs = prof["flux_col_inverse"][1]
        alg = 6.0 * 16.0 * nq * nm * nrhs
        ach = alg / (t * 1e-3) / 1e9
        out.update(kernel="flux pipeline (3 passes)", ms_per_step=t, bound="hbm", achieved=ach,
                   peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                   work="6 fields x (Nz+1) levels x band modes x 16 B per RHS")
    else:
        ach = work / (ms * 1e-3) / 1e9
        out.update(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s",
                   frac=ach / PEAK_HBM_GBS, work="bytes moved per launch")
    out["traffic"] = None
    return out


def cpu_baseline(cfg, st, steps: int, gpu_one_step):
    import numpy as np

    from oracle import nhqg_oracle as O

    cores = len(os.sched_getaffinity(0))
    O.set_workers(cores)
    o = O.build_operators(cfg)
    s = O.OState(st.q_hat.copy(), st.w_hat.copy(), st.th_hat.copy(), st.th_bar.copy())
    t0 = time.perf_counter()
    ref = s
    for _ in range(steps):
        ref = O.imex_step(ref, o)
    dt = time.perf_counter() - t0
    par = None
    if gpu_one_step is not None and steps == 1:
        par = {}
        for k, a, b in (("q_hat", gpu_one_step.q_hat, ref.q), ("w_hat", gpu_one_step.w_hat, ref.w),
                        ("th_hat", gpu_one_step.th_hat, ref.th),
                        ("th_bar", gpu_one_step.th_bar, ref.thb)):
            par[k] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    base = {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} step(s) of the 256^2x256 workload from the same initial state; "
                      f"oracle operator build excluded; scipy.fft workers={cores}, "
                      f"BLAS default threads; CPU: {model}"}
    return base, par


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--nz", type=int, default=256, help="override Nz (testing only)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        return 2

    import numpy as np

    from nhqg_amd import Stepper, make_grid
    from nhqg_amd.init import polar_config_256, polar_state_256

    cfg = polar_config_256(Nz=args.nz)
    g = make_grid(cfg)
    st0 = polar_state_256(g, seed=0)
    sp = Stepper(g)
    sp.set_state(st0)
    lay = sp.ctx.layout

    # one-step result for the full-size parity check
    sp.run(1)
    one = sp.state()
    sp.ctx.set_state(*st0)

    sp.run(args.warmup)
    sp.sync()
    t0 = time.perf_counter()
    sp.run_async(args.steps)
    sp.sync()
    wall = time.perf_counter() - t0
    dev_ms = sp.last_ms()
    diag = sp.diagnostics()
    prof = sp.ctx.profile_step()
    # plain 1 GiB device copy: practical HBM ceiling and the known-size
    # dispatch that calibrates the FETCH_SIZE / WRITE_SIZE counters
    copy_gbps = sp.ctx.stream_copy(1 << 30, reps=5)

    res = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain FP64 ARS222 balanced_sbp2_pc step, Nx=Ny=256, Nz=256, Ra=100, "
                        "Pr=1, L=24 onset wavelengths, polar trap gamma=2.2e-3 + sponge, "
                        "flux advection, 2/3 rule, dt=2.5e-5",
            "nx": g.Nx, "nz": g.Nz, "band_modes": lay.nm, "stored_modes": g.Nx * g.Nk,
            "parallelism": "1 GPU",
        },
        "roofline": roofline(prof, g, lay),
        "hbm_copy_gbps": copy_gbps,
        "kernels": {k: {"ms_per_step": v[0], "launches": v[1]} for k, v in prof.items()},
        "diagnostics_after": {k: diag[k] for k in ("KE_tot", "max_speed", "Nusselt_dealiased")},
    }
    if not args.no_cpu:
        base, par = cpu_baseline(cfg, st0, args.cpu_steps, one)
        res["cpu_baseline"] = base
        res["parity_one_step_rel_err"] = par
    else:
        res["cpu_baseline"] = None
    if not np.isfinite(res["value"]):
        return 1
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
