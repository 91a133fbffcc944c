"""Learned vertical-tail closure for coarse Nz (no counterpart in the
reference, which documents the undamped enstrophy tail of q at coarse Nz and
offers only a vertical cutoff for w and theta, nhqg/solver.py:177-191).

A per-column network (nhqg_amd/csrc/closure.hip) maps the top 16 Chebyshev
coefficients of q at a coarse Nz to a tendency on the same coefficients.  It
is trained on filtered coarse/fine pairs that the plain solver produces at a
finer Nz from the same state:

    s_c   = F(s_f)                       (Chebyshev truncation of the fine state)
    r     = (F(Step_fine(s_f)) - Step_coarse(s_c)) / dt   on q levels n0 .. n0+15

with input x = s_c.q[n0 .. n0+15] per horizontal mode, real and imaginary
parts as separate samples.  The network is odd (bias-free tanh layers,
NHQG_MLP_ODD): a horizontal shift by half a wavelength maps a column to its
negative, so under statistical homogeneity the best per-column predictor is
odd, and oddness keeps the ky = 0 columns Hermitian.  Training, the gradient
and inference run on the GPU with the project's own kernels; the plain fine
run is the fallback and the yardstick.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from nhqg_amd import _native
from nhqg_amd.grid import Grid
from nhqg_amd.solver import State, Stepper

N_TOP = 16
WIDTH = 128


def filter_state(s: State, gf: Grid, gc: Grid) -> State:
    """Fine state -> coarse Nz by truncating Chebyshev series: q and th_bar
    directly, w and theta through their Chebyshev coefficients and the coarse
    Galerkin projection (tau rows of the coarse basis)."""
    nqc = gc.Nz + 1
    q = np.array(s.q_hat[:nqc], dtype=np.complex128)
    wc = np.tensordot(gf.w_stencil, s.w_hat, axes=1)[:nqc]
    tc = np.tensordot(gf.dirichlet_stencil, s.th_hat, axes=1)[:nqc]
    w = np.tensordot(gc.w_pinv @ gc.proj_w, wc, axes=1)
    th = np.tensordot(gc.dirichlet_pinv @ gc.proj_dirichlet, tc, axes=1)
    thb = gc.proj_dirichlet @ np.asarray(s.th_bar[:nqc], dtype=np.float64)
    return State(q, w, th, thb)


def band_columns(f: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """(levels, Nx, Nk) complex -> (2 * band modes, levels) real samples."""
    idx = np.flatnonzero(mask.ravel() > 0)
    c = f.reshape(f.shape[0], -1)[:, idx]
    return np.concatenate([c.real.T, c.imag.T], axis=0)


@dataclass
class PairSet:
    x: np.ndarray      # (n, 16) coarse q top coefficients
    t: np.ndarray      # (n, 16) missing tendency
    dt: float
    n0: int


def make_pairs(gf: Grid, gc: Grid, state0: State, spinup: int, snapshots: int, every: int,
               seed: int = 0, n_samples: int | None = None, horizon: int = 1) -> PairSet:
    """Fine run from state0: after `spinup` steps, `snapshots` snapshots
    `every` steps apart; optionally subsampled to n_samples.

    horizon = 1: at each snapshot, one fine and one coarse step from the
    filtered state; x = the filtered state's top q levels, target = the
    one-step missing tendency.
    horizon = n > 1: the coarse model runs n steps from the filtered
    snapshot; x = the coarse state's top q levels after those n steps (the
    states the closure meets in a coarse run, tail pile-up included),
    target = (filtered fine state n steps on - coarse state) / (n dt), the
    mean tendency that would have kept the coarse run on the filtered fine
    trajectory.  The fine run then advances `every` steps (every >= n)."""
    if horizon > 1:
        return _relaxation_pairs(gf, gc, state0, spinup, snapshots, every, seed, n_samples,
                                 horizon)
    n0 = gc.Nz + 1 - N_TOP
    fine = Stepper(gf)
    fine.set_state(state0)
    if spinup:
        fine.run(spinup)
    coarse_ctx = gc.context(full=False)
    xs, ts = [], []
    for k in range(snapshots):
        if k:
            fine.run(every)
        sf = fine.state()
        sc = filter_state(sf, gf, gc)
        coarse_ctx.set_state(*sc)
        coarse_ctx.step(1)
        qc1 = coarse_ctx.get_state()[0]
        fine_ctx = gf.context(full=False)
        fine_ctx.set_state(*sf)
        fine_ctx.step(1)
        sf1 = State(*fine_ctx.get_state())
        fine.ctx.set_state(*sf)   # continue the fine run from sf
        qf1 = filter_state(sf1, gf, gc).q_hat
        r = (qf1[n0:] - qc1[n0:]) / float(gf.dt)
        xs.append(band_columns(sc.q_hat[n0:], gc.mask_23))
        ts.append(band_columns(r, gc.mask_23))
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    if n_samples is not None and n_samples < x.shape[0]:
        pick = np.random.default_rng(seed).choice(x.shape[0], n_samples, replace=False)
        x, t = x[pick], t[pick]
    return PairSet(x, t, float(gf.dt), n0)


def _relaxation_pairs(gf, gc, state0, spinup, snapshots, every, seed, n_samples, horizon):
    n0 = gc.Nz + 1 - N_TOP
    if every < horizon:
        raise ValueError("every must be >= horizon")
    fine = Stepper(gf)
    fine.set_state(state0)
    if spinup:
        fine.run(spinup)
    coarse = Stepper(gc)
    xs, ts = [], []
    for _ in range(snapshots):
        sc = filter_state(fine.state(), gf, gc)
        coarse.set_state(sc)
        coarse.run(horizon)
        qc = coarse.state().q_hat
        fine.run(horizon)
        qf = filter_state(fine.state(), gf, gc).q_hat
        if every > horizon:
            fine.run(every - horizon)
        r = (qf[n0:] - qc[n0:]) / (horizon * float(gc.dt))
        xs.append(band_columns(qc[n0:], gc.mask_23))
        ts.append(band_columns(r, gc.mask_23))
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    if n_samples is not None and n_samples < x.shape[0]:
        pick = np.random.default_rng(seed).choice(x.shape[0], n_samples, replace=False)
        x, t = x[pick], t[pick]
    return PairSet(x, t, float(gc.dt), n0)


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights and zero biases, flat in the kernel's order."""
    rng = np.random.default_rng(seed)

    def glorot(fo, fi):
        lim = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-lim, lim, size=(fo, fi))

    return np.concatenate([glorot(WIDTH, N_TOP).ravel(), np.zeros(WIDTH),
                           glorot(WIDTH, WIDTH).ravel(), np.zeros(WIDTH),
                           glorot(N_TOP, WIDTH).ravel(), np.zeros(N_TOP)]).astype(np.float32)


@dataclass
class Closure:
    """A trained network and its scales.  ``gain`` multiplies the tendency
    the network adds in the step: relaxation pairs fit the mean correction
    over their horizon, i.e. a relaxation time of the horizon, and the gain
    shortens it (chosen on a validation run by ``select_gain``; 1 applies the
    fitted correction as it is)."""
    mlp: "_native.MLP"
    n0: int
    in_scale: float
    out_scale: float
    loss: np.ndarray
    gain: float = 1.0

    def predict(self, x: np.ndarray) -> np.ndarray:
        """Tendency for raw inputs (FP32 weights), in batches the device
        network accepts."""
        x = self.in_scale * np.asarray(x)
        n = self.mlp.max_batch
        y = [self.mlp.forward(x[i:i + n]) for i in range(0, x.shape[0], n)]
        return self.out_scale * np.concatenate(y).astype(np.float64)

    def attach(self, grid: Grid) -> None:
        grid.attach_closure(self.mlp, self.n0, self.in_scale, self.out_scale * self.gain)

    def save(self, path: str, **meta) -> None:
        """FP32 parameters, scales and the top-level offset as .npz (loaded
        without pickle)."""
        np.savez(path, params=self.mlp.get_params(), n0=self.n0, in_scale=self.in_scale,
                 out_scale=self.out_scale, gain=self.gain, loss=np.asarray(self.loss, np.float32),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @staticmethod
    def load(path: str, nz: int | None = None) -> "Closure":
        """A saved closure on a new device MLP; nz places it on the top 16
        levels of another Nz (default: the saved offset)."""
        with np.load(path) as d:
            mlp = _native.MLP(N_TOP, WIDTH, N_TOP, max_batch=4096, odd=True)
            mlp.set_params(d["params"])
            n0 = int(d["n0"]) if nz is None else nz + 1 - N_TOP
            gain = float(d["gain"]) if "gain" in d.files else 1.0
            return Closure(mlp, n0, float(d["in_scale"]), float(d["out_scale"]),
                           np.array(d["loss"]), gain)


def train(pairs: PairSet, steps: int = 3000, batch: int = 512, lr: float = 1e-3,
          seed: int = 0) -> Closure:
    """Adam on the GPU over the normalised pair set (inputs by their rms,
    targets by theirs)."""
    in_scale = 1.0 / max(float(np.sqrt(np.mean(pairs.x ** 2))), 1e-300)
    out_scale = max(float(np.sqrt(np.mean(pairs.t ** 2))), 1e-300)
    mlp = _native.MLP(N_TOP, WIDTH, N_TOP, max_batch=max(batch, 4096), odd=True)
    mlp.set_params(init_params(seed))
    perm = np.random.default_rng(seed + 1).permutation(pairs.x.shape[0])
    mlp.set_data((pairs.x[perm] * in_scale).astype(np.float32),
                 (pairs.t[perm] / out_scale).astype(np.float32))
    loss = mlp.train(steps, batch, lr)
    return Closure(mlp, pairs.n0, in_scale, out_scale, loss)


def tendency_gain(c: Closure, pairs: PairSet) -> dict:
    """Error of the coarse one-step tendency on held-out pairs with and
    without the closure (relative to the missing tendency)."""
    pred = c.predict(pairs.x)
    base = float(np.sqrt(np.mean(pairs.t ** 2)))
    resid = float(np.sqrt(np.mean((pairs.t - pred) ** 2)))
    return {"rms_missing_tendency": base, "rms_error_with_closure": resid,
            "gain": 1.0 - resid / base}


def psi_ke_spectrum(q_hat: np.ndarray, g: Grid) -> np.ndarray:
    """psi-KE per Chebyshev degree: sum_k k^2 |psi_n,k|^2 w_k / Nx^4."""
    Nk = q_hat.shape[-1]
    wr = np.full(Nk, 2.0)
    wr[0] = 1.0
    if g.Nx % 2 == 0:
        wr[-1] = 1.0
    psi = -q_hat * g.inv_denom[None]
    return np.sum(g.ksq[None] * np.abs(psi) ** 2 * wr[None, None], axis=(1, 2)) / float(g.Nx) ** 4


def _filtered_diagnostics(gd: Grid, gf: Grid, sf: State) -> dict:
    """Diagnostics of the fine state restricted to the coarse degrees
    n <= Nz_coarse (evaluated on the coarse grid's device context)."""
    from nhqg_amd.diagnostics import unpack

    ctx = gd.context(full=False)
    ctx.set_state(*filter_state(sf, gf, gd))
    return unpack(ctx.diagnostics(), gd.Nz + 1)


def _spectrum_errors(e_c: np.ndarray, e_f: np.ndarray) -> dict:
    top = slice(e_c.size - N_TOP, e_c.size)
    return {"psi_ke_spectrum": float(np.abs(e_c - e_f).sum() / e_f.sum()),
            "psi_ke_top16": float(np.abs(e_c[top] - e_f[top]).sum() / e_f[top].sum())}


def judge(cfg_f, cfg_c, state0: State, n_steps: int, every: int, closure: Closure | None = None
          ) -> dict:
    """Coarse run (plain, and with the closure when given) against the plain
    fine run from the same state.

    Every `every` steps: psi-KE per Chebyshev degree n <= Nz_coarse
    (relative L1 error over all coarse degrees and over the top 16),
    max_speed of the fine state restricted to n <= Nz_coarse against the
    coarse run's, and Nu_d of each run.  Returns the series and the means
    over the checkpoints."""
    from nhqg_amd.grid import make_grid

    gf = make_grid(cfg_f, shells="band")
    runs = {"coarse": make_grid(cfg_c, shells="band")}
    if closure is not None:
        runs["coarse+closure"] = make_grid(cfg_c, shells="band")
    gd = make_grid(cfg_c, shells="band")
    fine = Stepper(gf)
    fine.set_state(state0)
    s0 = fine.state()
    steppers = {}
    for name, g in runs.items():
        sp = Stepper(g)
        sp.set_state(filter_state(s0, gf, g))
        if name == "coarse+closure":
            closure.attach(g)
        steppers[name] = sp
    nqc = cfg_c.Nz + 1
    series = {"step": []}
    done = 0
    while done < n_steps:
        n = min(every, n_steps - done)
        fine.run(n)
        for sp in steppers.values():
            sp.run(n)
        done += n
        sf = fine.state()
        e_f = psi_ke_spectrum(sf.q_hat, gf)[:nqc]
        fd = fine.diagnostics()
        rd = _filtered_diagnostics(gd, gf, sf)
        series["step"].append(done)
        series.setdefault("fine.Nu_d", []).append(fd["Nusselt_dealiased"])
        series.setdefault("fine.max_speed_resolved", []).append(rd["max_speed"])
        series.setdefault("fine.KE_tot", []).append(fd["KE_tot"])
        for name, sp in steppers.items():
            cs = sp.state()
            cd = sp.diagnostics()
            errs = _spectrum_errors(psi_ke_spectrum(cs.q_hat, runs[name]), e_f)
            series.setdefault(f"{name}.Nu_d", []).append(cd["Nusselt_dealiased"])
            series.setdefault(f"{name}.max_speed", []).append(cd["max_speed"])
            series.setdefault(f"{name}.KE_tot", []).append(cd["KE_tot"])
            for k, v in errs.items():
                series.setdefault(f"{name}.{k}_err", []).append(v)
    out = {"series": series, "summary": {}}
    nu_f = np.array(series["fine.Nu_d"])
    ms_f = np.array(series["fine.max_speed_resolved"])
    for name in steppers:
        nu = np.array(series[f"{name}.Nu_d"])
        ms = np.array(series[f"{name}.max_speed"])
        out["summary"][name] = {
            "psi_ke_spectrum_err_final": series[f"{name}.psi_ke_spectrum_err"][-1],
            "psi_ke_spectrum_err_mean": float(np.mean(series[f"{name}.psi_ke_spectrum_err"])),
            "psi_ke_top16_err_final": series[f"{name}.psi_ke_top16_err"][-1],
            "psi_ke_top16_err_mean": float(np.mean(series[f"{name}.psi_ke_top16_err"])),
            "Nu_d_rel_err_final": float(abs(nu[-1] - nu_f[-1]) / abs(nu_f[-1])),
            "Nu_d_rel_err_mean": float(np.mean(np.abs(nu - nu_f) / np.abs(nu_f))),
            "Nu_d_minus_1_rel_err_final": float(abs(nu[-1] - nu_f[-1]) / abs(nu_f[-1] - 1.0)),
            "max_speed_rel_err_final": float(abs(ms[-1] - ms_f[-1]) / ms_f[-1]),
            "max_speed_rel_err_mean": float(np.mean(np.abs(ms - ms_f) / ms_f)),
        }
    for g in (gf, gd, *runs.values()):
        g.close()
    return out


def select_gain(cfg_f, cfg_c, state0: State, c: Closure, gains, n_steps: int, every: int
                ) -> dict:
    """Validation run for the closure's gain: one plain fine run from state0
    and one coarse run per candidate gain against it.  Scores each gain by
    the larger of its mean and final top-16 psi-KE error, the two figures
    the main run reports, and returns the scores with the best gain."""
    from nhqg_amd.grid import make_grid

    gf = make_grid(cfg_f, shells="band")
    fine = Stepper(gf)
    fine.set_state(state0)
    sf0 = fine.state()
    runs = []
    for b in gains:
        g = make_grid(cfg_c, shells="band")
        sp = Stepper(g)
        sp.set_state(filter_state(sf0, gf, g))
        g.attach_closure(c.mlp, c.n0, c.in_scale, c.out_scale * float(b))
        runs.append((float(b), g, sp, []))
    done = 0
    while done < n_steps:
        n = min(every, n_steps - done)
        fine.run(n)
        for _, _, sp, _ in runs:
            sp.run(n)
        done += n
        e_f = psi_ke_spectrum(fine.state().q_hat, gf)[:cfg_c.Nz + 1]
        for _, g, sp, errs in runs:
            errs.append(_spectrum_errors(psi_ke_spectrum(sp.state().q_hat, g), e_f)["psi_ke_top16"])
    scores = {b: {"top16_mean": float(np.mean(e)), "top16_final": float(e[-1]),
                  "score": max(float(np.mean(e)), float(e[-1]))} for b, _, _, e in runs}
    for _, g, sp, _ in runs:
        sp.close()
        g.close()
    fine.close()
    gf.close()
    best = min(scores, key=lambda b: scores[b]["score"])
    return {"gain": best, "scores": scores}


def main(argv=None) -> int:
    """Train on pairs from one seeded fine run, choose the gain on a run from
    a second seed, judge on a run from a third; prints one JSON object."""
    import argparse
    import json
    import time

    from nhqg_amd.grid import make_grid
    from nhqg_amd.init import add_barotropic_flow, polar_config_256, thermal_noise_state

    ap = argparse.ArgumentParser(prog="python -m nhqg_amd.closure")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz-fine", type=int, default=256)
    ap.add_argument("--nz-coarse", type=int, default=64)
    ap.add_argument("--theta-amp", type=float, default=1e-2)
    ap.add_argument("--u-rms", type=float, default=15.0)
    ap.add_argument("--pairs", type=int, default=10000)
    ap.add_argument("--spinup", type=int, default=500)
    ap.add_argument("--snapshots", type=int, default=10)
    ap.add_argument("--snap-every", type=int, default=450)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed-train", type=int, default=0)
    ap.add_argument("--seed-judge", type=int, default=10)
    ap.add_argument("--seed-val", type=int, default=20, help="seed of the gain-selection run")
    ap.add_argument("--gains", type=str, default="1,2,3,4,5,6,8,10,12",
                    help="candidate gains (comma separated); a single value skips the selection")
    ap.add_argument("--save", type=str, default=None, help="write the trained closure (.npz)")
    ap.add_argument("--horizon", type=int, default=400,
                    help="1: one-step pairs; n > 1: n-step relaxation pairs from coarse runs")
    args = ap.parse_args(argv)

    cfg_f = polar_config_256(Nx=args.nx, Nz=args.nz_fine)
    cfg_c = polar_config_256(Nx=args.nx, Nz=args.nz_coarse)

    def seeded(g, seed):
        s = thermal_noise_state(g, seed=seed, amplitude=args.theta_amp)
        return add_barotropic_flow(s, g, seed=seed + 1, u_rms=args.u_rms)

    t0 = time.perf_counter()
    gf, gc = make_grid(cfg_f, shells="band"), make_grid(cfg_c, shells="band")
    pairs = make_pairs(gf, gc, seeded(gf, args.seed_train), spinup=args.spinup,
                       snapshots=args.snapshots, every=args.snap_every, seed=args.seed_train,
                       n_samples=args.pairs, horizon=args.horizon)
    t_pairs = time.perf_counter() - t0
    print(f"pairs {pairs.x.shape[0]} in {t_pairs:.1f} s", flush=True)
    gf.close()
    gc.close()
    t0 = time.perf_counter()
    c = train(pairs, steps=args.train_steps, batch=512, lr=args.lr, seed=args.seed_train)
    t_train = time.perf_counter() - t0
    print(f"trained in {t_train:.1f} s, loss {c.loss[:20].mean():.3e} -> {c.loss[-100:].mean():.3e}",
          flush=True)
    gains = [float(x) for x in args.gains.split(",")]
    t0 = time.perf_counter()
    if len(gains) > 1:
        gv = make_grid(cfg_f, shells="band")
        s_val = seeded(gv, args.seed_val)
        gv.close()
        sel = select_gain(cfg_f, cfg_c, s_val, c, gains, args.steps, args.every)
    else:
        sel = {"gain": gains[0], "scores": {}}
    c.gain = sel["gain"]
    t_select = time.perf_counter() - t0
    print(f"gain {c.gain} selected in {t_select:.1f} s: "
          + json.dumps({str(k): v for k, v in sel["scores"].items()}), flush=True)
    if args.save:
        c.save(args.save, nx=args.nx, nz_fine=args.nz_fine, nz_coarse=args.nz_coarse,
               horizon=args.horizon, pairs=pairs.x.shape[0])
    gj = make_grid(cfg_f, shells="band")
    s_judge = seeded(gj, args.seed_judge)
    gj.close()
    t0 = time.perf_counter()
    res = judge(cfg_f, cfg_c, s_judge, args.steps, args.every, closure=c)
    t_judge = time.perf_counter() - t0
    res["config"] = vars(args)
    res["timing_s"] = {"pairs": t_pairs, "train": t_train, "select_gain": t_select,
                       "judge": t_judge}
    res["gain_selection"] = {"gain": c.gain, "seed": args.seed_val,
                             "scores": {str(k): v for k, v in sel["scores"].items()}}
    res["train_loss"] = {"first20": float(c.loss[:20].mean()),
                         "last100": float(c.loss[-100:].mean())}
    print(json.dumps(res["summary"], indent=1), flush=True)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
