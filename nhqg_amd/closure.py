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
               seed: int = 0, n_samples: int | None = None) -> PairSet:
    """Fine run from state0: after `spinup` steps, `snapshots` snapshots
    `every` steps apart; at each, one fine and one coarse step from the
    filtered state give a pair set; optionally subsampled to n_samples."""
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
    mlp: "_native.MLP"
    n0: int
    in_scale: float
    out_scale: float
    loss: np.ndarray

    def predict(self, x: np.ndarray) -> np.ndarray:
        """Tendency for raw inputs (FP32 weights)."""
        return self.out_scale * self.mlp.forward(self.in_scale * np.asarray(x)).astype(np.float64)

    def attach(self, grid: Grid) -> None:
        grid.context(full=False).set_closure(self.mlp, self.n0, self.in_scale, self.out_scale)


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
