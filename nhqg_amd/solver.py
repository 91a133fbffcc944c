"""Time step: ``State`` and ``imex_step(state, grid)`` on the MI355X.

Mirrors the reference call signature and state layout
(nhqg/solver.py:33-38, 1228-1242): ``State(q_hat, w_hat, th_hat, th_bar)``
with q_hat (Nz+1, Nx, Nx//2+1) complex128, w_hat / th_hat (Nz-1, Nx,
Nx//2+1) complex128 and th_bar (Nz+1,) float64.  ``imex_step`` has copy
semantics (inputs are never mutated, like the reference); ``Stepper`` keeps
the state resident on the device and replays the captured step graph, the
counterpart of the reference's jit + lax.scan chunk
(scripts/run_polar.py:338-343, nhqg/solver.py:1385-1416).

Layouts: a band-limited state (zero outside the 2/3 mask) runs in the
band-compact device layout, which is exact for such states because the
step keeps out-of-band modes identically zero.  Any other state runs its
first step in the full layout (out-of-band modes take part in the products
and horizontal means exactly as in the reference, and the final
sanitation zeroes them), after which it is band-limited.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from nhqg_amd.grid import Grid
from nhqg_amd import _native


class State(NamedTuple):
    q_hat: np.ndarray
    w_hat: np.ndarray
    th_hat: np.ndarray
    th_bar: np.ndarray


def is_band_limited(state: State, grid: Grid) -> bool:
    out = grid.mask_23 == 0.0
    for f in (state.q_hat, state.w_hat, state.th_hat):
        if np.any(np.asarray(f)[:, out] != 0):
            return False
    return True


def _as_state(s) -> State:
    return State(np.asarray(s[0], dtype=np.complex128), np.asarray(s[1], dtype=np.complex128),
                 np.asarray(s[2], dtype=np.complex128), np.asarray(s[3], dtype=np.float64))


def imex_step(state: State, grid: Grid) -> State:
    """One ARS(2,2,2) step (balanced_sbp2_pc under evolve_mean) followed by
    sanitation, computed on the GPU (nhqg/solver.py:1228-1242)."""
    st = _as_state(state)
    ctx = grid.context(full=not is_band_limited(st, grid))
    ctx.set_state(*st)
    ctx.step(1)
    return State(*ctx.get_state())


def rk4_step(state: State, grid: Grid) -> State:
    """One explicit RK4 step of the full RHS (explicit + implicit tendency),
    the reference's validator (nhqg/solver.py:1276-1336), on the GPU: stage
    combinations, the RK4-only post-step dissipation exp(-rate dt), th_bar's
    Dirichlet projection and sanitation.  thermal_closure='fixed_conduction'
    configs only: the evolve_mean RHS with the mean-temperature coupling is
    not built (NativeError, NHQG_ERR_UNSUPPORTED)."""
    st = _as_state(state)
    ctx = grid.context(full=not is_band_limited(st, grid))
    ctx.set_state(*st)
    ctx.apply(_native.OP_RK4_STEP)
    return State(*ctx.get_state())


def sanitize_state(state: State, grid: Grid) -> State:
    """Hermitian projection of the self-conjugate rfft2 columns and 2/3 band
    limit (nhqg/solver.py:107-146), on the host."""
    Nx = grid.Nx
    neg = (-np.arange(Nx)) % Nx
    out = []
    for f in (state.q_hat, state.w_hat, state.th_hat):
        f = np.array(f, dtype=np.complex128, copy=True)
        cols = [0, f.shape[-1] - 1] if Nx % 2 == 0 else [0]
        for c in cols:
            col = f[:, :, c].copy()
            f[:, :, c] = 0.5 * (col + np.conj(col[:, neg]))
        out.append(f * grid.mask_23[None])
    return State(out[0], out[1], out[2], np.array(state.th_bar, dtype=np.float64))


class Stepper:
    """Device-resident stepping: ``set_state`` once, ``run(n)`` many times."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self.ctx = None
        self.steps = 0

    def set_state(self, state: State) -> None:
        st = _as_state(state)
        if is_band_limited(st, self.grid):
            self.ctx = self.grid.context(full=False)
            self.ctx.set_state(*st)
        else:
            full = self.grid.context(full=True)
            full.set_state(*st)
            full.step(1)
            self.steps += 1
            self.ctx = self.grid.context(full=False)
            self.ctx.set_state(*full.get_state())

    def run(self, n: int) -> None:
        if self.ctx is None:
            raise RuntimeError("Stepper.run before set_state")
        self.ctx.step(int(n))
        self.steps += int(n)

    def run_async(self, n: int) -> None:
        self.ctx.step_async(int(n))
        self.steps += int(n)

    def sync(self) -> None:
        self.ctx.sync()

    def last_ms(self) -> float:
        return self.ctx.last_step_ms()

    def state(self) -> State:
        return State(*self.ctx.get_state())

    def diagnostics(self) -> dict:
        from nhqg_amd.diagnostics import unpack

        return unpack(self.ctx.diagnostics(), self.grid.Nz + 1)


def run(grid: Grid, state: State, n_steps: int, save_interval: int, callback=None):
    """Counterpart of the reference's ``run`` (nhqg/solver.py:1385-1416):
    sanitize the incoming state, step n_steps on the device, call
    ``callback(state, step, t)`` every save_interval steps."""
    st = sanitize_state(_as_state(state), grid)
    sp = Stepper(grid)
    sp.set_state(st)
    snaps = []
    done = 0
    for _ in range(n_steps // save_interval):
        sp.run(save_interval)
        done += save_interval
        s = sp.state()
        t = done * float(grid.dt)
        if callback is not None:
            callback(s, done, t)
        snaps.append((t, s))
    rem = n_steps % save_interval
    if rem:
        sp.run(rem)
    return sp.state(), snaps
