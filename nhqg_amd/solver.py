"""Time step: ``State`` and ``imex_step(state, grid)`` on the MI355X.

Mirrors the reference call signature and state layout
(nhqg/solver.py:33-38, 1228-1242): ``State(q_hat, w_hat, th_hat, th_bar)``
with q_hat (Nz+1, Nx, Nx//2+1) complex128, w_hat / th_hat (Nz-1, Nx,
Nx//2+1) complex128 and th_bar (Nz+1,) float64.  ``imex_step`` has copy
semantics (inputs are never mutated, like the reference); ``Stepper`` keeps
the state resident on the device and replays the captured step graph, the
counterpart of the reference's jit + lax.scan chunk
(scripts/run_polar.py:338-343, nhqg/solver.py:1385-1416).

Layouts: a band-limited state (zero outside the 2/3 mask) with Hermitian
self-conjugate columns runs in the band-compact device layout, which is
exact for such states because the step keeps out-of-band modes identically
zero.  Any other state runs its first step in the full layout (out-of-band
and anti-Hermitian content take part in the products and horizontal means
exactly as in the reference, and the final sanitation removes them), after
which it is band-limited and Hermitian.
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
    """True when q, w, theta are zero outside the 2/3 band (always, under
    the 3/2 rule, which keeps every mode)."""
    if grid.rule32:
        return True
    out = grid.mask_23 == 0.0
    for f in (state.q_hat, state.w_hat, state.th_hat):
        if np.any(np.asarray(f)[:, out] != 0):
            return False
    return True


def is_hermitian(state: State, grid: Grid) -> bool:
    """True when the self-conjugate rfft2 columns (ky = 0 and, for even Nx,
    ky = Nx/2) of q, w, theta equal their own Hermitian projection, as every
    sanitized state does exactly."""
    neg = (-np.arange(grid.Nx)) % grid.Nx
    cols = [0, grid.Nk - 1] if grid.Nx % 2 == 0 else [0]
    for f in (state.q_hat, state.w_hat, state.th_hat):
        f = np.asarray(f)
        for c in cols:
            if np.any(f[:, :, c] != np.conj(f[:, neg, c])):
                return False
    return True


def needs_full_layout(state: State, grid: Grid) -> bool:
    """The band-compact device layout assumes a band-limited state with
    Hermitian self-conjugate columns; any other state takes its next step in
    the full layout, where out-of-band and anti-Hermitian content enter the
    products and means exactly as in the reference and the step's final
    sanitation removes them.  Under the 3/2 rule every context holds the
    full layout."""
    if grid.rule32:
        return False
    return not (is_band_limited(state, grid) and is_hermitian(state, grid))


def _as_state(s) -> State:
    return State(np.asarray(s[0], dtype=np.complex128), np.asarray(s[1], dtype=np.complex128),
                 np.asarray(s[2], dtype=np.complex128), np.asarray(s[3], dtype=np.float64))


def imex_step(state: State, grid: Grid) -> State:
    """One ARS(2,2,2) step (balanced_sbp2_pc under evolve_mean) followed by
    sanitation, computed on the GPU (nhqg/solver.py:1228-1242).  Runs on the
    grid's scratch context; a Stepper's resident state is not touched."""
    st = _as_state(state)
    ctx = grid.context(full=needs_full_layout(st, grid))
    ctx.set_state(*st)
    ctx.step(1)
    return State(*ctx.get_state())


def rk4_step(state: State, grid: Grid) -> State:
    """One explicit RK4 step of the full RHS (explicit + implicit tendency),
    the reference's validator (nhqg/solver.py:1276-1336), on the GPU: stage
    combinations, the RK4-only post-step dissipation exp(-rate dt), th_bar's
    Dirichlet projection and sanitation."""
    st = _as_state(state)
    ctx = grid.context(full=needs_full_layout(st, grid))
    ctx.set_state(*st)
    ctx.apply(_native.OP_RK4_STEP)
    return State(*ctx.get_state())


def sanitize_state(state: State, grid: Grid) -> State:
    """Hermitian projection of the self-conjugate rfft2 columns and, under
    the 2/3 rule, the band limit (nhqg/solver.py:107-146), on the host."""
    Nx = grid.Nx
    neg = (-np.arange(Nx)) % Nx
    out = []
    for f in (state.q_hat, state.w_hat, state.th_hat):
        f = np.array(f, dtype=np.complex128, copy=True)
        cols = [0, f.shape[-1] - 1] if Nx % 2 == 0 else [0]
        for c in cols:
            col = f[:, :, c].copy()
            f[:, :, c] = 0.5 * (col + np.conj(col[:, neg]))
        out.append(f if grid.rule32 else f * grid.mask_23[None])
    return State(out[0], out[1], out[2], np.array(state.th_bar, dtype=np.float64))


class Stepper:
    """Device-resident stepping: ``set_state`` once, ``run(n)`` many times.

    Each Stepper owns its device context(s), so several Steppers on one grid
    and the per-call imex_step / rk4_step never overwrite each other's state.
    A state that needs the full layout (not band-limited, or not Hermitian
    on the self-conjugate columns) is held in a full-layout context and its
    first step runs there; ``run(n)`` counts that step as the first of its n,
    so ``steps`` and every caller's step count stay exact."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self.ctx = None
        self._full = None
        self._pending_full = False
        self.steps = 0

    def _band(self):
        if self.ctx is None:
            self.ctx = self.grid.new_context(full=False)
        return self.ctx

    def set_state(self, state: State) -> None:
        st = _as_state(state)
        if needs_full_layout(st, self.grid):
            if self._full is None:
                self._full = self.grid.new_context(full=True)
            self._full.set_state(*st)
            self._pending_full = True
            self._band()
        else:
            self._band().set_state(*st)
            self._pending_full = False

    def _leave_full(self, n: int) -> int:
        """Run the pending full-layout step; returns the steps left of n."""
        if not self._pending_full or n <= 0:
            return n
        self._full.step(1)
        self.ctx.set_state(*self._full.get_state())
        self._pending_full = False
        return n - 1

    def run(self, n: int) -> None:
        if self.ctx is None:
            raise RuntimeError("Stepper.run before set_state")
        n = int(n)
        rest = self._leave_full(n)
        if rest:
            self.ctx.step(rest)
        self.steps += max(n, 0)

    def run_async(self, n: int) -> None:
        if self.ctx is None:
            raise RuntimeError("Stepper.run_async before set_state")
        n = int(n)
        rest = self._leave_full(n)
        if rest:
            self.ctx.step_async(rest)
        self.steps += max(n, 0)

    def sync(self) -> None:
        self.ctx.sync()

    def last_ms(self) -> float:
        return self.ctx.last_step_ms()

    def _current(self):
        return self._full if self._pending_full else self.ctx

    def state(self) -> State:
        return State(*self._current().get_state())

    def diagnostics(self) -> dict:
        from nhqg_amd.diagnostics import device_diagnostics

        return device_diagnostics(self._current(), self.grid)

    def upsample_from(self, other: "Stepper") -> None:
        """This Stepper's state <- other's resident state spectrally upsampled
        to this grid on the device (same Nz, Nx at least other's; the method
        of scripts/upsample_checkpoint_horiz.py:27-59), then sanitized."""
        self._band().upsample_from(other._current())
        self._pending_full = False

    def init_barotropic(self, seed: int, k_peak: float, k_width: float, amplitude: float,
                        confine_radius: float | None = None) -> None:
        """The driver's annular barotropic vorticity initial state
        (scripts/run_polar.py:115-158) built on the device: the white noise
        is numpy's default_rng(seed).standard_normal((Nx, Nx)), as the driver
        draws it; filtering, taper, normalisation and placement run in
        nhqg_init_barotropic."""
        noise = np.random.default_rng(seed).standard_normal((self.grid.Nx, self.grid.Nx))
        self._band().init_barotropic(noise, self.grid.L, k_peak, k_width, amplitude,
                                     confine_radius)
        self._pending_full = False

    def inject_cyclones(self, amplitude: float, radius: float, count: int = 1,
                        rng: np.random.Generator | None = None) -> None:
        """Gaussian cyclones added to the resident state's barotropic PV at
        positions drawn as the driver draws them (rng.uniform(0, L, 2) per
        cyclone, scripts/run_polar.py:91-112), on the device."""
        if self.ctx is None:
            raise RuntimeError("Stepper.inject_cyclones before a state is set")
        rng = np.random.default_rng() if rng is None else rng
        L = self.grid.L
        centers = np.array([rng.uniform(0.0, L, size=2) for _ in range(int(count))])
        if centers.size:
            self._current().inject_cyclones(centers, L, amplitude, radius)

    def close(self) -> None:
        for c in (self.ctx, self._full):
            if c is not None:
                c.close()
        self.ctx = self._full = None


def run(grid: Grid, state: State, n_steps: int, save_interval: int, use_imex: bool = True,
        callback=None):
    """Counterpart of the reference's ``run`` (nhqg/solver.py:1385-1416), same
    arguments: sanitize the incoming state, advance n_steps with imex_step
    (use_imex) or rk4_step on the device, call ``callback(state, step, t)``
    and keep a snapshot every save_interval steps."""
    st = sanitize_state(_as_state(state), grid)
    snaps = []
    done = 0
    if use_imex:
        sp = Stepper(grid)
        sp.set_state(st)
        advance, current = sp.run, sp.state
    else:
        ctx = grid.new_context(full=False)
        ctx.set_state(*st)

        def advance(n):
            for _ in range(n):
                ctx.apply(_native.OP_RK4_STEP)

        def current():
            return State(*ctx.get_state())

    for _ in range(n_steps // save_interval):
        advance(save_interval)
        done += save_interval
        s = current()
        t = done * float(grid.dt)
        if callback is not None:
            callback(s, done, t)
        snaps.append((t, s))
    rem = n_steps % save_interval
    if rem:
        advance(rem)
    return current(), snaps
