"""Judged diagnostics on the device (nhqg/diagnostics.py:597-675).

KE_tot / KE_bt / enstrophy / rms values / Nusselt use discrete Parseval
sums over the stored modes with the rfft2 weights of
nhqg/diagnostics.py:48-54; ``Nusselt_dealiased`` uses the solver-consistent
SBP flux profile (diagnostics.py:206-212 -> solver.py:370-374);
``max_speed`` transforms u, v to the Nx grid at every CGL level.
"""

from __future__ import annotations

import numpy as np

from nhqg_amd import _native


def unpack(vec: np.ndarray, nq: int) -> dict:
    d = {name: float(vec[i]) for i, name in enumerate(_native.DIAG_NAMES)}
    s = _native.N_DIAG_SCALARS
    d["q_vert_spec"] = vec[s:s + nq].copy()
    d["w_vert_spec"] = vec[s + nq:s + 2 * nq].copy()
    d["th_vert_spec"] = vec[s + 2 * nq:s + 3 * nq].copy()
    return d


def compute_diagnostics(state, grid) -> dict:
    """Diagnostics of a host state (nhqg/diagnostics.py:597), on the grid's
    scratch context of the layout the state needs."""
    from nhqg_amd.solver import _as_state, needs_full_layout

    st = _as_state(state)
    ctx = grid.context(full=needs_full_layout(st, grid))
    ctx.set_state(*st)
    return unpack(ctx.diagnostics(), grid.Nz + 1)
