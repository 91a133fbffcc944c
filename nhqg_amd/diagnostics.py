"""Diagnostics on the device (compute_diagnostics, nhqg/diagnostics.py:597-710).

KE_tot / KE_bt / enstrophy / rms values / Nusselt use discrete Parseval
sums over the stored modes with the rfft2 weights of
nhqg/diagnostics.py:48-54; ``Nusselt_dealiased`` uses the solver-consistent
flux profile (diagnostics.py:206-212: the SBP grid's for the
balanced_sbp2_pc exchange, solver.py:370-374) with the padded-grid means
the reference uses there; ``max_speed``, ``max_w``, ``max_theta`` and
``max_tw`` transform to the Nx grid at every CGL level; the horizontal
shell spectra bin the device's per-mode depth integrals on the host; the
SBP exchange monitor is diagnostics.py:334-369.
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


def shell_bin(mode_quantity: np.ndarray, ksq: np.ndarray, grid) -> tuple:
    """Sum a per-mode quantity into |k| shells of width 2 pi / L:
    bin i holds i dk <= |k| < (i + 1) dk, bins up to the grid's largest |k|
    (nhqg/diagnostics.py:56-77)."""
    dk = 2.0 * np.pi / float(grid.L)
    n_bins = int(float(np.sqrt(np.max(grid.ksq)) / dk)) + 1
    k_mag = np.sqrt(ksq)
    b = np.floor(k_mag / dk).astype(np.int64)
    b = np.where(k_mag < b * dk, b - 1, b)
    b = np.where(k_mag >= (b + 1) * dk, b + 1, b)
    spec = np.zeros(n_bins)
    ok = (b >= 0) & (b < n_bins)
    np.add.at(spec, b[ok], mode_quantity[ok])
    return np.arange(n_bins) * dk + dk / 2, spec


def horizontal_spectra(ctx, grid) -> dict:
    """q_horiz_spec, w_horiz_spec, th_horiz_spec and k_bins of
    compute_diagnostics (nhqg/diagnostics.py:99-108, 683-685): the device's
    depth-integrated mode powers, rfft-weighted and binned over |k|."""
    p = ctx.mode_power()
    lay = ctx.layout
    ky_i = np.arange(lay.nm) // lay.nkx
    wr = np.full(lay.nm, 2.0)
    if grid.Nk > 2:
        wr[(ky_i == 0) | (ky_i == grid.Nk - 1)] = 1.0
    else:
        wr[:] = 1.0
    out = {}
    for name, row in zip(("q", "w", "th"), p):
        k_bins, spec = shell_bin(row * wr, lay.tables["ksq"], grid)
        out[f"{name}_horiz_spec"] = spec
    out["k_bins"] = k_bins
    return out


def device_diagnostics(ctx, grid) -> dict:
    d = unpack(ctx.diagnostics(), grid.Nz + 1)
    d.update(horizontal_spectra(ctx, grid))
    return d


def compute_diagnostics(state, grid) -> dict:
    """Diagnostics of a host state (nhqg/diagnostics.py:597), on the grid's
    scratch context of the layout the state needs."""
    from nhqg_amd.solver import _as_state, needs_full_layout

    st = _as_state(state)
    ctx = grid.context(full=needs_full_layout(st, grid))
    ctx.set_state(*st)
    return device_diagnostics(ctx, grid)
