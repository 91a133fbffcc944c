"""Reference-side binding: the reference's imex_step / rk4_step on the MI355X.

This module is the stub a maintainer adds to the reference tree (as
``nhqg/mi355x.py``) to route its hot path to ``libnhqg.so``.  It works on
the reference's own objects:

* ``Grid`` (nhqg/grid.py:29-141), the NamedTuple that ``make_grid(cfg)``
  returns, holding jax (or numpy) arrays.  Every operator the device needs
  is read from its fields; the reference Grid carries no ``NHQGConfig``,
  so the scalars of the per-shell implicit factors (dt, the hyperviscous
  order and coefficients, drag, Ra, sigma) are recovered from its own
  arrays (``diss_rate_*``, ``Ra_sigma``, ``sigma``, ``dt``) and checked
  against them.
* ``State`` (nhqg/solver.py:33-38): the result is returned as the caller's
  State type, as numpy arrays (``jnp.asarray`` them if a jax array is
  wanted).

Dispatch a maintainer would add to nhqg/solver.py::imex_step
(nhqg/solver.py:1228-1242)::

    from nhqg import mi355x
    def imex_step(state, grid):
        if mi355x.enabled() and mi355x.supported(grid):
            return mi355x.step(state, grid)
        ...  # the JAX steppers, unchanged

Input that is not band-limited or not Hermitian on the self-conjugate
columns takes its step in the full device layout (nhqg_amd.solver), so the
binding reproduces the reference on any input, like nhqg_amd.imex_step.
"""

from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np

from nhqg_amd.grid import Grid
from nhqg_amd import solver as _solver

# Reference Grid fields read by this binding (nhqg/grid.py:29-141)
ARRAY_FIELDS = (
    "Z", "xi", "cc_weights", "G_Z", "G_Z2", "V", "V_inv", "dirichlet_stencil",
    "dirichlet_pinv", "w_stencil", "w_pinv", "proj_dirichlet", "proj_neumann", "proj_w",
    "Z_sbp", "sbp_H", "sbp_D1", "sbp_L", "cgl_to_sbp", "sbp_to_cgl", "kx", "ky", "ksq",
    "inv_denom", "mask_23", "diss_rate_q", "diss_rate_w", "diss_rate_th", "inv_alpha_q",
    "inv_alpha_th")
OPTIONAL_FIELDS = ("map_w_to_th", "map_th_to_w", "eta_hat", "sponge_phys", "sponge_phys_pad")
SCALAR_FIELDS = ("beta", "Ra_sigma", "sigma", "L", "Ld_inv_sq", "dt", "gamma_imex",
                 "mean_temp_eps_sq")
STATIC_FIELDS = ("Nx", "Nk", "Nz", "Nz_gal", "Npad", "thermal_closure", "q_boundary", "w_bc_top",
                 "nonlinear_advection", "vertical_cutoff_n", "imex_scheme",
                 "vertical_dealiasing", "horizontal_dealiasing",
                 "mean_exchange_discretization", "sbp_transfer_mode",
                 "sbp_corrector_substeps")


def enabled() -> bool:
    """The routing switch a maintainer would read (NHQG_BACKEND=mi355x)."""
    return os.environ.get("NHQG_BACKEND", "") == "mi355x"


def unsupported_reason(ref) -> str | None:
    """Why the device path does not cover this reference Grid, or None."""
    want = {"vertical_dealiasing": "none", "q_boundary": "none", "imex_scheme": "ars222",
            "sbp_transfer_mode": "interp"}
    if ref.horizontal_dealiasing not in ("23_rule", "32_rule"):
        return f"horizontal_dealiasing={ref.horizontal_dealiasing!r}"
    for name, v in want.items():
        if getattr(ref, name) != v:
            return f"{name}={getattr(ref, name)!r}"
    if ref.mean_exchange_discretization not in ("legacy", "balanced_sbp2_pc"):
        return f"mean_exchange_discretization={ref.mean_exchange_discretization!r}"
    if ref.Nx < 16 or ref.Nx > 1024 or ref.Nx & (ref.Nx - 1):
        return f"Nx={ref.Nx}"
    if ref.Npad != 3 * ref.Nx // 2:
        return f"Npad={ref.Npad}"
    if not 3 <= ref.Nz <= 256:
        return f"Nz={ref.Nz}"
    if np.asarray(ref.dt).dtype != np.float64:
        return "float32 grid"
    return None


def supported(ref) -> bool:
    return unsupported_reason(ref) is None


def _power_law(rate: np.ndarray, ksq: np.ndarray):
    """(coefficient, order, offset) with rate = coefficient * ksq**order +
    offset, recovered from the k = 0, (1, 0) and (2, 0) modes and checked on
    every mode (grid.py:669-677)."""
    off = float(rate[0, 0])
    r1, r2 = float(rate[1, 0]) - off, float(rate[2, 0]) - off
    if r1 == 0.0 and r2 == 0.0:
        coef, order = 0.0, 1
    else:
        order = int(round(np.log(r2 / r1) / np.log(ksq[2, 0] / ksq[1, 0])))
        coef = r1 / float(ksq[1, 0]) ** order
    model = coef * ksq ** order + off
    if not np.allclose(model, rate, rtol=1e-12, atol=1e-300):
        raise ValueError("dissipation rates are not of the form nu |k|^(2p) + drag")
    return coef, order, off


def _rates(ref, ksq):
    """The config scalars of nhqg/grid.py:395-402 recovered from the Grid."""
    nu_q, p_q, drag = _power_law(np.asarray(ref.diss_rate_q, np.float64), ksq)
    nu_w, p_w, _ = _power_law(np.asarray(ref.diss_rate_w, np.float64), ksq)
    nu_t, p_t, _ = _power_law(np.asarray(ref.diss_rate_th, np.float64), ksq)
    orders = {p for p, nu in ((p_q, nu_q), (p_w, nu_w), (p_t, nu_t)) if nu != 0.0}
    if len(orders) > 1:
        raise ValueError("dissipation rates disagree on the hyperviscous order")
    sigma = float(np.asarray(ref.sigma))
    return SimpleNamespace(dt=float(np.asarray(ref.dt)), hyper_order=orders.pop() if orders else 1,
                           nu_q=nu_q, drag=drag, nu_w=nu_w, nu_theta=nu_t * sigma, sigma=sigma,
                           Ra_tilde=float(np.asarray(ref.Ra_sigma)) * sigma, L=float(np.asarray(ref.L)))


def grid_from_reference(ref) -> Grid:
    """This package's Grid, holding the reference Grid's own arrays."""
    why = unsupported_reason(ref)
    if why is not None:
        raise NotImplementedError(f"the MI355X step is not built for {why}")
    kw = {f: np.array(np.asarray(getattr(ref, f)), dtype=np.float64) for f in ARRAY_FIELDS}
    for f in OPTIONAL_FIELDS:
        v = getattr(ref, f)
        kw[f] = None if v is None else np.array(np.asarray(v))
    if kw["eta_hat"] is not None:
        kw["eta_hat"] = kw["eta_hat"].astype(np.complex128)
    kw.update({f: float(np.asarray(getattr(ref, f))) for f in SCALAR_FIELDS})
    kw.update({f: getattr(ref, f) for f in STATIC_FIELDS})
    ksq = kw["ksq"]
    kw["ksq_idx"] = np.asarray(ref.ksq_idx).astype(np.int32)
    kb = None
    if kw["map_w_to_th"] is not None:
        kb = kw["map_th_to_w"] @ kw["map_w_to_th"]
    g = Grid(cfg=_rates(ref, ksq), buoyancy_K=kb, **kw)
    g._shell_values = np.unique(np.round(ksq.ravel(), decimals=8))
    g._shell_mode = "all"
    return g


_CACHE: dict = {}


def _grid_for(ref) -> Grid:
    key = id(ref)
    hit = _CACHE.get(key)
    if hit is None or hit[0] is not ref:
        hit = (ref, grid_from_reference(ref))
        _CACHE[key] = hit
    return hit[1]


def _as_reference(state, new):
    return type(state)(*new) if hasattr(type(state), "_fields") else new


def step(state, ref):
    """imex_step(state, grid) on the device (nhqg/solver.py:1228-1242)."""
    new = _solver.imex_step(_solver.State(*(np.asarray(x) for x in state)), _grid_for(ref))
    return _as_reference(state, new)


def rk4_step(state, ref):
    """rk4_step(state, grid) on the device (nhqg/solver.py:1276-1336)."""
    new = _solver.rk4_step(_solver.State(*(np.asarray(x) for x in state)), _grid_for(ref))
    return _as_reference(state, new)


def run(ref, state, n_steps: int, save_interval: int, use_imex: bool = True, callback=None):
    """run(grid, state, n_steps, save_interval, use_imex, callback)
    (nhqg/solver.py:1385-1416) with the state resident on the device."""
    final, snaps = _solver.run(_grid_for(ref), _solver.State(*(np.asarray(x) for x in state)),
                               n_steps, save_interval, use_imex,
                               None if callback is None else
                               (lambda s, k, t: callback(_as_reference(state, s), k, t)))
    return _as_reference(state, final), [(t, _as_reference(state, s)) for t, s in snaps]
