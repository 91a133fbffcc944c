"""A stand-in for the reference's ``Grid`` NamedTuple (nhqg/grid.py:29-141):
the same 79 field names in the same order, filled either from this
package's grid or from the oracle's operators.  Fields the device path
does not read (the work-grid, dealiasing and mass matrices of other
exchange modes) are None; the binding under test must not need them."""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from oracle import nhqg_oracle as O

FIELDS = (
    "Z", "xi", "cc_weights", "G_Z", "G_Z2", "V", "V_inv", "V_dealias", "V_dealias_inv",
    "V_exchange", "V_exchange_inv", "dirichlet_stencil", "dirichlet_pinv", "w_stencil",
    "w_pinv", "map_w_to_th", "map_th_to_w", "V_exchange_dirichlet", "G_exchange",
    "exchange_weights", "mean_mass", "mean_mass_inv", "theta_mass", "theta_mass_inv", "Z_sbp",
    "sbp_H", "sbp_D1", "sbp_L", "cgl_to_sbp", "sbp_to_cgl", "kx", "ky", "ksq", "inv_denom",
    "mask_23", "diss_q", "diss_w", "diss_th", "diss_rate_q", "diss_rate_w", "diss_rate_th",
    "inv_alpha_q", "inv_alpha_th", "imex_inv", "q_solve", "ksq_idx", "proj_dirichlet",
    "proj_neumann", "proj_w", "beta", "eta_hat", "sponge_phys", "sponge_phys_pad", "Ra_sigma",
    "sigma", "L", "Ld_inv_sq", "dt", "gamma_imex", "mean_temp_eps_sq", "Nx", "Nk", "Nz",
    "Nz_gal", "Nz_dealias", "Nz_exchange", "Npad", "thermal_closure", "q_boundary", "w_bc_top",
    "nonlinear_advection", "vertical_cutoff_n", "imex_scheme", "vertical_dealiasing",
    "horizontal_dealiasing", "mean_exchange_discretization", "sbp_transfer_mode",
    "sbp_corrector_substeps", "imex_matmul_chunk")

RefGrid = NamedTuple("RefGrid", [(f, object) for f in FIELDS])


def _statics(cfg):
    return dict(Nx=cfg.Nx, Nk=cfg.Nx // 2 + 1, Nz=cfg.Nz, Nz_gal=cfg.Nz - 1, Nz_dealias=cfg.Nz,
                Nz_exchange=cfg.Nz + 1, Npad=3 * cfg.Nx // 2,
                thermal_closure=cfg.thermal_closure, q_boundary=cfg.q_boundary,
                w_bc_top=cfg.w_bc_top, nonlinear_advection=cfg.nonlinear_advection,
                vertical_cutoff_n=cfg.vertical_cutoff_n, imex_scheme=cfg.imex_scheme,
                vertical_dealiasing=cfg.vertical_dealiasing,
                horizontal_dealiasing=cfg.horizontal_dealiasing,
                mean_exchange_discretization=cfg.mean_exchange_discretization,
                sbp_transfer_mode=cfg.sbp_transfer_mode,
                sbp_corrector_substeps=cfg.sbp_corrector_substeps,
                imex_matmul_chunk=cfg.imex_matmul_chunk)


def from_package_grid(g, cfg) -> RefGrid:
    """The reference-shaped tuple holding this package's grid arrays."""
    d = {f: None for f in FIELDS}
    for f in FIELDS:
        if hasattr(g, f):
            d[f] = getattr(g, f)
    d.update(_statics(cfg))
    for f, rate in (("diss_q", g.diss_rate_q), ("diss_w", g.diss_rate_w),
                    ("diss_th", g.diss_rate_th)):
        d[f] = np.exp(-rate * cfg.dt)
    d["dt"] = np.float64(cfg.dt)
    return RefGrid(**d)


def from_oracle(cfg) -> RefGrid:
    """The reference-shaped tuple holding the oracle's restatement of
    make_grid (nhqg/grid.py:439-789)."""
    o = O.build_operators(cfg)
    N = cfg.Nz
    ep, em = np.ones(N + 1), (-1.0) ** np.arange(N + 1)
    d = {f: None for f in FIELDS}
    d.update(Z=o.Z, xi=o.xi, cc_weights=o.cc, G_Z=o.G_Z, G_Z2=o.G_Z2, V=o.V, V_inv=o.V_inv,
             dirichlet_stencil=o.S_d, dirichlet_pinv=o.d_pinv, w_stencil=o.w_stencil,
             w_pinv=o.w_pinv, map_w_to_th=o.map_w_to_th, map_th_to_w=o.map_th_to_w,
             Z_sbp=o.Z_sbp, sbp_H=o.sbp_H, sbp_D1=o.sbp_D1, sbp_L=o.sbp_L, cgl_to_sbp=o.c2s,
             sbp_to_cgl=o.s2c, kx=o.kx, ky=o.ky, ksq=o.ksq, inv_denom=o.inv_denom,
             mask_23=o.mask, diss_q=np.exp(-o.diss_q * cfg.dt), diss_w=np.exp(-o.diss_w * cfg.dt),
             diss_th=np.exp(-o.diss_th * cfg.dt), diss_rate_q=o.diss_q, diss_rate_w=o.diss_w,
             diss_rate_th=o.diss_th, inv_alpha_q=o.inv_alpha_q, inv_alpha_th=o.inv_alpha_th,
             imex_inv=o.imex_inv, ksq_idx=o.ksq_idx.astype(np.int32), proj_dirichlet=o.proj_d,
             proj_neumann=O.tau_projection(np.stack([ep @ o.G_Z, em @ o.G_Z]), N),
             proj_w=o.proj_w, beta=np.float64(cfg.beta), eta_hat=o.eta_hat,
             sponge_phys=o.sponge_phys, sponge_phys_pad=o.sponge_phys_pad,
             Ra_sigma=np.float64(o.Ra_sigma), sigma=np.float64(cfg.sigma), L=np.float64(cfg.L),
             Ld_inv_sq=np.float64(o.Ld_inv_sq), dt=np.float64(cfg.dt),
             gamma_imex=np.float64(o.gamma_imex), mean_temp_eps_sq=np.float64(o.mu))
    d.update(_statics(cfg))
    return RefGrid(**d)
