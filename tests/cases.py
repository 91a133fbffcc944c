"""Shared configurations and states for the tests (seeded, small)."""

from __future__ import annotations

import numpy as np

from nhqg_amd import NHQGConfig
from oracle import nhqg_oracle as O

FIXTURE_CFG = dict(Nx=32, Nz=16, L=20.0, Ra_tilde=100.0, sigma=1.0, beta=0.0, Ld=float("inf"),
                   dt=1e-3, float_dtype="float64", nu_q=1.0, nu_w=1.0, nu_theta=1.0,
                   hyper_order=1, nonlinear_advection="flux", horizontal_dealiasing="23_rule",
                   thermal_closure="evolve_mean",
                   mean_exchange_discretization="balanced_sbp2_pc", sbp_corrector_substeps=2)


def fixture_cfg(**kw):
    d = dict(FIXTURE_CFG)
    d.update(kw)
    return NHQGConfig(**d)


def trap_sponge_cfg(**kw):
    d = dict(FIXTURE_CFG)
    d.update(Nx=128, Nz=12, L=20.0, dt=5e-4, gamma=5e-3, trap_sharpness=6.0, sponge_rate=5.0,
             sponge_sharpness=8.0, sbp_corrector_substeps=4)
    d.update(kw)
    return NHQGConfig(**d)


def neumann_cfg(**kw):
    d = dict(FIXTURE_CFG)
    d.update(w_bc_top="neumann")
    d.update(kw)
    return NHQGConfig(**d)


def active_state(cfg, seed=0, amp=1e-2):
    """Seeded band-limited state with nonzero q', w, theta and th_bar, built
    from the JAX-compatible draw of the reference's make_initial_state."""
    o = O.build_operators(cfg, with_shells=False)
    s = O.jax_initial_state(o, seed, amp, True)
    w = 0.5 * s.q[1:-1] * np.exp(1j)
    th = 0.3 * s.q[:-2] + 0.1 * s.q[2:]
    # project onto the Galerkin bases so boundary conditions hold
    w = O.vmat(o.w_pinv, O.vmat(o.proj_w, O.vmat(o.V_inv, O.vmat(o.V, np.concatenate(
        [w[:1] * 0, w, w[:1] * 0])))))
    th = O.vmat(o.d_pinv, O.vmat(o.proj_d, np.concatenate([th[:1] * 0, th, th[:1] * 0])))
    z = np.linspace(0.0, 1.0, cfg.Nz + 1)
    thb = o.V_inv @ (1e-3 * np.sin(np.pi * o.Z) * (1 + 0.5 * np.cos(3 * np.pi * o.Z)))
    thb = o.proj_d @ thb
    st = O.sanitize(O.OState(s.q, w, th, thb), o)
    return st


def oracle_reordered(fn, *args):
    """Run an oracle function with every vertical contraction summed in a
    different (equally valid) order: the spread against the plain oracle is
    the roundoff noise floor of the computation."""
    orig = O.vmat

    def split(M, f):
        M = np.ascontiguousarray(M)
        h = M.shape[1] // 2 + 1
        return orig(np.ascontiguousarray(M[:, h:]), f[h:]) + orig(np.ascontiguousarray(M[:, :h]), f[:h])

    O.vmat = split
    try:
        return fn(*args)
    finally:
        O.vmat = orig
