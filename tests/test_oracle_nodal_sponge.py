"""The sponge of w and theta on their nodal rows (the device's flux-form
pipeline at Nx = 256) against the reference's coefficient-level sponge.

The reference applies the sponge to the Galerkin coefficient levels of q',
w and theta (nhqg/solver.py:591-597, nhqg/spectral.py:54-60).  The device's
256-point row pass instead multiplies the nodal w and theta rows it already
holds by sigma and lets the back transforms Pw, Pd carry the result to
coefficients.  The sponge acts on every level alike, so it commutes with the
vertical maps (Pw S(VSw w) = S(w)); the two routes differ by roundoff only.
These checks, run on the oracle, bound that roundoff and show why q' keeps
the coefficient route: its nodal values carry the large barotropic part,
whose roundoff in the small high Chebyshev coefficients d/dZ amplifies.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import nhqg_oracle as O
from tests.cases import active_state, trap_sponge_cfg


def _nodal(o):
    """The three nodal routes (oracle notation), by field."""
    sp = O.sponge_tendency
    return {
        "q": lambda f: O.vmat(o.V_inv, sp(O.vmat(o.V, f), o)),
        "w": lambda f: O.vmat(o.w_pinv, O.vmat(o.proj_w, O.vmat(o.V_inv, sp(
            O.vmat(o.V, O.vmat(o.w_stencil, f)), o)))),
        "th": lambda f: O.vmat(o.d_pinv, O.vmat(o.proj_d, O.vmat(o.V_inv, sp(
            O.vmat(o.V, O.vmat(o.S_d, f)), o)))),
    }


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


RULES = ("23_rule", "32_rule")


@pytest.mark.parametrize("rule", RULES)
def test_nodal_sponge_of_w_and_theta_matches_the_coefficient_route(rule):
    cfg = trap_sponge_cfg(horizontal_dealiasing=rule)
    o = O.build_operators(cfg)
    s = active_state(cfg)
    nod = _nodal(o)
    for name, f in (("w", s.w), ("th", s.th)):
        ref = O.sponge_tendency(f, o)
        assert np.abs(ref).max() > 0.0
        assert _rel(nod[name](f), ref) <= 1e-13, name


@pytest.mark.parametrize("rule", RULES)
def test_step_with_nodal_w_theta_sponge_tracks_the_reference_step(rule):
    """One full imex_step with the w and theta sponges on the nodal route
    (q' on its coefficient levels, as on the device) within 1e-13 of the
    reference's step, per field, under either dealiasing rule."""
    cfg = trap_sponge_cfg(horizontal_dealiasing=rule)
    o = O.build_operators(cfg)
    s = active_state(cfg)
    ref = O.imex_step(s, o)
    nod = _nodal(o)
    orig = O.sponge_tendency
    calls = []

    def routed(f, o_):
        # explicit_rhs asks for q', w, theta in that order
        k = len(calls) % 3
        calls.append(k)
        return orig(f, o_) if k == 0 else nod["w" if k == 1 else "th"](f)

    O.sponge_tendency = routed
    try:
        out = O.imex_step(s, o)
    finally:
        O.sponge_tendency = orig
    assert calls and len(calls) % 3 == 0
    for k, a, b in (("q", out.q, ref.q), ("w", out.w, ref.w), ("th", out.th, ref.th),
                    ("thb", out.thb, ref.thb)):
        assert _rel(a, b) <= 1e-13, k
