"""The SBP substep cycle on boundary-zeroed input is a five-diagonal level
stencil (CPU).

One SBP substep transfer maps the updated SBP-grid values to Galerkin theta
and back: Psbp = T2S S2T with T2S = c2s V S_d and S2T = Pd V_inv s2c
(the reference's nhqg/solver.py:885-906, the device's corrector).  The
substep update zeroes the two boundary rows first, and on values that
vanish at both ends the Dirichlet projection Pd is the identity on the CGL
interpolant, so Psbp acts as c2s s2c, the composition of the two linear
interpolations between the CGL and SBP grids: half-width 2.  The device
keeps those five diagonals (create_impl) when every other entry of Psbp Z
is below 1e-12 of its largest, and runs each substep transfer but the last
as a stencil.  These checks pin both facts on the grids the tests use.
"""

import numpy as np
import pytest

from nhqg_amd import make_grid
from nhqg_amd.init import polar_config_256
from tests import cases

CONFIGS = {
    "fixture": cases.fixture_cfg,
    "trap_sponge": cases.trap_sponge_cfg,
    "neumann": cases.neumann_cfg,
    "trap_sponge_rule32": lambda: cases.trap_sponge_cfg(horizontal_dealiasing="32_rule"),
    "polar_256x64": lambda: polar_config_256(Nx=256, Nz=64, dt=2.5e-5),
}


def _cycle(g):
    c2s, s2c = g.cgl_to_sbp, g.sbp_to_cgl
    t2s = c2s @ g.V @ g.dirichlet_stencil
    s2t = g.dirichlet_pinv @ g.proj_dirichlet @ g.V_inv @ s2c
    return t2s @ s2t, c2s @ s2c


@pytest.mark.parametrize("name", list(CONFIGS))
def test_sbp_cycle_is_five_diagonal_on_interior_columns(name):
    g = make_grid(CONFIGS[name](), shells="band")
    P, M = _cycle(g)
    n = P.shape[0]
    i, j = np.indices(P.shape)
    interior = (j >= 1) & (j <= n - 2)
    mx = np.abs(P).max()
    off = np.abs(P[interior & (np.abs(i - j) > 2)]).max()
    assert off <= 1e-12 * mx, (name, off / mx)
    # ... and equal there to the two interpolations composed
    assert np.abs((P - M)[:, 1:n - 1]).max() <= 1e-13 * mx
    assert np.abs(M[np.abs(i - j) > 2]).max() == 0.0


def test_the_boundary_columns_are_not_banded():
    """Without the boundary zeroing the cycle is dense in its first and last
    columns (the projection corrects nonzero end values), so the stencil is
    only valid after the substep update's Z, as the device applies it."""
    g = make_grid(cases.fixture_cfg(), shells="band")
    P, _ = _cycle(g)
    n = P.shape[0]
    far = np.abs(np.arange(n) - 0) > 2
    assert np.abs(P[far, 0]).max() > 1e-6
