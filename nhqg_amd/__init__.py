"""MI355X-native NHQG time step (ARS(2,2,2), balanced_sbp2_pc) on gfx950.

Drop-in for the reference's ``make_grid(cfg)`` / ``imex_step(state, grid)``
with the same ``NHQGConfig`` fields and ``State`` layout.
"""

from nhqg_amd.config import NHQGConfig
from nhqg_amd.grid import Grid, make_grid
from nhqg_amd.solver import (State, Stepper, imex_step, is_band_limited, rk4_step, run,
                             sanitize_state)
from nhqg_amd.diagnostics import compute_diagnostics

__all__ = [
    "NHQGConfig", "Grid", "make_grid", "State", "Stepper", "imex_step", "is_band_limited",
    "rk4_step", "run", "sanitize_state", "compute_diagnostics",
]
