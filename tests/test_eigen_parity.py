"""The eigen-basis shell solve's decomposition of B0 taken per parity block
(grid._eigen_by_parity): an exact decomposition with pure-parity
eigenvectors, even block first, so that Q and Q^-1 are exact parity blocks
(the structure the split contractions use)."""

import numpy as np
import pytest

from nhqg_amd import make_grid
from nhqg_amd.grid import _eigen_by_parity, shell_matrix_B0
from tests import cases


@pytest.mark.parametrize("nz", [16, 33, 64])
def test_eigen_by_parity_decomposes_b0(nz):
    g = make_grid(cases.fixture_cfg(Nx=16, Nz=nz))
    B0 = shell_matrix_B0(g)
    n = B0.shape[0]
    par = np.arange(n) % 2
    # B0 couples only equal parities (the reflection z -> -z)
    assert np.all(B0[par[:, None] != par[None, :]] == 0.0)
    lam, Q, Qi = _eigen_by_parity(B0)
    assert np.abs(Q @ Qi - np.eye(n)).max() < 1e-10
    assert np.abs(Q @ np.diag(lam) @ Qi - B0).max() < 1e-9 * np.abs(B0).max()
    ne = (n + 1) // 2
    lab = (np.arange(n) >= ne).astype(int)     # eigen index -> parity block
    assert np.all(Q[par[:, None] != lab[None, :]] == 0.0)
    assert np.all(Qi[lab[:, None] != par[None, :]] == 0.0)
    # every eigenpair: B0 q = lam q
    assert np.abs(B0 @ Q - Q * lam[None, :]).max() < 1e-9 * np.abs(B0).max() * np.abs(Q).max()


def test_eigen_by_parity_falls_back_for_mixed_parity():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((9, 9)) + 9 * np.eye(9)
    A = A @ A.T
    lam, Q, Qi = _eigen_by_parity(A)
    assert np.abs(Q @ np.diag(lam) @ Qi - A).max() < 1e-10 * np.abs(A).max()
