"""Host-side logic without a GPU: adaptive Krylov budgets, model MPOs."""
import math

import numpy as np
import pytest

from mpskit_amd.algorithms import AdaptiveKrylov
from mpskit_amd import models as M


def test_adaptive_krylov_cold_start_and_stall():
    ak = AdaptiveKrylov()
    tol, kd, mi = ak.adapt_solver(decay_rate=0.0, g_global=1.0)
    assert (kd, mi) == (3, 1) and tol == 1e-2          # cold start: minimal budget, tol clamped to tol_max
    tol, kd, mi = ak.adapt_solver(decay_rate=1.0, g_global=1e-5)
    assert (kd, mi) == (20, 1) and tol == pytest.approx(1e-6)   # stalled bond: maximal budget


def test_adaptive_krylov_predicted_budget():
    """nmatvecs = round(log(g_local / tol) / log(1 / rho)) (dynamictols.jl:224-233)."""
    ak = AdaptiveKrylov()
    tol, kd, mi = ak.adapt_solver(decay_rate=0.5, g_local=1e-2, g_global=1e-6, eps_trunc=1e-4)
    assert tol == pytest.approx(1e-5)                  # max(0.1 * 1e-4, 0.1 * 1e-6)
    assert kd == round(math.log(1e-2 / 1e-5) / math.log(2.0))
    tol, kd, mi = ak.adapt_solver(decay_rate=0.9, g_local=1.0, g_global=1e-12)
    assert tol == 1e-12 and kd == 20


def test_finite_chain_boundaries():
    Ws = M.xxz(6, 0.5)
    assert Ws[0].shape == (1, 2, 2, 5) and Ws[-1].shape == (5, 2, 2, 1)
    assert all(W.shape == (5, 2, 2, 5) for W in Ws[1:-1])
    Wj = M.j1j2_bulk(1.0, 0.24)
    assert Wj.shape == (8, 2, 2, 8)
    # pass-through identities of the next-nearest-neighbour channels in the A block,
    # upper triangular in channel order
    for k in range(3):
        assert np.array_equal(Wj[1 + k, :, :, 4 + k], np.eye(2))
    for g in range(8):
        for k in range(g):
            assert not np.any(Wj[g, :, :, k]), (g, k)
