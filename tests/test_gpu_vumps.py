"""VUMPS with infinite MPO-Hamiltonian environments on the HIP path.

Energies per site against closed forms: gapped TFIM (unit cells 1 and 2) to
1e-10, the Majumdar-Ghosh point of the J1-J2 chain (-3/8), and the critical
Heisenberg chain (1/4 - ln 2) to the finite-D accuracy the reference's own
example reaches (docs/src/examples/quantum1d/4.xxz-heisenberg/index.md:561).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    from mpskit_amd import vumps
    return mpskit_amd, vumps


@pytest.mark.parametrize("cell", [1, 2])
def test_vumps_tfim_gapped(mk, cell):
    m, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * cell)
    psi = V.InfiniteMPS.random(cell, 2, 16, seed=1)
    print("vumps tfim cell", cell, flush=True)
    psi, envs, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-11, maxiter=60, verbosity=2))
    e = hist[-1]["energy"]
    print(f"e={e:.15f} exact={B.tfim_infinite_energy(g):.15f} eps={eps:.2e}", flush=True)
    assert abs(e - B.tfim_infinite_energy(g)) < 1e-10


def test_vumps_majumdar_ghosh(mk):
    m, V = mk
    H = V.InfiniteHamiltonian([m.models.j1j2_bulk(1.0, 0.5)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 8, seed=2)
    print("vumps j1j2 MG", flush=True)
    psi, envs, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-9, maxiter=80, verbosity=2))
    e = hist[-1]["energy"]
    assert abs(e - B.MAJUMDAR_GHOSH) < 1e-8, e


def test_vumps_heisenberg_critical(mk):
    m, V = mk
    H = V.InfiniteHamiltonian([m.models.xxz_bulk(1.0)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 48, seed=3)
    print("vumps heisenberg", flush=True)
    psi, envs, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-8, maxiter=80, verbosity=2))
    e = hist[-1]["energy"]
    # D = 48 on a critical chain: variational, within 2e-5 of 1/4 - ln 2
    assert B.HEISENBERG_INFINITE - 1e-9 < e < B.HEISENBERG_INFINITE + 2e-5, e


@pytest.mark.parametrize("precision", ["bf16x3"])
def test_vumps_mixed_matches_fp64(mk, precision):
    m, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 16, seed=1)
    psi, envs, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-11, maxiter=60, precision=precision,
                                                                    verbosity=2))
    assert abs(hist[-1]["energy"] - B.tfim_infinite_energy(g)) < 1e-10
