"""changebonds with OptimalExpand on infinite MPS (HIP dAC2 applies).

The expansion must leave the state unchanged (same energy / partition function
per site right after it) and grow every bond by the requested amount; after
re-converging, the larger bond dimension reaches the closed forms: the
reference's own leading_boundary check (2D Ising at beta_c, D = 10 -> 13,
2.5337 per site to 1e-3, test/algorithms/groundstate.jl:456-483) and the TFIM
energy per site (Pfeuty) with VUMPS.
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
    from mpskit_amd import boundary, changebonds, vumps
    return mpskit_amd, boundary, changebonds, vumps


def test_optimal_expand_ising_boundary(mk):
    m, Bd, Cb, V = mk
    beta = float(np.log(1 + np.sqrt(2)) / 2)
    O = Bd.InfiniteMPO([m.models.classical_ising(beta)])
    psi = V.InfiniteMPS.random(1, 2, 10, seed=0)
    print("boundary D=10 -> 13", flush=True)
    psi, envs, _, hist = Bd.leading_boundary(psi, O, Bd.BoundaryVUMPS(tol=1e-4, maxiter=100, verbosity=1))
    before = Bd.expectation_value_mpo(psi, O, envs)
    psi = Cb.changebonds(psi, O, Cb.OptimalExpand(trunc_dim=3), envs)
    assert tuple(psi.AL[0].shape) == (13, 2, 13)
    after = Bd.expectation_value_mpo(psi, O, Bd.BoundaryEnvironments(psi, O))
    assert abs(after - before) < 1e-8 * abs(before)
    psi, envs, _, hist = Bd.leading_boundary(psi, O, Bd.BoundaryVUMPS(tol=1e-4, maxiter=100, verbosity=1))
    assert tuple(psi.AL[0].shape) == (13, 2, 13)
    assert abs(hist[-1]["value"] - 2.5337) < 1e-3


@pytest.mark.parametrize("cell", [1, 2])
def test_optimal_expand_vumps_tfim(mk, cell):
    m, Bd, Cb, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * cell)
    psi = V.InfiniteMPS.random(cell, 2, 6, seed=1)
    print("vumps tfim D=6 -> 12, cell", cell, flush=True)
    psi, envs, _, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-8, maxiter=60, verbosity=1))
    e_before = V.energy_per_site(psi, H, envs)
    # the growth is capped by the local complement, d D - D = 6 here
    psi = Cb.changebonds(psi, H, Cb.OptimalExpand(trunc_dim=8), envs)
    assert all(tuple(a.shape) == (12, 2, 12) for a in psi.AL)
    e_after = V.energy_per_site(psi, H, V.InfiniteEnvironments(psi, H, tol=1e-12))
    assert abs(e_after - e_before) < 1e-9
    psi, envs, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-11, maxiter=60, verbosity=1))
    assert abs(hist[-1]["energy"] - B.tfim_infinite_energy(g)) < 1e-10


def test_svdcut_finite_matches_schmidt_truncation(mk):
    """Cutting a converged L = 12 TFIM ground state to D = 8 keeps the state to
    its discarded weight; a cut above the current bonds changes nothing."""
    import torch
    m, Bd, Cb, V = mk
    L, g = 12, 1.0
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 64, seed=2)
    print("svdcut finite L=12", flush=True)
    psi, _, _, _ = m.find_groundstate(psi, H, m.DMRG2(tol=1e-12, maxiter=8, trunc_dim=64, verbosity=0))
    v0 = psi.to_dense().cpu().numpy()
    same = Cb.svdcut_finite(Cb._copy_finite(psi), Cb.SvdCut(trunc_dim=64))
    assert abs(abs(np.dot(same.to_dense().cpu().numpy(), v0)) - 1.0) < 1e-12
    cut = Cb.svdcut_finite(Cb._copy_finite(psi), Cb.SvdCut(trunc_dim=8))
    assert max(cut.bond_dims()) == 8
    v = cut.to_dense().cpu().numpy()
    assert abs(np.linalg.norm(v) - 1.0) < 1e-12
    assert 1.0 - abs(np.dot(v, v0)) < 1e-4


def test_svdcut_infinite_tfim(mk):
    m, Bd, Cb, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)])
    psi = V.InfiniteMPS.random(1, 2, 16, seed=1)
    print("svdcut infinite D=16 -> 6", flush=True)
    psi, envs, _, _ = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-10, maxiter=60, verbosity=0))
    e16 = V.energy_per_site(psi, H, envs)
    cut = Cb.svdcut_infinite(psi, Cb.SvdCut(trunc_dim=6))
    assert tuple(cut.AL[0].shape) == (6, 2, 6)
    e6 = V.energy_per_site(cut, H, V.InfiniteEnvironments(cut, H, tol=1e-12))
    assert e6 >= B.tfim_infinite_energy(g) - 1e-12
    assert abs(e6 - e16) < 1e-6


def test_optimal_expand_finite_with_single_site_dmrg(mk):
    """Single-site DMRG keeps its bonds; finite OptimalExpand grows them without
    changing the state, and alternating the two reaches the full bond dimension
    and the open-TFIM closed form."""
    m, Bd, Cb, V = mk
    L, g = 10, 1.0
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 2, seed=7)
    envs = m.environments(psi, H)
    print("finite expansion with DMRG1", flush=True)
    full = [min(2 ** k, 2 ** (L - k)) for k in range(L + 1)]
    for _ in range(10):
        psi, envs, _, hist = m.find_groundstate(psi, H, m.DMRG(tol=1e-12, maxiter=6, verbosity=0), envs)
        if psi.bond_dims() == full:
            break
        before = m.expectation_value(psi, H, envs)
        dims = psi.bond_dims()
        psi = Cb.changebonds_finite(psi, H, Cb.OptimalExpand(trunc_dim=8), envs)
        print("bond dims", psi.bond_dims(), flush=True)
        assert abs(m.expectation_value(psi, H, envs) - before) < 1e-10
        assert all(b >= a for a, b in zip(dims, psi.bond_dims())) and psi.bond_dims() != dims
    psi, envs, _, hist = m.find_groundstate(psi, H, m.DMRG(tol=1e-12, maxiter=10, verbosity=0), envs)
    assert psi.bond_dims() == full
    assert abs(hist[-1]["energy"] - B.tfim_open_energy(L, g)) / L < 1e-10


def test_randexpand_infinite_and_finite_preserve_the_state(mk):
    m, Bd, Cb, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 6, seed=1)
    print("randexpand", flush=True)
    psi, envs, _, _ = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-8, maxiter=60, verbosity=0))
    e0 = V.energy_per_site(psi, H, envs)
    big = Cb.randexpand_infinite(psi, Cb.RandExpand(trunc_dim=3, seed=2))
    assert all(tuple(a.shape) == (9, 2, 9) for a in big.AL)
    assert abs(V.energy_per_site(big, H, V.InfiniteEnvironments(big, H, tol=1e-12)) - e0) < 1e-9
    L = 10
    Hf = m.FiniteMPOHamiltonian(m.models.tfim(L, 1.0))
    f = m.FiniteMPS.random(L, 2, 4, seed=3)
    envs_f = m.environments(f, Hf)
    e_f = m.expectation_value(f, Hf, envs_f)
    v0 = f.to_dense().cpu().numpy()
    f = Cb.changebonds_finite(f, None, Cb.RandExpand(trunc_dim=2, seed=4), None)
    assert max(f.bond_dims()) == 6
    v1 = f.to_dense().cpu().numpy()
    assert abs(abs(np.dot(v0, v1)) / (np.linalg.norm(v0) * np.linalg.norm(v1)) - 1.0) < 1e-12
    assert abs(m.expectation_value(f, Hf, m.environments(f, Hf)) - e_f) < 1e-10
