"""DMRG2 on the HIP path: ground-state energies against closed forms and exact
diagonalisation; mixed-precision sweeps against the float64 sweep.

Tolerances: energy per site 1e-10 where the MPS is exact or
converged; the L=20 Heisenberg energy of the reference's own test
(test/algorithms/groundstate.jl:172-174) at its atol 1e-4.
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
    return mpskit_amd


def run(mk, Ws, L, D, precision="fp64", sweeps=10, seed=0, exact=False, tol=1e-12):
    H = mk.FiniteMPOHamiltonian(Ws)
    psi = mk.FiniteMPS.random(L, 2, D, seed=seed)
    alg = mk.DMRG2(tol=tol, maxiter=sweeps, trunc_dim=D, precision=precision, exact_sweeps=exact, verbosity=2)
    psi, envs, eps, hist = mk.find_groundstate(psi, H, alg)
    return hist[-1]["energy"], hist


def test_tfim_small_exact_bond(mk):
    L, g = 12, 1.0
    print("dmrg tfim L=12", flush=True)
    E, hist = run(mk, mk.models.tfim(L, g), L, 64, sweeps=6)
    assert abs(E - B.tfim_open_energy(L, g)) / L < 1e-10, (E, B.tfim_open_energy(L, g))


def test_xxz_vs_exact_diagonalisation(mk):
    L, delta = 10, 0.7
    e_ed = np.linalg.eigvalsh(B.dense_xxz(L, delta).toarray())[0]
    print("dmrg xxz L=10", flush=True)
    E, _ = run(mk, mk.models.xxz(L, delta), L, 32, sweeps=8)
    assert abs(E - e_ed) / L < 1e-10, (E, e_ed)


def test_j1j2_vs_exact_diagonalisation(mk):
    """Exercises the CA / AB (A-block) Jordan terms inside a sweep."""
    L = 10
    e_ed = np.linalg.eigvalsh(B.dense_j1j2(L, 1.0, 0.24).toarray())[0]
    print("dmrg j1j2 L=10", flush=True)
    E, _ = run(mk, mk.models.j1j2(L, 1.0, 0.24), L, 32, sweeps=8)
    assert abs(E - e_ed) / L < 1e-10, (E, e_ed)


def test_heisenberg_L20_reference_energy(mk):
    """test/algorithms/groundstate.jl:151-175: E = -8.6824724 for the L=20 S=1/2 chain."""
    print("dmrg heisenberg L=20", flush=True)
    E, _ = run(mk, mk.models.xxz(20, 1.0), 20, 40, sweeps=8, tol=1e-10)
    assert abs(E - B.HEISENBERG_L20_OBC) < 1e-4, E


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_tfim_L32_mixed_matches_fp64(mk, precision):
    """TFIM L=32, g=1, D=64, 10 sweeps: float64 and mixed sweeps both reach the
    free-fermion energy; per-site difference <= 1e-10."""
    L, g = 32, 1.0
    e_exact = B.tfim_open_energy(L, g)
    print("dmrg tfim L=32 fp64", flush=True)
    E64, _ = run(mk, mk.models.tfim(L, g), L, 64, sweeps=10, exact=True)
    print(f"dmrg tfim L=32 {precision}", flush=True)
    Em, _ = run(mk, mk.models.tfim(L, g), L, 64, precision=precision, sweeps=10, exact=True)
    print(f"E64={E64:.15f} Emixed={Em:.15f} exact={e_exact:.15f}", flush=True)
    assert abs(E64 - e_exact) / L < 1e-10
    assert abs(Em - E64) / L < 1e-10


def test_eigh_gauge_matches_svd_gauge(mk):
    """The Gram-eigh truncated split: isometries to 1e-13 and the same DMRG energy."""
    import torch
    from mpskit_amd.algorithms import truncated_split
    g = torch.Generator(device="cuda").manual_seed(0)
    M = torch.randn(600, 600, dtype=torch.float64, device="cuda", generator=g)
    M = M * torch.logspace(0, -9, 600, dtype=torch.float64, device="cuda")[None, :]
    for direction in ("right", "left"):
        A1, B1, S1, eps = truncated_split(M, direction, 300, 0.0, "eigh")
        A2, B2, S2, eps2 = truncated_split(M, direction, 300, 0.0, "svd")
        iso = A1 if direction == "right" else B1.t()
        err = float(torch.linalg.matrix_norm(iso.t() @ iso - torch.eye(300, dtype=torch.float64, device="cuda")))
        assert err < 1e-13, (direction, err)
        # same truncated matrix and discarded weight
        assert float(torch.linalg.matrix_norm(A1 @ B1 - A2 @ B2)) < 1e-10 * float(torch.linalg.matrix_norm(M))
        # Gram eigenvalues carry an absolute error ~ k * eps_mach * sigma_max^2
        smax2 = float(S2[0]) ** 2
        assert abs(eps ** 2 - eps2 ** 2) < 600 * 2.3e-16 * smax2
    L, g_ = 12, 1.0
    H = mk.FiniteMPOHamiltonian(mk.models.tfim(L, g_))
    psi = mk.FiniteMPS.random(L, 2, 64, seed=0)
    alg = mk.DMRG2(tol=1e-12, maxiter=6, trunc_dim=64, gauge="eigh", verbosity=0)
    _, _, _, hist = mk.find_groundstate(psi, H, alg)
    assert abs(hist[-1]["energy"] - B.tfim_open_energy(L, g_)) / L < 1e-10



@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_single_site_dmrg_tfim_full_bond(mk, precision):
    """DMRG (one site, dmrg.jl:62-170) keeps the start state's bonds: from a
    random state with full bond dimensions it reaches the open-TFIM closed form."""
    L, g = 12, 1.0
    H = mk.FiniteMPOHamiltonian(mk.models.tfim(L, g))
    psi = mk.FiniteMPS.random(L, 2, 64, seed=3)
    print("dmrg1 tfim L=12", precision, flush=True)
    psi, envs, eps, hist = mk.find_groundstate(psi, H, mk.DMRG(tol=1e-10, maxiter=30, precision=precision,
                                                               verbosity=2))
    assert abs(hist[-1]["energy"] - B.tfim_open_energy(L, g)) / L < 1e-10, hist[-1]
    assert psi.bond_dims() == [min(2 ** k, 2 ** (L - k), 64) for k in range(L + 1)]


def test_single_site_dmrg_xxz_vs_exact_diagonalisation(mk):
    L, delta = 10, 0.7
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, delta))
    psi = mk.FiniteMPS.random(L, 2, 32, seed=4)
    print("dmrg1 xxz L=10", flush=True)
    psi, envs, eps, hist = mk.find_groundstate(psi, H, mk.DMRG(tol=1e-10, maxiter=30, verbosity=2))
    e_ed = np.linalg.eigvalsh(B.dense_xxz(L, delta).toarray())[0]
    assert abs(hist[-1]["energy"] - e_ed) / L < 1e-10
