"""Single-site IDMRG with growing infinite environments on the HIP path.

Energies per site against closed forms: gapped TFIM (unit cell 2) to 1e-10 in
float64 and with BF16x3 applies, and the Majumdar-Ghosh point of the J1-J2
chain (-3/8); the converged IDMRG state agrees with VUMPS at the same bond
dimension.
"""
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    from mpskit_amd import idmrg, vumps
    return mpskit_amd, idmrg, vumps


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_idmrg_tfim_gapped(mk, precision):
    m, I, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 16, seed=1)
    print("idmrg tfim", precision, flush=True)
    psi, envs, eps, hist = I.find_groundstate_idmrg(psi, H, I.IDMRG(tol=1e-11, maxiter=300, precision=precision,
                                                                    verbosity=2))
    e = hist[-1]["energy"]
    print(f"e={e:.15f} exact={B.tfim_infinite_energy(g):.15f} eps={eps:.2e} iters={len(hist)}", flush=True)
    assert eps <= 1e-11
    assert abs(e - B.tfim_infinite_energy(g)) < 1e-10


def test_idmrg_majumdar_ghosh_matches_vumps(mk):
    m, I, V = mk
    H = V.InfiniteHamiltonian([m.models.j1j2_bulk(1.0, 0.5)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 8, seed=2)
    print("idmrg j1j2 MG", flush=True)
    psi, envs, eps, hist = I.find_groundstate_idmrg(psi, H, I.IDMRG(tol=1e-10, maxiter=300, verbosity=2))
    e = hist[-1]["energy"]
    assert abs(e - B.MAJUMDAR_GHOSH) < 1e-8, e
    # the converged state is a VUMPS fixed point too: a VUMPS run from it stays put
    psi2, envs2, eps2, hist2 = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-9, maxiter=5, verbosity=2))
    assert abs(hist2[-1]["energy"] - e) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_idmrg2_tfim_gapped_grows_bond(mk, precision):
    m, I, V = mk
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 4, seed=1)
    print("idmrg2 tfim", precision, flush=True)
    psi, envs, eps, hist = I.find_groundstate_idmrg2(
        psi, H, I.IDMRG2(trunc_dim=16, tol=1e-10, maxiter=300, precision=precision, verbosity=2))
    e = hist[-1]["energy"]
    print(f"e={e:.15f} exact={B.tfim_infinite_energy(g):.15f} eps={eps:.2e} D={hist[-1]['D']}", flush=True)
    assert hist[0]["D"] > 4 and hist[-1]["D"] <= 16
    assert abs(e - B.tfim_infinite_energy(g)) < 1e-9


def test_idmrg2_majumdar_ghosh(mk):
    m, I, V = mk
    H = V.InfiniteHamiltonian([m.models.j1j2_bulk(1.0, 0.5)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 2, seed=2)
    print("idmrg2 j1j2 MG", flush=True)
    psi, envs, eps, hist = I.find_groundstate_idmrg2(psi, H, I.IDMRG2(trunc_dim=8, tol=1e-10, maxiter=300,
                                                                      verbosity=2))
    assert abs(hist[-1]["energy"] - B.MAJUMDAR_GHOSH) < 1e-8, hist[-1]["energy"]
