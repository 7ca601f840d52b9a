"""Excited states by penalised DMRG2 (FiniteExcited) on the HIP path.

The projector derivative's vector v = GLo . AC2(psi_i) . GRo is checked against
the brute-force P^T |psi_i> (P the dense isometry of the optimised state at the
bond, oracle/brute.py); the first two excited energies of an open TFIM chain
against exact diagonalisation of the Kronecker-built Hamiltonian.
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
    from mpskit_amd import excited
    return mpskit_amd, excited


@pytest.mark.parametrize("pos", [0, 3, 6])
def test_projected_ac2_matches_dense_isometry(mk, pos):
    m, X = mk
    L = 8
    cur = m.FiniteMPS.random(L, 2, 8, seed=1)
    ket = m.FiniteMPS.random(L, 2, 6, seed=2)
    cur.move_center(pos)
    print("projected_ac2 pos", pos, flush=True)
    v = X.OverlapEnvironments(ket).projected_ac2(pos, cur)
    ALs = [None if a is None else a.cpu().numpy() for a in cur.AL]
    ARs = [None if a is None else a.cpu().numpy() for a in cur.AR]
    P = B.isometry_ac2(ALs, ARs, pos, 2)
    ref = P.T @ ket.to_dense().cpu().numpy()
    got = v.permute(3, 2, 1, 0).reshape(-1).cpu().numpy()     # column-major flattening
    assert np.linalg.norm(got - ref) < 1e-12 * max(np.linalg.norm(ref), 1.0)


def test_lanczos_min_matches_eigh(mk):
    import torch
    _, X = mk
    rng = np.random.default_rng(7)
    A = rng.standard_normal((200, 200))
    A = A + A.T
    At = torch.from_numpy(A).cuda()
    lam, x, info = X.lanczos_min(lambda v: At @ v, torch.from_numpy(rng.standard_normal(200)).cuda(), tol=1e-11)
    ev, U = np.linalg.eigh(A)
    assert abs(lam - ev[0]) < 1e-10
    assert abs(abs(float(np.dot(x.cpu().numpy(), U[:, 0]))) - 1.0) < 1e-10


def test_tfim_first_excitations_vs_exact_diagonalisation(mk):
    m, X = mk
    L, g = 10, 1.5
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 32, seed=0)
    print("ground state L=10 g=1.5", flush=True)
    psi, _, _, _ = m.find_groundstate(psi, H, m.DMRG2(tol=1e-12, maxiter=10, trunc_dim=32, verbosity=0))
    print("excitations num=2", flush=True)
    Es, states = X.excitations(H, X.FiniteExcited(weight=10.0, trunc_dim=32, tol=1e-11, verbosity=2), psi, num=2)
    ev = np.linalg.eigvalsh(B.dense_tfim(L, g).toarray())
    assert abs(Es[0] - ev[1]) < 1e-8, (Es, ev[:3])
    assert abs(Es[1] - ev[2]) < 1e-8, (Es, ev[:3])
    # the excited states are orthogonal to the ground state and to each other
    vg = psi.to_dense().cpu().numpy()
    v1, v2 = (s.to_dense().cpu().numpy() for s in states)
    for a, b in ((vg, v1), (vg, v2), (v1, v2)):
        assert abs(np.dot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b)) < 1e-6
