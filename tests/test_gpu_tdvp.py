"""Imaginary-time TDVP / TDVP2 on the HIP path.

At full bond dimension (bonds min(2^k, 2^(L-k))) the projector-splitting sweep
reproduces exp(-T H) psi0 exactly, for any step: the state after T is compared
with the dense exp(-T H) psi0 (scipy expm of the Kronecker-built TFIM,
oracle/brute.py) to 1e-11, so a dropped, mis-timed or sign-flipped forward or
backward step, or a wrong environment, fails.  The
Lanczos exponential itself is checked against expm on a dense symmetric
operator.  Long imaginary-time TDVP2 runs reach the open-TFIM closed-form
ground energy (an exact eigenstate is a fixed point of the sweep).
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
    from mpskit_amd import tdvp
    return mpskit_amd, tdvp


def test_exponentiate_matches_expm(mk):
    import torch
    from scipy.linalg import expm
    _, Td = mk
    rng = np.random.default_rng(3)
    n = 300
    A = rng.standard_normal((n, n))
    A = (A + A.T) / np.sqrt(n)
    At = torch.from_numpy(A).cuda()
    x0 = rng.standard_normal(n)
    print("exponentiate n=300", flush=True)
    for tau, kd in ((-0.3, 30), (-6.0, 10)):     # the second needs substeps
        y, info = Td.exponentiate(lambda v: At @ v, tau, torch.from_numpy(x0).cuda(), tol=1e-13, krylovdim=kd)
        ref = expm(tau * A) @ x0
        assert np.linalg.norm(y.cpu().numpy() - ref) < 1e-11 * np.linalg.norm(ref), (tau, info)
    assert info["substeps"] > 1


def _dense_error(m, Td, alg, T_total, nsteps, L=8, g=1.2):
    from scipy.linalg import expm
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 2 ** (L // 2), seed=4)
    v0 = psi.to_dense().cpu().numpy()
    psi, _, _ = Td.time_evolve(psi, H, T_total / nsteps, nsteps, alg)
    v = psi.to_dense().cpu().numpy()
    ref = expm(-T_total * B.dense_tfim(L, g).toarray()) @ v0
    return np.linalg.norm(v - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("kind", ["TDVP", "TDVP2"])
def test_tdvp_full_bond_dimension_against_dense_exponential(mk, kind):
    m, Td = mk
    alg = Td.TDVP() if kind == "TDVP" else Td.TDVP2(trunc_dim=64)
    print("tdvp", kind, "L=8 dense comparison", flush=True)
    e1 = _dense_error(m, Td, alg, 0.2, 2)
    e2 = _dense_error(m, Td, alg, 0.4, 8)
    print(kind, "errors", e1, e2, flush=True)
    assert e1 < 1e-11 and e2 < 1e-11


def test_tdvp2_imaginary_time_ground_state(mk):
    m, Td = mk
    L, g = 16, 2.0
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 16, seed=5)
    print("tdvp2 imaginary time L=16 g=2", flush=True)
    psi, envs, hist = Td.time_evolve(psi, H, 0.1, 60, Td.TDVP2(trunc_dim=16, trunc_err=1e-12), normalize=True)
    E = m.expectation_value(psi, H, envs)
    assert abs(E - B.tfim_open_energy(L, g)) / L < 1e-9, (E, B.tfim_open_energy(L, g))


def test_tdvp_imaginary_time_keeps_the_norm_decay(mk):
    """Without normalisation the norm carries exp(-dt E): after the state has
    relaxed, |psi(t+dt)| / |psi(t)| = exp(-dt E0)."""
    import torch
    m, Td = mk
    L, g = 12, 2.0
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 16, seed=6)
    print("tdvp norm decay L=12", flush=True)
    psi, envs, _ = Td.time_evolve(psi, H, 0.1, 60, Td.TDVP2(trunc_dim=32), normalize=True)
    n0 = float(torch.linalg.vector_norm(psi.AC))
    psi, envs, _ = Td.time_evolve(psi, H, 0.05, 1, Td.TDVP())
    n1 = float(torch.linalg.vector_norm(psi.AC))
    E0 = B.tfim_open_energy(L, g)
    assert abs(np.log(n1 / n0) + 0.05 * E0) < 1e-8


def test_infinite_tdvp_imaginary_time_reaches_the_ground_state(mk):
    """Imaginary-time TDVP on an InfiniteMPS: a VUMPS fixed point is a fixed point
    of the step, so long runs reach the infinite-TFIM closed form (Pfeuty)."""
    m, Td = mk
    from mpskit_amd import vumps as V
    g = 0.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)])
    psi = V.InfiniteMPS.random(1, 2, 12, seed=3)
    print("infinite tdvp tfim D=12", flush=True)
    envs = None
    for _ in range(40):
        psi, envs, _ = Td.timestep_infinite(psi, H, 0.5, Td.TDVP(), envs)
    assert abs(V.energy_per_site(psi, H, envs) - B.tfim_infinite_energy(g)) < 1e-9
