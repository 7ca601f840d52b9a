"""leading_boundary of the square-lattice Ising transfer MPO on the HIP path.

The per-site dominant eigenvalue is the partition function per site, so the
boundary MPS is checked against Onsager's closed form: to 1e-10 in ln Z away
from the critical point (FP64 VUMPS, BF16x3 VUMPS with FP64 refinement, VOMPS,
unit cells 1 and 2), and at beta_c to the 1e-3 the reference's own test
accepts at D = 10 (test/algorithms/groundstate.jl:456-483).
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
    from mpskit_amd import boundary, vumps
    return mpskit_amd, boundary, vumps


def _run(mk, beta, D, alg, cell=1, seed=0):
    m, Bd, V = mk
    O = Bd.InfiniteMPO([m.models.classical_ising(beta)] * cell)
    psi = V.InfiniteMPS.random(cell, 2, D, seed=seed)
    print("leading_boundary", type(alg).__name__, alg.precision, "beta", beta, "D", D, "cell", cell, flush=True)
    psi, envs, eps, hist = Bd.leading_boundary(psi, O, alg)
    return np.log(hist[-1]["value"]) / cell, eps, hist


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_vumps_boundary_ising_disordered(mk, precision):
    _, Bd, _ = mk
    beta = 0.3
    logz, eps, hist = _run(mk, beta, 16, Bd.BoundaryVUMPS(tol=1e-11, maxiter=60, verbosity=2, precision=precision))
    assert eps <= 1e-11
    assert abs(logz - B.ising2d_log_partition(beta)) < 1e-10
    if precision != "fp64":
        assert hist[-1]["mixed_ops"] > 0


def test_vumps_boundary_ising_two_site_cell(mk):
    # beta = 0.3 as for the one-site cell: closer to beta_c (0.35 already) the
    # Galerkin error of both cells creeps below 1e-10 only slowly at D = 16
    _, Bd, _ = mk
    beta = 0.3
    logz, eps, _ = _run(mk, beta, 16, Bd.BoundaryVUMPS(tol=1e-11, maxiter=60, verbosity=2), cell=2)
    assert eps <= 1e-11
    assert abs(logz - B.ising2d_log_partition(beta)) < 1e-10


def test_vomps_boundary_ising(mk):
    _, Bd, _ = mk
    beta = 0.3
    logz, eps, _ = _run(mk, beta, 16, Bd.VOMPS(tol=1e-10, maxiter=300, verbosity=1))
    assert eps <= 1e-10
    assert abs(logz - B.ising2d_log_partition(beta)) < 1e-9


def test_vumps_boundary_ising_critical(mk):
    _, Bd, _ = mk
    beta = float(np.log(1 + np.sqrt(2)) / 2)
    logz, eps, _ = _run(mk, beta, 10, Bd.BoundaryVUMPS(tol=1e-4, maxiter=100, verbosity=1))
    assert abs(np.exp(logz) - 2.5337) < 1e-3
    assert abs(logz - B.ising2d_log_partition(beta)) < 1e-3


def test_approximate_infinite_rotation_mpo(mk):
    """An on-site orthogonal rotation R as a w = 1 MPO maps phi to an MPS of the
    same bond dimension; approximate must reproduce it: <sigma>_psi equals
    <R^T sigma R>_phi."""
    m, Bd, V = mk
    from mpskit_amd import observables as Ob
    t = 0.4
    R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    O = Bd.InfiniteMPO([R[None, :, :, None]])
    phi = V.InfiniteMPS.random(1, 2, 8, seed=11)
    psi = V.InfiniteMPS.random(1, 2, 8, seed=12)
    print("approximate infinite rotation", flush=True)
    psi, envs, eps, hist = Bd.approximate_infinite(psi, O, phi, Bd.VOMPS(tol=1e-11, maxiter=100, verbosity=1))
    assert eps <= 1e-10
    for P in (B.PZ, B.PX):
        want = Ob.expectation_value_local(phi, 0, R.T @ P @ R)
        assert abs(Ob.expectation_value_local(psi, 0, P) - want) < 1e-10


def test_approximate_infinite_returns_the_boundary_fixed_point(mk):
    """The Ising transfer MPO applied to its own boundary MPS is that MPS again
    (up to normalisation): same entanglement spectrum."""
    m, Bd, V = mk
    from mpskit_amd import observables as Ob
    O = Bd.InfiniteMPO([m.models.classical_ising(0.3)])
    phi = V.InfiniteMPS.random(1, 2, 12, seed=0)
    print("approximate infinite boundary fixed point", flush=True)
    phi, _, _, _ = Bd.leading_boundary(phi, O, Bd.BoundaryVUMPS(tol=1e-12, maxiter=80, verbosity=0))
    psi = V.InfiniteMPS.random(1, 2, 12, seed=13)
    psi, envs, eps, hist = Bd.approximate_infinite(psi, O, phi, Bd.VOMPS(tol=1e-10, maxiter=200, verbosity=1))
    assert eps <= 1e-10
    s_phi, s_psi = Ob.entanglement_spectrum(phi, 0), Ob.entanglement_spectrum(psi, 0)
    assert float(torch_max_abs(s_phi - s_psi)) < 1e-7


def torch_max_abs(t):
    return t.abs().max()
