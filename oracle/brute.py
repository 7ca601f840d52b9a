"""Brute-force pins for the oracle: dense Hamiltonians, small MPS, closed forms.

TEST INFRASTRUCTURE ONLY (same rule as ``oracle.oracle``): used by ``tests/``
and ``__graft_entry__.smoke()`` as the checker, never by the product path.

Everything here is built without the index strings of ``oracle.oracle``:
Hamiltonians are sums of Kronecker products of spin matrices, states are
dense vectors contracted from MPS tensors site by site, and projected
operators are ``P_bra^T H P_ket`` with explicit isometries.  A dropped term,
a wrong sign or a transposed leg in any contraction restated in
``oracle.oracle`` changes ``P^T H P x`` and fails the comparison.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# spin-1/2 operators (real): S^z, S^+, S^-, and Pauli X, Z
SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])
SM = SP.T.copy()
SX = 0.5 * (SP + SM)
PX = 2.0 * SX
PZ = 2.0 * SZ
I2 = np.eye(2)


def _site_op(op, i, L):
    """Dense-sparse embedding of a one-site operator (site 0 is most significant)."""
    m = sp.identity(1, format="csr")
    for j in range(L):
        m = sp.kron(m, sp.csr_matrix(op) if j == i else sp.identity(op.shape[0]), format="csr")
    return m


def bond_sum(L, terms, dist=1):
    """sum_i sum_(coef, a, b) coef * a_i b_{i+dist} for an open chain."""
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L - dist):
        for coef, a, b in terms:
            H = H + coef * (_site_op(a, i, L) @ _site_op(b, i + dist, L))
    return H


def field_sum(L, coef, op):
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L):
        H = H + coef * _site_op(op, i, L)
    return H


def dense_tfim(L, g):
    """H = -sum sigma^z sigma^z - g sum sigma^x (test/setup/testsetup.jl:102-119)."""
    return bond_sum(L, [(-1.0, PZ, PZ)]) + field_sum(L, -g, PX)


def dense_xxz(L, delta):
    """H = sum S^x S^x + S^y S^y + delta S^z S^z = sum (S+S- + S-S+)/2 + delta SzSz."""
    return bond_sum(L, [(0.5, SP, SM), (0.5, SM, SP), (delta, SZ, SZ)])


def dense_j1j2(L, j1, j2):
    t = [(0.5, SP, SM), (0.5, SM, SP), (1.0, SZ, SZ)]
    return j1 * bond_sum(L, t, 1) + j2 * bond_sum(L, t, 2)


def dense_heisenberg_range(L, js):
    """H = sum_r js[r-1] sum_i S_i.S_{i+r} for an open chain."""
    t = [(0.5, SP, SM), (0.5, SM, SP), (1.0, SZ, SZ)]
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for r, j in enumerate(js, start=1):
        H = H + j * bond_sum(L, t, r)
    return H


def dense_exp_ising(L, J, lam, g):
    """H = sum_{i<j} J lam^(j-i-1) sigma^z_i sigma^z_j - g sum sigma^x_i (open chain)."""
    H = field_sum(L, -g, PX)
    for r in range(1, L):
        H = H + J * lam ** (r - 1) * bond_sum(L, [(1.0, PZ, PZ)], r)
    return H


def mpo_to_dense(Ws):
    """Contract a finite MPO (W[g,b,h,k], boundary bonds of dimension 1) to a matrix."""
    T = Ws[0][0]                                  # (b, h, k)
    T = T.reshape(T.shape[0], T.shape[1], T.shape[2])
    for W in Ws[1:]:
        # T[B, H, g] W[g, b, h, k] -> T[(B b), (H h), k]
        nb, nh = T.shape[0], T.shape[1]
        T = np.einsum("BHg,gbhk->BbHhk", T, W).reshape(nb * W.shape[1], nh * W.shape[2], W.shape[3])
    return T[:, :, 0]


# ---------------------------------------------------------------------------
# Small MPS utilities
# ---------------------------------------------------------------------------

def random_mps(L, d, Dmax, rng):
    """Random open-boundary MPS tensors A[l,p,r] with bonds min(d^k, d^(L-k), Dmax)."""
    dims = [min(d ** k, d ** (L - k), Dmax) for k in range(L + 1)]
    return [rng.standard_normal((dims[k], d, dims[k + 1])) for k in range(L)]


def left_canonical(As):
    """QR sweep left->right; returns (ALs, norm) with the state normalised."""
    out, R = [], np.eye(1)
    for A in As:
        A = np.einsum("lm,mpr->lpr", R, A)
        l, p, r = A.shape
        Q, R = np.linalg.qr(A.reshape(l * p, r, order="F"))
        k = Q.shape[1]
        out.append(Q.reshape(l, p, k, order="F"))
    nrm = R[0, 0]
    out[-1] = out[-1] * np.sign(nrm)
    return out


def right_canonical(As):
    out, R = [], np.eye(1)
    for A in reversed(As):
        A = np.einsum("lpm,mr->lpr", A, R)
        l, p, r = A.shape
        Q, Rt = np.linalg.qr(A.reshape(l, p * r, order="F").T)
        k = Q.shape[1]
        out.append(Q.T.reshape(k, p, r, order="F"))
        R = Rt.T
    out.reverse()
    out[0] = out[0] * np.sign(R[0, 0])
    return out


def mps_to_dense(As):
    psi = As[0][0]                                 # (p, r)
    for A in As[1:]:
        psi = np.einsum("Pm,mpr->Ppr", psi, A).reshape(-1, A.shape[2])
    return psi[:, 0]


def left_block(ALs):
    """Matrix U[(s_1..s_k), r] of the left block (site 1 most significant)."""
    U = np.ones((1, 1))
    for A in ALs:
        U = np.einsum("Sl,lpr->Spr", U, A).reshape(-1, A.shape[2])
    return U


def right_block(ARs):
    """Matrix V[l, (s_k..s_L)] of the right block."""
    V = np.ones((1, 1))
    for A in reversed(ARs):
        V = np.einsum("lpr,rS->lpS", A, V).reshape(A.shape[0], -1)
    return V


def _fortran_cols(T, nrow_axes):
    """Rows: C-order over the first ``nrow_axes`` axes; columns: Fortran order
    over the rest (first column axis fastest), matching the library layout."""
    nrow = int(np.prod(T.shape[:nrow_axes]))
    ncol_axes = T.ndim - nrow_axes
    perm = list(range(nrow_axes)) + list(range(T.ndim - 1, nrow_axes - 1, -1))
    return T.transpose(perm).reshape(nrow, -1)


def isometry_ac2(ALs, ARs, j, d):
    """P[(s), (f,h,i,jj)] for the two-site centre at sites j, j+1 (0-based)."""
    U = left_block(ALs[:j])          # (S_l, f)
    V = right_block(ARs[j + 2:])     # (i, S_r)
    T = np.einsum("Sf,hH,jJ,iR->ShjRfHiJ", U, np.eye(d), np.eye(d), V)
    return _fortran_cols(T, 4)


def isometry_ac(ALs, ARs, j, d):
    """P[(s), (f,h,i)] for the one-site centre at site j (0-based)."""
    U = left_block(ALs[:j])
    V = right_block(ARs[j + 1:])
    T = np.einsum("Sf,hH,iR->ShRfHi", U, np.eye(d), V)
    return _fortran_cols(T, 3)


def isometry_c(ALs, ARs, j):
    """P[(s), (f,i)] for the bond left of site j (0-based): left block = sites < j."""
    U = left_block(ALs[:j])
    V = right_block(ARs[j:])
    T = np.einsum("Sf,iR->SRfi", U, V)
    return _fortran_cols(T, 2)


# ---------------------------------------------------------------------------
# Closed-form energies
# ---------------------------------------------------------------------------

def tfim_open_energy(L, g):
    """Open-chain TFIM ground energy -sum_k s_k(M), M bidiagonal (M_ii=g, M_i,i+1=1)
    (Lieb-Schultz-Mattis / Pfeuty)."""
    M = np.diag(np.full(L, float(g))) + np.diag(np.ones(L - 1), 1)
    return -np.linalg.svd(M, compute_uv=False).sum()


TFIM_INFINITE_G1 = -4.0 / np.pi
HEISENBERG_INFINITE = 0.25 - np.log(2.0)
MAJUMDAR_GHOSH = -0.375
HEISENBERG_L20_OBC = -8.6824724   # test/algorithms/groundstate.jl:172-174 (atol 1e-4)


def tfim_infinite_energy(g):
    """Infinite TFIM (H = -sum s^z s^z - g sum s^x) ground energy per site,
    e0 = -(1/pi) int_0^pi sqrt(1 + g^2 - 2 g cos k) dk  (Pfeuty)."""
    from scipy.integrate import quad
    val, _ = quad(lambda k: np.sqrt(1.0 + g * g - 2.0 * g * np.cos(k)), 0.0, np.pi, epsabs=1e-14, epsrel=1e-14)
    return -val / np.pi


CATALAN = 0.915965594177219015054603514932384110774


def ising2d_log_partition(beta):
    """Square-lattice Ising model (J = 1, no field), ln Z per site (Onsager,
    single-integral form): ln(2 cosh 2b) + (1/pi) int_0^{pi/2}
    ln[(1 + sqrt(1 - k^2 sin^2 t)) / 2] dt with k = 2 sinh 2b / cosh^2 2b."""
    from scipy.integrate import quad
    k = 2.0 * np.sinh(2 * beta) / np.cosh(2 * beta) ** 2
    f = lambda t: np.log(0.5 * (1.0 + np.sqrt(max(1.0 - (k * np.sin(t)) ** 2, 0.0))))
    val, _ = quad(f, 0.0, np.pi / 2, epsabs=1e-14, epsrel=1e-14, limit=200)
    return np.log(2.0 * np.cosh(2 * beta)) + val / np.pi


def ising2d_log_partition_double(beta):
    """The same from Onsager's double integral: ln 2 + (1/(8 pi^2)) int int
    ln[cosh^2 2b - sinh 2b (cos t1 + cos t2)] dt1 dt2 over [0, 2 pi]^2 (an
    independent form: the inner integral is done in closed form,
    int_0^{2pi} ln(a - b cos t) dt = 2 pi ln((a + sqrt(a^2 - b^2)) / 2))."""
    from scipy.integrate import quad
    c2, s2 = np.cosh(2 * beta) ** 2, np.sinh(2 * beta)

    def inner(t1):
        a = c2 - s2 * np.cos(t1)
        return 2 * np.pi * np.log(0.5 * (a + np.sqrt(max(a * a - s2 * s2, 0.0))))

    val, _ = quad(inner, 0.0, 2 * np.pi, epsabs=1e-13, epsrel=1e-13, limit=200)
    return np.log(2.0) + val / (8 * np.pi ** 2)


def ising2d_torus_partition(beta, L):
    """Brute-force Z of an L x L periodic Ising lattice (2^(L*L) configurations)."""
    n = L * L
    cfg = ((np.arange(2 ** n)[:, None] >> np.arange(n)[None, :]) & 1) * 2 - 1
    s = cfg.reshape(-1, L, L)
    e = (s * np.roll(s, 1, axis=1)).sum(axis=(1, 2)) + (s * np.roll(s, 1, axis=2)).sum(axis=(1, 2))
    return float(np.exp(beta * e).sum())



def tfim_infinite_sx(g):
    """<sigma^x> per site of the infinite TFIM ground state, -d e0 / dg
    (Hellmann-Feynman on Pfeuty's e0): (1/pi) int_0^pi (g - cos k) / sqrt(1 + g^2 - 2 g cos k) dk."""
    from scipy.integrate import quad
    val, _ = quad(lambda k: (g - np.cos(k)) / np.sqrt(1.0 + g * g - 2.0 * g * np.cos(k)), 0.0, np.pi,
                  epsabs=1e-14, epsrel=1e-14)
    return val / np.pi
