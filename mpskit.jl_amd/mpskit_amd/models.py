"""Jordan-block MPO builders for the chains this path is exercised on.

Each site tensor is a dense float64 array ``W[g, b, h, k]`` of shape
``(w_left, d, d, w_right)`` (legs: left MPO bond, physical out, physical in,
right MPO bond), in upper-triangular Jordan form

    [ 1  C  D ]
    [ .  A  B ]
    [ .  .  1 ]

(``JordanMPOTensor``, jordanmpotensor.jl:1-69).  Channel 0 of a bond is "not
started", the last channel is "finished", the channels in between carry the
``A`` block.  For finite chains the first tensor keeps only row 0 and the last
tensor only the last column, as ``FiniteMPOHamiltonian`` does
(mpohamiltonian.jl:439-519); a bond of dimension 1 on the left then means
"not started" and on the right "finished".

Operator matrices are ``O[b, h] = <b|O|h>`` (row = physical out).
"""
from __future__ import annotations

import numpy as np

SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
SPLUS = np.array([[0.0, 1.0], [0.0, 0.0]])
SMINUS = SPLUS.T.copy()
SIGMA_X = SPLUS + SMINUS
SIGMA_Z = 2.0 * SZ
ID2 = np.eye(2)


def _jordan(d, C_ops, B_ops, D_op, A_blocks=()):
    """Assemble a bulk Jordan tensor from lists of start/end operators.

    ``C_ops[k]`` starts channel k+1, ``B_ops[k]`` ends channel k+1, ``D_op`` is
    the on-site term, ``A_blocks`` is a list of (row_channel, col_channel, op).
    """
    nmid = len(C_ops)
    w = nmid + 2
    W = np.zeros((w, d, d, w))
    W[0, :, :, 0] = np.eye(d)
    W[w - 1, :, :, w - 1] = np.eye(d)
    for k, op in enumerate(C_ops):
        W[0, :, :, k + 1] = op
    for k, op in enumerate(B_ops):
        W[k + 1, :, :, w - 1] = op
    W[0, :, :, w - 1] = D_op
    for r, c, op in A_blocks:
        W[r, :, :, c] = op
    return W


def finite_chain(Wbulk, L):
    """Finite-chain MPO: first site keeps row 0, last site keeps the last column."""
    Ws = [Wbulk.copy() for _ in range(L)]
    Ws[0] = Ws[0][:1].copy()
    Ws[-1] = Ws[-1][..., -1:].copy()
    if L == 1:
        Ws[0] = Wbulk[:1, :, :, -1:].copy()
    return Ws


def tfim_bulk(g=1.0):
    """H = -sum sigma^z_i sigma^z_{i+1} - g sum sigma^x_i  (w = 3, d = 2)."""
    return _jordan(2, [-SIGMA_Z], [SIGMA_Z], -g * SIGMA_X)


def xxz_bulk(delta=1.0):
    """H = sum (S+S- + S-S+)/2 + delta Sz Sz  (w = 5, d = 2), channels (S+, S-, Sz)."""
    return _jordan(2, [0.5 * SPLUS, 0.5 * SMINUS, delta * SZ], [SMINUS, SPLUS, SZ], np.zeros((2, 2)))


def heisenberg_range_bulk(js):
    """H = sum_r js[r-1] sum_i S_i.S_{i+r}, r = 1..R  (levels 1,3,..,3,1 -> w = 3R + 2, d = 2).

    Channel block j (j = 0..R-1, channels 3j+1..3j+3) carries a pair term that
    still has R-j sites to go: the start operator enters block j with coupling
    js[R-1-j], an identity in the A block passes block j on to block j+1, and
    only the last block ends the term.  Couplings only go to higher channels
    (upper triangular), the level order the environment fixed points rely on.
    For R >= 3 the A blocks of two neighbouring sites connect (block 0 -> 1 -> 2),
    which is the continuing-continuing term of a two-site derivative
    (hamiltonian_derivatives.jl:243-270, :352-362)."""
    R = len(js)
    a = [SPLUS / np.sqrt(2.0), SMINUS / np.sqrt(2.0), SZ]
    b = [SMINUS / np.sqrt(2.0), SPLUS / np.sqrt(2.0), SZ]
    C, B, A = [], [], []
    for j in range(R):
        C += [js[R - 1 - j] * o for o in a]
        B += (b if j == R - 1 else [np.zeros((2, 2))] * 3)
        if j + 1 < R:
            A += [(1 + 3 * j + k, 4 + 3 * j + k, np.eye(2)) for k in range(3)]
    return _jordan(2, C, B, np.zeros((2, 2)), A)


def j1j2_bulk(j1=1.0, j2=0.24):
    """H = j1 sum S_i.S_{i+1} + j2 sum S_i.S_{i+2}  (levels 1,3,3,1 -> w = 8, d = 2).

    Channels 1-3 start a next-nearest-neighbour term, which an identity in the A
    block passes to channels 4-6; channels 4-6 also start the nearest-neighbour
    term, and both end there."""
    return heisenberg_range_bulk([j1, j2])


def exp_ising_bulk(J=1.0, lam=0.5, g=1.0):
    """H = sum_{i<j} J lam^(j-i-1) sigma^z_i sigma^z_j - g sum sigma^x_i  (w = 3, d = 2).

    The single middle channel feeds itself through A = lam * 1, so the A blocks of
    any two neighbouring bulk sites connect: an infinite-range (exponentially
    decaying) interaction in three channels."""
    return _jordan(2, [SIGMA_Z], [J * SIGMA_Z], -g * SIGMA_X, [(1, 1, lam * ID2)])


def tfim(L, g=1.0):
    return finite_chain(tfim_bulk(g), L)


def xxz(L, delta=1.0):
    return finite_chain(xxz_bulk(delta), L)


def j1j2(L, j1=1.0, j2=0.24):
    return finite_chain(j1j2_bulk(j1, j2), L)


def heisenberg_range(L, js):
    return finite_chain(heisenberg_range_bulk(js), L)


def exp_ising(L, J=1.0, lam=0.5, g=1.0):
    return finite_chain(exp_ising_bulk(J, lam, g), L)


def classical_ising(beta=float(np.log(1.0 + np.sqrt(2.0)) / 2)):
    """Row-to-row transfer MPO tensor of the square-lattice Ising model (J = 1),
    ``O[g, b, h, k]`` (left, out, in, right): a delta tensor on the site with the
    symmetric square root of the bond Boltzmann matrix [[e^b, e^-b], [e^-b, e^b]]
    on each of its four legs (classical_ising_tensors / classical_ising,
    test/setup/testsetup.jl:193-214, 232-234).  The per-site dominant eigenvalue
    of the resulting InfiniteMPO is the partition function per site."""
    t = np.array([[np.exp(beta), np.exp(-beta)], [np.exp(-beta), np.exp(beta)]])
    ev, U = np.linalg.eigh(t)
    nt = U @ np.diag(np.sqrt(ev)) @ U.T
    return np.einsum("gs,bs,hs,ks->gbhk", nt, nt, nt, nt)

