"""U(1)-symmetric finite MPS and two-site DMRG on the charge-block path.

TensorKit stores a U(1)-symmetric tensor as a list of dense blocks, one per
charge sector (``TensorMap`` over ``U1Space``; the reference's derivative
applies run one GEMM per fusion-tree block, utility.jl:285-302).  Here a
U(1) tensor is a dense array in a charge-sorted basis: every bond space is a
concatenation of charge sectors (``U1Bond``), and only charge-conserving blocks
are ever nonzero.  The two-site derivative is prepared densely and then told
its charge blocks (``mk_op_set_charge_blocks``), so every apply computes only
those blocks, all in one grouped persistent GEMM; local eigensolves run through
``BatchedDerivative`` (hipGraph-captured Krylov steps).  Truncation splits the
two-site center per middle-charge block (batched float64 eigensolves of the
blocks' Gram matrices) and keeps the largest singular values over all blocks,
as TensorKit's truncated SVD of a block-sparse tensor does (orthoview.jl:401-420).

Charges are numbers of up spins; the physical basis of the spin-1/2 models is
(up, down), so the physical charges are (1, 0).  Interior bond sectors are
padded to multiples of 8 states with zero vectors, which keeps the sector
offsets of mixed-precision operands 16-byte aligned; the boundary bonds have
one state.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import tensors as T
from ._lib import check, load
from .derivatives import BatchedDerivative, JordanMPO_AC2_Hamiltonian, prepare_operator
from .environments import FiniteEnvironments
from .algorithms import expectation_value

PHYS_Q = (1, 0)   # (up, down) basis of models.xxz / models.tfim-type spin-1/2 chains


def pad8(n: int) -> int:
    return (n + 7) // 8 * 8


@dataclass
class U1Bond:
    """Charge-sorted bond space: sector s has charge q[s] and dims[s] states."""
    q: list
    dims: list

    @property
    def D(self) -> int:
        return int(sum(self.dims))

    @property
    def offsets(self):
        return np.concatenate([[0], np.cumsum(self.dims)]).astype(np.int64)

    def charge_of_index(self, device):
        return torch.repeat_interleave(torch.tensor(self.q, device=device), torch.tensor(self.dims, device=device))


def set_charge_blocks(op, left: U1Bond, right: U1Bond, phys_q=PHYS_Q):
    """Attach the charge blocks of a two-site operator (mk_op_set_charge_blocks)."""
    lo, ro = left.offsets, right.offsets
    lq, rq = np.asarray(left.q, dtype=np.int32), np.asarray(right.q, dtype=np.int32)
    pq = np.asarray(phys_q, dtype=np.int32)
    check(load().mk_op_set_charge_blocks(op.handle, len(left.q), lo.ctypes.data_as(C.POINTER(C.c_int64)),
                                         lq.ctypes.data_as(C.POINTER(C.c_int)), len(right.q),
                                         ro.ctypes.data_as(C.POINTER(C.c_int64)),
                                         rq.ctypes.data_as(C.POINTER(C.c_int)),
                                         pq.ctypes.data_as(C.POINTER(C.c_int))))
    xe, ye, fl = C.c_int64(), C.c_int64(), C.c_double()
    check(load().mk_op_info(op.handle, C.byref(xe), C.byref(ye), C.byref(fl)))
    op.flops = fl.value
    return op


class CompactLayout:
    """Compact charge-block storage of a charge-blocked two-site operator
    (``mk_op_compact_charge_blocks``): ``gather`` takes a dense column-major
    (Dl, d1, Dr, d2) tensor to the compact vector of its charge blocks,
    ``scatter`` puts a compact vector back into a dense tensor (zero outside the
    blocks)."""

    def __init__(self, op, shape):
        nb = C.c_int()
        check(load().mk_op_charge_blocks(op.handle, None, 0, C.byref(nb)))
        buf = (C.c_int64 * (5 * nb.value))()
        check(load().mk_op_charge_blocks(op.handle, buf, nb.value, C.byref(nb)))
        blocks = np.frombuffer(buf, dtype=np.int64).reshape(nb.value, 5)
        Dl, d1, Dr, d2 = shape
        M = Dl * d1
        idx = []
        for r0, m, c0, n, _ in blocks:
            idx.append((np.arange(r0, r0 + m)[:, None] + M * np.arange(c0, c0 + n)[None, :]).ravel(order="F"))
        self.blocks = blocks
        self.shape = tuple(shape)
        self.idx = torch.from_numpy(np.concatenate(idx)).to(T.device())
        self.n = int(self.idx.numel())

    def gather(self, x):
        return T.as_colmajor(x).permute(3, 2, 1, 0).reshape(-1)[self.idx].contiguous()

    def scatter(self, xc):
        out = T.zeros(self.shape)
        out.permute(3, 2, 1, 0).reshape(-1)[self.idx] = xc.reshape(-1)
        return out


def compact_charge_blocks(op, shape):
    """Switch a charge-blocked operator to compact storage; returns its layout
    (the operator's x and y become the compact vectors)."""
    check(load().mk_op_compact_charge_blocks(op.handle))
    lay = CompactLayout(op, shape)
    op.x_shape = op.y_shape = (lay.n,)
    return lay


def _block_rows(left: U1Bond, phys_q, c):
    """Fused (l, p) indices of middle charge c, l fastest (column-major fusion)."""
    idx = []
    o = left.offsets
    D = left.D
    for h, qp in enumerate(phys_q):
        for s, q in enumerate(left.q):
            if q + qp == c:
                idx.append(np.arange(o[s], o[s + 1]) + D * h)
    return np.concatenate(idx) if idx else np.zeros(0, dtype=np.int64)


def _block_cols(right: U1Bond, phys_q, c):
    """Fused (r, p) indices (r fastest) with right charge c + q(p)."""
    idx = []
    o = right.offsets
    D = right.D
    for j, qp in enumerate(phys_q):
        for s, q in enumerate(right.q):
            if q == c + qp:
                idx.append(np.arange(o[s], o[s + 1]) + D * j)
    return np.concatenate(idx) if idx else np.zeros(0, dtype=np.int64)


def _middle_charges(left: U1Bond, right: U1Bond, phys_q):
    a = {q + p for q in left.q for p in phys_q}
    b = {q - p for q in right.q for p in phys_q}
    return sorted(a & b)


def _batched_gram_eigh(blocks):
    """Eigen-decompositions of the Gram matrices of a list of matrices (rows side):
    per block (eigenvalues descending, eigenvectors), batched over size groups."""
    from .algorithms import gram_eigh_grouped
    return gram_eigh_grouped(blocks)


class U1FiniteMPS:
    """Mixed-canonical U(1)-symmetric finite MPS (dense charge-sorted storage)."""

    def __init__(self, bonds, AL, AR, AC, center):
        self.bonds = bonds
        self.L = len(bonds) - 1
        self.AL, self.AR, self.AC, self.center = AL, AR, AC, center

    # -- construction --------------------------------------------------------
    @classmethod
    def random(cls, L, n_up, sector_dims, phys_q=PHYS_Q, seed=0, device=None):
        """Random state with n_up up spins.  ``sector_dims(k, q)`` gives the
        requested dimension of charge q on bond k (capped by the number of
        states with that charge, padded to multiples of 8 on interior bonds)."""
        device = device or T.device()
        bonds = []
        for k in range(L + 1):
            qs, ds = [], []
            for q in range(max(0, n_up - (L - k)), min(k, n_up) + 1):
                cap = min(math.comb(k, q), math.comb(L - k, n_up - q))
                d = min(cap, int(sector_dims(k, q)))
                if d <= 0:
                    continue
                qs.append(q)
                ds.append(d if k in (0, L) else pad8(d))
            bonds.append(U1Bond(qs, ds))
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        # right-orthonormal sites from the right: blockwise QR of each site's transpose
        AR = [None] * L
        for j in range(L - 1, 0, -1):
            AR[j] = _random_right_isometry(bonds[j], bonds[j + 1], phys_q, g, device)
        lq = bonds[0]
        A0 = torch.zeros((lq.D, len(phys_q), bonds[1].D), dtype=torch.float64, device=device)
        for c in _middle_charges(lq, bonds[1], phys_q):
            rows, cols = _block_rows(lq, phys_q, c), _block_cols_single(bonds[1], c)
            A0f = T.fuse(A0, 2)
            blk = torch.randn((len(rows), len(cols)), dtype=torch.float64, device=device, generator=g)
            A0f[torch.as_tensor(rows, device=device)[:, None], torch.as_tensor(cols, device=device)[None, :]] = blk
            A0 = T.as_colmajor(T.unfuse(A0f, A0.shape))
        A0 = A0 / torch.linalg.vector_norm(A0)
        return cls(bonds, [None] * L, AR, T.as_colmajor(A0), 0)

    # -- access ----------------------------------------------------------------
    def AC2(self, pos):
        if self.center == pos:
            A, B = self.AC, self.AR[pos + 1]
        elif self.center == pos + 1:
            A, B = self.AL[pos], self.AC
        else:
            raise ValueError("AC2: center must sit on the bond")
        return T.as_colmajor(torch.einsum("fhm,mjr->fhrj", A, B))

    def bond_dims(self):
        return [b.D for b in self.bonds]


def _block_cols_single(right: U1Bond, c):
    o = right.offsets
    for s, q in enumerate(right.q):
        if q == c:
            return np.arange(o[s], o[s + 1])
    return np.zeros(0, dtype=np.int64)


def _random_right_isometry(left: U1Bond, right: U1Bond, phys_q, g, device):
    """Random site tensor A[l, p, r] whose (l) x (p, r) blocks have orthonormal rows."""
    Dl, d, Dr = left.D, len(phys_q), right.D
    A = torch.zeros((Dl, d, Dr), dtype=torch.float64, device=device)
    Af = T.fuse(A, 1)   # Dl x (d Dr), columns (p, r) with p fastest
    for s, q in enumerate(left.q):
        rows = np.arange(left.offsets[s], left.offsets[s + 1])
        cols = []
        for p, qp in enumerate(phys_q):
            for t, qr in enumerate(right.q):
                if qr == q + qp:
                    cols.append(p + d * np.arange(right.offsets[t], right.offsets[t + 1]))
        if not cols:
            continue
        cols = np.concatenate(cols)
        blk = torch.randn((len(cols), len(rows)), dtype=torch.float64, device=device, generator=g)
        Q, _ = torch.linalg.qr(blk)   # columns orthonormal (rank min)
        R = torch.zeros((len(rows), len(cols)), dtype=torch.float64, device=device)
        R[:Q.shape[1], :] = Q.t()
        Af[torch.as_tensor(rows, device=device)[:, None], torch.as_tensor(cols, device=device)[None, :]] = R
    return T.as_colmajor(T.unfuse(Af, (Dl, d, Dr)))


def u1_split(M, left: U1Bond, right: U1Bond, direction, trunc_dim, phys_q=PHYS_Q, trunc_err=0.0):
    """Truncated split of the fused two-site center M ((Dl d) x (Dr d)) per middle
    charge.  Returns (A, B, new_bond, eps): direction "right" gives A a left
    isometry, "left" gives B a right isometry; the other factor is the center."""
    cs = _middle_charges(left, right, phys_q)
    dev = M.device
    rows = [torch.as_tensor(_block_rows(left, phys_q, c), device=dev) for c in cs]
    cols = [torch.as_tensor(_block_cols(right, phys_q, c), device=dev) for c in cs]
    blocks = [M[r[:, None], cl[None, :]] for r, cl in zip(rows, cols)]
    side = [b if direction == "right" else b.t() for b in blocks]
    keep = [i for i, b in enumerate(side) if b.shape[0] > 0 and b.shape[1] > 0]
    eig = _batched_gram_eigh([side[i] for i in keep])
    lam = [e[0] for e in eig]
    Q = [e[1] for e in eig]
    # global truncation over all blocks
    allv = torch.cat([lam[k][:side[i].shape[0]] for k, i in enumerate(keep)])
    order = torch.argsort(allv, descending=True)
    kmax = min(trunc_dim, int(allv.numel()))
    if trunc_err > 0.0:
        sv = allv[order]
        tail = torch.sqrt(torch.flip(torch.cumsum(torch.flip(sv, [0]), 0), [0]))
        kmax = max(1, min(kmax, int((tail > trunc_err).sum().item())))
    thr = float(allv[order[kmax - 1]]) if kmax > 0 else float("inf")
    eps = float(torch.sqrt(allv[order[kmax:]].sum())) if kmax < allv.numel() else 0.0
    counts = []
    taken = 0
    for k, i in enumerate(keep):
        v = lam[k][:side[i].shape[0]]
        n = int((v >= thr).sum().item()) if thr > 0 else int((v > 0).sum().item())
        counts.append(n)
        taken += n
    # ties at the threshold may overshoot: trim from the smallest-eigenvalue sectors
    while taken > kmax:
        k = min((k for k in range(len(keep)) if counts[k] > 0), key=lambda k: float(lam[k][counts[k] - 1]))
        counts[k] -= 1
        taken -= 1
    new_q, new_d, parts = [], [], []
    for k, i in enumerate(keep):
        n = counts[k]
        if n == 0:
            continue
        new_q.append(cs[i])
        new_d.append(pad8(n))
        parts.append((i, k, n))
    bond = U1Bond(new_q, new_d)
    Dn = bond.D
    offs = bond.offsets
    m, n_ = M.shape
    if direction == "right":
        A = torch.zeros((m, Dn), dtype=torch.float64, device=dev)
        B = torch.zeros((Dn, n_), dtype=torch.float64, device=dev)
        for s, (i, k, n) in enumerate(parts):
            r = rows[i]
            U = Q[k][:r.numel(), :n]
            A[r[:, None], torch.arange(offs[s], offs[s] + n, device=dev)[None, :]] = U
            B[offs[s]:offs[s] + n][:, cols[i]] = U.t() @ blocks[i]
        return A, B, bond, eps
    A = torch.zeros((m, Dn), dtype=torch.float64, device=dev)
    B = torch.zeros((Dn, n_), dtype=torch.float64, device=dev)
    for s, (i, k, n) in enumerate(parts):
        cl = cols[i]
        V = Q[k][:cl.numel(), :n]
        B[torch.arange(offs[s], offs[s] + n, device=dev)[:, None], cl[None, :]] = V.t()
        A[rows[i][:, None], torch.arange(offs[s], offs[s] + n, device=dev)[None, :]] = blocks[i] @ V
    return A, B, bond, eps


def u1_gauge2(psi: U1FiniteMPS, pos, direction, ac2, trunc_dim, trunc_err=0.0, phys_q=PHYS_Q):
    Dl, d1, Dr, d2 = ac2.shape
    M = T.fuse(ac2, 2)
    A, B, bond, eps = u1_split(M, psi.bonds[pos], psi.bonds[pos + 2], direction, trunc_dim, phys_q, trunc_err)
    k = bond.D
    psi.bonds[pos + 1] = bond
    if direction == "right":
        psi.AL[pos] = T.as_colmajor(T.unfuse(A, (Dl, d1, k)))
        c = T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1)
        psi.AC = T.as_colmajor(c / torch.linalg.vector_norm(c))
        psi.AR[pos] = psi.AR[pos + 1] = None
        psi.center = pos + 1
    else:
        psi.AR[pos + 1] = T.as_colmajor(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1))
        c = T.unfuse(A, (Dl, d1, k))
        psi.AC = T.as_colmajor(c / torch.linalg.vector_norm(c))
        psi.AL[pos] = psi.AL[pos + 1] = None
        psi.center = pos
    return eps


def _aligned(b: U1Bond):
    return all(int(o) % 8 == 0 for o in b.offsets)


def find_groundstate_u1(psi: U1FiniteMPS, H, trunc_dim, sweeps=2, precision="fp64", tol=1e-10, krylovdim=20,
                        maxiter=50, use_graph=True, phys_q=PHYS_Q, verbosity=0, storage="compact"):
    """U(1) DMRG2 sweeps (dmrg.jl:286-354) on the charge-block path.  The center
    starts at site 0 (as built by U1FiniteMPS.random).  Returns history
    (per sweep: energy, time, solve time, applies, bond dims at the middle).

    ``storage`` "compact" solves every local problem on the compact vector of
    its charge blocks (``mk_op_compact_charge_blocks``); "dense" keeps the dense
    charge-sorted two-site vector and solves through the batch path
    (hipGraph-captured Krylov steps)."""
    L = psi.L
    envs = FiniteEnvironments(psi, H)
    precs = ("fp64",) if precision == "fp64" else ("fp64", precision)
    # right sweep 0..L-2, left sweep L-2..0: the center is always on the bond
    bonds = [(p, "right") for p in range(L - 1)] + [(p, "left") for p in range(L - 2, -1, -1)]
    hist = []
    for it in range(1, sweeps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ts, nops, flops = 0.0, 0, 0.0
        for pos, direction in bonds:
            GL, GR = envs.leftenv(pos, psi), envs.rightenv(pos + 1, psi)
            op = prepare_operator(JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs)
            lb, rb = psi.bonds[pos], psi.bonds[pos + 2]
            set_charge_blocks(op, lb, rb, phys_q)
            prec = precision if (_aligned(lb) and _aligned(rb)) else "fp64"
            x0 = psi.AC2(pos)
            if storage == "compact":
                lay = compact_charge_blocks(op, x0.shape)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                _, xc, info = op.eigsolve(lay.gather(x0), krylovdim=krylovdim, maxiter=maxiter, tol=tol,
                                          precision=prec)
                torch.cuda.synchronize()
                ts += time.perf_counter() - t1
                nops += info["numops"]
                u1_gauge2(psi, pos, direction, lay.scatter(xc), trunc_dim, 0.0, phys_q)
            else:
                Bt = BatchedDerivative([op])
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                _, vecs, infos = Bt.eigsolve([x0], krylovdim=krylovdim, maxiter=maxiter, tol=tol, precision=prec,
                                             use_graph=use_graph)
                torch.cuda.synchronize()
                ts += time.perf_counter() - t1
                nops += infos[0]["numops"]
                u1_gauge2(psi, pos, direction, vecs[0], trunc_dim, 0.0, phys_q)
                Bt.free()
            op.free()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        E = expectation_value(psi, H, envs)
        mid = psi.bonds[L // 2]
        hist.append(dict(iter=it, energy=E, time=dt, solve_time=ts, numops=nops, mid_bond=(list(mid.q),
                                                                                           list(mid.dims))))
        if verbosity:
            print(f"U1 DMRG2 {it}: E = {E:.15f}  time = {dt:.2f}s (solves {ts:.2f}s)  applies = {nops}  "
                  f"D_mid = {mid.D} over {len(mid.q)} sectors", flush=True)
    return hist
