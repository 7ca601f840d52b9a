"""Bond expansion and truncation (changebonds with OptimalExpand / SvdCut) on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  OptimalExpand  src/algorithms/changebonds/optimalexpand.jl:37-117 (Zauner-Stauber
                 et al., PRB 97 (2018)): per bond, the two-site update dAC2(AC AR)
                 projected onto the null spaces of AL and AR, its dominant
                 ``trunc_dim`` singular directions added to the bond
  _expand!       src/algorithms/changebonds/changebonds.jl:59-84: AL -> [AL expansion;
                 0 0], AR -> [AR 0; expansion 0], C -> [C 0; 0 0] (the state is unchanged)
  finite         src/algorithms/changebonds/optimalexpand.jl:119-198 (changebond! :right per
                 bond, left to right): AC -> [AC 0] re-gauged, AR -> [AR; added rows]
  SvdCut         src/algorithms/changebonds/svdcut.jl:21-41,98-120: truncated SVD of every bond
                 matrix C in the mixed-canonical gauge, right to left for a finite
                 MPS; for an infinite MPS AL[i] U, U^T AL[i+1] and a new gauge fix

The two-site update is the HIP library's dAC2 apply: the Jordan operator for an
InfiniteHamiltonian (environments of vumps.InfiniteEnvironments), the generic
one for an InfiniteMPO (boundary.BoundaryEnvironments).  Null spaces and the
small SVD are float64 torch.linalg (rocSOLVER) calls.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import tensors as T
from .derivatives import JordanMPO_AC2_Hamiltonian, MPO_AC2_Hamiltonian, prepare_operator
from .states import FiniteMPS, lq_right, qr_left
from .vumps import InfiniteMPS


@dataclass
class RandExpand:
    """RandExpand(; trunc = truncrank(trunc_dim)) (randexpand.jl:29-35): at most
    ``trunc_dim`` random directions from the null spaces of AL and AR added to
    every bond (no Hamiltonian involved)."""
    trunc_dim: int = 1
    seed: int = 0


@dataclass
class OptimalExpand:
    """OptimalExpand(; trunc = truncrank(trunc_dim)): at most ``trunc_dim``
    directions are added to every bond."""
    trunc_dim: int = 1


def _expand_infinite(psi, ALx, ARx):
    """_expand! (changebonds.jl:59-84): ALx[i] are columns added to AL[i], ARx[i]
    rows added to AR[i]; the new directions carry zero weight in C."""
    N = psi.N
    AL, AR, C = [None] * N, [None] * N, [None] * N
    for i in range(N):
        kl, kr = ALx[i - 1].shape[2], ALx[i].shape[2]     # added to the left / right bond of site i
        Dl, d, Dr = psi.AL[i].shape
        a = T.zeros((Dl + kl, d, Dr + kr))
        a[:Dl, :, :Dr] = psi.AL[i]
        a[:Dl, :, Dr:] = ALx[i]
        AL[i] = a
        b = T.zeros((Dl + kl, d, Dr + kr))
        b[:Dl, :, :Dr] = psi.AR[i]
        b[Dl:, :, :Dr] = ARx[i]
        AR[i] = b
        c = T.zeros((Dr + kr, Dr + kr))
        c[:Dr, :Dr] = psi.C[i]
        C[i] = c
    return InfiniteMPS(AL, AR, C)


def _left_null(AL):
    """Orthonormal complement of the columns of AL as an (l p) x r isometry."""
    M = T.fuse(AL, 2)
    m, r = M.shape
    Q, _ = torch.linalg.qr(M, mode="complete")
    return Q[:, r:]


def _right_null(AR):
    """Orthonormal complement of the rows of AR as an l x (p r) co-isometry, as rows."""
    M = T.fuse(AR, 1)
    l, n = M.shape
    Q, _ = torch.linalg.qr(M.t(), mode="complete")
    return Q[:, l:].t()


def changebonds(psi: InfiniteMPS, H, alg: OptimalExpand, envs):
    """``changebonds(psi, H, OptimalExpand(...), envs)`` for an InfiniteHamiltonian
    (vumps.InfiniteHamiltonian) or an InfiniteMPO (boundary.InfiniteMPO); returns
    the expanded state (environments must be recomputed by the caller)."""
    N = psi.N
    jordan = hasattr(H, "W")
    ALx, ARx = [None] * N, [None] * N          # ALx[i]: added columns of AL[i]; ARx[i]: added rows of AR[i]
    for i in range(N):
        j = (i + 1) % N
        ac2 = T.as_colmajor(torch.einsum("fhm,mjr->fhrj", psi.AC[i], psi.AR[j]))
        if jordan:
            desc = JordanMPO_AC2_Hamiltonian(envs.GLs[i], H.W[i], H.W[j], envs.GRs[j])
        else:
            desc = MPO_AC2_Hamiltonian(envs.GLs[i], H.O[i], H.O[j], envs.GRs[j])
        op = prepare_operator(desc)
        y = op(ac2)
        op.free()
        # rows (f, h), columns (j, r) with j fastest, as the fused AR[j]
        Y = T.fuse(T.as_colmajor(y.permute(0, 1, 3, 2)), 2)
        VL, VR = _left_null(psi.AL[i]), _right_null(psi.AR[j])
        k = min(alg.trunc_dim, VL.shape[1], VR.shape[0])
        Dl, d, Dr = psi.AL[i].shape
        d2, Dr2 = psi.AR[j].shape[1], psi.AR[j].shape[2]
        if k == 0:
            ALx[i] = T.zeros((Dl, d, 0))
            ARx[j] = T.zeros((0, d2, Dr2))
            continue
        X = VL.t() @ Y @ VR.t()
        nrm = float(torch.linalg.matrix_norm(X))
        if nrm > torch.finfo(torch.float64).eps ** 0.75:
            X = X / nrm
        U, _, Vh = torch.linalg.svd(X, full_matrices=False)
        ALx[i] = T.as_colmajor(T.unfuse(VL @ U[:, :k], (Dl, d, k)))
        ARx[j] = T.as_colmajor(T.unfuse(Vh[:k] @ VR, (k, d2, Dr2)))
    return _expand_infinite(psi, ALx, ARx)


@dataclass
class SvdCut:
    """SvdCut(; trunc = truncrank(trunc_dim) & truncerror(trunc_err)): keep at most
    ``trunc_dim`` singular values per bond and drop a tail of 2-norm <= ``trunc_err``."""
    trunc_dim: int = 64
    trunc_err: float = 0.0


def _keep(S, alg: SvdCut):
    k = min(alg.trunc_dim, int(S.shape[0]))
    if alg.trunc_err > 0.0:
        tail = torch.sqrt(torch.flip(torch.cumsum(torch.flip(S * S, [0]), 0), [0]))
        k = max(1, min(k, int((tail > alg.trunc_err).sum().item())))
    return k


def svdcut_finite(psi: FiniteMPS, alg: SvdCut) -> FiniteMPS:
    """changebonds!(psi::FiniteMPS, SvdCut(...)) (svdcut.jl:32-41), in place:
    bonds L-2 .. 0 in turn, each bond matrix from an LQ of the center."""
    for b in range(psi.L - 2, -1, -1):
        psi.move_center(b + 1)
        C, AR = lq_right(psi.AC, householder=True)
        U, S, Vh = torch.linalg.svd(C, full_matrices=False)
        k = _keep(S, alg)
        psi.AR[b + 1] = T.as_colmajor(torch.einsum("km,mpr->kpr", Vh[:k], AR))
        ac = torch.einsum("lpm,mk->lpk", psi.AL[b], U[:, :k] * S[:k])
        psi.AL[b] = None
        psi.AC = T.as_colmajor(ac / torch.linalg.vector_norm(ac))
        psi.center = b
    return psi


def svdcut_infinite(psi: InfiniteMPS, alg: SvdCut) -> InfiniteMPS:
    """changebonds(psi::InfiniteMPS, SvdCut(...)) (svdcut.jl:98-120): truncate each
    bond in the singular basis of its C, then gauge-fix the new tensors."""
    N = psi.N
    A = [a.clone() for a in psi.AL]
    for i in range(N):
        U, S, _ = torch.linalg.svd(psi.C[i], full_matrices=False)
        Uk = U[:, : _keep(S, alg)]
        A[i] = T.as_colmajor(torch.einsum("lpm,mk->lpk", A[i], Uk))
        j = (i + 1) % N
        A[j] = T.as_colmajor(torch.einsum("mk,mpr->kpr", Uk, A[j]))
    return InfiniteMPS.from_tensors(A)


def _copy_finite(psi: FiniteMPS) -> FiniteMPS:
    return FiniteMPS([T.as_colmajor(psi.tensor_at(j).clone()) for j in range(psi.L)])


def _complement_cols(M):
    """Orthonormal basis of the complement of range(M) (numerical rank by SVD)."""
    U, S, _ = torch.linalg.svd(M, full_matrices=True)
    r = int((S > S[0] * M.shape[0] * torch.finfo(torch.float64).eps).sum()) if S.numel() else 0
    return U[:, r:]


def changebonds_finite(psi: FiniteMPS, H, alg, envs):
    """changebonds!(psi::FiniteMPS, H, OptimalExpand(...), envs) (optimalexpand.jl:119-198):
    every bond, left to right, grows by at most ``trunc_dim`` directions taken from
    the two-site update projected onto the complement of the current state; the
    state is unchanged (the new directions carry zero weight).  ``envs`` are the
    FiniteEnvironments of (psi, H); the center ends on the last site.  With
    ``alg`` a RandExpand the directions come from a random update instead (H and
    envs unused)."""
    gen = torch.Generator(device=psi.AC.device).manual_seed(getattr(alg, "seed", 0))
    for i in range(psi.L - 1):
        psi.move_center(i)
        left, right = psi.AC, psi.AR[i + 1]
        Dl, d, D = left.shape
        d2, Dr = right.shape[1], right.shape[2]
        NL = _complement_cols(T.fuse(left, 2))
        NR = _complement_cols(T.fuse(right, 1).t()).t()
        k = min(alg.trunc_dim, NL.shape[1], NR.shape[0])
        if k > 0 and isinstance(alg, RandExpand):
            # randexpand.jl:79-120: a randomised two-site update restricted to the complement
            g2 = torch.randn(NL.shape[1], NR.shape[0], generator=gen, dtype=torch.float64, device=NL.device)
            nrm = float(torch.linalg.matrix_norm(g2))
        elif k > 0:
            op = prepare_operator(JordanMPO_AC2_Hamiltonian(envs.leftenv(i, psi), H[i], H[i + 1],
                                                            envs.rightenv(i + 1, psi)))
            y = op(psi.AC2(i))
            op.free()
            Y = T.fuse(T.as_colmajor(y.permute(0, 1, 3, 2)), 2)
            g2 = NL.t() @ Y @ NR.t()
            nrm = float(torch.linalg.matrix_norm(g2))
            if nrm <= torch.finfo(torch.float64).eps ** 0.75:
                k = 0
        if k == 0:
            AL, C = qr_left(left, householder=True)
            psi.AL[i] = AL
            nxt = torch.einsum("lm,mpr->lpr", C, right)
        else:
            _, _, Vh = torch.linalg.svd(g2 / nrm, full_matrices=False)
            ar_re = T.unfuse(Vh[:k] @ NR, (k, d2, Dr))
            ext = T.zeros((Dl, d, D + k))
            ext[:, :, :D] = left
            AL, C = qr_left(ext, householder=True)
            nar = T.zeros((D + k, d2, Dr))
            nar[:D] = right
            nar[D:] = ar_re
            psi.AL[i] = AL
            nxt = torch.einsum("lm,mpr->lpr", C, nar)
        psi.AR[i + 1] = None
        psi.AC = T.as_colmajor(nxt / torch.linalg.vector_norm(nxt))
        psi.center = i + 1
    return psi


def randexpand_infinite(psi: InfiniteMPS, alg: RandExpand) -> InfiniteMPS:
    """changebonds!(psi::InfiniteMPS, RandExpand(...)) (randexpand.jl:37-57):
    random isometries into the left null space of AL[i] and the right null space
    of AR[i+1]."""
    g = torch.Generator(device=psi.AL[0].device).manual_seed(alg.seed)
    N = psi.N
    ALx, ARx = [None] * N, [None] * N
    for i in range(N):
        j = (i + 1) % N
        VL, VR = _left_null(psi.AL[i]), _right_null(psi.AR[j])
        k = min(alg.trunc_dim, VL.shape[1], VR.shape[0])
        Dl, d, _ = psi.AL[i].shape
        d2, Dr2 = psi.AR[j].shape[1], psi.AR[j].shape[2]
        XL, _ = torch.linalg.qr(torch.randn(VL.shape[1], max(k, 1), generator=g, dtype=torch.float64,
                                            device=VL.device))
        XR, _ = torch.linalg.qr(torch.randn(VR.shape[0], max(k, 1), generator=g, dtype=torch.float64,
                                            device=VR.device))
        ALx[i] = T.as_colmajor(T.unfuse(VL @ XL[:, :k], (Dl, d, k)))
        ARx[j] = T.as_colmajor(T.unfuse(XR[:, :k].t() @ VR, (k, d2, Dr2)))
    return _expand_infinite(psi, ALx, ARx)

