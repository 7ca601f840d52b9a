"""Infinite MPS, MPO-Hamiltonian environment fixed points and VUMPS on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  InfiniteMPS gauge fixing   src/states/ortho.jl:113-340 (uniform_leftorth!/rightorth!)
  regauge!                   src/states/ortho.jl:178-184
  environments (Hamiltonian) src/environments/infinite_envs.jl:178-346
  VUMPS                      src/algorithms/groundstate/vumps.jl:56-192
  energy                     src/algorithms/expval.jl:150-162
  Galerkin error             src/algorithms/toolbox.jl:63-79

Every transfer application and every derivative apply / eigensolve is the HIP
library's.  The small Krylov solvers around the transfer matrices (GMRES for
the environment levels, Arnoldi for the gauge fixed points) keep their basis on
the GPU in float64; only the projected (Hessenberg) problems are solved on the
host.  Channel order of the dense MPO must be upper triangular (Jordan form),
which the environment recursion relies on, as the reference's level order does.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import tensors as T
from .krylov import cgs2
from .derivatives import JordanMPO_AC_Hamiltonian, MPO_AC_Hamiltonian, MPO_C_Hamiltonian, prepare_operator
from .states import lq_right, pos_qr, qr_left
from .transfer import prepare_transfer_left, prepare_transfer_right, transfer_left, transfer_right


def _dot(a, b):
    return torch.dot(a.reshape(-1), b.reshape(-1))


def _norm(a):
    return torch.linalg.vector_norm(a)


# ---------------------------------------------------------------------------
# small Krylov solvers over GPU vectors
# ---------------------------------------------------------------------------

def _flat(t):
    """Column-major tensor -> contiguous 1-D view of its storage (copy otherwise)."""
    return T.as_colmajor(t).permute(*reversed(range(t.dim()))).reshape(-1)


def _unflat(f, shape):
    """Inverse of _flat: a column-major view of the 1-D tensor ``f``."""
    return f.reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))


def gmres(apply, b, x0, tol, krylovdim=30, maxiter=50, apply_inner=None, eta=0.0):
    """Solve apply(x) = b (KrylovKit.linsolve with GMRES, used at infinite_envs.jl:236-324).
    Returns (x, info) with info = dict(normres, numops, converged).  The Krylov
    basis is one (krylovdim + 1) x n GPU matrix, orthogonalised by two classical
    Gram-Schmidt passes per step with one readback of the Hessenberg column.

    With ``apply_inner`` (a cheaper, less accurate version of the same map, e.g.
    its FP32 / BF16x3 apply) the Krylov steps use it and every restart
    recomputes the residual with ``apply``: each cycle is one step of iterative
    refinement, the cycle ends once its own residual is below ``eta`` times
    the cycle's starting residual, and convergence is judged on the residual
    of ``apply`` alone."""
    inner = apply if apply_inner is None else apply_inner
    shape = tuple(b.shape)
    bf = _flat(b)
    x = _flat(x0).clone()
    r = bf - _flat(apply(_unflat(x, shape)))
    beta = float(torch.linalg.vector_norm(r))
    ops = 1
    bnorm = max(float(torch.linalg.vector_norm(bf)), 1e-300)
    Vb = torch.empty((krylovdim + 1, bf.numel()), dtype=torch.float64, device=bf.device)
    for _ in range(maxiter):
        if beta <= tol:
            return _unflat(x, shape), dict(normres=beta, numops=ops, converged=True)
        Vb[0] = r / beta
        H = np.zeros((krylovdim + 1, krylovdim))
        y = None
        goal = max(tol, eta * beta)
        for j in range(krylovdim):
            w = _flat(inner(_unflat(Vb[j], shape)))
            ops += 1
            hcol = cgs2(Vb, j + 1, w)
            H[: j + 1, j] = hcol[:-1]
            H[j + 1, j] = hcol[-1]
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y, *_ = np.linalg.lstsq(H[: j + 2, : j + 1], e1, rcond=None)
            res = float(np.linalg.norm(H[: j + 2, : j + 1] @ y - e1))
            if res <= goal or H[j + 1, j] <= 1e-14 * bnorm:
                break
            Vb[j + 1] = w / H[j + 1, j]
        yt = torch.from_numpy(np.ascontiguousarray(y)).to(Vb.device)
        x = x + Vb[: yt.numel()].t() @ yt
        r = bf - _flat(apply(_unflat(x, shape)))
        ops += 1
        beta = float(torch.linalg.vector_norm(r))
    return _unflat(x, shape), dict(normres=beta, numops=ops, converged=beta <= tol)


def arnoldi_dominant(apply, v0, tol=1e-12, krylovdim=20, maxiter=20, relative=False, info=None):
    """Dominant (largest |lambda|) eigenvector of a real map, restarted Arnoldi
    (KrylovKit.eigsolve(..., :LM) as used by fixedpoint in ortho.jl:307-313 and
    infinite_envs.jl:132-134).  Eager: the Ritz pair is tested after every
    Krylov step, so a converged start vector costs one apply.  The residual
    |H[j+1, j] s_j| is absolute, or relative to |lambda| with ``relative``.
    Ritz values within 1e-6 of the largest modulus count as tied; of those the
    one with the largest real part is taken (KrylovKit leaves such ties to its
    sort order), so a transfer map returns its positive fixed point.
    The basis is one (krylovdim + 1) x n GPU matrix; each step orthogonalises
    with two classical Gram-Schmidt passes (GEMVs) and reads H back once."""
    shape = tuple(v0.shape)
    x = _flat(v0)
    n = x.numel()
    Vb = torch.empty((krylovdim + 1, n), dtype=torch.float64, device=x.device)
    Vb[0] = x / torch.linalg.vector_norm(x)
    lam = 0.0
    ops = 0
    res = math.inf

    def done(res, lam, hnext):
        return res <= (tol * max(abs(lam), 1e-300) if relative else tol) or hnext <= 1e-14

    for _ in range(maxiter):
        H = np.zeros((krylovdim + 1, krylovdim))
        for j in range(krylovdim):
            w = _flat(apply(_unflat(Vb[j], shape)))
            ops += 1
            hcol = cgs2(Vb, j + 1, w)
            H[: j + 1, j] = hcol[:-1]
            H[j + 1, j] = hcol[-1]
            ev, evec = np.linalg.eig(H[: j + 1, : j + 1])
            # :LM with ties on the largest modulus (peripheral spectrum of a
            # non-injective MPS, e.g. a complex pair beside the real fixed point)
            # resolved towards the largest real part
            peri = np.flatnonzero(np.abs(ev) >= np.abs(ev).max() * (1 - 1e-6))
            i = int(peri[np.argmax(np.real(ev[peri]))])
            s = evec[:, i]
            s = s * np.exp(-1j * np.angle(s[int(np.argmax(np.abs(s)))]))
            lam = float(np.real(ev[i]))
            res = abs(H[j + 1, j] * s[-1]) / np.linalg.norm(s)
            if done(res, lam, H[j + 1, j]):
                break
            Vb[j + 1] = w / H[j + 1, j]
        sr = torch.from_numpy(np.ascontiguousarray(np.real(s))).to(Vb.device)
        x = Vb[: sr.numel()].t() @ sr
        x = x / torch.linalg.vector_norm(x)
        if done(res, lam, H[j + 1, j]):
            break
        Vb[0] = x
    if info is not None:
        info.update(numops=ops, normres=float(res))
    return lam, _unflat(x, shape)


# ---------------------------------------------------------------------------
# InfiniteMPS
# ---------------------------------------------------------------------------

def _ctimes(C, A):
    return T.as_colmajor(torch.einsum("lm,mpr->lpr", C, A))


def _timesc(A, C):
    return T.as_colmajor(torch.einsum("lpm,mr->lpr", A, C))


def _pos_qr(M):
    return pos_qr(M)


def uniform_leftorth(A, C0, tol=1e-14, maxiter=200, eig_miniter=10):
    """AL[i] C[i] = C[i-1] A[i] (uniform_leftorth!, ortho.jl:215-272).  The error
    of an iteration is the change of C over the QR sweep, after the fixed-point
    step (ortho.jl:241-244)."""
    N = len(A)
    C = [None] * N
    C[-1] = C0 / _norm(C0)
    AL = [None] * N
    eps = math.inf
    for it in range(maxiter):
        if it >= eig_miniter:
            _, vec = arnoldi_dominant(lambda v: _left_cell(v, A, AL), C[-1], tol=max(eps * eps, 1e-15),
                                      krylovdim=30, maxiter=200)
            _, R = _pos_qr(vec)
            C[-1] = R
        C_before = C[-1]
        for i in range(N):
            AL[i], C[i] = qr_left(_ctimes(C[i - 1], A[i]))
        C[-1] = C[-1] / _norm(C[-1])
        eps = float(_norm(C_before - C[-1]))
        if eps < tol:
            break
    return AL, C


def uniform_rightorth(A, C0, tol=1e-14, maxiter=200, eig_miniter=10):
    """C[i-1] AR[i] = A[i] C[i] (uniform_rightorth!, ortho.jl:274-340).  The
    error of an iteration is the change of C over the LQ sweep, after the
    fixed-point step (ortho.jl:308-312)."""
    N = len(A)
    C = [None] * N
    C[-1] = C0 / _norm(C0)
    AR = [None] * N
    eps = math.inf
    for it in range(maxiter):
        if it >= eig_miniter:
            _, vec = arnoldi_dominant(lambda v: _right_cell(v, A, AR), C[-1], tol=max(eps * eps, 1e-15),
                                      krylovdim=30, maxiter=200)
            Qt, Rt = _pos_qr(vec.t())
            C[-1] = Rt.t()
        C_before = C[-1]
        for i in range(N - 1, -1, -1):
            Lm, AR[i] = lq_right(_timesc(A[i], C[i]))
            C[i - 1] = Lm
        C[-1] = C[-1] / _norm(C[-1])
        eps = float(_norm(C_before - C[-1]))
        if eps < tol:
            break
    return AR, C


def _left_cell(v, A, AL):
    for i in range(len(A)):
        v = transfer_left(v, A[i], AL[i])
    return v


def _right_cell(v, A, AR):
    for i in range(len(A) - 1, -1, -1):
        v = transfer_right(v, A[i], AR[i])
    return v


class InfiniteMPS:
    """Uniform MPS with unit cell N: AL, AR, C (bond right of each site), AC = AL C."""

    def __init__(self, AL, AR, C):
        self.AL, self.AR, self.C = AL, AR, C
        self.N = len(AL)
        self.AC = [_timesc(AL[i], C[i]) for i in range(self.N)]

    @classmethod
    def from_tensors(cls, A, tol=1e-14):
        """gaugefix!(psi, A) with order :LR (ortho.jl:113-145)."""
        D = A[0].shape[0]
        C0 = torch.eye(D, dtype=torch.float64, device=A[0].device)
        AL, C = uniform_leftorth(A, C0, tol)
        AR, C = uniform_rightorth(AL, C[-1], tol)
        return cls(AL, AR, C)

    @classmethod
    def random(cls, N, d, D, seed=0):
        rng = np.random.default_rng(seed)
        A = [T.from_numpy(rng.standard_normal((D, d, D))) for _ in range(N)]
        return cls.from_tensors(A)


def regauge(AC, C):
    """AL = Q_AC Q_C^T from positive QRs of AC and C (regauge!, ortho.jl:178-184)."""
    Q_AC, _ = qr_left(AC)
    Q_C, _ = _pos_qr(C)
    return _timesc(Q_AC, Q_C.t())


# ---------------------------------------------------------------------------
# Environments of an infinite Jordan-form MPO Hamiltonian
# ---------------------------------------------------------------------------

class InfiniteHamiltonian:
    """Unit cell of dense Jordan-form MPO tensors W[g,b,h,k] (upper triangular channels)."""

    def __init__(self, Ws):
        self.host = [np.asarray(W, dtype=np.float64) for W in Ws]
        self.W = [T.from_numpy(W) for W in self.host]
        self.N = len(Ws)
        self.w = self.host[0].shape[0]
        for W in self.host:
            for g in range(self.w):
                for k in range(g):
                    if np.any(W[g, :, :, k]):
                        raise ValueError("MPO channels must be upper triangular (Jordan order)")

    def diag_kind(self, i):
        blocks = [W[i, :, :, i] for W in self.host]
        if all(np.array_equal(b, np.eye(b.shape[0])) for b in blocks):
            return "identity"
        if all(not np.any(b) for b in blocks):
            return "empty"
        return "diagonal"


def _chan(G, lo, hi):
    return T.as_colmajor(G[:, lo:hi, :])


# inner-cycle target of the refined GMRES, relative to the cycle's starting
# residual: about the relative error of one FP32 / BF16x3 transfer
_ETA = {"fp32": 1e-5, "bf16x3": 1e-5}


def _fixed_point_solve(apply, b, x0, tol, precision):
    """GMRES on apply(x, "fp64") x = b; in FP32 / BF16x3 the Krylov steps use
    apply(x, precision) and every restart recomputes the FP64 residual."""
    if precision == "fp64":
        return gmres(lambda x: apply(x, "fp64"), b, x0, tol)
    return gmres(lambda x: apply(x, "fp64"), b, x0, tol, maxiter=200,
                 apply_inner=lambda x: apply(x, precision), eta=_ETA[precision])


def _cell_ops(prepare, As, Ws, precision):
    precs = ("fp64",) if precision == "fp64" else ("fp64", precision)
    return [prepare(A, A, O=W, precisions=precs) for A, W in zip(As, Ws)]


class InfiniteEnvironments:
    """GLs[i]: left environment of site i (D x w x D); GRs[i]: right environment of
    site i (bond between i and i+1).  Solved per channel in Jordan order.

    The fixed-point equations of the identity and diagonal channels
    (infinite_envs.jl:207-346) are solved by GMRES over prepared transfer
    matrices of the unit cell; with ``precision`` "fp32" or "bf16x3" their
    Krylov steps run the transfer GEMMs in that precision while every restart
    takes the FP64 residual, so the solution meets the same FP64 tolerance."""

    def __init__(self, psi, H, tol=1e-12, precision="fp64"):
        self.GLs, self.GRs = None, None
        self.tol = tol
        self.precision = precision
        self.recalculate(psi, H)

    def recalculate(self, psi, H, tol=None):
        tol = self.tol if tol is None else tol
        self.GLs = self._left(psi, H, tol, self.precision)
        self.GRs = self._right(psi, H, tol, self.precision)
        return self

    @staticmethod
    def _left(psi, H, tol, precision="fp64"):
        N, w = psi.N, H.w
        D = [psi.AL[i].shape[0] for i in range(N)]
        GL = [T.zeros((D[i], w, D[i])) for i in range(N)]
        GL[0][:, 0, :] = torch.eye(D[0], dtype=torch.float64, device=GL[0].device)

        def cycle(i):
            for s in range(N):
                y = transfer_left(_chan(GL[s], 0, i + 1), T.as_colmajor(H.W[s][: i + 1, :, :, i:i + 1]),
                                  psi.AL[s], psi.AL[s])
                GL[(s + 1) % N][:, i, :] = y[:, 0, :]

        if N > 1:
            cycle(0)
        rho_r = psi.C[N - 1] @ psi.C[N - 1].t()
        for i in range(1, w):
            prev = GL[0][:, i, :].clone()
            GL[0][:, i, :] = 0.0
            cycle(i)
            b = T.as_colmajor(GL[0][:, i, :].clone())
            kind = H.diag_kind(i)
            if kind == "identity":
                I0 = torch.eye(D[0], dtype=torch.float64, device=b.device)
                ops = _cell_ops(prepare_transfer_left, psi.AL, [None] * N, precision)

                def apply(x, p):
                    y = x
                    for op in ops:
                        y = op(y, precision=p)
                    y = y - torch.trace(rho_r @ y) * I0        # regularize!(v, r_LL, l_LL)
                    return x - y

                x, _ = _fixed_point_solve(apply, b, T.as_colmajor(prev), tol, precision)
                for op in ops:
                    op.free()
                GL[0][:, i, :] = x
                if N > 1:
                    cycle(i)
                for s in range(N):
                    r = psi.C[s - 1] @ psi.C[s - 1].t()
                    g = GL[s][:, i, :]
                    GL[s][:, i, :] = g - torch.trace(g @ r) * torch.eye(g.shape[0], dtype=g.dtype, device=g.device)
            elif kind == "diagonal":
                ops = _cell_ops(prepare_transfer_left, psi.AL,
                                [T.as_colmajor(H.W[s][i:i + 1, :, :, i:i + 1]) for s in range(N)], precision)

                def apply(x, p):
                    y = x[:, None, :]
                    for op in ops:
                        y = op(T.as_colmajor(y), precision=p)
                    return x - y[:, 0, :]

                x, _ = _fixed_point_solve(apply, b, T.as_colmajor(prev), tol, precision)
                for op in ops:
                    op.free()
                GL[0][:, i, :] = x
                if N > 1:
                    cycle(i)
            elif N > 1:
                cycle(i)
        return GL

    @staticmethod
    def _right(psi, H, tol, precision="fp64"):
        N, w = psi.N, H.w
        D = [psi.AR[i].shape[2] for i in range(N)]
        GR = [T.zeros((D[i], w, D[i])) for i in range(N)]
        GR[N - 1][:, w - 1, :] = torch.eye(D[N - 1], dtype=torch.float64, device=GR[0].device)

        def cycle(i):
            for s in range(N - 1, -1, -1):
                y = transfer_right(_chan(GR[s], i, w), T.as_colmajor(H.W[s][i:i + 1, :, :, i:]),
                                   psi.AR[s], psi.AR[s])
                GR[s - 1][:, i, :] = y[:, 0, :]

        if N > 1:
            cycle(w - 1)
        rho_l = psi.C[N - 1].t() @ psi.C[N - 1]
        for i in range(w - 2, -1, -1):
            prev = GR[N - 1][:, i, :].clone()
            GR[N - 1][:, i, :] = 0.0
            cycle(i)
            b = T.as_colmajor(GR[N - 1][:, i, :].clone())
            kind = H.diag_kind(i)
            if kind == "identity":
                I0 = torch.eye(D[N - 1], dtype=torch.float64, device=b.device)
                ops = _cell_ops(prepare_transfer_right, psi.AR[::-1], [None] * N, precision)

                def apply(x, p):
                    y = x
                    for op in ops:
                        y = op(y, precision=p)
                    y = y - torch.trace(rho_l @ y) * I0        # regularize!(v, l_RR, r_RR)
                    return x - y

                x, _ = _fixed_point_solve(apply, b, T.as_colmajor(prev), tol, precision)
                for op in ops:
                    op.free()
                GR[N - 1][:, i, :] = x
                if N > 1:
                    cycle(i)
                for s in range(N):
                    l = psi.C[s].t() @ psi.C[s]
                    g = GR[s][:, i, :]
                    GR[s][:, i, :] = g - torch.trace(g @ l) * torch.eye(g.shape[0], dtype=g.dtype, device=g.device)
            elif kind == "diagonal":
                ops = _cell_ops(prepare_transfer_right, psi.AR[::-1],
                                [T.as_colmajor(H.W[s][i:i + 1, :, :, i:i + 1]) for s in range(N - 1, -1, -1)],
                                precision)

                def apply(x, p):
                    y = x[:, None, :]
                    for op in ops:
                        y = op(T.as_colmajor(y), precision=p)
                    return x - y[:, 0, :]

                x, _ = _fixed_point_solve(apply, b, T.as_colmajor(prev), tol, precision)
                for op in ops:
                    op.free()
                GR[N - 1][:, i, :] = x
                if N > 1:
                    cycle(i)
            elif N > 1:
                cycle(i)
        return GR


def energy_per_site(psi, H, envs):
    """sum over the cell of <AC| GL W[:, end] GR[end] |AC> (expval.jl:150-162) / N."""
    e = 0.0
    for s in range(psi.N):
        W_end = T.as_colmajor(H.W[s][:, :, :, -1:])
        GR_end = T.as_colmajor(envs.GRs[s][:, -1:, :])
        op = prepare_operator(MPO_AC_Hamiltonian(envs.GLs[s], W_end, GR_end))
        e += float(_dot(psi.AC[s], op(psi.AC[s])))
        op.free()
    return e / psi.N


def calc_galerkin(psi, H, envs):
    """max over sites of |(1 - AL AL^T) normalize(dAC(AC))| (toolbox.jl:63-79)."""
    worst = 0.0
    for s in range(psi.N):
        op = prepare_operator(JordanMPO_AC_Hamiltonian(envs.GLs[s], H.W[s], envs.GRs[s]))
        y = op(psi.AC[s])
        y = y / _norm(y)
        Xm, Ym = T.fuse(psi.AL[s], 2), T.fuse(y, 2)
        worst = max(worst, float(_norm(Ym - Xm @ (Xm.t() @ Ym))))
        op.free()
    return worst


@dataclass
class VUMPS:
    """VUMPS (vumps.jl:23-55); ``precision`` applies to the AC / C eigensolver
    applies and to the transfer GEMMs of the environment fixed-point solves
    (both refined to their FP64 tolerances).

    The inner solvers take the reference's dynamic tolerances
    (utility/defaults.jl:23-33, utility/dynamictols.jl:86-89): at iteration
    ``it`` with Galerkin error eps each uses clamp(factor * eps / sqrt(it),
    tol_min, tol_max) with factor 1e-3 for the eigensolves, 1e-6 for the
    gauge fixing and 1e-4 for the environments; the eigensolver is Lanczos
    with krylovdim 30 and up to 200 restarts."""
    tol: float = 1e-10
    maxiter: int = 100
    verbosity: int = 1
    precision: str = "fp64"
    tol_factor: float = 1e-3
    gauge_tol_factor: float = 1e-6
    envs_tol_factor: float = 1e-4
    tol_min: float = 1e-14
    tol_max: float = 1e-4
    krylovdim: int = 30
    eig_maxiter: int = 200

    def dynamic_tol(self, factor, it, eps):
        return min(max(factor * eps / math.sqrt(it), self.tol_min), self.tol_max)


def find_groundstate_vumps(psi: InfiniteMPS, H: InfiniteHamiltonian, alg: VUMPS):
    """``find_groundstate(psi::InfiniteMPS, H, VUMPS())`` (vumps.jl:56-125)."""
    envs = InfiniteEnvironments(psi, H, tol=1e-10, precision=alg.precision)
    eps = calc_galerkin(psi, H, envs)
    history = []
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tol = alg.dynamic_tol(alg.tol_factor, it, eps)
        new_AL = []
        for s in range(psi.N):
            # localupdate (_localupdate_vumps_step!, vumps.jl:156-173)
            Hac = prepare_operator(JordanMPO_AC_Hamiltonian(envs.GLs[s], H.W[s], envs.GRs[s]), precisions=precs)
            _, AC, _ = Hac.eigsolve(psi.AC[s], krylovdim=alg.krylovdim, maxiter=alg.eig_maxiter, tol=tol,
                                    precision=alg.precision)
            Hac.free()
            Hc = prepare_operator(MPO_C_Hamiltonian(envs.GLs[(s + 1) % psi.N], envs.GRs[s]), precisions=precs)
            _, C, _ = Hc.eigsolve(psi.C[s], krylovdim=alg.krylovdim, maxiter=alg.eig_maxiter, tol=tol,
                                  precision=alg.precision)
            Hc.free()
            new_AL.append(regauge(AC, C))
        # gauge_step! (vumps.jl:175-183): right-orthogonalise the new AL, AC = AL C
        AR, C = uniform_rightorth(new_AL, psi.C[-1], tol=alg.dynamic_tol(alg.gauge_tol_factor, it, eps))
        psi = InfiniteMPS(new_AL, AR, C)
        envs.recalculate(psi, H, tol=alg.dynamic_tol(alg.envs_tol_factor, it, eps))
        eps = calc_galerkin(psi, H, envs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        e = energy_per_site(psi, H, envs)
        history.append(dict(iter=it, energy=e, eps=eps, time=dt))
        if alg.verbosity >= 2:
            print(f"VUMPS {it:3d}: e = {e:.15f}  err = {eps:.3e}  time = {dt:.3f}s", flush=True)
        if eps <= alg.tol:
            break
    return psi, envs, eps, history
