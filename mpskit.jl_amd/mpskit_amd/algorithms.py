"""DMRG2 driver on the HIP path (src/algorithms/groundstate/dmrg.jl).

Reproduces the reference's control flow: per bond, build the Jordan two-site
derivative from lazily updated environments, measure the local Galerkin error,
solve the local eigenproblem with an adaptive Krylov budget
(``AdaptiveKrylov``, src/utility/dynamictols.jl:204-235), split the result by a
truncated SVD (``gauge2!``, src/states/orthoview.jl:401-420) and sweep
left-right-left until the sweep-wide error reaches the tolerance or the
truncation error (dmrg.jl:286-354).  Contractions, environment updates and the
Krylov solver are the HIP library's; QR/SVD of the gauge steps are float64 on
the GPU (rocSOLVER via torch.linalg); energies are float64.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import torch

from . import tensors as T
from .derivatives import JordanMPO_AC2_Hamiltonian, JordanMPO_AC_Hamiltonian, prepare_operator
from .environments import FiniteEnvironments
from .states import FiniteMPS, lq_right, pos_qr, qr_left


@dataclass
class AdaptiveKrylov:
    """Per-bond Krylov budget controller (dynamictols.jl:182-235)."""
    krylovdim_min: int = 3
    krylovdim_max: int = 20
    iter_min: int = 1
    iter_max: int = 1
    tol_min: float = 1e-12
    tol_max: float = 1e-2
    truncation_factor: float = 0.1
    tol_factor: float = 0.1

    def adapt_solver(self, decay_rate=0.0, g_local=0.0, g_global=0.0, eps_trunc=0.0):
        """Returns (tol, krylovdim, maxiter) as ``adapt_solver`` (dynamictols.jl:204-235)."""
        tol = min(max(max(self.truncation_factor * eps_trunc, self.tol_factor * g_global), self.tol_min),
                  self.tol_max)
        rho = min(max(decay_rate, 0.0), 1.0)
        if not (0.0 < rho < 1.0):
            if rho == 0.0:
                return tol, self.krylovdim_min, self.iter_min
            return tol, self.krylovdim_max, self.iter_max
        R = max(g_local, self.tol_min) / tol
        nmatvecs = int(round(math.log(R) / math.log(1.0 / rho))) if R > 0 else 0
        krylovdim = min(max(nmatvecs, self.krylovdim_min), self.krylovdim_max)
        maxiter = min(max(-(-5 * nmatvecs // (2 * krylovdim)) - 3, self.iter_min), self.iter_max)
        return tol, krylovdim, maxiter


@dataclass
class Lanczos:
    """Fixed Krylov settings (KrylovKit.Lanczos as used by fixedpoint, fixedpoint.jl:15-18)."""
    tol: float = 1e-10
    krylovdim: int = 30
    maxiter: int = 100

    def adapt_solver(self, **kw):
        return self.tol, self.krylovdim, self.maxiter


@dataclass
class Davidson:
    """Local solver by Davidson iteration (``mk_davidson``) instead of Lanczos.

    ``mode``: "none" (residual corrections), "diag" (diagonal Davidson
    corrections) or "mlp" (corrections from ``mlp``, a trained
    ``MLPPreconditioner``); ``collect`` appends (features, ideal correction)
    pairs of every solve's iterates to ``mlp``'s training set.  Convergence is
    the float64 residual; non-converged solves continue with Lanczos.
    """
    tol: float = 1e-10
    maxdim: int = 20
    maxiter: int = 100
    mode: str = "none"
    mlp: object = None
    collect: bool = False
    stats: list = field(default_factory=list)   # (numops, converged) per solve

    def adapt_solver(self, **kw):
        return self.tol, self.maxdim, self.maxiter


@dataclass
class DMRG2:
    """Two-site DMRG (dmrg.jl:171-209).  ``trunc_dim`` is the maximal bond
    dimension kept by the truncated SVD (``truncrank``) and ``trunc_err`` an
    optional cut on discarded weight; ``precision`` selects the arithmetic of
    the Krylov applies ("fp64", "fp32", "bf16x3"; residuals stay float64)."""
    tol: float = 1e-10
    maxiter: int = 100
    verbosity: int = 1
    alg_eigsolve: object = field(default_factory=AdaptiveKrylov)
    trunc_dim: int = 64
    trunc_err: float = 0.0
    precision: str = "fp64"
    exact_sweeps: bool = False   # run exactly `maxiter` sweeps (no early stop)
    gauge: str = "auto"          # truncated split: "svd", "eigh" or "auto" (see truncated_split)


def project_complement(Y, X):
    """Y - X (X^T Y) for a left isometry X (utility.jl:22)."""
    Xm = T.fuse(X, 2)
    Ym = T.fuse(Y, 2)
    return Ym - Xm @ (Xm.t() @ Ym)


def truncated_split(M, direction, trunc_dim, trunc_err, method="auto"):
    """Rank-k split M ~ A B of a float64 matrix for the two-site gauge step.

    direction "right": A is a left isometry (the new AL), B the new center;
    direction "left":  B is a right isometry (the new AR), A the new center.
    Returns (A, B, s, eps) with s the kept singular values and eps the 2-norm of
    the discarded ones.  method "svd" uses the thin SVD (MatrixAlgebraKit's
    svd_trunc in the reference, orthoview.jl:401-420); "eigh" takes the
    isometry from the eigenvectors of the Gram matrix on the isometry's side
    (M M^T for "right", M^T M for "left"; eigh returns orthonormal vectors to
    working precision) and the center as the exact product with it, so no
    singular value is ever divided by; "auto" uses eigh from 512 rows on, where
    the GPU SVD is tens of times slower than the symmetric eigensolver.

    "auto" also recognises splits that discard nothing (trunc_err = 0 and
    trunc_dim >= min(m, n)): every singular direction is kept, so any isometry
    with the right span gives the same state (the basis inside the bond is a
    gauge choice).  If the isometry's side is the smaller one, the isometry is
    square and the identity serves; otherwise it is the Q of a positive QR (LQ
    for "left") of M.  S is None for these splits and eps = 0.
    """
    m, n = M.shape
    if method == "auto" and trunc_err <= 0.0 and trunc_dim >= min(m, n):
        iso = m if direction == "right" else n
        if iso <= min(m, n):
            I = torch.eye(iso, dtype=M.dtype, device=M.device)
            return (I, M, None, 0.0) if direction == "right" else (M, I, None, 0.0)
        if direction == "right":
            Q, R = pos_qr(M)
            return Q, R, None, 0.0
        Qt, Rt = pos_qr(M.t())
        return Rt.t(), Qt.t(), None, 0.0
    if method == "auto":
        method = "eigh" if min(m, n) >= 512 else "svd"
    if method == "svd":
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
        lam = S * S
    else:
        G = M @ M.t() if direction == "right" else M.t() @ M
        lam, Q = torch.linalg.eigh(0.5 * (G + G.t()))
        lam, Q = torch.flip(lam, [0]).clamp_min(0.0), torch.flip(Q, [1])
        return _split_from_gram(M, direction, lam, Q, trunc_dim, trunc_err)
    k = min(trunc_dim, int(lam.shape[0]))
    if trunc_err > 0.0:
        tail = torch.sqrt(torch.flip(torch.cumsum(torch.flip(lam, [0]), 0), [0]))
        keep = int((tail > trunc_err).sum().item())
        k = max(1, min(k, keep))
    eps = float(torch.sqrt(lam[k:].sum())) if k < lam.shape[0] else 0.0
    U, S, Vh = U[:, :k], S[:k], Vh[:k]
    return (U, S[:, None] * Vh, S, eps) if direction == "right" else (U * S[None, :], Vh, S, eps)


def _split_from_gram(M, direction, lam, Q, trunc_dim, trunc_err):
    """Rank-k split of M from the eigen-decomposition (lam descending, >= 0; Q the
    eigenvectors) of its Gram matrix on the isometry's side: the isometry is
    Q[:, :k] and the center the exact product with it."""
    S = torch.sqrt(lam)
    k = min(trunc_dim, int(lam.shape[0]))
    if trunc_err > 0.0:
        tail = torch.sqrt(torch.flip(torch.cumsum(torch.flip(lam, [0]), 0), [0]))
        keep = int((tail > trunc_err).sum().item())
        k = max(1, min(k, keep))
    eps = float(torch.sqrt(lam[k:].sum())) if k < lam.shape[0] else 0.0
    Q = Q[:, :k].contiguous()
    if direction == "right":
        return Q, Q.t() @ M, S[:k], eps
    return M @ Q, Q.t().contiguous(), S[:k], eps


def _eigh_batch_ms(n, b):
    """Model of one batched float64 eigh of b matrices of order n (ms), fitted to
    rocSOLVER on one MI355X: a per-call cost growing with n (launch-bound
    tridiagonalisation) plus a per-matrix cost."""
    return 2.0 + 0.02 * n + b * 10.5 * (n / 2048.0) ** 2.5


def gram_eigh_grouped(mats):
    """Eigen-decompositions of the Gram matrices M M^T of many matrices: a list
    of (lam, Q) per matrix, eigenvalues descending (clamped at 0) and their
    eigenvectors.  The matrices, sorted by row count, are cut into groups by a
    dynamic programme over the batched-eigh cost model (per-call cost against
    padding waste); every group is one batched float64 eigh over copies padded to
    its largest size.  The padding carries -1 on its diagonal, so its eigenvalues
    sort below every eigenvalue of the (positive semidefinite) Gram matrices and
    the first r eigenvectors of an r-row matrix have no component in its padding."""
    order = sorted(range(len(mats)), key=lambda i: -mats[i].shape[0])
    sz = [mats[i].shape[0] for i in order]
    m = len(order)
    best, cut = [0.0] + [float("inf")] * m, [0] * (m + 1)
    for e in range(1, m + 1):
        for b in range(e):
            c = best[b] + _eigh_batch_ms(sz[b], e - b)
            if c < best[e]:
                best[e], cut[e] = c, b
    groups, e = [], m
    while e > 0:
        groups.append(order[cut[e]:e])
        e = cut[e]
    out = [None] * len(mats)
    for group in groups:
        r0 = mats[group[0]].shape[0]
        G = torch.zeros((len(group), r0, r0), dtype=torch.float64, device=mats[group[0]].device)
        for b, i in enumerate(group):
            r = mats[i].shape[0]
            g = mats[i] @ mats[i].t()
            G[b, :r, :r] = 0.5 * (g + g.t())
            if r < r0:
                G[b, r:, r:].diagonal().fill_(-1.0)
        lam, Q = torch.linalg.eigh(G)
        lam, Q = torch.flip(lam, [1]), torch.flip(Q, [2])
        for b, i in enumerate(group):
            r = mats[i].shape[0]
            out[i] = (lam[b, :r].clamp_min(0.0), Q[b, :r, :r])
    return out


def truncated_split_batched(Ms, direction, trunc_dim, trunc_err):
    """``truncated_split(..., method="eigh")`` of many matrices at once, the Gram
    matrices decomposed by :func:`gram_eigh_grouped`.  ``trunc_dim`` is one bond
    cap for all matrices or a list with one cap per matrix."""
    sides = [M if direction == "right" else M.t() for M in Ms]
    caps = list(trunc_dim) if isinstance(trunc_dim, (list, tuple)) else [trunc_dim] * len(Ms)
    if len(caps) != len(Ms):
        raise ValueError("truncated_split_batched: one trunc_dim per matrix")
    return [_split_from_gram(M, direction, lam, Q, k, trunc_err)
            for M, (lam, Q), k in zip(Ms, gram_eigh_grouped(sides), caps)]


def gauge2(psi: FiniteMPS, pos, direction, ac2, trunc_dim, trunc_err, method="auto", split=None):
    """Truncated split of the two-site center (``gauge2!``, orthoview.jl:401-420)
    with ``normalize = true``; returns the discarded-weight 2-norm.  ``split``:
    a precomputed ``truncated_split`` result of the fused center."""
    Dl, d1, Dr, d2 = ac2.shape
    M = T.fuse(ac2, 2)
    A, B, S, eps = split if split is not None else truncated_split(M, direction, trunc_dim, trunc_err, method)
    k = A.shape[1]
    # rows of B run over (r, p2) with r fastest -> tensor[s, p2, r]
    if direction == "right":
        psi.AL[pos] = T.as_colmajor(T.unfuse(A, (Dl, d1, k)))
        c = T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1)
        psi.AC = T.as_colmajor(c / torch.linalg.vector_norm(c))
        psi.AR[pos] = None
        psi.AR[pos + 1] = None
        psi.center = pos + 1
    else:
        psi.AR[pos + 1] = T.as_colmajor(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1))
        c = T.unfuse(A, (Dl, d1, k))
        psi.AC = T.as_colmajor(c / torch.linalg.vector_norm(c))
        psi.AL[pos] = None
        psi.AL[pos + 1] = None
        psi.center = pos
    return eps


def expectation_value(psi: FiniteMPS, H, envs: FiniteEnvironments):
    """<psi|H|psi> = <AC| dAC |AC> at the center (expval.jl:96-104, mpohamiltonian.jl:922-936)."""
    c = psi.center
    op = prepare_operator(JordanMPO_AC_Hamiltonian(envs.leftenv(c, psi), H[c], envs.rightenv(c, psi)))
    y = op(psi.AC)
    return float(torch.dot(psi.AC.reshape(-1), y.reshape(-1)) / torch.dot(psi.AC.reshape(-1), psi.AC.reshape(-1)))


def local_update(pos, direction, psi, H, alg: DMRG2, envs, eps_global, eps_trunc, decay_rate, stats):
    """``local_update!(::DMRG2)`` (dmrg.jl:211-244)."""
    psi.move_center(pos if direction == "right" else pos + 1)
    GL = envs.leftenv(pos, psi)
    GR = envs.rightenv(pos + 1, psi)
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    Heff = prepare_operator(JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs)
    ac2 = psi.AC2(pos)
    y = Heff(ac2)
    y = y / torch.linalg.vector_norm(y)
    Dl, d1, Dr = psi.AC.shape
    if psi.center == pos and Dl * d1 <= Dr:
        # the left isometry of the center is square: its complement is empty
        eps_local = 0.0
    else:
        # the center of a sweep is numerically rank deficient (Schmidt values
        # below roundoff), where Cholesky QR cannot succeed: Householder directly
        ALp = psi.AL[pos] if psi.center == pos + 1 else qr_left(psi.AC, householder=True)[0]
        eps_local = float(torch.linalg.vector_norm(project_complement(y, ALp)))
    tol, krylovdim, maxiter = alg.alg_eigsolve.adapt_solver(decay_rate=decay_rate, g_local=eps_local,
                                                            g_global=eps_global, eps_trunc=eps_trunc)
    krylovdim = max(2, min(krylovdim, ac2.numel()))
    es = alg.alg_eigsolve
    if isinstance(es, Davidson):
        from .learned import davidson
        _, newac2, info = davidson(Heff, ac2, maxdim=krylovdim, maxiter=maxiter, tol=tol, mlp=es.mlp, mode=es.mode,
                                   collect=es.collect)
        es.stats.append((info["numops"], info["converged"]))
    else:
        _, newac2, info = Heff.eigsolve(ac2, krylovdim=krylovdim, maxiter=maxiter, tol=tol,
                                        precision=alg.precision)
    stats["numops"] += info["numops"] + 1
    eps_trunc = gauge2(psi, pos, direction, newac2, alg.trunc_dim, alg.trunc_err, alg.gauge)
    numops = max(1, info["numops"])
    decay = (info["normres"] / eps_local) ** (1.0 / numops) if eps_local > 0 else 0.0
    decay_rate = min(max(decay, 1e-3), 0.999)
    Heff.free()
    return eps_local, eps_trunc, decay_rate


@dataclass
class DMRG:
    """Single-site DMRG (dmrg.jl:62-170): one-site Jordan dAC eigensolves and a
    QR / LQ gauge step that moves the center (no truncation, no expansion, so
    the bond dimensions of the start state are kept)."""
    tol: float = 1e-10
    maxiter: int = 100
    verbosity: int = 1
    alg_eigsolve: object = field(default_factory=AdaptiveKrylov)
    precision: str = "fp64"
    exact_sweeps: bool = False


def local_update1(pos, direction, psi, H, alg: DMRG, envs, eps_global, decay_rate, stats):
    """``local_update!(::DMRG)`` (dmrg.jl:124-158): Galerkin error of the site,
    dAC eigensolve, gauge with ``normalize = true``."""
    psi.move_center(pos)
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    Heff = prepare_operator(JordanMPO_AC_Hamiltonian(envs.leftenv(pos, psi), H[pos], envs.rightenv(pos, psi)),
                            precisions=precs)
    y = Heff(psi.AC)
    y = y / torch.linalg.vector_norm(y)
    # calc_galerkin(site, ...) (toolbox.jl:63-70): zero when the left isometry of
    # the center is square; else Householder QR (the center of a converging
    # sweep is numerically rank deficient)
    Dl, d1, Dr = psi.AC.shape
    eps_local = 0.0 if Dl * d1 <= Dr else float(
        torch.linalg.vector_norm(project_complement(y, qr_left(psi.AC, householder=True)[0])))
    tol, krylovdim, maxiter = alg.alg_eigsolve.adapt_solver(decay_rate=decay_rate, g_local=eps_local,
                                                            g_global=eps_global, eps_trunc=0.0)
    krylovdim = max(2, min(krylovdim, psi.AC.numel()))
    _, ac, info = Heff.eigsolve(psi.AC, krylovdim=krylovdim, maxiter=maxiter, tol=tol, precision=alg.precision)
    Heff.free()
    stats["numops"] += info["numops"] + 1
    ac = T.as_colmajor(ac)
    if direction == "right":
        AL, C = qr_left(ac, householder=True)
        psi.AL[pos] = AL
        nxt = torch.einsum("lm,mpr->lpr", C, psi.AR[pos + 1])
        psi.AR[pos + 1] = None
        psi.center = pos + 1
    else:
        C, AR = lq_right(ac, householder=True)
        psi.AR[pos] = AR
        nxt = torch.einsum("lpm,mr->lpr", psi.AL[pos - 1], C)
        psi.AL[pos - 1] = None
        psi.center = pos - 1
    psi.AC = T.as_colmajor(nxt / torch.linalg.vector_norm(nxt))
    numops = max(1, info["numops"])
    decay = (info["normres"] / eps_local) ** (1.0 / numops) if eps_local > 0 else 0.0
    return eps_local, min(max(decay, 1e-3), 0.999)


def _find_groundstate_dmrg1(psi: FiniteMPS, H, alg: DMRG, envs: FiniteEnvironments | None):
    L = psi.L
    envs = envs or FiniteEnvironments(psi, H)
    eps_locals = [1.0] * L
    decay_rates = [0.0] * L
    eps_global = 1.0
    history = []
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = {"numops": 0}
        # _sweep_ranges(::DMRG): sites 1..L-1 to the right, L..2 to the left (1-based)
        for pos, direction in [(p, "right") for p in range(L - 1)] + [(p, "left") for p in range(L - 1, 0, -1)]:
            eps_locals[pos], decay_rates[pos] = local_update1(pos, direction, psi, H, alg, envs, eps_global,
                                                              decay_rates[pos], stats)
            eps_global = max(eps_locals)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        E = expectation_value(psi, H, envs)
        history.append(dict(iter=it, energy=E, eps=eps_global, trunc=0.0, time=dt, numops=stats["numops"]))
        if alg.verbosity >= 2:
            print(f"DMRG {it:3d}: E = {E:.15f}  err = {eps_global:.3e}  time = {dt:.3f}s", flush=True)
        if not alg.exact_sweeps and eps_global <= alg.tol:
            break
    return psi, envs, eps_global, history


def find_groundstate(psi: FiniteMPS, H, alg, envs: FiniteEnvironments | None = None):
    """``find_groundstate(psi, H, DMRG2(...) | DMRG(...))`` (dmrg.jl:278-356):
    returns (psi, envs, eps, history)."""
    if isinstance(alg, DMRG):
        return _find_groundstate_dmrg1(psi, H, alg, envs)
    L = psi.L
    envs = envs or FiniteEnvironments(psi, H)
    n = L - 1
    eps_locals = [1.0] * n
    eps_truncs = [0.0] * n
    decay_rates = [0.0] * n
    eps_global = max(eps_locals)
    history = []
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = {"numops": 0}
        for pos in list(range(0, L - 1)):
            eps_locals[pos], eps_truncs[pos], decay_rates[pos] = local_update(
                pos, "right", psi, H, alg, envs, eps_global, eps_truncs[pos], decay_rates[pos], stats)
            eps_global = max(eps_locals)
        for pos in list(range(L - 3, -1, -1)):
            eps_locals[pos], eps_truncs[pos], decay_rates[pos] = local_update(
                pos, "left", psi, H, alg, envs, eps_global, eps_truncs[pos], decay_rates[pos], stats)
            eps_global = max(eps_locals)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        E = expectation_value(psi, H, envs)
        history.append(dict(iter=it, energy=E, eps=eps_global, trunc=max(eps_truncs), time=dt,
                            numops=stats["numops"]))
        if alg.verbosity >= 2:
            print(f"DMRG2 {it:3d}: E = {E:.15f}  err = {eps_global:.3e}  trunc = {max(eps_truncs):.2e}  "
                  f"time = {dt:.3f}s", flush=True)
        if not alg.exact_sweeps and eps_global <= max(alg.tol, max(eps_truncs)):
            break
    return psi, envs, eps_global, history


def find_groundstate_batched(psis, Hs, alg: DMRG2, solver: Lanczos | None = None, use_graph: bool = True,
                             verbosity: int = 0, on_bond=None):
    """DMRG2 sweeps of many chains of one length in lockstep (a parameter sweep).

    Every bond step prepares the two-site Jordan operators of all instances,
    solves their local problems with one batched eigensolve (``BatchedDerivative``:
    one grouped launch per Krylov step for all instances) and splits every
    instance's solution as ``local_update`` does (dmrg.jl:211-244).  ``solver``
    fixes the Krylov settings shared by the batch (default Lanczos(tol=1e-10,
    krylovdim=20, maxiter=100)).  Returns (psis, envs, history); history holds per
    sweep the energies, the wall time and the applies of the slowest instance.
    ``on_bond(sweep, pos, direction, ops, xs)``, if given, sees every bond's
    prepared operators and start vectors before the batched solve (its time is
    excluded from the sweep time).  ``alg.trunc_dim`` may be a list with one
    bond cap per instance (uneven instances of a parameter sweep).
    """
    from .derivatives import BatchedDerivative
    solver = solver or Lanczos(tol=1e-10, krylovdim=20, maxiter=100)
    n = len(psis)
    L = psis[0].L
    if any(p.L != L for p in psis) or len(Hs) != n:
        raise ValueError("find_groundstate_batched: instances must be chains of one length")
    caps = list(alg.trunc_dim) if isinstance(alg.trunc_dim, (list, tuple)) else [alg.trunc_dim] * n
    if len(caps) != n:
        raise ValueError("find_groundstate_batched: one trunc_dim per instance")
    envs = [FiniteEnvironments(p, H) for p, H in zip(psis, Hs)]
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    bonds = [(p, "right") for p in range(L - 1)] + [(p, "left") for p in range(L - 3, -1, -1)]
    history = []
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        numops = 0
        t_solve = 0.0
        for pos, direction in bonds:
            ops, xs = [], []
            for psi, H, env in zip(psis, Hs, envs):
                psi.move_center(pos if direction == "right" else pos + 1)
                GL, GR = env.leftenv(pos, psi), env.rightenv(pos + 1, psi)
                ops.append(prepare_operator(JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs))
                xs.append(psi.AC2(pos))
            if on_bond is not None:
                torch.cuda.synchronize()
                th = time.perf_counter()
                on_bond(it, pos, direction, ops, xs)
                torch.cuda.synchronize()
                t0 += time.perf_counter() - th
            Bt = BatchedDerivative(ops)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            _, vecs, infos = Bt.eigsolve(xs, krylovdim=solver.krylovdim, maxiter=solver.maxiter, tol=solver.tol,
                                         precision=alg.precision, use_graph=use_graph)
            t_solve += time.perf_counter() - ts
            numops += max(i["numops"] for i in infos)
            if alg.gauge == "svd":
                splits = [None] * n
            else:
                splits = truncated_split_batched([T.fuse(v, 2) for v in vecs], direction, caps, alg.trunc_err)
            for psi, v, sp, cap in zip(psis, vecs, splits, caps):
                gauge2(psi, pos, direction, v, cap, alg.trunc_err, alg.gauge, split=sp)
            Bt.free()
            for op in ops:
                op.free()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        energies = [expectation_value(p, H, e) for p, H, e in zip(psis, Hs, envs)]
        history.append(dict(iter=it, energies=energies, time=dt, solve_time=t_solve, numops=numops))
        if verbosity >= 1:
            print(f"batched DMRG2 {it:3d}: {n} instances  time = {dt:.3f}s (solves {t_solve:.3f}s)  "
                  f"applies = {numops}", flush=True)
    return psis, envs, history

