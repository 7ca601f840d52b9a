"""Effective-Hamiltonian (derivative) operators on the HIP path.

Mirrors MPSKit's operator family (src/algorithms/derivatives/):

  DerivativeOperator           derivatives.jl:8-11      ``H(x)``, ``H * x``
  MPO_C_Hamiltonian            mpo_derivatives.jl:23-26
  MPO_AC_Hamiltonian           mpo_derivatives.jl:28-31
  MPO_AC2_Hamiltonian          mpo_derivatives.jl:33-36
  PrecomputedDerivative        mpo_derivatives.jl:168-188 (prepared form)
  JordanMPO_AC_Hamiltonian     hamiltonian_derivatives.jl:15-107
  JordanMPO_AC2_Hamiltonian    hamiltonian_derivatives.jl:168-369
  prepare_operator             prepare_operator!! (derivatives.jl:236-246)
  JordanMPO_AC2_Terms, JordanMPO_AC_Terms
                               constructed (unprepared) Jordan operators,
                               prepared by the one-argument prepare_operator!!

Every operator is prepared once into an immutable ``mk_op`` handle; applying
it enqueues the HIP kernels on torch's current stream.  Inputs/outputs are
column-major float64 GPU tensors in the reference's leg order (see
``tensors``).  ``precision`` selects the bulk arithmetic of an apply:
"fp64" (reference arithmetic), "fp32" (FP32 MFMA) or "bf16x3" (split-BF16
MFMA with FP32 accumulation); inputs and outputs are float64 in every mode.
"""
from __future__ import annotations

import ctypes as C
import sys
from typing import Iterable

import torch

from . import tensors as T
from ._lib import FP64, PREC_NAMES, SpaceMismatch, check, load
from .context import default_context

PrecisionArg = str | int


def _prec(p: PrecisionArg) -> int:
    if isinstance(p, str):
        return PREC_NAMES[p.lower()]
    return int(p)


def _mask(precisions: Iterable[PrecisionArg] | PrecisionArg | None) -> int:
    if precisions is None:
        return 1 << FP64
    if isinstance(precisions, (str, int)):
        precisions = [precisions]
    m = 1 << FP64
    for p in precisions:
        m |= 1 << _prec(p)
    return m


def _req(t: torch.Tensor, shape, name: str) -> torch.Tensor:
    if tuple(t.shape) != tuple(shape):
        raise SpaceMismatch(1, f"{name}: shape {tuple(t.shape)} != expected {tuple(shape)}")
    return T.as_colmajor(t)


class DerivativeOperator:
    """Abstract supertype (derivatives.jl:8-11): callable, and ``H * x == H(x)``."""

    def __call__(self, x, *args, **kw):  # pragma: no cover - abstract
        raise NotImplementedError

    def __mul__(self, x):
        return self(x)


class PreparedDerivative(DerivativeOperator):
    """A prepared operator handle (``PrecomputedDerivative`` / prepared Jordan operator)."""

    def __init__(self, handle: C.c_void_p, ctx, x_shape, y_shape, kind: str, precisions: int):
        self.handle = handle
        self.ctx = ctx
        self.x_shape = tuple(int(n) for n in x_shape)
        self.y_shape = tuple(int(n) for n in y_shape)
        self.kind = kind
        self.precisions = precisions
        xe, ye, fl = C.c_int64(), C.c_int64(), C.c_double()
        check(load().mk_op_info(handle, C.byref(xe), C.byref(ye), C.byref(fl)))
        self.flops = fl.value

    def __call__(self, x: torch.Tensor, precision: PrecisionArg = "fp64", out: torch.Tensor | None = None,
                 alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
        x = _req(x, self.x_shape, "x")
        if out is None:
            out = T.empty(self.y_shape, x.device)
            beta = 0.0
        else:
            if tuple(out.shape) != self.y_shape or not T.is_colmajor(out):
                raise SpaceMismatch(1, "out: wrong shape or layout")
        self.ctx.sync_stream()
        check(load().mk_apply(self.handle, T.ptr(x), T.ptr(out), float(alpha), float(beta), _prec(precision)))
        return out

    def eigsolve(self, x0: torch.Tensor, krylovdim: int = 30, maxiter: int = 100, tol: float = 1e-12,
                 precision: PrecisionArg = "fp64"):
        """Lowest (:SR) eigenpair, ``fixedpoint(H, x0, :SR, Lanczos(...))`` (fixedpoint.jl:15-18).

        Returns (eigenvalue, eigenvector, info) with info = dict(normres, numops, converged).
        """
        x = _req(x0, self.x_shape, "x0").clone()
        self.ctx.sync_stream()
        ev, rn, nops, conv = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        check(load().mk_eigsolve(self.handle, T.ptr(x), int(krylovdim), int(maxiter), float(tol),
                                 _prec(precision), C.byref(ev), C.byref(rn), C.byref(nops), C.byref(conv)))
        return ev.value, x, dict(normres=rn.value, numops=nops.value, converged=bool(conv.value))

    def free(self) -> None:
        h = getattr(self, "handle", None)
        if sys.is_finalizing():      # the HIP runtime may already be torn down
            return
        if h is not None and h.value:
            load().mk_op_free(h)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchedDerivative:
    """Many prepared square operators applied and solved together (``mk_batch_*``).

    The reference applies one operator per call (derivatives.jl:8-11) and runs
    one local eigensolve per ``fixedpoint`` (fixedpoint.jl:15-18); a batch runs
    the applies of all instances as one grouped launch and their Lanczos solves
    in lockstep.  Vectors are packed into one segmented device buffer
    (``pack``/``unpack``); the operators must stay alive while the batch does.
    """

    def __init__(self, ops):
        self.ops = list(ops)
        if not self.ops:
            raise ValueError("BatchedDerivative: no operators")
        for op in self.ops:
            if op.x_shape != op.y_shape:
                raise SpaceMismatch(1, "BatchedDerivative: operators must be square")
        self.ctx = self.ops[0].ctx
        n = len(self.ops)
        arr = (C.c_void_p * n)(*[op.handle for op in self.ops])
        h = C.c_void_p()
        check(load().mk_batch_create(self.ctx.handle, n, arr, C.byref(h)))
        self.handle = h
        offs = (C.c_int64 * n)()
        tot, grp = C.c_int64(), C.c_int()
        check(load().mk_batch_layout(h, offs, C.byref(tot), C.byref(grp)))
        self.offsets = [int(o) for o in offs]
        self.total = int(tot.value)
        self.grouped = bool(grp.value)

    def pack(self, xs) -> torch.Tensor:
        flat = torch.zeros(self.total, dtype=torch.float64, device=xs[0].device)
        for op, o, x in zip(self.ops, self.offsets, xs):
            x = _req(x, op.x_shape, "x")
            n = x.numel()
            flat[o:o + n].copy_(x.permute(*reversed(range(x.dim()))).reshape(-1))
        return flat

    def unpack(self, flat: torch.Tensor):
        out = []
        for op, o in zip(self.ops, self.offsets):
            n = 1
            for d in op.x_shape:
                n *= d
            out.append(T.as_colmajor(flat[o:o + n].reshape(tuple(reversed(op.x_shape)))
                                     .permute(*reversed(range(len(op.x_shape))))))
        return out

    def apply(self, flat: torch.Tensor, precision: PrecisionArg = "fp64") -> torch.Tensor:
        """y = H x for every instance of a packed vector."""
        if flat.numel() != self.total or flat.dtype != torch.float64 or not flat.is_contiguous():
            raise SpaceMismatch(1, "BatchedDerivative.apply: wrong packed vector")
        y = torch.zeros_like(flat)
        self.ctx.sync_stream()
        check(load().mk_batch_apply(self.handle, T.ptr(flat), T.ptr(y), _prec(precision)))
        return y

    def eigsolve(self, xs, krylovdim: int = 30, maxiter: int = 100, tol: float = 1e-12,
                 precision: PrecisionArg = "fp64", use_graph: bool = True):
        """Lowest eigenpairs of all instances; returns (eigenvalues, eigenvectors, infos)."""
        flat = self.pack(xs)
        n = len(self.ops)
        ev, rn = (C.c_double * n)(), (C.c_double * n)()
        nops, conv = (C.c_int * n)(), (C.c_int * n)()
        self.ctx.sync_stream()
        check(load().mk_batch_eigsolve(self.handle, T.ptr(flat), int(krylovdim), int(maxiter), float(tol),
                                       _prec(precision), int(bool(use_graph)), ev, rn, nops, conv))
        infos = [dict(normres=rn[i], numops=nops[i], converged=bool(conv[i])) for i in range(n)]
        return [ev[i] for i in range(n)], self.unpack(flat), infos

    def free(self) -> None:
        h = getattr(self, "handle", None)
        if sys.is_finalizing():
            return
        if h is not None and h.value:
            load().mk_batch_free(h)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def autotune_stats():
    """(decisions, cache hits) of the grouped-GEMM tile autotuner in this process."""
    d, h = C.c_int(), C.c_int()
    check(load().mk_autotune_stats(C.byref(d), C.byref(h)))
    return d.value, h.value


# ---------------------------------------------------------------------------
# Generic MPO derivatives (mpo_derivatives.jl)
# ---------------------------------------------------------------------------

class MPODerivativeOperator(DerivativeOperator):
    """``MPODerivativeOperator`` (mpo_derivatives.jl:10-16): environments plus the
    MPO tensors in between.  Its raw apply prepares on first use and reuses the
    prepared form (the same contraction, associated as L * (x * R))."""

    def __init__(self, leftenv, operators, rightenv):
        self.leftenv = leftenv
        self.operators = tuple(operators)
        self.rightenv = rightenv
        self._prepared = None

    def __call__(self, x, precision: PrecisionArg = "fp64", **kw):
        if self._prepared is None:
            self._prepared = prepare_operator(self)
        return self._prepared(x, precision, **kw)


def MPO_C_Hamiltonian(GL, GR):
    return MPODerivativeOperator(GL, (), GR)


def MPO_AC_Hamiltonian(GL, W, GR):
    return MPODerivativeOperator(GL, (W,), GR)


def MPO_AC2_Hamiltonian(GL, W1, W2, GR):
    return MPODerivativeOperator(GL, (W1, W2), GR)


class JordanMPO_AC_Hamiltonian(DerivativeOperator):
    """``JordanMPO_AC_Hamiltonian`` (hamiltonian_derivatives.jl:15-107): bra == ket."""

    def __init__(self, GL, W, GR):
        self.GL, self.W, self.GR = GL, W, GR
        self._prepared = None

    def __call__(self, x, precision: PrecisionArg = "fp64", **kw):
        if self._prepared is None:
            self._prepared = prepare_operator(self)
        return self._prepared(x, precision, **kw)


class JordanMPO_AC2_Hamiltonian(DerivativeOperator):
    """``JordanMPO_AC2_Hamiltonian`` (hamiltonian_derivatives.jl:168-369): bra == ket."""

    def __init__(self, GL, W1, W2, GR):
        self.GL, self.W1, self.W2, self.GR = GL, W1, W2, GR
        self._prepared = None

    def __call__(self, x, precision: PrecisionArg = "fp64", **kw):
        if self._prepared is None:
            self._prepared = prepare_operator(self)
        return self._prepared(x, precision, **kw)


class JordanMPO_AC2_Terms(DerivativeOperator):
    """A constructed, not yet prepared ``JordanMPO_AC2_Hamiltonian``: its fields
    (hamiltonian_derivatives.jl:168-200) as column-major device tensors, None
    for ``missing``.  Layouts: II[c,i], IC[e,c,j,i], ID[e,j], CB[b,e,h,j],
    CA[b,e,c,h,j,i], AB[a,b,e,f,h,j], BE[a,b,f,h], DE[b,h], EE[a,f]; AA is the
    tuple (GL, W1, W2, GR) of the generic operator over the connected channels.
    ``dims`` = (Dl, d1, Dr, d2).  ``prepare_operator`` folds the terms as the
    reference's one-argument ``prepare_operator!!`` does (:371-468)."""

    KEYS = ("II", "IC", "ID", "CB", "CA", "AB", "BE", "DE", "EE")

    def __init__(self, dims, AA=None, **terms):
        unknown = set(terms) - set(self.KEYS)
        if unknown:
            raise TypeError(f"JordanMPO_AC2_Terms: unknown terms {sorted(unknown)}")
        self.dims = tuple(int(n) for n in dims)
        self.terms = {k: terms.get(k) for k in self.KEYS}
        self.AA = AA
        self._prepared = None

    def __call__(self, x, precision: PrecisionArg = "fp64", **kw):
        if self._prepared is None:
            self._prepared = prepare_operator(self)
        return self._prepared(x, precision, **kw)


class JordanMPO_AC_Terms(DerivativeOperator):
    """A constructed, not yet prepared ``JordanMPO_AC_Hamiltonian`` (fields at
    hamiltonian_derivatives.jl:15-32): D[b,h], I[i,c], E[a,f], C[b,c,h,i],
    B[a,b,f,h], and A = (GL, W, GR) of the generic operator over the middle
    channels.  ``dims`` = (Dl, d, Dr)."""

    KEYS = ("D", "I", "E", "C", "B")

    def __init__(self, dims, A=None, **terms):
        unknown = set(terms) - set(self.KEYS)
        if unknown:
            raise TypeError(f"JordanMPO_AC_Terms: unknown terms {sorted(unknown)}")
        self.dims = tuple(int(n) for n in dims)
        self.terms = {k: terms.get(k) for k in self.KEYS}
        self.A = A
        self._prepared = None

    def __call__(self, x, precision: PrecisionArg = "fp64", **kw):
        if self._prepared is None:
            self._prepared = prepare_operator(self)
        return self._prepared(x, precision, **kw)


def _prepare_terms(H, ctx, mask) -> PreparedDerivative:
    from ._lib import JordanAC2Terms, JordanACTerms
    lib = load()
    h = C.c_void_p()
    keep = []                                  # column-major copies stay alive during the call

    def ptr(t, shape, name):
        if t is None:
            return None
        t = _req(t, shape, name)
        keep.append(t)
        return T.ptr(t)

    if isinstance(H, JordanMPO_AC2_Terms):
        Dl, d1, Dr, d2 = H.dims
        shapes = dict(II=(Dr, Dr), IC=(d2, Dr, d2, Dr), ID=(d2, d2), CB=(d1, d2, d1, d2),
                      CA=(d1, d2, Dr, d1, d2, Dr), AB=(Dl, d1, d2, Dl, d1, d2), BE=(Dl, d1, Dl, d1), DE=(d1, d1),
                      EE=(Dl, Dl))
        s = JordanAC2Terms()
        for k in H.KEYS:
            setattr(s, k, ptr(H.terms[k], shapes[k], k))
        if H.AA is not None:
            GL, W1, W2, GR = H.AA
            wl, wm, wr = GL.shape[1], W1.shape[3], GR.shape[1]
            s.AA_GL = ptr(GL, (Dl, wl, Dl), "AA GL")
            s.AA_W1 = ptr(W1, (wl, d1, d1, wm), "AA W1")
            s.AA_W2 = ptr(W2, (wm, d2, d2, wr), "AA W2")
            s.AA_GR = ptr(GR, (Dr, wr, Dr), "AA GR")
            s.aa_wl, s.aa_wm, s.aa_wr = wl, wm, wr
        check(lib.mk_prepare_ac2_jordan_terms(ctx.handle, C.byref(s), Dl, d1, Dr, d2, mask, C.byref(h)))
        return PreparedDerivative(h, ctx, (Dl, d1, Dr, d2), (Dl, d1, Dr, d2), "JordanAC2", mask)
    Dl, d, Dr = H.dims
    shapes = dict(D=(d, d), I=(Dr, Dr), E=(Dl, Dl), C=(d, Dr, d, Dr), B=(Dl, d, Dl, d))
    s = JordanACTerms()
    for k in H.KEYS:
        setattr(s, k, ptr(H.terms[k], shapes[k], k))
    if H.A is not None:
        GL, W, GR = H.A
        wl, wr = GL.shape[1], GR.shape[1]
        s.A_GL = ptr(GL, (Dl, wl, Dl), "A GL")
        s.A_W = ptr(W, (wl, d, d, wr), "A W")
        s.A_GR = ptr(GR, (Dr, wr, Dr), "A GR")
        s.a_wl, s.a_wr = wl, wr
    check(lib.mk_prepare_ac_jordan_terms(ctx.handle, C.byref(s), Dl, d, Dr, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, (Dl, d, Dr), (Dl, d, Dr), "JordanAC", mask)


def prepare_operator(H, precisions=None) -> PreparedDerivative:
    """``prepare_operator!!`` (derivatives.jl:236-246, mpo_derivatives.jl:216-247,
    hamiltonian_derivatives.jl:109-157, :371-468)."""
    if isinstance(H, PreparedDerivative):
        return H
    lib = load()
    ctx = default_context()
    mask = _mask(precisions)
    h = C.c_void_p()
    if isinstance(H, (JordanMPO_AC2_Terms, JordanMPO_AC_Terms)):
        return _prepare_terms(H, ctx, mask)
    if isinstance(H, MPODerivativeOperator):
        GL, GR = H.leftenv, H.rightenv
        if GL.dim() != 3 or GR.dim() != 3:
            raise SpaceMismatch(1, "environments must be D x w x D tensors")
        Da, wl, Df = GL.shape
        Di, wr, Dc = GR.shape
        GL, GR = T.as_colmajor(GL), T.as_colmajor(GR)
        n = len(H.operators)
        if n == 0:
            if wl != wr:
                raise SpaceMismatch(1, "C derivative: MPO bond of GL and GR differ")
            check(lib.mk_prepare_c(ctx.handle, T.ptr(GL), T.ptr(GR), Da, Df, wl, Di, Dc, mask, C.byref(h)))
            return PreparedDerivative(h, ctx, (Df, Di), (Da, Dc), "C", mask)
        if n == 1:
            W = _req(H.operators[0], (wl, H.operators[0].shape[1], H.operators[0].shape[1], wr), "W")
            d = W.shape[1]
            check(lib.mk_prepare_ac_generic(ctx.handle, T.ptr(GL), T.ptr(W), T.ptr(GR), Da, Df, wl, d, wr, Di, Dc,
                                            mask, C.byref(h)))
            return PreparedDerivative(h, ctx, (Df, d, Di), (Da, d, Dc), "AC", mask)
        if n == 2:
            W1, W2 = H.operators
            wm = W1.shape[3]
            d1, d2 = W1.shape[1], W2.shape[1]
            W1 = _req(W1, (wl, d1, d1, wm), "W1")
            W2 = _req(W2, (wm, d2, d2, wr), "W2")
            check(lib.mk_prepare_ac2_generic(ctx.handle, T.ptr(GL), T.ptr(W1), T.ptr(W2), T.ptr(GR), Da, Df, wl, d1,
                                             wm, d2, wr, Di, Dc, mask, C.byref(h)))
            return PreparedDerivative(h, ctx, (Df, d1, Di, d2), (Da, d1, Dc, d2), "AC2", mask)
        raise ValueError("MPODerivativeOperator with more than two MPO tensors")
    if isinstance(H, JordanMPO_AC_Hamiltonian):
        Dl, wl, _ = H.GL.shape
        Dr, wr, _ = H.GR.shape
        d = H.W.shape[1]
        GL = _req(H.GL, (Dl, wl, Dl), "GL")
        GR = _req(H.GR, (Dr, wr, Dr), "GR")
        W = _req(H.W, (wl, d, d, wr), "W")
        check(lib.mk_prepare_ac_jordan(ctx.handle, T.ptr(GL), T.ptr(W), T.ptr(GR), Dl, wl, d, wr, Dr, mask,
                                       C.byref(h)))
        return PreparedDerivative(h, ctx, (Dl, d, Dr), (Dl, d, Dr), "JordanAC", mask)
    if isinstance(H, JordanMPO_AC2_Hamiltonian):
        Dl, wl, _ = H.GL.shape
        Dr, wr, _ = H.GR.shape
        wm, d1, d2 = H.W1.shape[3], H.W1.shape[1], H.W2.shape[1]
        GL = _req(H.GL, (Dl, wl, Dl), "GL")
        GR = _req(H.GR, (Dr, wr, Dr), "GR")
        W1 = _req(H.W1, (wl, d1, d1, wm), "W1")
        W2 = _req(H.W2, (wm, d2, d2, wr), "W2")
        check(lib.mk_prepare_ac2_jordan(ctx.handle, T.ptr(GL), T.ptr(W1), T.ptr(W2), T.ptr(GR), Dl, wl, d1, wm, d2,
                                        wr, Dr, mask, C.byref(h)))
        return PreparedDerivative(h, ctx, (Dl, d1, Dr, d2), (Dl, d1, Dr, d2), "JordanAC2", mask)
    raise TypeError(f"cannot prepare {type(H).__name__}")


def gemm(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor | None = None, alpha: float = 1.0, beta: float = 0.0,
         transa: bool = False, transb: bool = False, precision: PrecisionArg = "fp64") -> torch.Tensor:
    """Column-major C = alpha op(A) op(B) + beta C through the library's MFMA kernels."""
    A, B = T.as_colmajor(A), T.as_colmajor(B)
    M, K = (A.shape[1], A.shape[0]) if transa else (A.shape[0], A.shape[1])
    K2, N = (B.shape[1], B.shape[0]) if transb else (B.shape[0], B.shape[1])
    if K != K2:
        raise SpaceMismatch(1, f"gemm: inner dimensions {K} != {K2}")
    if C is None:
        C = T.empty((M, N), A.device)
        beta = 0.0
    ctx = default_context()
    check(load().mk_gemm(ctx.handle, int(transa), int(transb), M, N, K, float(alpha), T.ptr(A), max(A.shape[0], 1),
                         T.ptr(B), max(B.shape[0], 1), float(beta), T.ptr(C), max(C.shape[0], 1), _prec(precision)))
    return C
