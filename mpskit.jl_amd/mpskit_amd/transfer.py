"""Transfer matrices on the HIP path (src/transfermatrix/transfer.jl, transfermatrix.jl).

``transfer_left(v, A, Abar)`` / ``transfer_right(v, A, Abar)`` for bond tensors,
``transfer_left(x, O, A, Abar)`` / ``transfer_right(v, O, A, Abar)`` for MPO
environments, ``TransferMatrix(above, middle, below)`` with ``tm(v)``,
``v * tm`` (flipped, i.e. a left action), products and ``regularize``.
"""
from __future__ import annotations

from . import tensors as T
from ._lib import SpaceMismatch, check, load
from .context import default_context


def _shape_check(cond, msg):
    if not cond:
        raise SpaceMismatch(1, msg)


def transfer_left(v, *args):
    """transfer_left(v, A, Abar) (transfer.jl:18-31) or transfer_left(x, O, A, Abar)
    (transfer.jl:125-127); ``O is None`` falls back to the MPS-only form."""
    if len(args) == 3:
        O, A, Ab = args
        if O is None:
            return transfer_left(v, A, Ab)
        return _tl_mpo(v, O, A, Ab)
    A, Ab = args
    return _tl_mps(v, A, Ab)


def transfer_right(v, *args):
    """transfer_right(v, A, Abar) (transfer.jl:43-55) or transfer_right(v, O, A, Abar)
    (transfer.jl:128-130)."""
    if len(args) == 3:
        O, A, Ab = args
        if O is None:
            return transfer_right(v, A, Ab)
        return _tr_mpo(v, O, A, Ab)
    A, Ab = args
    return _tr_mps(v, A, Ab)


def prepare_transfer_left(A, Ab, O=None, precisions=None):
    """The map x -> transfer_left(x, [O,] A, Ab) prepared for repeated applies
    (``mk_prepare_transfer_left``), e.g. the fixed-point solves of the infinite
    environments (infinite_envs.jl:207-346): a ``PreparedDerivative`` whose
    call takes ``precision`` ("fp64", "fp32", "bf16x3").  x is (Dm, Dn) for an
    MPS transfer (O None) and (Dm, wl, Dn) for an MPO transfer."""
    import ctypes as C
    from .derivatives import PreparedDerivative, _mask
    A, Ab = T.as_colmajor(A), T.as_colmajor(Ab)
    Dn, d, Dc = A.shape
    Dm, d2, Da = Ab.shape
    _shape_check(d == d2, "prepare_transfer_left: physical spaces differ")
    if O is None:
        wl = wr = 1
        xs, ys = (Dm, Dn), (Da, Dc)
    else:
        O = T.as_colmajor(O)
        wl, dO, dO2, wr = O.shape
        _shape_check(dO == d == dO2, "prepare_transfer_left: MPO physical spaces differ")
        xs, ys = (Dm, wl, Dn), (Da, wr, Dc)
    ctx, h, mask = default_context(), C.c_void_p(), _mask(precisions)
    check(load().mk_prepare_transfer_left(ctx.handle, None if O is None else T.ptr(O), T.ptr(A), T.ptr(Ab),
                                          Dm, wl, Dn, d, wr, Dc, Da, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, xs, ys, "TransferLeft", mask)


def prepare_transfer_right(A, Ab, O=None, precisions=None):
    """The map v -> transfer_right(v, [O,] A, Ab) prepared for repeated applies
    (``mk_prepare_transfer_right``); v is (Di, Do) for an MPS transfer and
    (Di, wr, Do) for an MPO transfer."""
    import ctypes as C
    from .derivatives import PreparedDerivative, _mask
    A, Ab = T.as_colmajor(A), T.as_colmajor(Ab)
    Da, d, Di = A.shape
    Dc, d2, Do = Ab.shape
    _shape_check(d == d2, "prepare_transfer_right: physical spaces differ")
    if O is None:
        wl = wr = 1
        xs, ys = (Di, Do), (Da, Dc)
    else:
        O = T.as_colmajor(O)
        wl, dO, dO2, wr = O.shape
        _shape_check(dO == d == dO2, "prepare_transfer_right: MPO physical spaces differ")
        xs, ys = (Di, wr, Do), (Da, wl, Dc)
    ctx, h, mask = default_context(), C.c_void_p(), _mask(precisions)
    check(load().mk_prepare_transfer_right(ctx.handle, None if O is None else T.ptr(O), T.ptr(A), T.ptr(Ab),
                                           Da, d, Di, wl, wr, Dc, Do, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, xs, ys, "TransferRight", mask)


def _tl_mps(v, A, Ab):
    Dm, Dn = v.shape
    Dn2, d, Dc = A.shape
    Dm2, d2, Da = Ab.shape
    _shape_check(Dn == Dn2 and Dm == Dm2 and d == d2, "transfer_left: incompatible spaces")
    v, A, Ab = T.as_colmajor(v), T.as_colmajor(A), T.as_colmajor(Ab)
    y = T.empty((Da, Dc), v.device)
    ctx = default_context()
    check(load().mk_transfer_left_mps(ctx.handle, T.ptr(v), T.ptr(A), T.ptr(Ab), Dm, Dn, d, Dc, Da, T.ptr(y)))
    return y


def _tr_mps(v, A, Ab):
    Da, d, Di = A.shape
    Db, d2, Dj = Ab.shape
    _shape_check(tuple(v.shape) == (Di, Dj) and d == d2, "transfer_right: incompatible spaces")
    v, A, Ab = T.as_colmajor(v), T.as_colmajor(A), T.as_colmajor(Ab)
    y = T.empty((Da, Db), v.device)
    ctx = default_context()
    check(load().mk_transfer_right_mps(ctx.handle, T.ptr(v), T.ptr(A), T.ptr(Ab), Da, d, Di, Db, Dj, T.ptr(y)))
    return y


def _tl_mpo(x, O, A, Ab):
    Dm, wl, Dn = x.shape
    wl2, d, _, wr = O.shape
    Dn2, d2, Dc = A.shape
    Dm2, d3, Da = Ab.shape
    _shape_check(wl == wl2 and Dn == Dn2 and Dm == Dm2 and d == d2 == d3, "transfer_left(mpo): incompatible spaces")
    x, O, A, Ab = (T.as_colmajor(t) for t in (x, O, A, Ab))
    y = T.empty((Da, wr, Dc), x.device)
    ctx = default_context()
    check(load().mk_transfer_left_mpo(ctx.handle, T.ptr(x), T.ptr(O), T.ptr(A), T.ptr(Ab), Dm, wl, Dn, d, wr, Dc,
                                      Da, T.ptr(y)))
    return y


def _tr_mpo(v, O, A, Ab):
    Da, d, Di = A.shape
    wl, d2, _, wr = O.shape
    Dc, d3, Do = Ab.shape
    _shape_check(tuple(v.shape) == (Di, wr, Do) and d == d2 == d3, "transfer_right(mpo): incompatible spaces")
    v, O, A, Ab = (T.as_colmajor(t) for t in (v, O, A, Ab))
    y = T.empty((Da, wl, Dc), v.device)
    ctx = default_context()
    check(load().mk_transfer_right_mpo(ctx.handle, T.ptr(v), T.ptr(O), T.ptr(A), T.ptr(Ab), Da, d, Di, wl, wr, Dc,
                                       Do, T.ptr(y)))
    return y


def regularize_(v, lvec, rvec):
    """In-place ``regularize!`` (transfermatrix.jl:70-80); v is D1 x D2 or D1 x w x D2."""
    v3 = v if v.dim() == 3 else v.unsqueeze(1)
    D1, w, D2 = v3.shape
    _shape_check(tuple(lvec.shape) == (D2, D1) and tuple(rvec.shape) == (D1, D2), "regularize: incompatible spaces")
    if not T.is_colmajor(v3):
        raise ValueError("regularize_: v must be column-major")
    lvec, rvec = T.as_colmajor(lvec), T.as_colmajor(rvec)
    ctx = default_context()
    check(load().mk_regularize(ctx.handle, v3.data_ptr(), T.ptr(lvec), T.ptr(rvec), D1, w, D2))
    return v


class AbstractTransferMatrix:
    def __mul__(self, vec):
        return self(vec)

    def __rmul__(self, vec):
        return self.flip()(vec)


class SingleTransferMatrix(AbstractTransferMatrix):
    """``SingleTransferMatrix`` (transfermatrix.jl:4-10)."""

    def __init__(self, above, middle, below, isflipped=False):
        self.above, self.middle, self.below, self.isflipped = above, middle, below, isflipped

    def flip(self):
        return SingleTransferMatrix(self.above, self.middle, self.below, not self.isflipped)

    def __call__(self, vec):
        if self.isflipped:
            return transfer_left(vec, self.middle, self.above, self.below)
        return transfer_right(vec, self.middle, self.above, self.below)


class ProductTransferMatrix(AbstractTransferMatrix):
    """``ProductTransferMatrix`` (transfermatrix.jl:12-31): applied right to left."""

    def __init__(self, tms):
        self.tms = list(tms)

    def flip(self):
        return ProductTransferMatrix([t.flip() for t in reversed(self.tms)])

    def __call__(self, vec):
        for t in reversed(self.tms):
            vec = t(vec)
        return vec


class RegTransferMatrix(AbstractTransferMatrix):
    """``RegTransferMatrix`` (transfermatrix.jl:33-38): project out after each application."""

    def __init__(self, tm, lvec, rvec):
        self.tm, self.lvec, self.rvec = tm, lvec, rvec

    def flip(self):
        return RegTransferMatrix(self.tm.flip(), self.rvec, self.lvec)

    def __call__(self, vec):
        return regularize_(self.tm(vec), self.lvec, self.rvec)


def TransferMatrix(above, middle=None, below=None, isflipped=False):
    """``TransferMatrix(a[, b], c)`` (transfermatrix.jl:57-66); lists build products."""
    if below is None:
        below = above
    if isinstance(above, (list, tuple)):
        mids = middle if isinstance(middle, (list, tuple)) else [middle] * len(above)
        tot = ProductTransferMatrix([SingleTransferMatrix(a, m, b) for a, m, b in zip(above, mids, below)])
        return tot.flip() if isflipped else tot
    return SingleTransferMatrix(above, middle, below, isflipped)


def regularize(tm, lvec, rvec):
    return RegTransferMatrix(tm, lvec, rvec)
