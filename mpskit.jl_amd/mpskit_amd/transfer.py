"""Transfer matrices on the HIP path (src/transfermatrix/transfer.jl, transfermatrix.jl).

``transfer_left(v, A, Abar)`` / ``transfer_right(v, A, Abar)`` for bond tensors,
``transfer_left(x, O, A, Abar)`` / ``transfer_right(v, O, A, Abar)`` for MPO
environments, ``TransferMatrix(above, middle, below)`` with ``tm(v)``,
``v * tm`` (flipped, i.e. a left action), products and ``regularize``.

The two functions take every argument pattern of the reference's method
table (transfer.jl:18-161), told apart by tensor rank as there: an excited
(quasiparticle) upper tensor A is (Dl, d, De, Dr), the reference's
A[l p; e r]; legs of v between its first and last pass through; equal-rank
A / Abar with several physical legs (density matrices, MPO pairs) contract
them all.  Braidings are trivial (bosonic spaces).
"""
from __future__ import annotations

import math

from . import tensors as T
from ._lib import SpaceMismatch, check, load
from .context import default_context


def _shape_check(cond, msg):
    if not cond:
        raise SpaceMismatch(1, msg)


def _fuse_phys(A):
    """(Dl, p1, ..., pk, Dr) -> (Dl, p1 ... pk, Dr), column-major."""
    return T.reshape(A, (A.shape[0], math.prod(A.shape[1:-1]), A.shape[-1])) if A.dim() != 3 else A


def transfer_left(v, *args):
    """transfer_left(v, [O,] A, Abar) with the reference's dispatch:

    - v (Dm, Dn), A, Abar MPS tensors: transfer.jl:18-31; several physical
      legs on both (density matrices, MPO pairs :111-113) are contracted;
    - v (Dm, g..., Dn): the legs g pass through (transfer.jl:18-31 with
      N1 > 1, :76-79, :85-91, :101-104);
    - A excited (Dn, d, De, Dc), Abar (Dm, d, Da): transfer.jl:58-60, :66-69;
    - with an MPO O (wl, d, d, wr): x (Dm, wl, Dn) (:125-127), A excited
      (:133-135), x (Dm, wl, De, Dn) with an excitation leg (:141-143), A and
      Abar density-matrix tensors (Dn, d, dq, Dc) with O on the first
      physical leg (:150-155); ``O is None`` is the MPS-only form (:122)."""
    if len(args) == 3:
        O, A, Ab = args
        if O is None:
            return transfer_left(v, A, Ab)
        if A.dim() == 4 and Ab.dim() == 3 and v.dim() == 3:
            return _tl_excited(v, O, A, Ab)
        if A.dim() == 3 and Ab.dim() == 3 and v.dim() == 4:
            return _tl_leg(v, O, A, Ab)
        if A.dim() == 4 and Ab.dim() == 4 and v.dim() == 3:
            return _tl_mpo(v, _mpo_extend(O, 1, A.shape[2]), _fuse_phys(A), _fuse_phys(Ab))
        _shape_check(v.dim() == 3 and A.dim() == 3 and Ab.dim() == 3, "transfer_left(mpo): unsupported ranks")
        return _tl_mpo(v, O, A, Ab)
    A, Ab = args
    if A.dim() == Ab.dim() + 1 == 4:
        _shape_check(v.dim() in (2, 3), "transfer_left(excited): v must be (Dm, Dn) or (Dm, g, Dn)")
        if v.dim() == 2:
            y = _tl_excited(T.reshape(v, (v.shape[0], 1, v.shape[1])), None, A, Ab)
            return T.reshape(y, (y.shape[0], y.shape[2], y.shape[3]))
        return _tl_excited(v, None, A, Ab)
    _shape_check(A.dim() == Ab.dim() and A.dim() >= 3 and v.dim() >= 2, "transfer_left: unsupported ranks")
    A, Ab = _fuse_phys(A), _fuse_phys(Ab)
    if v.dim() == 2:
        return _tl_mps(v, A, Ab)
    g = v.shape[1:-1]
    x = T.reshape(v, (v.shape[0], math.prod(g), 1, v.shape[-1]))
    y = _tl_leg(x, None, A, Ab)
    return T.reshape(y, (y.shape[0], *g, y.shape[-1]))


def transfer_right(v, *args):
    """transfer_right(v, [O,] A, Abar), mirror of :func:`transfer_left`:
    transfer.jl:43-55 (passing legs of v: :80-83, :92-99, :105-108; MPO pairs
    :114-116), excited A (Da, d, De, Di): :61-63, :70-73; with an MPO:
    :128-130, excited A :136-138, v (Di, wr, De, Dj) :144-148, density
    matrices :156-161."""
    if len(args) == 3:
        O, A, Ab = args
        if O is None:
            return transfer_right(v, A, Ab)
        if A.dim() == 4 and Ab.dim() == 3 and v.dim() == 3:
            return _tr_excited(v, O, A, Ab)
        if A.dim() == 3 and Ab.dim() == 3 and v.dim() == 4:
            return _tr_leg(v, O, A, Ab)
        if A.dim() == 4 and Ab.dim() == 4 and v.dim() == 3:
            return _tr_mpo(v, _mpo_extend(O, 1, A.shape[2]), _fuse_phys(A), _fuse_phys(Ab))
        _shape_check(v.dim() == 3 and A.dim() == 3 and Ab.dim() == 3, "transfer_right(mpo): unsupported ranks")
        return _tr_mpo(v, O, A, Ab)
    A, Ab = args
    if A.dim() == Ab.dim() + 1 == 4:
        _shape_check(v.dim() in (2, 3), "transfer_right(excited): v must be (Di, Dj) or (Di, g, Dj)")
        if v.dim() == 2:
            y = _tr_excited(T.reshape(v, (v.shape[0], 1, v.shape[1])), None, A, Ab)
            return T.reshape(y, (y.shape[0], y.shape[2], y.shape[3]))
        return _tr_excited(v, None, A, Ab)
    _shape_check(A.dim() == Ab.dim() and A.dim() >= 3 and v.dim() >= 2, "transfer_right: unsupported ranks")
    A, Ab = _fuse_phys(A), _fuse_phys(Ab)
    if v.dim() == 2:
        return _tr_mps(v, A, Ab)
    g = v.shape[1:-1]
    x = T.reshape(v, (v.shape[0], math.prod(g), 1, v.shape[-1]))
    y = _tr_leg(x, None, A, Ab)
    return T.reshape(y, (y.shape[0], *g, y.shape[-1]))


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


def _mpo_dims(O, x_leg, what):
    """(wl, d, wr) of O, or of the identity on a passenger leg of size x_leg."""
    if O is None:
        return x_leg, None, x_leg
    wl, d, d2, wr = O.shape
    _shape_check(d == d2, what + ": MPO physical spaces differ")
    return wl, d, wr


def _tl_excited(x, O, A, Ab):
    """y[a,b,e,c] = x[m,g,n] A[n,h,e,c] O[g,k,h,b] conj(Ab[m,k,a])
    (``mk_transfer_left_excited``)."""
    Dm, wl, Dn = x.shape
    Dn2, d, De, Dc = A.shape
    Dm2, d2, Da = Ab.shape
    wl2, dO, wr = _mpo_dims(O, wl, "transfer_left(excited)")
    _shape_check(Dn == Dn2 and Dm == Dm2 and d == d2 and wl == wl2 and dO in (None, d),
                 "transfer_left(excited): incompatible spaces")
    x, A, Ab = T.as_colmajor(x), T.as_colmajor(A), T.as_colmajor(Ab)
    O = None if O is None else T.as_colmajor(O)
    y = T.empty((Da, wr, De, Dc), x.device)
    check(load().mk_transfer_left_excited(default_context().handle, T.ptr(x), None if O is None else T.ptr(O),
                                          T.ptr(A), T.ptr(Ab), Dm, wl, Dn, d, De, wr, Dc, Da, T.ptr(y)))
    return y


def _tr_excited(v, O, A, Ab):
    """y[a,b,e,c] = A[a,h,e,i] O[b,k,h,l] conj(Ab[c,k,o]) v[i,l,o]
    (``mk_transfer_right_excited``)."""
    Di, wr, Dj = v.shape
    Da, d, De, Di2 = A.shape
    Db, d2, Dj2 = Ab.shape
    wl, dO, wr2 = _mpo_dims(O, wr, "transfer_right(excited)")
    _shape_check(Di == Di2 and Dj == Dj2 and d == d2 and wr == wr2 and dO in (None, d),
                 "transfer_right(excited): incompatible spaces")
    v, A, Ab = T.as_colmajor(v), T.as_colmajor(A), T.as_colmajor(Ab)
    O = None if O is None else T.as_colmajor(O)
    y = T.empty((Da, wl, De, Db), v.device)
    check(load().mk_transfer_right_excited(default_context().handle, T.ptr(v), None if O is None else T.ptr(O),
                                           T.ptr(A), T.ptr(Ab), Da, d, De, Di, wl, wr, Db, Dj, T.ptr(y)))
    return y


def _tl_leg(x, O, A, Ab):
    """y[a,b,e,c] = x[m,g,e,n] A[n,h,c] O[g,k,h,b] conj(Ab[m,k,a])
    (``mk_transfer_left_leg``)."""
    Dm, wl, De, Dn = x.shape
    Dn2, d, Dc = A.shape
    Dm2, d2, Da = Ab.shape
    wl2, dO, wr = _mpo_dims(O, wl, "transfer_left(leg)")
    _shape_check(Dn == Dn2 and Dm == Dm2 and d == d2 and wl == wl2 and dO in (None, d),
                 "transfer_left(leg): incompatible spaces")
    x, A, Ab = T.as_colmajor(x), T.as_colmajor(A), T.as_colmajor(Ab)
    O = None if O is None else T.as_colmajor(O)
    y = T.empty((Da, wr, De, Dc), x.device)
    check(load().mk_transfer_left_leg(default_context().handle, T.ptr(x), None if O is None else T.ptr(O),
                                      T.ptr(A), T.ptr(Ab), Dm, wl, De, Dn, d, wr, Dc, Da, T.ptr(y)))
    return y


def _tr_leg(v, O, A, Ab):
    """y[a,b,e,c] = A[a,h,i] O[b,k,h,l] conj(Ab[c,k,o]) v[i,l,e,o]
    (``mk_transfer_right_leg``)."""
    Di, wr, De, Dj = v.shape
    Da, d, Di2 = A.shape
    Db, d2, Dj2 = Ab.shape
    wl, dO, wr2 = _mpo_dims(O, wr, "transfer_right(leg)")
    _shape_check(Di == Di2 and Dj == Dj2 and d == d2 and wr == wr2 and dO in (None, d),
                 "transfer_right(leg): incompatible spaces")
    v, A, Ab = T.as_colmajor(v), T.as_colmajor(A), T.as_colmajor(Ab)
    O = None if O is None else T.as_colmajor(O)
    y = T.empty((Da, wl, De, Db), v.device)
    check(load().mk_transfer_right_leg(default_context().handle, T.ptr(v), None if O is None else T.ptr(O),
                                       T.ptr(A), T.ptr(Ab), Da, d, Di, wl, wr, De, Db, Dj, T.ptr(y)))
    return y


def _mpo_extend(O, ne, nq):
    """O widened by identities on a virtual (ne) and a physical (nq) leg
    (``mk_mpo_extend``): out[(w,e),(k,q),(h,r),(v,f)] = O[w,k,h,v] d(e,f) d(q,r)."""
    O = T.as_colmajor(O)
    wl, d, _, wr = O.shape
    out = T.empty((wl * ne, d * nq, d * nq, wr * ne), O.device)
    check(load().mk_mpo_extend(default_context().handle, T.ptr(O), wl, d, wr, ne, nq, T.ptr(out)))
    return out


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
