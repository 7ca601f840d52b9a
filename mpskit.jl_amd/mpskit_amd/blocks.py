"""Block-stored environments and Jordan MPO tensors at the C ABI.

MPSKit's MPOHamiltonian path keeps every environment as a ``BlockTensorMap``
over a ``SumSpace`` on its MPO leg, one dense TensorMap per Jordan level
(src/environments/abstract_envs.jl:32-58), and every MPO tensor as a
``JordanMPOTensor`` holding its nonzero blocks, identity blocks as scalars
(src/operators/jordanmpotensor.jl:44-69).  These helpers pass those layouts to
the library as they are: a list of per-level device tensors (level L of shape
D1 x w_L x D2, channels in order) and a list of MPO blocks.

    prepare_ac2_jordan_levels   mk_prepare_ac2_jordan_levels
    prepare_ac_jordan_levels    mk_prepare_ac_jordan_levels
    prepare_c_levels            mk_prepare_c_levels
    transfer_left_levels        mk_transfer_left_mpo_levels
    transfer_right_levels       mk_transfer_right_mpo_levels
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import tensors as T
from ._lib import EnvLevels, JordanMPO, check, load
from .context import default_context
from .derivatives import PreparedDerivative, _mask


class _Keep:
    """ctypes structures plus everything their pointers refer to."""

    def __init__(self):
        self.refs = []

    def add(self, *objs):
        self.refs.extend(objs)
        return objs[0] if len(objs) == 1 else objs


def split_levels(G, widths):
    """Column-major copies of the channel ranges of a dense D1 x w x D2 tensor."""
    out, o = [], 0
    for w in widths:
        out.append(T.as_colmajor(G[:, o:o + w, :].clone()))
        o += w
    if o != G.shape[1]:
        raise ValueError("split_levels: widths do not cover the MPO leg")
    return out


def env_levels(levels, keep: _Keep) -> EnvLevels:
    levels = [T.as_colmajor(t) for t in levels]
    n = len(levels)
    ptrs = (C.c_void_p * n)(*[T.ptr(t) for t in levels])
    widths = (C.c_int64 * n)(*[int(t.shape[1]) for t in levels])
    keep.add(levels, ptrs, widths)
    s = EnvLevels()
    s.n = n
    s.ptr = C.cast(ptrs, C.POINTER(C.c_void_p))
    s.width = C.cast(widths, C.POINTER(C.c_int64))
    keep.add(s)
    return s


def jordan_mpo(W, keep: _Keep) -> JordanMPO:
    """Block list of a dense Jordan MPO tensor W[g,b,h,k] (host array): every
    nonzero d x d block, with blocks that are multiples of the identity passed
    as scalars (the JordanMPOTensor storage)."""
    W = np.asarray(W, dtype=np.float64)
    wl, d, _, wr = W.shape
    rows, cols, blocks, scalars = [], [], [], []
    for g in range(wl):
        for k in range(wr):
            b = W[g, :, :, k]
            if not np.any(b):
                continue
            rows.append(g)
            cols.append(k)
            if np.array_equal(b, b[0, 0] * np.eye(d)):
                blocks.append(None)
                scalars.append(float(b[0, 0]))
            else:
                t = T.from_numpy(b)
                keep.add(t)
                blocks.append(T.ptr(t))
                scalars.append(0.0)
    n = len(rows)
    ra = (C.c_int * n)(*rows)
    ca = (C.c_int * n)(*cols)
    ba = (C.c_void_p * n)(*blocks)
    sa = (C.c_double * n)(*scalars)
    keep.add(ra, ca, ba, sa)
    s = JordanMPO()
    s.wl, s.d, s.wr, s.nblocks = wl, d, wr, n
    s.row = C.cast(ra, C.POINTER(C.c_int))
    s.col = C.cast(ca, C.POINTER(C.c_int))
    s.block = C.cast(ba, C.POINTER(C.c_void_p))
    s.scalar = C.cast(sa, C.POINTER(C.c_double))
    keep.add(s)
    return s


def prepare_ac2_jordan_levels(GL_levels, W1, W2, GR_levels, precisions=None) -> PreparedDerivative:
    ctx, keep, h = default_context(), _Keep(), C.c_void_p()
    Dl, Dr = int(GL_levels[0].shape[0]), int(GR_levels[0].shape[0])
    d1, d2 = np.asarray(W1).shape[1], np.asarray(W2).shape[1]
    mask = _mask(precisions)
    check(load().mk_prepare_ac2_jordan_levels(ctx.handle, C.byref(env_levels(GL_levels, keep)),
                                              C.byref(jordan_mpo(W1, keep)), C.byref(jordan_mpo(W2, keep)),
                                              C.byref(env_levels(GR_levels, keep)), Dl, Dr, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, (Dl, d1, Dr, d2), (Dl, d1, Dr, d2), "JordanAC2", mask)


def prepare_ac_jordan_levels(GL_levels, W, GR_levels, precisions=None) -> PreparedDerivative:
    ctx, keep, h = default_context(), _Keep(), C.c_void_p()
    Dl, Dr = int(GL_levels[0].shape[0]), int(GR_levels[0].shape[0])
    d = np.asarray(W).shape[1]
    mask = _mask(precisions)
    check(load().mk_prepare_ac_jordan_levels(ctx.handle, C.byref(env_levels(GL_levels, keep)),
                                             C.byref(jordan_mpo(W, keep)), C.byref(env_levels(GR_levels, keep)),
                                             Dl, Dr, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, (Dl, d, Dr), (Dl, d, Dr), "JordanAC", mask)


def prepare_c_levels(GL_levels, GR_levels, precisions=None) -> PreparedDerivative:
    ctx, keep, h = default_context(), _Keep(), C.c_void_p()
    Da, Df = int(GL_levels[0].shape[0]), int(GL_levels[0].shape[2])
    Di, Dc = int(GR_levels[0].shape[0]), int(GR_levels[0].shape[2])
    mask = _mask(precisions)
    check(load().mk_prepare_c_levels(ctx.handle, C.byref(env_levels(GL_levels, keep)),
                                     C.byref(env_levels(GR_levels, keep)), Da, Df, Di, Dc, mask, C.byref(h)))
    return PreparedDerivative(h, ctx, (Df, Di), (Da, Dc), "C", mask)


def transfer_left_levels(x_levels, W, A, Ab, out_widths):
    """Block-stored MPO transfer_left (transfer.jl:125-127): returns the levels of
    y (Da x w_L x Dc), one per entry of ``out_widths``."""
    ctx, keep = default_context(), _Keep()
    A, Ab = T.as_colmajor(A), T.as_colmajor(Ab)
    Dm, Dn = int(x_levels[0].shape[0]), int(x_levels[0].shape[2])
    Dc, Da = int(A.shape[2]), int(Ab.shape[2])
    ys = [T.empty((Da, int(w), Dc), A.device) for w in out_widths]
    check(load().mk_transfer_left_mpo_levels(ctx.handle, C.byref(env_levels(x_levels, keep)),
                                             C.byref(jordan_mpo(W, keep)), T.ptr(A), T.ptr(Ab), Dm, Dn, Dc, Da,
                                             C.byref(env_levels(ys, keep))))
    return ys


def transfer_right_levels(v_levels, W, A, Ab, out_widths):
    """Block-stored MPO transfer_right (transfer.jl:128-130)."""
    ctx, keep = default_context(), _Keep()
    A, Ab = T.as_colmajor(A), T.as_colmajor(Ab)
    Di, Do = int(v_levels[0].shape[0]), int(v_levels[0].shape[2])
    Da, Dc = int(A.shape[0]), int(Ab.shape[0])
    ys = [T.empty((Da, int(w), Dc), A.device) for w in out_widths]
    check(load().mk_transfer_right_mpo_levels(ctx.handle, C.byref(env_levels(v_levels, keep)),
                                              C.byref(jordan_mpo(W, keep)), T.ptr(A), T.ptr(Ab), Da, Di, Dc, Do,
                                              C.byref(env_levels(ys, keep))))
    return ys
