"""The orthogonalisation step shared by the host Krylov solvers.

The Arnoldi, Lanczos and GMRES loops around the library (environment and gauge
fixed points, transfer spectra, ED deflation, TDVP's exponential; KrylovKit's
``ClassicalGramSchmidt2`` in the reference) keep their basis as one
(kmax + 1) x n GPU matrix, one basis vector per row.  Each Krylov step
orthogonalises the new vector against the first k rows with two classical
Gram-Schmidt passes and reads the Hessenberg column back: ``cgs2`` does that
through ``mk_krylov_orthogonalize`` (the library's fused Gram-Schmidt kernels:
three reads of the basis, one host sync).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, load
from .context import default_context


def cgs2(Vb: torch.Tensor, k: int, w: torch.Tensor) -> np.ndarray:
    """w <- w - V^T (c1 + c2) in place for V = Vb[:k] (c1 = V w, c2 = V (w - V^T c1));
    returns the host column [c1 + c2, |w|] of length k + 1.  Up to 32 basis
    vectors through the library, more through GEMVs."""
    if not (Vb.is_cuda and Vb.is_contiguous() and w.is_contiguous() and w.dim() == 1 and k >= 1
            and Vb.dtype == torch.float64 and w.dtype == torch.float64 and Vb.shape[1] == w.numel()):
        raise ValueError("cgs2: needs contiguous float64 GPU tensors of matching length and k >= 1")
    if k > 32:
        # beyond the fused kernels' register tiles: the same passes as GEMVs
        Vj = Vb[:k]
        h1 = Vj @ w
        w -= Vj.t() @ h1
        h2 = Vj @ w
        w -= Vj.t() @ h2
        return torch.cat([h1 + h2, torch.linalg.vector_norm(w)[None]]).cpu().numpy()
    ctx = default_context()
    ctx.sync_stream()
    h = (C.c_double * (k + 1))()
    check(load().mk_krylov_orthogonalize(ctx.handle, C.c_void_p(Vb.data_ptr()), int(Vb.shape[1]), int(k),
                                         C.c_void_p(w.data_ptr()), int(w.numel()), h))
    return np.frombuffer(h, dtype=np.float64, count=k + 1).copy()


def potrf_(G: torch.Tensor) -> tuple[torch.Tensor, int]:
    """Lower Cholesky factor of a symmetric positive definite float64 GPU matrix,
    computed in place in G's memory by the library (``mk_potrf``).  G's storage
    is read as a column-major matrix, which for a symmetric G is G itself.
    Returns (L, info) as ``torch.linalg.cholesky_ex`` does: L a view of G's
    memory (its strict upper triangle zero), info 0 on success, else the 1-based
    column of the first non-positive pivot."""
    n = G.shape[0]
    if not (G.is_cuda and G.dtype == torch.float64 and G.dim() == 2 and G.shape[1] == n and G.is_contiguous()):
        raise ValueError("potrf_: needs a contiguous square float64 GPU matrix")
    ctx = default_context()
    ctx.sync_stream()
    info = C.c_int(0)
    check(load().mk_potrf(ctx.handle, C.c_void_p(G.data_ptr()), int(n), int(max(n, 1)), C.byref(info)))
    # the column-major factor L in row-major storage reads as L^T
    return G.t(), int(info.value)
