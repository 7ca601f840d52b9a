"""Column-major (Julia / TensorKit layout) float64 device tensors.

A Trivial-symmetry ``TensorMap`` stores one column-major array with axes
``(codomain..., domain...)``.  Here such a tensor is a ``torch.float64`` CUDA
tensor whose *shape* lists the legs in that order and whose strides are
column-major, so ``t[a, g, f]`` indexes like Julia and ``t.data_ptr()`` is the
buffer the C ABI expects.  Torch is used only for device memory and streams.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

DTYPE = torch.float64


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("mpskit_amd needs a ROCm GPU (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def colmajor_strides(shape: Sequence[int]) -> tuple:
    s, out = 1, []
    for n in shape:
        out.append(s)
        s *= max(int(n), 1)
    return tuple(out)


def is_colmajor(t: torch.Tensor) -> bool:
    if t.numel() <= 1:
        return True
    for n, st, want in zip(t.shape, t.stride(), colmajor_strides(t.shape)):
        if n > 1 and st != want:
            return False
    return True


def empty(shape: Sequence[int], dev=None) -> torch.Tensor:
    """Uninitialised column-major float64 tensor on the GPU."""
    shape = tuple(int(n) for n in shape)
    dev = dev or device()
    t = torch.empty(tuple(reversed(shape)), dtype=DTYPE, device=dev)
    return t.permute(*reversed(range(len(shape)))) if len(shape) > 1 else t


def zeros(shape: Sequence[int], dev=None) -> torch.Tensor:
    t = empty(shape, dev)
    t.zero_()
    return t


def as_colmajor(t: torch.Tensor) -> torch.Tensor:
    """Return ``t`` itself if it is a column-major float64 GPU tensor, else a copy."""
    if t.dtype == DTYPE and t.is_cuda and is_colmajor(t):
        return t
    out = empty(t.shape, t.device if t.is_cuda else None)
    out.copy_(t)
    return out


def from_numpy(a: np.ndarray, dev=None) -> torch.Tensor:
    a = np.asarray(a, dtype=np.float64)
    out = empty(a.shape, dev)
    out.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return out


def to_numpy(t: torch.Tensor) -> np.ndarray:
    """Host copy with the same index order (Fortran-ordered array)."""
    return np.asfortranarray(t.detach().cpu().numpy())


def randn(shape: Sequence[int], generator: torch.Generator | None = None, dev=None) -> torch.Tensor:
    out = empty(shape, dev)
    out.copy_(torch.randn(tuple(int(n) for n in shape), dtype=DTYPE, device=out.device, generator=generator))
    return out


def ptr(t: torch.Tensor) -> int:
    if not (t.is_cuda and t.dtype == DTYPE and is_colmajor(t)):
        raise ValueError("expected a column-major float64 GPU tensor")
    return t.data_ptr()


def fuse(t: torch.Tensor, rows: int) -> torch.Tensor:
    """View the column-major tensor as a (prod(first `rows` legs)) x (rest) matrix."""
    m = int(np.prod(t.shape[:rows])) if rows else 1
    n = int(np.prod(t.shape[rows:])) if rows < t.dim() else 1
    return t.permute(*reversed(range(t.dim()))).reshape(n, m).t()


def reshape(t: torch.Tensor, shape: Sequence[int]) -> torch.Tensor:
    """Column-major reshape (legs fused or split with the first index fastest);
    a view of a column-major tensor, else of a column-major copy."""
    t = as_colmajor(t)
    shape = tuple(int(n) for n in shape)
    r = t.permute(*reversed(range(t.dim()))).reshape(tuple(reversed(shape)))
    return r.permute(*reversed(range(len(shape)))) if len(shape) > 1 else r


def unfuse(mat: torch.Tensor, shape: Sequence[int]) -> torch.Tensor:
    """Inverse of :func:`fuse` for a column-major matrix view."""
    shape = tuple(int(n) for n in shape)
    return mat.t().reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))
