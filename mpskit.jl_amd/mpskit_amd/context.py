"""Library context: one per (device, host thread), bound to torch's current stream."""
from __future__ import annotations

import ctypes as C
import sys
import threading

import torch

from ._lib import check, load

_tls = threading.local()


class Context:
    """Owns an ``mk_ctx`` (device + stream-ordered scratch arena).

    Calls are enqueued on torch's current stream, so library kernels and torch
    operations on the same stream are ordered.  Contexts are per host thread,
    which is the library's concurrency rule (operators are immutable after
    prepare; the scratch arena is not shared across threads).
    """

    def __init__(self, device: int | None = None):
        lib = load()
        if not torch.cuda.is_available():
            raise RuntimeError("mpskit_amd needs a ROCm GPU (no CPU fallback)")
        self.device = torch.cuda.current_device() if device is None else int(device)
        h = C.c_void_p()
        check(lib.mk_ctx_create(self.device, C.byref(h)))
        self.handle = h
        self._lib = lib

    def sync_stream(self) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(self._lib.mk_ctx_set_stream(self.handle, C.c_void_p(s)))

    def synchronize(self) -> None:
        check(self._lib.mk_ctx_synchronize(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if sys.is_finalizing():      # the HIP runtime may already be torn down
            return
        if h is not None and h.value:
            try:
                self._lib.mk_ctx_destroy(h)
            except Exception:
                pass
            self.handle = None


def default_context() -> Context:
    dev = torch.cuda.current_device()
    ctxs = getattr(_tls, "ctxs", None)
    if ctxs is None:
        ctxs = _tls.ctxs = {}
    if dev not in ctxs:
        ctxs[dev] = Context(dev)
    ctx = ctxs[dev]
    ctx.sync_stream()
    return ctx
