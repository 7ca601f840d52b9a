"""Learned Davidson preconditioner (``mk_mlp_*``, ``mk_davidson``).

The reference solves every local eigenproblem with KrylovKit's Lanczos
(``fixedpoint``, fixedpoint.jl:15-18).  ``davidson`` solves the same problem
by Davidson iteration; its correction vectors come from the residual (mode
"none"), the diagonal Davidson correction ("diag") or a small network
(``MLPPreconditioner``, mode "mlp") trained on the device on (residual
features, ideal correction) pairs collected from earlier solves.  Convergence
is judged by the float64 residual in every mode, and a solve that does not
converge continues with the library's Lanczos.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import torch

from . import tensors as T
from ._lib import check, load
from .context import default_context

BLOCK = 256
IN, HID, OUT = 2 * BLOCK, 256, BLOCK
MODES = {"none": 0, "diag": 1, "mlp": 2}


def param_count() -> int:
    n = C.c_int64()
    check(load().mk_mlp_param_count(C.byref(n)))
    return n.value


class MLPPreconditioner:
    """3-layer FP32 MLP (512 -> 256 -> 256 -> 256, tanh) on 256-element residual blocks."""

    def __init__(self, seed: int = 0, capacity: int = 200_000):
        self.ctx = default_context()
        h = C.c_void_p()
        check(load().mk_mlp_create(self.ctx.handle, C.c_uint64(seed), int(capacity), C.byref(h)))
        self.handle = h

    # parameters as a flat float32 numpy array [W0 b0 W1 b1 W2 b2], W row-major (out x in)
    def get_params(self) -> np.ndarray:
        p = np.empty(param_count(), dtype=np.float32)
        check(load().mk_mlp_get_params(self.handle, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != param_count():
            raise ValueError("set_params: wrong parameter count")
        check(load().mk_mlp_set_params(self.handle, p.ctypes.data_as(C.c_void_p)))

    @staticmethod
    def _f32(X: torch.Tensor, cols: int) -> torch.Tensor:
        X = X.to(torch.float32).contiguous()
        if not X.is_cuda or X.dim() != 2 or X.shape[1] != cols:
            raise ValueError(f"expected a row-major float32 GPU matrix with {cols} columns")
        return X

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Y (rows x 256) = MLP(X (rows x 512)) for row-major float32 GPU matrices."""
        X = self._f32(X, IN)
        Y = torch.empty((X.shape[0], OUT), dtype=torch.float32, device=X.device)
        self.ctx.sync_stream()
        check(load().mk_mlp_forward(self.handle, C.c_void_p(X.data_ptr()), X.shape[0], C.c_void_p(Y.data_ptr())))
        torch.cuda.synchronize()
        return Y

    def gradient(self, X: torch.Tensor, Tgt: torch.Tensor):
        """(loss, flat gradient) of the mean squared error to Tgt."""
        X, Tgt = self._f32(X, IN), self._f32(Tgt, OUT)
        if Tgt.shape[0] != X.shape[0]:
            raise ValueError("gradient: X and targets have different row counts")
        g = np.empty(param_count(), dtype=np.float32)
        loss = C.c_float()
        self.ctx.sync_stream()
        check(load().mk_mlp_gradient(self.handle, C.c_void_p(X.data_ptr()), C.c_void_p(Tgt.data_ptr()), X.shape[0],
                                     g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return loss.value, g

    @property
    def dataset_size(self) -> int:
        n = C.c_int64()
        check(load().mk_mlp_dataset_size(self.handle, C.byref(n)))
        return n.value

    def train(self, steps: int = 500, batch: int = 2048, lr: float = 1e-3, seed: int = 0, rows=None):
        """Adam on random minibatches of the collected pairs (rows = (begin, end) or all);
        returns (first, last) minibatch loss."""
        r0, r1 = rows if rows is not None else (0, -1)
        a, b = C.c_float(), C.c_float()
        self.ctx.sync_stream()
        check(load().mk_mlp_train(self.handle, int(steps), int(batch), float(lr), C.c_uint64(seed), int(r0), int(r1),
                                  C.byref(a), C.byref(b)))
        return a.value, b.value

    def evaluate(self, rows) -> float:
        """Loss of the current parameters on collected rows [begin, end)."""
        loss = C.c_float()
        self.ctx.sync_stream()
        check(load().mk_mlp_eval(self.handle, int(rows[0]), int(rows[1]), C.byref(loss)))
        return loss.value

    def fit(self, rounds: int = 8, steps: int = 250, batch: int = 2048, lr: float = 1e-4, val_frac: float = 0.2,
            seed: int = 0):
        """Train on the first (1 - val_frac) of the collected rows with early stopping on
        the last val_frac (the rows of the latest solves): the parameters with the lowest
        validation loss are kept.  Returns dict(initial, best, history) of validation losses."""
        n = self.dataset_size
        nv = max(1, int(round(n * val_frac)))
        tr, va = (0, n - nv), (n - nv, n)
        if tr[1] <= 0:
            raise ValueError("fit: not enough collected rows")
        best_p, best = self.get_params(), self.evaluate(va)
        init, hist = best, []
        for k in range(rounds):
            self.train(steps=steps, batch=batch, lr=lr, seed=seed + k, rows=tr)
            v = self.evaluate(va)
            hist.append(v)
            if v < best:
                best, best_p = v, self.get_params()
        self.set_params(best_p)
        return dict(initial=init, best=best, history=hist)

    def select(self, problems, rounds: int = 6, steps: int = 200, batch: int = 2048, lr: float = 1e-4,
               tol: float = 1e-10, maxdim: int = 20, seed: int = 0):
        """Train in rounds on all collected rows and keep the parameters whose Davidson
        solves of the held-out ``problems`` (a list of (prepared operator, start vector))
        need the fewest applies to a float64 residual <= tol; the parameters held
        before training are a candidate too.  Returns dict(initial, best, history) of
        apply counts."""
        def score():
            tot = 0
            for op, x0 in problems:
                _, _, info = davidson(op, x0, maxdim=maxdim, maxiter=400, tol=tol, mlp=self, mode="mlp")
                tot += info["numops"] if info["converged"] else 10 ** 9
            return tot
        best_p, best = self.get_params(), score()
        init, hist = best, []
        for k in range(rounds):
            self.train(steps=steps, batch=batch, lr=lr, seed=seed + k)
            v = score()
            hist.append(v)
            if v < best:
                best, best_p = v, self.get_params()
        self.set_params(best_p)
        return dict(initial=init, best=best, history=hist)

    def free(self) -> None:
        h = getattr(self, "handle", None)
        if sys.is_finalizing():
            return
        if h is not None and h.value:
            load().mk_mlp_free(h)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def davidson(op, x0: torch.Tensor, maxdim: int = 20, maxiter: int = 100, tol: float = 1e-10,
             mlp: MLPPreconditioner | None = None, mode: str = "none", collect: bool = False):
    """Lowest eigenpair of a prepared operator by Davidson iteration.

    Returns (eigenvalue, eigenvector, info) with info = dict(normres, numops, converged).
    """
    if tuple(x0.shape) != op.x_shape:
        raise ValueError("davidson: x0 has the wrong shape")
    x = T.as_colmajor(x0).clone()
    ev, rn, nops, conv = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    op.ctx.sync_stream()
    check(load().mk_davidson(op.handle, T.ptr(x), int(maxdim), int(maxiter), float(tol),
                             mlp.handle if mlp is not None else None, MODES[mode], int(bool(collect)),
                             C.byref(ev), C.byref(rn), C.byref(nops), C.byref(conv)))
    return ev.value, x, dict(normres=rn.value, numops=nops.value, converged=bool(conv.value))
