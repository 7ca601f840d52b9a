"""float64 restatement of the learned preconditioner's network (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker of the HIP kernels; the product path never imports it.

The network has no counterpart in the reference: it is this library's
learned Davidson preconditioner.  Layout of the flat parameter vector:
[W0 b0 W1 b1 W2 b2], W_l row-major (out x in), layers 512 -> 256 -> 256 -> 256,
tanh after the first two.  The network output y refines the diagonal
correction q = X[:, 256:]: prediction P = q (1 + y).  Loss: mean over rows of
|T_r| (1 - cos(P_r, T_r)).
"""
import numpy as np

SIZES = [(512, 256), (256, 256), (256, 256)]


def unpack(p):
    out, o = [], 0
    for fi, fo in SIZES:
        W = p[o:o + fi * fo].reshape(fo, fi)
        o += fi * fo
        b = p[o:o + fo]
        o += fo
        out.append((W, b))
    return out


def forward(p, X):
    (W0, b0), (W1, b1), (W2, b2) = unpack(np.asarray(p, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    Y1 = np.tanh(X @ W0.T + b0)
    Y2 = np.tanh(Y1 @ W1.T + b1)
    return Y2 @ W2.T + b2, (Y1, Y2)


def loss_and_grad(p, X, T):
    """(loss, flat gradient) by backpropagation."""
    p = np.asarray(p, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    (W0, b0), (W1, b1), (W2, b2) = unpack(p)
    Y3, (Y1, Y2) = forward(p, X)
    q = X[:, 256:]
    P = q * (1.0 + Y3)
    rows = P.shape[0]
    nP = np.linalg.norm(P, axis=1, keepdims=True)
    nT = np.linalg.norm(T, axis=1, keepdims=True)
    pt = np.sum(P * T, axis=1, keepdims=True)
    ok = (nP > 0) & (nT > 0)
    nPs, nTs = np.where(ok, nP, 1.0), np.where(ok, nT, 1.0)
    loss_r = np.where(ok, nT * (1.0 - pt / (nPs * nTs)), 0.0)
    dP = np.where(ok, -(T - (pt / nPs ** 2) * P) / nPs, 0.0)
    D3 = dP * q / rows
    gW2, gb2 = D3.T @ Y2, D3.sum(0)
    D2 = (D3 @ W2) * (1.0 - Y2 ** 2)
    gW1, gb1 = D2.T @ Y1, D2.sum(0)
    D1 = (D2 @ W1) * (1.0 - Y1 ** 2)
    gW0, gb0 = D1.T @ X, D1.sum(0)
    g = np.concatenate([gW0.ravel(), gb0, gW1.ravel(), gb1, gW2.ravel(), gb2])
    return float(loss_r.sum() / rows), g
