"""The learned preconditioner's float64 oracle (oracle/mlp.py), pinned by
central finite differences of its own loss; a planted backpropagation mistake
(the q factor of the output gradient dropped) fails the same check."""
import numpy as np

from oracle import mlp as O


def _fd_check(loss_and_grad, p, X, T, idx, h=1e-6):
    _, g = loss_and_grad(p, X, T)
    worst = 0.0
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (O.loss_and_grad(p + e, X, T)[0] - O.loss_and_grad(p - e, X, T)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / (abs(fd) + 1e-7))
    return worst


def _case(seed=0, rows=9):
    rng = np.random.default_rng(seed)
    n = sum(a * b + b for a, b in O.SIZES)
    p = rng.standard_normal(n) * 0.05
    X = rng.standard_normal((rows, 512))
    T = rng.standard_normal((rows, 256))
    # every layer, weights and biases
    starts = np.cumsum([0] + [a * b + b for a, b in O.SIZES])
    idx = np.concatenate([rng.choice(np.arange(starts[k], starts[k + 1]), 12, replace=False) for k in range(3)])
    return p, X, T, idx


def test_oracle_gradient_matches_finite_differences():
    p, X, T, idx = _case()
    assert _fd_check(O.loss_and_grad, p, X, T, idx) < 1e-4


def test_planted_backprop_mistake_is_caught():
    p, X, T, idx = _case(seed=1)

    def wrong(p, X, T):
        # forgets the q factor of dloss/dY
        Y3, (Y1, Y2) = O.forward(p, X)
        q = X[:, 256:]
        P = q * (1.0 + Y3)
        rows = P.shape[0]
        nP = np.linalg.norm(P, axis=1, keepdims=True)
        pt = np.sum(P * T, axis=1, keepdims=True)
        D3 = -(T - (pt / nP ** 2) * P) / nP / rows
        (W0, b0), (W1, b1), (W2, b2) = O.unpack(p)
        D2 = (D3 @ W2) * (1 - Y2 ** 2)
        D1 = (D2 @ W1) * (1 - Y1 ** 2)
        g = np.concatenate([(D1.T @ X).ravel(), D1.sum(0), (D2.T @ Y1).ravel(), D2.sum(0), (D3.T @ Y2).ravel(),
                            D3.sum(0)])
        return 0.0, g

    assert _fd_check(wrong, p, X, T, idx) > 1e-2


def test_zero_last_layer_is_the_diagonal_correction():
    p, X, T, _ = _case(seed=2)
    n0 = 512 * 256 + 256 + 256 * 256 + 256
    p[n0:] = 0.0
    Y, _ = O.forward(p, X)
    assert np.all(Y == 0.0)
