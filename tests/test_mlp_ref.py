"""The CPU reference of the matrix-core perceptron (oracle/mlp_ref.py),
pinned independently of the GPU: with the BF16 rounding off, its gradients
are the derivatives of its loss (central finite differences in FP64), and
its GELU derivative is the derivative of its GELU. A dropped term, a wrong
transpose or a missing activation derivative fails these."""
import numpy as np

from oracle import mlp_ref


def _loss(Ws, bs, X, T, w):
    Y, _ = mlp_ref.forward(Ws, bs, X, rounding=False)
    return float(np.sum(w[:, None] * (Y - T) ** 2) / Y.size)


def test_gelu_derivative():
    z = np.linspace(-4, 4, 101)
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp_ref.gelu_d(z))) < 1e-8


def test_gradients_are_the_loss_derivatives():
    rng = np.random.default_rng(5)
    dims = [5, 7, 6, 3]
    Ws = [rng.normal(size=(dims[l], dims[l + 1])) / np.sqrt(dims[l]) for l in range(3)]
    bs = [0.1 * rng.normal(size=dims[l + 1]) for l in range(3)]
    X = rng.normal(size=(11, 5))
    T = rng.normal(size=(11, 3))
    w = rng.uniform(0.5, 2.0, 11)
    loss, gW, gb = mlp_ref.gradients(Ws, bs, X, T, w, rounding=False)
    assert abs(loss - _loss(Ws, bs, X, T, w)) < 1e-14
    h = 1e-6
    for l in range(3):
        for idx in [(0, 0), (dims[l] - 1, dims[l + 1] - 1), (dims[l] // 2, 1)]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (_loss(Wp, bs, X, T, w) - _loss(Wm, bs, X, T, w)) / (2 * h)
            assert abs(fd - gW[l][idx]) < 1e-7 * max(1.0, abs(fd)), (l, idx)
        for j in range(dims[l + 1]):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][j] += h
            bm[l][j] -= h
            fd = (_loss(Ws, bp, X, T, w) - _loss(Ws, bm, X, T, w)) / (2 * h)
            assert abs(fd - gb[l][j]) < 1e-7 * max(1.0, abs(fd)), (l, j)
