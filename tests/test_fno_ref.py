"""The CPU reference of the Fourier neural operator (oracle/fno_ref.py),
pinned by what the mathematics fixes: its spectral part (the kernels'
transform matrices and mode layout) equals numpy's two-dimensional FFT
truncated to the kept modes, and its gradients are the derivatives of its
loss (central differences)."""
import numpy as np

from oracle import fno_ref


def test_spectral_part_is_the_truncated_fft():
    rng = np.random.default_rng(3)
    m, w = 4, 3
    hv = rng.normal(size=(2, 64, 64, w))
    Wm = np.zeros((2 * m, m, 2 * w, 2 * w))
    Wm[:, :] = np.eye(2 * w)
    # the transform tables are FP32, as the kernels hold them: agreement to their rounding
    S, _ = fno_ref.spectral(hv, Wm, m, rounding=False)
    want = fno_ref.spectral_fft(hv, m)
    assert np.max(np.abs(S - want)) < 1e-6 * np.max(np.abs(want))
    # a single mode passes through with its amplitude and phase
    x = np.arange(64)
    one = np.cos(2 * np.pi * (3 * x[None, :] - 2 * x[:, None]) / 64 + 0.4)[None, :, :, None] * np.ones((1, 1, 1, w))
    S1, _ = fno_ref.spectral(one, Wm, m, rounding=False)
    assert np.max(np.abs(S1 - one)) < 1e-6
    # a mode outside the kept set is removed
    out = np.cos(2 * np.pi * 5 * x / 64)[None, None, :, None] * np.ones((1, 64, 1, w))
    S2, _ = fno_ref.spectral(out, Wm, m, rounding=False)
    assert np.max(np.abs(S2)) < 1e-6


def _setup(seed=5):
    rng = np.random.default_rng(seed)
    cin, cout, w, m, L, ph = 3, 4, 4, 2, 2, 4
    lay = fno_ref.layout(cin, cout, w, m, L, ph)
    Pv = 0.5 * rng.normal(size=lay["_size"])
    X = rng.normal(size=(2, 4096, cin))
    E = rng.normal(size=(2, 4096, cout))
    R = rng.normal(size=(2, 4096, cout))
    wb = np.array([0.7, 1.3])
    return lay, Pv, X, E, R, wb, (cin, cout, w, m, L, ph)


def _loss(Pv, X, E, R, wb, dims):
    out, _ = fno_ref.forward(Pv, X, *dims, rounding=False)
    return fno_ref.loss_and_dout(out, E, R, wb)[0]


def test_gradients_are_the_loss_derivatives():
    lay, Pv, X, E, R, wb, dims = _setup()
    loss, G = fno_ref.gradients(Pv, X, E, R, wb, *dims, rounding=False)
    assert abs(loss - _loss(Pv, X, E, R, wb, dims)) < 1e-12 * loss
    rng = np.random.default_rng(1)
    h = 1e-5
    for name, o, s in fno_ref.entries(lay):
        n = int(np.prod(s))
        for j in rng.choice(n, min(n, 3), replace=False):
            Pp, Pm = Pv.copy(), Pv.copy()
            Pp[o + j] += h
            Pm[o + j] -= h
            fd = (_loss(Pp, X, E, R, wb, dims) - _loss(Pm, X, E, R, wb, dims)) / (2 * h)
            assert abs(fd - G[o + j]) < 1e-6 * max(1.0, abs(fd)), (name, j, fd, G[o + j])


def test_loss_derivative_by_the_output():
    rng = np.random.default_rng(2)
    out = rng.normal(size=(2, 4096, 4))
    E = rng.normal(size=(2, 4096, 4))
    R = rng.normal(size=(2, 4096, 4))
    wb = np.array([1.0, 2.0])
    _, dout = fno_ref.loss_and_dout(out, E, R, wb)
    h = 1e-6
    for idx in [(0, 5, 0), (1, 4000, 3), (1, 17, 1)]:
        op, om = out.copy(), out.copy()
        op[idx] += h
        om[idx] -= h
        fd = (fno_ref.loss_and_dout(op, E, R, wb)[0] - fno_ref.loss_and_dout(om, E, R, wb)[0]) / (2 * h)
        assert abs(fd - dout[idx]) < 1e-5 * abs(fd) + 2e-9  # the sum's rounding over 2h
