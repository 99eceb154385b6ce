"""The Fourier neural operator (felis_amd/csrc/felis_fno.hip) on the GPU
against its CPU reference (oracle/fno_ref.py, pinned in
tests/test_fno_ref.py): outputs and gradients with BF16 operands rounded at
the same points, and Adam reducing the loss on a learnable map.

Tolerances: the kernels accumulate in FP32, the reference in FP64; an
intermediate value that lands near a BF16 rounding boundary can round to
the neighbouring value in one of them (one BF16 step, 2^-8 relative), and
one such value among a map's kept modes moves the whole map's output by
about that step times its weight. Outputs therefore agree to 5e-3 of their
norm and each parameter block's gradient to 2e-2 of its norm; a layout or
transform mistake moves them by O(1)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
from oracle import fno_ref  # noqa: E402

CONFIGS = {"small": dict(cin=5, cout=4, w=8, m=4, L=2, ph=8, B=3),
           "specified": dict(cin=24, cout=6, w=32, m=8, L=4, ph=64, B=2)}


def _net(c, seed=3):
    rng = np.random.default_rng(seed)
    net = fl.FNO(c["cin"], c["cout"], c["w"], c["m"], c["L"], c["ph"], capacity=c["B"], seed=seed)
    P = net.params()
    lay = fno_ref.layout(c["cin"], c["cout"], c["w"], c["m"], c["L"], c["ph"])
    # biases and mode weights large enough that every path carries signal
    for name, o, s in fno_ref.entries(lay):
        n = int(np.prod(s))
        if name.startswith("Wm"):
            P[o:o + n] = rng.normal(size=n) * 4.0 / c["w"]
        elif name[0] in "bc" or name == "lb":
            P[o:o + n] = 0.1 * rng.normal(size=n)
    net.set_params(P.astype(np.float32))
    mu = rng.normal(size=c["cin"]).astype(np.float32) * 0.1
    isd = rng.uniform(0.5, 2.0, c["cin"]).astype(np.float32)
    net.set_input_affine(mu, isd)
    X = rng.normal(size=(c["B"], 4096, c["cin"])).astype(np.float32)
    return net, P.astype(np.float32), mu, isd, X, rng


def _dims(c):
    return c["cin"], c["cout"], c["w"], c["m"], c["L"], c["ph"]


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fno_forward_matches_reference(name):
    c = CONFIGS[name]
    net, P, mu, isd, X, _ = _net(c)
    Y = net.forward(X)
    want, _ = fno_ref.forward(P, X, *_dims(c), mu=mu, isd=isd, rounding=True)
    net.close()
    assert np.max(np.abs(want)) > 0
    assert np.linalg.norm(Y - want) < 5e-3 * np.linalg.norm(want)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fno_gradients_match_reference(name):
    c = CONFIGS[name]
    net, P, mu, isd, X, rng = _net(c)
    E = rng.normal(size=(c["B"], 4096, c["cout"])).astype(np.float32)
    R = rng.normal(size=(c["B"], 4096, c["cout"])).astype(np.float32)
    wb = rng.uniform(0.5, 1.5, c["B"]).astype(np.float32)
    loss = net.train_step(X, E, R, wb, lr=0.0)
    G = net.grad()
    net.close()
    lref, Gref = fno_ref.gradients(P, X, E, R, wb, *_dims(c), mu=mu, isd=isd, rounding=True)
    assert abs(loss - lref) < 1e-3 * lref
    lay = fno_ref.layout(*_dims(c))
    for bname, o, s in fno_ref.entries(lay):
        n = int(np.prod(s))
        g, gr = G[o:o + n], Gref[o:o + n]
        assert np.linalg.norm(gr) > 0, bname
        assert np.linalg.norm(g - gr) < 2e-2 * np.linalg.norm(gr), bname


def test_fno_training_reduces_loss():
    """A factor map that depends on the input maps through smooth spatial
    structure (a low-pass of channel 0 and a product of channels 1 and 2):
    Adam on the GPU brings the loss below a fifth of the untrained
    network's."""
    c = dict(cin=4, cout=2, w=16, m=4, L=2, ph=16, B=8)
    rng = np.random.default_rng(11)
    net = fl.FNO(c["cin"], c["cout"], c["w"], c["m"], c["L"], c["ph"], capacity=c["B"], seed=5)
    x = np.arange(64)

    def sample(B):
        X = np.zeros((B, 64, 64, c["cin"]), np.float32)
        for b in range(B):
            for ch in range(c["cin"]):
                kx, ky = rng.integers(0, 3, 2)
                ph_ = rng.uniform(0, 2 * np.pi)
                X[b, :, :, ch] = np.cos(2 * np.pi * (kx * x[None, :] + ky * x[:, None]) / 64 + ph_)
        o = 0.3 * X[..., 0] + 0.2 * X[..., 1] * X[..., 2]
        E = np.zeros((B, 64, 64, 2), np.float32)
        E[..., 0] = rng.normal(size=(B, 64, 64))
        E[..., 1] = rng.normal(size=(B, 64, 64))
        R = np.zeros_like(E)
        R[..., 0] = o * E[..., 0]
        R[..., 1] = o * E[..., 1]
        return X.reshape(B, 4096, -1), E.reshape(B, 4096, 2), R.reshape(B, 4096, 2), np.ones(B, np.float32)

    first = None
    losses = []
    for it in range(300):
        X, E, R, wb = sample(c["B"])
        losses.append(net.train_step(X, E, R, wb, lr=3e-3 if it < 200 else 1e-3))
        if first is None:
            first = losses[0]
    net.close()
    assert np.mean(losses[-20:]) < 0.2 * first, (first, np.mean(losses[-20:]))
