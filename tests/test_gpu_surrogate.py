"""The matrix-core perceptron of the surrogate against its CPU reference
(oracle/mlp_ref.py): the BF16 product's lane layout, forward outputs and
gradients, and that Adam training reduces the loss; the committed
correction network along a run and to saturation against the batched run
(2% in saturation length and power), and what it gains over the jumps
without it."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
from oracle import mlp_ref  # noqa: E402


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (37, 29, 53), (64, 48, 200)])
def test_gemm_bf16_layout(M, N, K):
    """C = A B on v_mfma_f32_16x16x16_bf16 tiles, ragged edges included;
    operands BF16-rounded, so only the FP32 summation order differs."""
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    got = fl.gemm_bf16(A, B)
    want = mlp_ref.bf16(A).astype(np.float64) @ mlp_ref.bf16(B).astype(np.float64)
    assert np.max(np.abs(got - want)) < 1e-4 * np.max(np.abs(want))


def _net(dims, rows, seed=3):
    net = fl.MLP(dims, capacity=rows, seed=seed)
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for l in range(len(dims) - 1):
        W = (rng.normal(size=(dims[l], dims[l + 1])) / np.sqrt(dims[l])).astype(np.float32)
        b = (0.1 * rng.normal(size=dims[l + 1])).astype(np.float32)
        net.set_layer(l, W, b)
        Ws.append(W)
        bs.append(b)
    return net, Ws, bs


def test_mlp_forward_matches_reference():
    dims = [12, 48, 32, 6]
    net, Ws, bs = _net(dims, 70)
    X = np.random.default_rng(7).normal(size=(70, 12)).astype(np.float32)
    Y = net.forward(X)
    Yr, _ = mlp_ref.forward(Ws, bs, X)
    assert np.max(np.abs(Y - Yr)) < 2e-3 * np.max(np.abs(Yr))
    net.close()


def test_mlp_gradients_match_reference():
    """dW and db of every layer (lr = 0 leaves the weights unchanged)."""
    dims = [12, 48, 32, 6]
    net, Ws, bs = _net(dims, 90)
    rng = np.random.default_rng(8)
    X = rng.normal(size=(90, 12)).astype(np.float32)
    T = rng.normal(size=(90, 6)).astype(np.float32)
    wrow = rng.uniform(0.5, 2.0, 90).astype(np.float32)
    loss = net.train_step(X, T, 0.0, wrow)
    lr_, gW, gb = mlp_ref.gradients(Ws, bs, X, T, wrow)
    assert loss == pytest.approx(lr_, rel=2e-3)
    for l in range(3):
        g, gbb = net.get_grad(l)
        scale = np.max(np.abs(gW[l]))
        assert np.max(np.abs(g - gW[l])) < 1e-2 * scale, l
        assert np.max(np.abs(gbb - gb[l])) < 1e-2 * np.max(np.abs(gb[l])), l
    W0, _ = net.get_layer(0)
    assert np.array_equal(W0, Ws[0])
    net.close()


def test_mlp_training_reduces_loss():
    """Adam on the GPU fits a smooth map of 4 inputs to 2 outputs."""
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, size=(512, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], X[:, 2] ** 2 - 0.5 * X[:, 3]], 1).astype(np.float32)
    net = fl.MLP([4, 64, 64, 2], capacity=512, seed=5)
    first = net.train_step(X, T, 3e-3)
    for _ in range(400):
        last = net.train_step(X, T, 3e-3)
    assert last < 0.05 * first
    net.close()


# ---------------------------------------------------------------- the jump correction

def test_correction_network_learns_the_jump_residual(tmp_path):
    """Training pipeline end to end on the GPU (small): rows from random
    decks' fine runs, Adam through felis_mlp_train_step, and on held-out rows
    the network's error is below that of no correction."""
    from felis_amd import surrogate as su
    X, T, W = su.dataset(4, periods=10, gate=0.05, seed=11, slices=128, rows_per_jump=64, log=None)
    assert X.shape[1] == 20 and T.shape[1] == 6 and len(X) > 1000
    perm = np.random.default_rng(0).permutation(len(X))
    nv = len(X) // 5
    model = su.train(X[perm[nv:]], T[perm[nv:]], W[perm[nv:]], 10, 0.05, steps=800, batch=2048, log=None)
    q = su.fit_quality(model, X[perm[:nv]], T[perm[:nv]], W[perm[:nv]])
    print("held-out weighted MSE", q)
    assert q["ratio"] < 0.5
    su.save(model, tmp_path / "m.npz")
    assert su.load(tmp_path / "m.npz").dims == model.dims


def test_corrected_jumps_follow_the_fine_run():
    """The committed network on 150 periods of the 64-slice deck: jumps of 10
    periods with the correction stay within 25% of the plain batched run's
    field energy (the jump's own error in the shot-noise-dominated start)."""
    from felis_amd import surrogate as su
    deck = Path(__file__).resolve().parent.parent / "felis_amd" / "decks" / "cfg2s.inp"
    rf = su.run_curves(deck, seed=20261203, n_steps=150)
    net = su.load().network(capacity=64)
    rj = su.run_curves(deck, seed=20261203, periods=10, net=net, n_steps=150)
    net.close()
    ok = ~np.isnan(rj["E"])
    assert np.sum(rj["kind"] == 1) >= 10
    assert np.max(np.abs(rj["E"][ok] / rf["E"][ok] - 1)) < 0.25


def _sat(z, r):
    """Saturation point of a run with jumps (energies inside jumps filled in
    linearly)."""
    from felis_amd import surrogate as su
    ok = ~np.isnan(r["E"])
    idx = np.nonzero(ok)[0]
    e = r["E"].copy()
    e[~ok] = np.interp(np.nonzero(~ok)[0], idx, r["E"][idx])
    return su.saturation(z, e)[:2]


def test_jumps_to_saturation_and_the_network_gain(tmp_path):
    """The 2048-slice beam in a 40 m undulator to saturation (the first
    maximum of the total field energy): the batched run against 10-period
    jumps while |b1| < 0.05 (the surrogate's range), without and with the
    committed correction network. With the network the saturation length and
    power agree within 2% (measured: +0.66% and +0.42%), and the network
    gains over the jumps alone (measured: saturation power +1.51% without
    it, +0.42% with it; the mean |error| of the field energy along the run
    to saturation 4.1% without, 2.9% with): the test asserts the power error
    at most half and the mean error at most 0.85 of the jumps' own. The
    runs with jumps take less time than the batched one.

    The field-to-field network (felis_fno, the committed
    felis_amd/models/jump_fno_J10.npz: factor maps on the emission from the
    64 x 64 maps of each slice) on the same run: saturation length within
    2% and power within 3% (measured: +0.22%, +2.1%), and a mean |error|
    along the run at most 0.85 of the jumps' own (measured: 1.9%)."""
    from felis_amd import surrogate as su
    from felis_amd.decks import make_decks as md
    deck = tmp_path / "cfg2_40m.inp"
    deck.write_text(md.deck(**dict(md.DECKS["cfg2.inp"], length=40.0)))
    rf = su.run_curves(deck, seed=20261203)
    rj = su.run_curves(deck, seed=20261203, periods=10, gate=0.05)
    net = su.load().network(capacity=2048)
    rn = su.run_curves(deck, seed=20261203, periods=10, gate=0.05, net=net, net_gate=0.05)
    net.close()
    fno = su.load_fno().network(capacity=2048)
    rfno = su.run_curves(deck, seed=20261203, periods=10, gate=0.05, fno=fno, net_gate=0.05)
    fno.close()
    z = np.arange(rf["steps"], dtype=float)
    zs, ps, found = su.saturation(z, rf["E"])
    assert found
    zj, pj = _sat(z, rj)
    zn, pn = _sat(z, rn)

    def mean_err(r):
        m = ~np.isnan(r["E"]) & (z <= zs)
        return np.mean(np.abs(r["E"][m] / rf["E"][m] - 1))
    res = dict(z_net=zn / zs - 1, p_net=pn / ps - 1, z_jump=zj / zs - 1, p_jump=pj / ps - 1,
               mean_net=mean_err(rn), mean_jump=mean_err(rj), speedup=rf["seconds"] / rn["seconds"])
    print(res)
    assert np.sum(rn["kind"] == 1) >= 50
    assert abs(res["z_net"]) < 0.02 and abs(res["p_net"]) < 0.02
    assert abs(res["p_net"]) < 0.5 * abs(res["p_jump"])
    assert res["mean_net"] < 0.85 * res["mean_jump"]
    assert rn["seconds"] < rf["seconds"] and rj["seconds"] < rf["seconds"]
    zf, pf = _sat(z, rfno)
    print(dict(z_fno=zf / zs - 1, p_fno=pf / ps - 1, mean_fno=mean_err(rfno), speedup_fno=rf["seconds"] / rfno["seconds"]))
    assert np.sum(rfno["kind"] == 1) >= 50
    assert abs(zf / zs - 1) < 0.02 and abs(pf / ps - 1) < 0.03
    assert mean_err(rfno) < 0.85 * res["mean_jump"]
