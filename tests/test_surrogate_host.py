"""Host logic of the learned jump correction (felis_amd/surrogate.py), on
the CPU: target alignment and form, the saturation point of a gain curve,
the training decks and the model file."""
import math

import numpy as np
import pytest

from felis_amd import surrogate as su


def _field_sums(a, b):
    """felis_jump_residual field sums of complex per-slice fields a, b
    [2 nh, slice_n, cells]."""
    f = np.zeros(a.shape[:2] + (4,))
    ab = (np.conj(a) * b).sum(-1)
    f[..., 0], f[..., 1] = ab.real, ab.imag
    f[..., 2] = (np.abs(a) ** 2).sum(-1)
    f[..., 3] = (np.abs(b) ** 2).sum(-1)
    return f


def test_gain_targets_polar_and_aligned():
    rng = np.random.default_rng(1)
    nh, sn, J = 3, 20, 4
    a = rng.normal(size=(2 * nh, sn, 9)) + 1j * rng.normal(size=(2 * nh, sn, 9))
    g = np.exp(rng.uniform(-0.1, 0.1, (nh, sn)) + 1j * rng.uniform(-0.5, 0.5, (nh, sn)))
    b = a * np.concatenate([g, g])[:, :, None]
    # the residual is read after a slippage of J slots: the field solved at
    # slot s sits at s + J
    a_post = np.zeros_like(a)
    b_post = np.zeros_like(b)
    a_post[:, J:] = a[:, :sn - J]
    b_post[:, J:] = b[:, :sn - J]
    T, e, ok = su.gain_targets(_field_sums(a_post, b_post), nh, J)
    assert ok[:sn - J].all() and not ok[sn - J:].any()
    for h in range(nh):
        assert np.allclose(T[:sn - J, 2 * h], np.log(np.abs(g[h, :sn - J])), atol=1e-12)
        assert np.allclose(T[:sn - J, 2 * h + 1], np.angle(g[h, :sn - J]), atol=1e-12)
    assert np.allclose(e[:sn - J], (np.abs(b[0, :sn - J]) ** 2 + np.abs(b[nh, :sn - J]) ** 2).sum(-1))


def test_gain_target_keeps_energy_where_shapes_differ():
    """b = a + a part orthogonal to a: |g|^2 is the energy ratio (the
    least-squares factor would be 1 and lose the orthogonal energy)."""
    nh, sn = 1, 3
    a = np.zeros((2, sn, 4), complex)
    b = np.zeros((2, sn, 4), complex)
    a[0, :, 0] = 1.0
    b[0, :, 0] = 1.0
    b[0, :, 1] = 0.5
    T, _, ok = su.gain_targets(_field_sums(a, b), nh, 0)
    assert ok.all()
    assert np.allclose(T[:, 0], 0.5 * math.log(1.25))
    assert np.allclose(T[:, 1], 0.0)


def test_saturation_point():
    z = np.linspace(0, 30, 301)
    P = np.exp(z / 3.0) / (1 + np.exp((z - 20) / 1.5))
    zs, Ps, found = su.saturation(z, P)
    k = int(np.argmax(P))
    assert found and zs == pytest.approx(z[k]) and Ps == pytest.approx(P[k])
    zs, Ps, found = su.saturation(z, np.exp(z / 3.0))
    assert not found and zs == z[-1]


def test_training_decks_carry_the_sampled_parameters(tmp_path):
    from felis_amd import _lib as fl
    decks = su.training_decks(3, seed=5, out_dir=tmp_path, slices=96)
    assert len(decks) == 3
    for path, p in decks:
        assert 2000 <= p["current"] <= 5000 and 0.5e-4 <= p["spread"] <= 2e-4 and 1e3 <= p["power"] <= 1e5
        assert float(fl.find_parameter("EBeam", 1, "peak_current", path)) == pytest.approx(p["current"])
        assert float(fl.find_parameter("Seed", 1, "power", path)) == pytest.approx(p["power"], rel=1e-5)
        assert float(fl.find_parameter("EBeam", 1, "energy_spread", path)) == pytest.approx(p["spread"] * 100,
                                                                                          rel=1e-5)
        deck = fl.Deck(path)
        d = deck.desc(fl.MIXED_FP32)
        assert (d.slice_n, d.nx, d.ny, d.n_harm) == (96, 129, 129, 3)
        deck.close()


def test_model_file_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    dims = [20, 8, 6]
    m = su.Model(dims=dims, W=[rng.normal(size=(20, 8)).astype(np.float32), rng.normal(size=(8, 6)).astype(np.float32)],
                 b=[np.zeros(8, np.float32), np.ones(6, np.float32)], mu=np.arange(20, dtype=np.float32),
                 inv_sd=np.ones(20, np.float32), periods=10, gate=0.05, meta=dict(rows=5))
    su.save(m, tmp_path / "m.npz")
    r = su.load(tmp_path / "m.npz")
    assert r.dims == dims and r.periods == 10 and r.gate == 0.05 and r.meta == dict(rows=5)
    for x, y in zip(m.W + m.b + [m.mu, m.inv_sd], r.W + r.b + [r.mu, r.inv_sd]):
        assert np.array_equal(x, y)


def test_emission_targets_aligned():
    """b = a + (e - 1) E for a planted e per (slice, harmonic): the targets
    recover (log|e|, arg e), read J slots after the field was solved."""
    rng = np.random.default_rng(4)
    nh, sn, J = 3, 16, 5
    a = rng.normal(size=(2 * nh, sn, 7)) + 1j * rng.normal(size=(2 * nh, sn, 7))
    E = rng.normal(size=(2 * nh, sn, 7)) + 1j * rng.normal(size=(2 * nh, sn, 7))
    e = np.exp(rng.uniform(-0.2, 0.2, (nh, sn)) + 1j * rng.uniform(-0.5, 0.5, (nh, sn)))
    b = a + (np.concatenate([e, e]) - 1)[:, :, None] * E
    fe = np.zeros((2 * nh, sn, 3))
    fe[:, J:, 0] = (np.conj(E) * (b - a)).sum(-1).real[:, :sn - J]
    fe[:, J:, 1] = (np.conj(E) * (b - a)).sum(-1).imag[:, :sn - J]
    fe[:, J:, 2] = (np.abs(E) ** 2).sum(-1)[:, :sn - J]
    f = np.ones((2 * nh, sn, 4))
    T, w, ok = su.emission_targets(fe, f, nh, J)
    assert ok[:sn - J].all() and not ok[sn - J:].any()
    for h in range(nh):
        assert np.allclose(T[:sn - J, 2 * h], np.log(np.abs(e[h, :sn - J])), atol=1e-12)
        assert np.allclose(T[:sn - J, 2 * h + 1], np.angle(e[h, :sn - J]), atol=1e-12)
