"""Field diagnostics of diag/felField.cpp on the live field (felis_field_diag,
felis_amd/diag.py), on the 64-slice cut of the 2048-slice workload after
three periods.

felField reads a field dump; here the same dump is written
(felis_save_field_hdf), read back by the independent reader, and felField's
loops (felField.cpp:126-160) are restated on it in numpy:
  - the device's slice power and fluence agree with them to 1e-12
    relative (FP64 sums in another order);
  - felField's own energy check holds: the power summed over slices times
    bucketZ/c equals the fluence summed over cells times dx dy, for every
    polarisation and harmonic (1e-12);
  - the text files carry the same numbers at felField's 6 digits.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
import h5read  # noqa: E402
from felis_amd import diag  # noqa: E402

ROOT = Path(__file__).resolve().parent.parent
DECK = ROOT / "felis_amd" / "decks" / "cfg2s.inp"
C_LIGHT = 299792458.0


@pytest.fixture(scope="module")
def state(tmp_path_factory):
    deck = fl.Deck(DECK)
    d = deck.desc(fl.MIXED_FP32)
    ctx = fl.Context(d)
    deck.load(ctx, seed=20261203)
    ctx.run(deck.lattice(0, 3), 0, 3)
    _, spot = deck.save_settings(d.n_species)
    tmp = tmp_path_factory.mktemp("diag")
    ctx.save_field_hdf(tmp / "field3.h5", spot)
    p, f = ctx.field_diag(spot)
    yield dict(d=d, ctx=ctx, dir=tmp, power=p, fluence=f, spot=spot)
    ctx.close()


def _felfield(path, d):
    """felField.cpp:126-160 restated on a field dump"""
    with h5read.File(path) as f:
        coefP = f.dataset("powerNorm")[0][0]
        coefI = f.dataset("intensityNorm")[0][0]
        harms = list(f.dataset("harmony")[0])
        P = np.zeros((2, d.n_harm, d.slice_n))
        I = np.zeros((2, d.n_harm, d.nx * d.ny))
        for h, H in enumerate(harms):
            for pol, name in enumerate(("Ux", "Uy")):
                v = f.dataset(f"{name}{H}")[0].reshape(d.slice_n, d.nx * d.ny, 2)
                a2 = v[..., 0] ** 2 + v[..., 1] ** 2          # |U|^2 [slice][cell]
                P[pol, h] = a2.sum(axis=1) * coefP
                I[pol, h] = a2.sum(axis=0) * coefI
    return P, I, coefP, coefI


def test_diag_matches_felfield_on_the_dump(state):
    d = state["d"]
    P, I, coefP, coefI = _felfield(state["dir"] / "field3.h5", d)
    assert (coefP, coefI) == state["ctx"].field_norms(state["spot"])
    scale_p = np.abs(P).max()
    scale_i = np.abs(I).max()
    assert scale_p > 0 and scale_i > 0
    assert np.max(np.abs(state["power"] - P)) <= 1e-12 * scale_p
    assert np.max(np.abs(state["fluence"] - I)) <= 1e-12 * scale_i


def test_felfield_energy_check(state):
    """felField.cpp:201-223: the two energies it prints agree."""
    d = state["d"]
    bucket = d.num_slice * d.lambda0
    e_power = state["power"].sum(axis=2) * bucket / C_LIGHT
    e_fluence = state["fluence"].sum(axis=2) * d.dx * d.dy
    assert np.all(np.abs(e_power - e_fluence) <= 1e-12 * np.abs(e_power).max())


def test_text_files(state, tmp_path):
    d = state["d"]
    diag.write_power(tmp_path / "power3", 3, d, state["power"])
    diag.write_cross(tmp_path / "crossP3", 3, d, state["fluence"])
    lines = (tmp_path / "power3").read_text().splitlines()
    assert lines[0] == "#z = %g[m]" % (3 * d.dz)
    assert lines[1].startswith("#s[m]         Px1[W]       Py1[W]       Px2[W]")
    rows = np.array([[float(t) for t in l.split()] for l in lines[2:]])
    assert rows.shape == (d.slice_n, 1 + 2 * d.n_harm)
    want = np.stack([state["power"][pol, h] for h in range(d.n_harm) for pol in (0, 1)], axis=1)
    assert np.allclose(rows[:, 1:], want, rtol=6e-6, atol=0)
    assert np.allclose(rows[:, 0], np.arange(d.slice_n) * d.num_slice * d.lambda0, rtol=6e-6)
    lines = (tmp_path / "crossP3").read_text().splitlines()
    body = [l for l in lines[2:] if l.strip()]
    assert len(body) == d.nx * d.ny
    vals = np.array([[float(t) for t in l.split()] for l in body])
    want = np.stack([state["fluence"][pol, h] for h in range(d.n_harm) for pol in (0, 1)], axis=1)
    assert np.allclose(vals[:, 2:], want, rtol=6e-6, atol=0)


def test_particle_and_static_text_files(state, tmp_path):
    """felParticle's and the Static savers' text files carry the live state's
    numbers at their 6 digits."""
    d, ctx = state["d"], state["ctx"]
    p = ctx.download_particles()
    so, sw = ctx.particle_layout()
    diag.write_felparticle(tmp_path / "0Particle3", 3, d, p, so, sw, 0, skip=97)
    rows = np.array([[float(t) for t in l.split()] for l in (tmp_path / "0Particle3").read_text().splitlines()[1:]])
    n = np.arange(0, so[d.slice_n], 97)
    sl = np.searchsorted(so, n, side="right") - 1
    s = (sl * 2 * np.pi * d.num_slice + p["theta"][n]) * d.lambda0 / (2 * np.pi)
    want = np.stack([s, p["x"][n], p["y"][n], p["gamma"][n], p["px"][n], p["py"][n], sw[sl]], axis=1)
    assert rows.shape == want.shape
    assert np.allclose(rows, want, rtol=6e-6, atol=1e-300)
    diag.write_fields_txt(tmp_path / "Power3", d, state["fluence"], ctx.field_norms(state["spot"])[1])
    body = [l for l in (tmp_path / "Power3").read_text().splitlines()[1:] if l.strip()]
    vals = np.array([[float(t) for t in l.split()] for l in body])
    U = ctx.download_field(fl.UX).reshape(d.n_harm, d.slice_n + 2, d.ny, d.nx)[:, 1:-1]
    raw = (np.abs(U) ** 2).sum(axis=1)            # [h][j][i]
    want = raw.transpose(0, 2, 1).reshape(d.n_harm, -1)  # rows i-major, j-minor
    # %g keeps 6 digits; below 1e-300 (subnormal doubles) the absolute floor
    bad = ~np.isclose(vals[:, 2], want[0], rtol=6e-6, atol=1e-300)
    assert not bad.any(), [(int(i), body[i], want[0][i]) for i in np.nonzero(bad)[0][:5]]
