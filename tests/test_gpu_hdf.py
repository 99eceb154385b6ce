"""HDF5 dump and restore of a context in the reference's file layout
(saveFieldHDF.cpp, saveParticleHDF.cpp, restoreFieldHDF.cpp,
restoreParticleHDF.cpp), on the 64-slice cut of the 2048-slice workload
(129 x 129, harmonics 1, 3, 5).

The reference savers need parallel HDF5 (H5Pset_fapl_mpio) and MPI, which
this image does not have, so no reference-written file exists to compare
with. The layout is pinned instead by what the reference's writers and its
diag/felField.cpp reader name and shape, read back through an independent
ctypes reader (tests/h5read.py) and h5dump:
  - dataset names, shapes, stored types (8-byte floats, 4-byte ints for the
    int metadata, 8-byte ints for the particle attributes) and values, bit
    for bit against the downloaded state;
  - save -> restore into a fresh context gives the same state bit for bit;
  - a restart from the files (a fresh context, as the reference restarts
    without loading a beam or seed) continues the run: every later
    period's field energy agrees with the uninterrupted run to 1e-6 (the
    deposit's atomic summation order);
  - a restore onto another grid is refused with the reference's checks.
"""
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
import h5read  # noqa: E402

ROOT = Path(__file__).resolve().parent.parent
DECK = ROOT / "felis_amd" / "decks" / "cfg2s.inp"
SEED = 20261203
# constants.h of the reference
E_MASS, C_LIGHT, E_CHARGE, Z0 = 9.10938291e-31, 299792458.0, 1.60217662e-19, 376.730313668


def _ctx(load=True, precision=None):
    deck = fl.Deck(DECK)
    d = deck.desc(fl.MIXED_FP32 if precision is None else precision)
    ctx = fl.Context(d)
    if load:
        deck.load(ctx, seed=SEED)
    return deck, d, ctx


def _slots(U, d):
    """slots 1..slice_n of a downloaded field as [h][slice][2N] doubles"""
    N = d.nx * d.ny
    a = np.asarray(U).reshape(d.n_harm, d.slice_n + 2, N)[:, 1:-1, :]
    out = np.empty(a.shape[:2] + (2 * N,))
    out[..., 0::2] = a.real
    out[..., 1::2] = a.imag
    return out


@pytest.fixture(scope="module")
def saved(tmp_path_factory):
    deck, d, ctx = _ctx()
    ctx.run(deck.lattice(0, 3), 0, 3)
    tmp = tmp_path_factory.mktemp("hdf")
    gamma0, spot = deck.save_settings(d.n_species)
    ctx.save_field_hdf(tmp / "field3.h5", spot)
    ctx.save_particles_hdf(tmp / "particle3.h5", gamma0)
    state = dict(Ux=ctx.download_field(fl.UX), Uy=ctx.download_field(fl.UY), part=ctx.download_particles(),
                 layout=ctx.particle_layout())
    yield dict(deck=deck, d=d, ctx=ctx, dir=tmp, state=state, gamma0=gamma0, spot=spot)
    ctx.close()


def test_field_file_layout(saved):
    d, st = saved["d"], saved["state"]
    N = d.nx * d.ny
    harms = [d.harmony[i] for i in range(d.n_harm)]
    meta_int = ["sliceN", "harmony", "numHarmony", "nx", "ny"]
    meta_dbl = ["minZ", "dz", "bucketZ", "powerNorm", "intensityNorm", "dx", "dy", "minX", "minY"]
    with h5read.File(saved["dir"] / "field3.h5") as f:
        fields = [f"{p}{H}" for H in harms for p in ("Ux", "Uy")]
        assert f.names() == sorted(fields + meta_int + meta_dbl)
        for p, U in (("Ux", st["Ux"]), ("Uy", st["Uy"])):
            ref = _slots(U, d)
            for h, H in enumerate(harms):
                v, cls, size = f.dataset(f"{p}{H}")
                assert (cls, size, v.shape) == ("float", 8, (d.slice_n, 2 * N))
                assert np.array_equal(v, ref[h])
        got = {}
        for k in meta_int:
            v, cls, size = f.dataset(k)
            assert (cls, size) == ("int", 4), k
            got[k] = v
        for k in meta_dbl:
            v, cls, size = f.dataset(k)
            assert (cls, size, v.shape) == ("float", 8, (1,)), k
            got[k] = v[0]
    assert list(got["harmony"]) == harms and got["numHarmony"][0] == d.n_harm
    assert (got["sliceN"][0], got["nx"][0], got["ny"][0]) == (d.slice_n, d.nx, d.ny)
    assert (got["minZ"], got["dz"], got["dx"], got["dy"], got["minX"], got["minY"]) == (
        d.min_z, d.dz, d.dx, d.dy, d.min_x, d.min_y)
    bucket = d.num_slice * d.lambda0
    coef = E_MASS * C_LIGHT ** 2 * d.ks / E_CHARGE
    assert got["bucketZ"] == bucket
    assert got["powerNorm"] == pytest.approx(coef ** 2 / (2 * Z0) * d.dx * d.dy, rel=1e-14)
    assert got["intensityNorm"] == pytest.approx(coef ** 2 / (2 * Z0) * bucket / C_LIGHT, rel=1e-14)


def test_particle_file_layout(saved):
    d, st = saved["d"], saved["state"]
    p = st["part"]
    seg_off, seg_w = st["layout"]
    sn = d.slice_n
    with h5read.File(saved["dir"] / "particle3.h5") as f:
        species = [str(s) for s in range(d.n_species)]
        meta = ["minZ", "dz", "bucketZ", "dPhi", "lambda0", "gamma0", "numData", "nSpecies", "sliceN"]
        assert f.names() == sorted(species + meta)
        for s in range(d.n_species):
            b, e = seg_off[s * sn], seg_off[(s + 1) * sn]
            rows, cls, size = f.dataset(str(s))
            assert (cls, size, rows.shape) == ("float", 8, (e - b, 8))
            cnt = np.diff(seg_off[s * sn:(s + 1) * sn + 1])
            sl = np.repeat(np.arange(sn), cnt)
            want = np.stack([p["theta"][b:e], p["x"][b:e], p["y"][b:e], p["gamma"][b:e], p["px"][b:e],
                             p["py"][b:e], sl.astype(float), np.repeat(seg_w[s * sn:(s + 1) * sn], cnt)], axis=1)
            assert np.array_equal(rows, want)
            for a in ("cntP", "totalCnt"):
                v, cls, size = f.attribute(str(s), a)
                assert (cls, size, list(v)) == ("int", 8, [e - b]), a
        vals = {k: f.dataset(k) for k in meta}
    for k in ("numData", "nSpecies", "sliceN"):
        assert vals[k][1:] == ("int", 4), k
    assert (vals["numData"][0][0], vals["nSpecies"][0][0], vals["sliceN"][0][0]) == (8, d.n_species, sn)
    assert vals["dPhi"][0][0] == 2 * math.pi * d.num_slice
    assert vals["lambda0"][0][0] == d.lambda0 and vals["bucketZ"][0][0] == d.lambda0 * d.num_slice
    assert (vals["minZ"][0][0], vals["dz"][0][0]) == (d.min_z, d.dz)
    assert np.array_equal(vals["gamma0"][0], saved["gamma0"])
    # the deck's beam energy: gamma0 = energy / mc2 + 1 (mc2 = 0.510998918 MeV)
    assert saved["gamma0"][0] == pytest.approx(d.gamma_ref, rel=1e-15)


def test_h5dump_reads_the_files(saved):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not Path(h5dump).exists():
        pytest.skip("h5dump is not installed")
    d = saved["d"]
    out = subprocess.run([h5dump, "-H", str(saved["dir"] / "field3.h5")], capture_output=True, text=True,
                         timeout=60).stdout
    assert 'DATASET "Ux1"' in out and "H5T_IEEE_F64LE" in out
    assert f"( {d.slice_n}, {2 * d.nx * d.ny} )" in out
    out = subprocess.run([h5dump, "-A", "-d", "/0", str(saved["dir"] / "particle3.h5")], capture_output=True,
                         text=True, timeout=60).stdout
    assert 'ATTRIBUTE "totalCnt"' in out and "H5T_STD_I64LE" in out


def test_restore_gives_the_saved_state(saved):
    deck, d, ctx = _ctx(load=False)
    ctx.restore_field_hdf(saved["dir"] / "field3.h5")
    ctx.restore_particles_hdf(saved["dir"] / "particle3.h5")
    st = saved["state"]
    assert np.array_equal(_slots(ctx.download_field(fl.UX), d), _slots(st["Ux"], d))
    assert np.array_equal(_slots(ctx.download_field(fl.UY), d), _slots(st["Uy"], d))
    so, sw = ctx.particle_layout()
    assert np.array_equal(so, st["layout"][0]) and np.array_equal(sw, st["layout"][1])
    got = ctx.download_particles()
    for k in ("x", "y", "px", "py", "theta", "gamma"):
        assert np.array_equal(got[k], st["part"][k]), k
    ctx.close()


def test_restart_continues_the_run(saved):
    """Six periods straight vs three, dump, restart in a fresh context, three
    more under the uninterrupted run's lattice table. (The reference's own
    restart, main.cpp:50-58, does not rerun updateK_quadG for the period
    before the restart, so its first solveField runs under the deck's initial
    lattice, here with no current: deck.lattice(3, 3) reproduces that, and
    its first entry differs from the continuation's.)"""
    deck, d, full = _ctx()
    nh = d.n_harm
    tab = deck.lattice(0, 6)
    e_full, _, _ = full.run(tab, 0, 6)
    full.close()
    _, _, ctx = _ctx(load=False)
    ctx.restore_field_hdf(saved["dir"] / "field3.h5")
    ctx.restore_particles_hdf(saved["dir"] / "particle3.h5")
    restart = deck.lattice(3, 3)
    assert restart[0].current_flag == 0 and tab[6].current_flag == 1
    assert all(bytes(a) == bytes(b) for a, b in zip(restart[1:], tab[7:]))
    e2, _, _ = ctx.run(tab[6:], 3, 3)
    ctx.close()
    E1 = e_full[3:, :nh].sum(1) + e_full[3:, nh:].sum(1)
    E2 = e2[:, :nh].sum(1) + e2[:, nh:].sum(1)
    assert np.all(np.abs(E2 - E1) <= 1e-6 * E1)


def test_restore_refuses_another_grid(saved, tmp_path):
    deck = fl.Deck(ROOT / "felis_amd" / "decks" / "cfg1.inp")
    d = deck.desc(fl.MIXED_FP32)
    ctx = fl.Context(d)
    with pytest.raises(fl.FelisError, match="nx is 129, the context's is 101"):
        ctx.restore_field_hdf(saved["dir"] / "field3.h5")
    with pytest.raises(fl.FelisError, match="cannot open"):
        ctx.restore_field_hdf(tmp_path / "missing.h5")
    ctx.close()


def test_empty_slices_round_trip(tmp_path):
    """A slice with no particles stays empty and the next slice keeps its own
    rows (the reference's range search would give it every earlier row too)."""
    deck, d, ctx = _ctx()
    so, sw = ctx.particle_layout()
    p = ctx.download_particles()
    cnt = np.diff(so)
    keep = np.ones(so[-1], bool)
    for s in (5, 6, 40):  # empty slices 5, 6 and 40 of species 0
        keep[so[s]:so[s + 1]] = False
        cnt[s] = 0
    so2 = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    ctx.upload_particles(so2, sw, *(p[k][keep] for k in ("x", "y", "px", "py", "theta", "gamma")))
    gamma0, _ = deck.save_settings(d.n_species)
    ctx.save_particles_hdf(tmp_path / "particle0.h5", gamma0)
    _, _, c2 = _ctx(load=False)
    c2.restore_particles_hdf(tmp_path / "particle0.h5")
    so3, sw3 = c2.particle_layout()
    assert np.array_equal(so3, so2)
    assert np.array_equal(sw3[cnt > 0], sw[cnt > 0]) and np.all(sw3[cnt == 0] == 0)
    got = c2.download_particles()
    for k in ("x", "theta", "gamma"):
        assert np.array_equal(got[k], p[k][keep]), k
    ctx.close()
    c2.close()
