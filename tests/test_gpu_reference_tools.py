"""The reference's own post-processing tools read this library's dumps.

diag/felField.cpp and diag/felParticle.cpp need only serial HDF5, so
`make -C oracle ref` compiles them from the reference's sources against the
image's HDF5 into oracle/_ref/ (test infrastructure, like the rest of
oracle/). Here the library writes field<it>.h5 and particle<it>.h5 after a
few periods, the reference's tools read them, and their text outputs are
compared with what the library computes on the device and writes in the same
formats (felis_amd/diag.py):
  - felField's power<it> (slice power per harmonic and polarisation) and
    crossP<it> (transverse fluence) of the single-harmonic steady-state deck:
    same header lines, numbers equal at the 6 digits the files carry;
  - felParticle's 0Particle<it> (longitudinal position, x, y, gamma, px, py,
    weight of every particle) of the steady-state deck and of the 64-slice
    time-dependent deck: the same text, line for line.
A reader that found a dataset, type, shape or metadata field other than where
the reference's saveFieldHDF.cpp / saveParticleHDF.cpp put it would fail or
print other numbers.

felField reads the `harmony` list into size_t elements with H5T_NATIVE_INT
(felField.cpp:66-68 and its restoreMeta), so for a dump of more than one
harmonic it packs two 32-bit entries into its first element: with harmonics
1, 3, 5 it looks up datasets "Ux12884901889", "Ux5" and "Ux0" and reports the
fifth harmonic's power in its second column pair, zeros elsewhere. That is the
reference's behaviour on its own files (saveFieldHDF.cpp:76 writes the list
as int); the test checks that felField does exactly that on ours, which
pins the list's element type and order.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
from felis_amd import diag  # noqa: E402

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"

needs_tools = pytest.mark.skipif(not (REF / "felField").exists() or not (REF / "felParticle").exists(),
                                 reason="oracle/_ref/felField, felParticle not built (make -C oracle ref)")


def _numbers(path):
    head, rows = [], []
    for line in Path(path).read_text().splitlines():
        if line.startswith("#"):
            head.append(line.rstrip())
        elif line.strip():
            rows.append([float(v) for v in line.split()])
    return head, np.array(rows)


@pytest.fixture(scope="module",
                params=[("cfg0.inp", fl.MIXED_FP32, 4), ("cfg0.inp", fl.PLAIN_FP64, 4), ("cfg2s.inp", fl.MIXED_FP32, 3)],
                ids=["steady-mixed", "steady-fp64", "64slice-mixed"])
def dumped(request, tmp_path_factory):
    name, prec, step = request.param
    deck = fl.Deck(ROOT / "felis_amd" / "decks" / name)
    d = deck.desc(prec)
    ctx = fl.Context(d)
    deck.load(ctx, seed=20261203)
    ctx.run(deck.lattice(0, step), 0, step)
    g0, spot = deck.save_settings(d.n_species)
    tmp = tmp_path_factory.mktemp("reftools")
    ctx.save_field_hdf(tmp / f"field{step}.h5", spot)
    ctx.save_particles_hdf(tmp / f"particle{step}.h5", g0)
    power, fluence = ctx.field_diag(spot)
    ours = tmp / "ours"
    ours.mkdir()
    diag.write_power(ours / f"power{step}", step, d, power)
    diag.write_cross(ours / f"crossP{step}", step, d, fluence)
    part = ctx.download_particles()
    seg_off, seg_w = deck.beam_layout(d.n_species)
    diag.write_felparticle(ours / f"0Particle{step}", step, d, part, seg_off, seg_w, 0)
    ctx.close()
    return dict(dir=tmp, ours=ours, step=step, d=d, power=np.asarray(power).reshape(2, d.n_harm, d.slice_n))


@needs_tools
def test_felfield_reads_the_field_dump(dumped):
    tmp, ours, step, d = dumped["dir"], dumped["ours"], dumped["step"], dumped["d"]
    r = subprocess.run([str(REF / "felField"), "1", str(step), str(step), "1"], cwd=tmp, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    if d.n_harm == 1:
        for name in (f"power{step}", f"crossP{step}"):
            h_ref, v_ref = _numbers(tmp / name)
            h_our, v_our = _numbers(ours / name)
            assert h_ref == h_our
            assert v_ref.shape == v_our.shape and v_ref.shape[0] > 0
            # %g keeps 6 significant digits: equal to half a unit in the last one
            np.testing.assert_allclose(v_our, v_ref, rtol=1e-5, atol=1e-300)
            assert np.any(v_ref[:, 1 if name.startswith("power") else 2:] > 0)
    else:
        assert [d.harmony[h] for h in range(d.n_harm)] == [1, 3, 5]
        assert "'Ux0'" in r.stderr
        _, v_ref = _numbers(tmp / f"power{step}")
        p5 = dumped["power"][0, 2]
        np.testing.assert_allclose(v_ref[:, 3], p5, rtol=1e-5, atol=1e-300)
        assert np.all(v_ref[:, [1, 2, 4, 5, 6]] == 0)
        assert np.any(p5 > 0)


@needs_tools
def test_felparticle_reads_the_particle_dump(dumped):
    tmp, ours, step = dumped["dir"], dumped["ours"], dumped["step"]
    r = subprocess.run([str(REF / "felParticle"), "3", str(step), str(step), "1", "1"], cwd=tmp,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = (tmp / f"0Particle{step}").read_text().splitlines()
    our = (ours / f"0Particle{step}").read_text().splitlines()
    assert len(ref) == len(our) > 1000
    assert ref == our
