"""The autotuner's launch shapes of the batched step, through the C-ABI on the
64-slice cut of the 2048-slice workload (129 x 129, harmonics 1, 3, 5).

Every tuning the autotuner may pick runs the same arithmetic in another
launch shape or order: the field energy of every period and the final field
agree with the default tuning's to 1e-5 relative (float atomics of the
deposit summed in another order; the deposit's LDS budget of 32 KiB forces
its harmonics into chunks, and at 4 KiB the windows go through LDS in strips
of node rows, one harmonic per pass; other deposit run lengths combine
other runs of particles; a field solve with 8 lanes' columns in LDS passes
the rest through scratch), and the particles' phases to 1e-4 rad.
A 4 KiB budget with 256 radial grids is raised to hold the radial source row
(257 complex doubles).
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
from felis_amd import tune  # noqa: E402

ROOT = Path(__file__).resolve().parent.parent
DECK = ROOT / "felis_amd" / "decks" / "cfg2s.inp"
SEED = 20261203
N = 40


def _run(t, nr=None):
    deck = fl.Deck(DECK)
    d = deck.desc(fl.MIXED_FP32)
    if nr is not None:
        d.nr = nr
    ctx = fl.Context(d)
    deck.load(ctx, seed=SEED)
    if t is not None:
        ctx.set_tuning(**t)
    e, b, _ = ctx.run(deck.lattice(0, N), 0, N)
    out = dict(e=e, b=b, U=ctx.download_field(fl.UX), theta=ctx.download_particles()["theta"])
    ctx.close()
    return out


@pytest.fixture(scope="module")
def ref():
    return _run(None)


@pytest.fixture(scope="module")
def ref_nr256():
    return _run(None, nr=256)


@pytest.mark.parametrize("t", [dict(step_graph=1, deposit_threads=1024, deposit_lds=0),
                               dict(step_graph=0, deposit_threads=256, deposit_lds=0),
                               dict(step_graph=0, deposit_threads=512, deposit_lds=65536),
                               dict(step_graph=1, deposit_threads=256, deposit_lds=32768),
                               dict(adi_threads=1024),
                               dict(step_graph=1, deposit_threads=512, deposit_lds=65536, adi_threads=1024),
                               dict(deposit_threads=512, deposit_lds=4096),
                               dict(deposit_run=4), dict(deposit_run=16, deposit_threads=256),
                               dict(adi_lanes=8), dict(adi_lanes=8, adi_threads=1024, step_graph=1)],
                         ids=["graph", "dep256", "dep512_64k", "graph_dep256_32k", "adi1024", "graph_dep512_adi1024",
                              "dep_window_past_lds", "run4", "run16_dep256", "adi_split8", "adi_split8_graph"])
def test_tunings_agree_with_default(ref, t):
    r = _run(t)
    assert tune._rel_energy_err(r["e"], ref["e"]) < 1e-5
    assert np.max(np.abs(r["U"] - ref["U"])) < 1e-5 * np.max(np.abs(ref["U"]))
    assert np.max(np.abs(r["theta"] - ref["theta"])) < 1e-4


def test_radial_row_fits_a_small_budget(ref_nr256):
    """256 radial grids (the largest the context accepts) at a 4 KiB deposit
    budget: the radial row (4112 bytes) does not fit 4 KiB, so the run raises
    the budget to hold it, and agrees with the default tuning."""
    r = _run(dict(deposit_threads=512, deposit_lds=4096), nr=256)
    assert tune._rel_energy_err(r["e"], ref_nr256["e"]) < 1e-5
    assert np.max(np.abs(r["U"] - ref_nr256["U"])) < 1e-5 * np.max(np.abs(ref_nr256["U"]))
    assert np.max(np.abs(r["theta"] - ref_nr256["theta"])) < 1e-4


def test_set_tuning_rejects_bad_shapes():
    deck = fl.Deck(DECK)
    ctx = fl.Context(deck.desc(fl.MIXED_FP32))
    with pytest.raises(fl.FelisError):
        ctx.set_tuning(deposit_threads=384)
    with pytest.raises(fl.FelisError):
        ctx.set_tuning(deposit_lds=-1)
    with pytest.raises(fl.FelisError):
        ctx.set_tuning(adi_threads=256)
    with pytest.raises(fl.FelisError):
        ctx.set_tuning(deposit_run=3)
    with pytest.raises(fl.FelisError):
        ctx.set_tuning(adi_lanes=16)
    ctx.set_tuning(step_graph=1, deposit_threads=512, deposit_lds=40000, adi_threads=1024, deposit_run=16, adi_lanes=8)
    assert ctx.tuning() == dict(step_graph=1, deposit_threads=512, deposit_lds=40000, adi_threads=1024, deposit_run=16,
                                adi_lanes=8)
    ctx.close()


def test_autotuner_picks_an_agreeing_candidate(tmp_path):
    e = tune.tune(DECK, periods=10, repeats=1, verbose=False)
    assert len(e["table"]) == len(list(tune.candidates()))
    assert all(r["ok"] for r in e["table"])
    assert e["ms_per_step"] <= e["default_ms_per_step"]
    path = tmp_path / "tuning.json"
    tune.write_cache(e, path)
    deck = fl.Deck(DECK)
    ctx = fl.Context(deck.desc(fl.MIXED_FP32))
    deck.load(ctx, seed=SEED)
    assert tune.apply(ctx, path) == e["best"]
    assert ctx.tuning() == e["best"]
    ctx.close()
