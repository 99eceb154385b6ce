"""The batched deposit in 64-bit fixed point (felis_set_deposit_mode
FELIS_DEPOSIT_FIXED), through the C-ABI.

solveField.cpp:246-359 adds the particles' contributions in particle order;
the batched deposit's float and double atomics add them in whatever order the
waves reach a cell, so two runs agree only to rounding. In the fixed-point
mode every contribution is rounded once to a multiple of a power-of-two
step and the sums are exact integers, so the whole step is bitwise the same
from run to run and for every deposit launch shape and LDS budget (one pass,
harmonic chunks, strips of node rows), with and without the captured step.
Checked on the 64-slice cut of the 2048-slice workload (harmonics 1, 3, 5,
one polarisation), on a helical two-polarisation case with an even harmonic
(the xi2 bound) and on three longitudinal space-charge modes; the
fixed-point runs agree with the atomic ones to the same 1e-5 as two
tunings do, and with the FP64 oracle to the batched path's tolerance.
"""
from pathlib import Path

import numpy as np
import pytest

import fel_cases as fc

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")
from felis_amd import tune  # noqa: E402

ROOT = Path(__file__).resolve().parent.parent
DECK = ROOT / "felis_amd" / "decks" / "cfg2s.inp"
SEED = 20261203
N = 40

TUNINGS = [None, dict(deposit_threads=256, deposit_lds=32768), dict(deposit_threads=512, deposit_lds=4096),
           dict(step_graph=1, deposit_threads=1024, deposit_lds=0)]


def _deck_run(mode, t):
    deck = fl.Deck(DECK)
    ctx = fl.Context(deck.desc(fl.MIXED_FP32))
    deck.load(ctx, seed=SEED)
    ctx.set_deposit_mode(mode)
    assert ctx.deposit_mode() == mode
    if t is not None:
        ctx.set_tuning(**t)
    e, b, tw = ctx.run(deck.lattice(0, N), 0, N)
    p = ctx.download_particles()
    out = dict(e=e, b=b, tw=tw, U=ctx.download_field(fl.UX), theta=p["theta"], gamma=p["gamma"])
    ctx.close()
    return out


def _case_run(case, mode, t, nsteps=12):
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    ctx.set_deposit_mode(mode)
    if t is not None:
        ctx.set_tuning(**t)
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2 * nsteps)]
    e, b, tw = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    p = ctx.download_particles()
    out = dict(e=e, b=b, tw=tw, U=ctx.download_field(fl.UX), Uy=ctx.download_field(fl.UY), theta=p["theta"],
               gamma=p["gamma"])
    ctx.close()
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def fixed_runs():
    return [_deck_run(fl.DEPOSIT_FIXED, t) for t in TUNINGS] + [_deck_run(fl.DEPOSIT_FIXED, None)]


def test_fixed_deposit_is_bitwise_reproducible(fixed_runs):
    """Energies, bunching, twiss, final field and particles after 40 periods:
    identical over four launch shapes and a repeat of the default."""
    for r in fixed_runs[1:]:
        _same(fixed_runs[0], r)


def test_fixed_deposit_agrees_with_the_atomic_one(fixed_runs):
    a = _deck_run(fl.DEPOSIT_ATOMIC, None)
    f = fixed_runs[0]
    assert tune._rel_energy_err(f["e"], a["e"]) < 1e-5
    assert np.max(np.abs(f["U"] - a["U"])) < 1e-5 * np.max(np.abs(a["U"]))
    assert np.max(np.abs(f["theta"] - a["theta"])) < 1e-4


CASES = {
    "helical_h12_two_pol": dict(harmonics=(1, 2), ue=0.3, und_type=2),
    "sc_modes_3": dict(harmonics=(1, 3), sc_l=3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fixed_deposit_cases(name):
    """Bitwise equal over launch shapes, and against the FP64 oracle to the
    batched path's tolerance (energy 1e-4, bunching 1e-3 of the weight)."""
    kw = dict(slice_n=6, ppslice=256, nx=33, ny=31, **CASES[name])
    case = fc.make_case(**kw)
    runs = [_case_run(fc.make_case(**kw), fl.DEPOSIT_FIXED, t) for t in TUNINGS]
    for r in runs[1:]:
        _same(runs[0], r)
    f = runs[0]
    if CASES[name].get("ue"):
        assert np.max(np.abs(f["Uy"])) > 0  # both polarisations live
    nsteps = 12
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2 * nsteps)]
    o = case.oracle_state()
    rc, e_o, b_o, _ = o.run(tab, 0, nsteps, True)
    assert rc == 0
    nh = len(kw["harmonics"])
    tot_g = f["e"][:, :nh].sum(1) + f["e"][:, nh:].sum(1)
    tot_o = e_o[:, :nh].sum(1) + e_o[:, nh:].sum(1)
    assert np.max(np.abs(tot_g - tot_o) / tot_o) < 1e-4
    bg = np.abs(f["b"][:, 0:2 * nh:2] + 1j * f["b"][:, 1:2 * nh:2]) / f["b"][:, 2 * nh:2 * nh + 1]
    assert np.max(np.abs(bg - b_o[:, :nh])) < 1e-3


def test_deposit_mode_rejects_other_values():
    deck = fl.Deck(DECK)
    ctx = fl.Context(deck.desc(fl.MIXED_FP32))
    assert ctx.deposit_mode() == fl.DEPOSIT_ATOMIC
    with pytest.raises(fl.FelisError):
        ctx.set_deposit_mode(2)
    assert ctx.deposit_mode() == fl.DEPOSIT_ATOMIC
    ctx.close()
