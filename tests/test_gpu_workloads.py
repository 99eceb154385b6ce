"""GPU: the workload decks end to end, the device loaders and the error word.

* The device seed loader (loadSeed.cpp:25-80) against the golden-pinned oracle
  on the steady-state and 64-slice grids, with the seed parameters the
  reference's parameterSetting.cpp derives (tests/golden/deck_desc.json).
* The steady-state deck (81 x 81, 8192 particles, 154 periods; main.cpp:92-156
  with no slippage) and the 256-slice deck (101 x 101, 4096 particles per
  slice, plain FP64, 30 periods) against the FP64 oracle loop.
* The 64-slice cut of the 2048-slice deck over its whole 770-period run:
  batched FP32/FP64 and plain FP64 against the FP64 oracle.
* The 2048-slice deck over its whole run: the batched path against the plain
  FP64 path on the GPU (oracle-matched above), integrated power of every
  harmonic at every period, bunching, and per-slice power at the end.
* The device error word: a NaN theta and a step past the dTheta guard give
  FELIS_ERR_NAN and FELIS_ERR_GUARD from felis_run and from the per-phase
  operators, in both precisions (particlePush.cpp:370-374,
  rearrangeParticles.cpp:105-108); every kernel exits normally.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np
import pytest

import fel_cases as fc
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")

ROOT = Path(__file__).resolve().parent.parent
DECKS = ROOT / "felis_amd" / "decks"
GOLD = Path(__file__).resolve().parent / "golden"
PRECS = [pytest.param(0, id="fp64"), pytest.param(1, id="mixed")]


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _tot(e, nh):
    return e[:, :nh].sum(1) + e[:, nh:].sum(1)


def _loaded(name, prec, seed=7):
    deck = fl.Deck(DECKS / name)
    d = deck.desc(prec)
    ctx = fl.Context(d)
    deck.load(ctx, seed=seed)
    return deck, d, ctx


def _oracle_of(deck, d, ctx):
    seg_off, seg_w = deck.beam_layout(d.n_species)
    return po.state_from_desc(d, seg_off, seg_w, ctx.download_particles(), ctx.download_field(fl.UX),
                              ctx.download_field(fl.UY))


def _bunching(b, nh):
    return np.abs(b[:, 0:2 * nh:2] + 1j * b[:, 1:2 * nh:2]) / b[:, 2 * nh:2 * nh + 1]


# ------------------------------------------------------------------ loaders

@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", ["cfg0.inp", "cfg2s.inp"])
def test_device_seed_loader_matches_oracle(name, prec):
    """felis_deck_load_seed against orc_load_seed_3d (pinned by the
    reference's own loadSeed.cpp) with the deck's derived seed parameters:
    FP64 to 1e-12, the complex64 field to its rounding."""
    s = json.loads((GOLD / "deck_desc.json").read_text())[name]["seed"]
    deck = fl.Deck(DECKS / name)
    d = deck.desc(prec)
    ctx = fl.Context(d)
    assert deck.L.felis_deck_load_seed(deck.h, ctx.h) == fl.OK
    gx, gy = ctx.download_field(fl.UX), ctx.download_field(fl.UY)
    st = po.state_from_desc(d, np.zeros(d.n_species * d.slice_n + 1, np.int64), np.ones(d.n_species * d.slice_n),
                            {k: np.zeros(0) for k in ("x", "y", "px", "py", "theta", "gamma")},
                            np.zeros_like(gx), np.zeros_like(gy))
    harm = [d.harmony[h] for h in range(d.n_harm)]
    po.load_seed_3d(st.params, st.Ux, st.Uy, harm.index(s["load_h"]), s["a0"], s["duration"], d.min_z, s["zR"],
                    s["focus"], s["spot_sig_r"], 0.0, s["laser_alpha"], s["laser_psi"])
    scale = np.max(np.abs(st.Ux)) + np.max(np.abs(st.Uy))
    assert scale > 0
    tol = 1e-12 if prec == 0 else 1e-7
    assert np.max(np.abs(gx - st.Ux)) <= tol * scale
    assert np.max(np.abs(gy - st.Uy)) <= tol * scale
    ctx.close()


# ---------------------------------------------------------------- workloads

@pytest.mark.parametrize("prec,tol", [pytest.param(0, 1e-9, id="fp64"), pytest.param(1, 1e-3, id="mixed")])
def test_steady_state_deck_whole_run(prec, tol):
    """The steady-state deck (Static mode: one slice, no slippage) over all 154
    periods of its 4 m undulator against the oracle loop from the same state:
    the integrated power at every period, the bunching and the twiss sums. The
    seed grows through the run, so the check covers the gain."""
    deck, d, ctx = _loaded("cfg0.inp", prec)
    assert d.mode == 1 and deck.max_step == 154
    n = deck.max_step
    tab = deck.lattice(0, n)
    st = _oracle_of(deck, d, ctx)
    e_g, b_g, t_g = ctx.run(tab, 0, n)
    po.set_threads(_threads())
    rc, e_o, b_o, t_o = st.run(po.lattice_table(tab), 0, n, False)
    assert rc == 0
    nh = d.n_harm
    pg, po_ = _tot(e_g, nh), _tot(e_o, nh)
    err = np.max(np.abs(pg - po_) / po_)
    print(f"steady state, {n} periods: max rel. power error {err:.3e}, gain {po_[-1] / po_[0]:.3f}")
    assert po_[-1] > 1.5 * po_[0]
    assert err < tol
    assert np.max(np.abs(_bunching(b_g, nh) - b_o[:, :nh])) < (1e-9 if prec == 0 else 1e-3)
    assert fc.rel_err(t_g, t_o) < (1e-9 if prec == 0 else 1e-4)
    ctx.close()


def test_256_slice_deck_plain_fp64():
    """The 256-slice, 101 x 101 deck (4096 particles per slice, flat-top
    current) on the plain FP64 path for 30 periods against the oracle loop:
    per-harmonic energies, bunching and twiss to 1e-9, and the per-slice
    field energies of the last period to 1e-9."""
    deck, d, ctx = _loaded("cfg1.inp", 0)
    assert d.slice_n == 256 and d.nx == 101
    n = 30
    tab = deck.lattice(0, n)
    st = _oracle_of(deck, d, ctx)
    e_g, b_g, t_g = ctx.run(tab, 0, n)
    po.set_threads(_threads())
    rc, e_o, b_o, t_o = st.run(po.lattice_table(tab), 0, n, True)
    assert rc == 0
    assert fc.rel_err(e_g, e_o) < 1e-9
    assert np.max(np.abs(_bunching(b_g, d.n_harm) - b_o[:, :d.n_harm])) < 1e-9
    assert fc.rel_err(t_g, t_o) < 1e-9
    ux = ctx.download_field(fl.UX)
    sl_g = np.sum(np.abs(ux[0, 1:-1]) ** 2, axis=(1, 2))
    sl_o = np.sum(np.abs(st.Ux[0, 1:-1]) ** 2, axis=(1, 2))
    assert fc.rel_err(sl_g, sl_o) < 1e-9
    ctx.close()


def test_64_slice_cut_whole_run_against_oracle():
    """The 64-slice cut of the 2048-slice deck (harmonics 1, 3, 5, 129 x
    129, Gaussian current) over the whole 20 m run (770 periods), from one
    initial state: the plain FP64 path to 1e-9 and the batched FP32/FP64 path
    to 1e-3 in the integrated power of every harmonic at every period, and in
    the per-slice power of the fundamental at the end."""
    deck, d, ctx_m = _loaded("cfg2s.inp", 1)
    n = deck.max_step
    assert n == 770 and d.n_harm == 3
    tab = deck.lattice(0, n)
    st = _oracle_of(deck, d, ctx_m)
    d0 = deck.desc(0)
    ctx_p = fl.Context(d0)
    p = ctx_m.download_particles()
    seg_off, seg_w = deck.beam_layout(d.n_species)
    ctx_p.upload_particles(seg_off, seg_w, p["x"], p["y"], p["px"], p["py"], p["theta"], p["gamma"])
    ctx_p.upload_field(fl.UX, st.Ux)
    ctx_p.upload_field(fl.UY, st.Uy)
    e_m, b_m, _ = ctx_m.run(tab, 0, n)
    e_p, b_p, _ = ctx_p.run(tab, 0, n)
    po.set_threads(_threads())
    rc, e_o, b_o, _ = st.run(po.lattice_table(tab), 0, n, True)
    assert rc == 0
    nh = d.n_harm
    for h in range(nh):
        # the harmonics start from zero: no deposit at step 0 (main.cpp:32, 127)
        ph_o = e_o[:, h] + e_o[:, nh + h]
        on = ph_o > 0
        assert on[2:].all()
        err_p = np.max(np.abs(e_p[on, h] + e_p[on, nh + h] - ph_o[on]) / ph_o[on])
        err_m = np.max(np.abs(e_m[on, h] + e_m[on, nh + h] - ph_o[on]) / ph_o[on])
        print(f"64-slice cut, harmonic {d.harmony[h]}: max rel. power error plain {err_p:.3e}, batched {err_m:.3e}")
        assert err_p < 1e-9
        assert err_m < 1e-3
    assert e_o[-1, 0] > 5 * e_o[0, 0]             # the seed is amplified along the run
    assert np.max(np.abs(_bunching(b_m, nh) - b_o[:, :nh])) < 2e-3
    sl_o = np.sum(np.abs(st.Ux[0, 1:-1]) ** 2, axis=(1, 2))
    sl_m = np.sum(np.abs(ctx_m.download_field(fl.UX)[0, 1:-1]) ** 2, axis=(1, 2))
    strong = sl_o > 1e-3 * sl_o.max()
    assert np.max(np.abs(sl_m[strong] - sl_o[strong]) / sl_o[strong]) < 1e-3
    ctx_m.close()
    ctx_p.close()


def test_2048_slice_gain_curve_batched_against_plain():
    """The 2048-slice deck (11.67 M particles, harmonics 1, 3, 5, 129 x 129)
    over its whole 770-period run from one initial state: the batched
    FP32/FP64 path against the plain FP64 path, both on the GPU. The
    integrated power of the fundamental agrees to 1e-3 at every period, the
    harmonics' to 1e-2, the bunching factors to 2e-3, and the per-slice power
    of the fundamental at the end to 1e-3 on every slice above 1e-3 of the
    peak."""
    deck, d, ctx_m = _loaded("cfg2.inp", 1)
    n = deck.max_step
    assert d.slice_n == 2048 and n == 770
    tab = deck.lattice(0, n)
    ctx_p = fl.Context(deck.desc(0))
    p = ctx_m.download_particles()
    seg_off, seg_w = deck.beam_layout(d.n_species)
    ctx_p.upload_particles(seg_off, seg_w, p["x"], p["y"], p["px"], p["py"], p["theta"], p["gamma"])
    del p
    ctx_p.upload_field(fl.UX, ctx_m.download_field(fl.UX))
    ctx_p.upload_field(fl.UY, ctx_m.download_field(fl.UY))
    e_m, b_m, _ = ctx_m.run(tab, 0, n)
    e_p, b_p, _ = ctx_p.run(tab, 0, n)
    nh = d.n_harm
    for h in range(nh):
        ref = e_p[:, h] + e_p[:, nh + h]
        on = ref > 0
        assert on[2:].all()
        err = np.abs(e_m[on, h] + e_m[on, nh + h] - ref[on]) / ref[on]
        print(f"2048 slices, harmonic {d.harmony[h]}: max rel. power error {err.max():.3e} "
              f"(period {int(err.argmax())}), final {err[-1]:.3e}")
        assert err.max() < (1e-3 if h == 0 else 1e-2)
    assert e_p[-1, 0] > 100 * e_p[0, 0]
    db = np.abs(_bunching(b_m, nh) - _bunching(b_p, nh))
    print(f"2048 slices: max bunching difference {db.max():.3e}")
    assert db.max() < 2e-3
    sm, sp = ctx_m.slice_energy()[0, 0], ctx_p.slice_energy()[0, 0]
    strong = sp > 1e-3 * sp.max()
    serr = np.max(np.abs(sm[strong] - sp[strong]) / sp[strong])
    print(f"2048 slices: max rel. slice power error at 20 m {serr:.3e} over {int(strong.sum())} slices")
    assert serr < 1e-3
    ctx_m.close()
    ctx_p.close()


# --------------------------------------------------------------- error word

def _error_case(kind):
    case = fc.make_case(slice_n=3, ppslice=64, harmonics=(1, 3))
    if kind == "nan":
        case.state["theta"][17] = np.nan
    else:
        # a field strong enough that the ponderomotive term alone moves theta
        # by more than 2 pi in one period (particlePush.cpp:370-374)
        case.state["Ux"][:, 2] *= 2e6
    return case


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,code", [("nan", -5), ("guard", -3)])
def test_error_word_from_the_loop(kind, code, prec):
    case = _error_case(kind)
    ctx = fc.gpu_context(case, prec)
    tab = [case.lattice()] * 5
    with pytest.raises(fl.FelisError) as e:
        ctx.run([fc.to_lib_lattice(l) for l in tab], 0, 2)
    assert e.value.code == code
    # the context stays usable: the word was cleared when it was read
    ctx.sync()
    ctx.close()


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,code", [("nan", -5), ("guard", -3)])
def test_error_word_from_the_operators(kind, code, prec):
    case = _error_case(kind)
    ctx = fc.gpu_context(case, prec)
    lat = fc.to_lib_lattice(case.lattice())
    ctx.solve_field(lat)
    ctx.transverse_push(lat)
    with pytest.raises(fl.FelisError) as e:
        ctx.push_theta_gamma(lat)
        ctx.periodic()
    assert e.value.code == code
    ctx.close()
