"""The product's deck side against the reference. CPU only (host code of
libfelis_hip.so; no GPU call).

* felis_deck_lattice: the lattice table the batched and plain loops consume,
  from a deck with a tapered undulator block, a sloped helical block in air
  and two FODO quad lists, against the sequence the reference's own
  FindParameters and updateK_quadG.cpp write for the same deck
  (tests/golden/update_k_quadg_deck.npz; tests/golden/make_golden.py).
* felis_deck_desc / felis_deck_steps: the derived Domain scalars of
  parameterSetting.cpp:71-77, 89-111, 168-175, 186-196, 236-240 for this deck
  and the five workload decks (tests/golden/deck_desc.json).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from felis_amd import _lib as fl

GOLD = Path(__file__).resolve().parent / "golden"
DECKS = Path(__file__).resolve().parent.parent / "felis_amd" / "decks"


def test_deck_lattice_table_matches_the_reference():
    """updateK_quadG.cpp:8-94 over 1200 periods x 2 half steps, as the
    product's parser and lattice expansion drive it: K (taper, quadratic taper,
    slope), ue, K0_alpha, ku, g (overlap-weighted quads of two lists), lambdaU,
    undulator type and the current / drift flags, bit for bit; entry 0 is the
    Domain state parameterSetting.cpp:168-174 leaves before the loop."""
    g = np.load(GOLD / "update_k_quadg_deck.npz", allow_pickle=False)
    deck = fl.Deck(GOLD / "deck_lattice.inp")
    n = g["seq"].shape[0] // 2
    tab = deck.lattice(0, n)
    rows = np.array([[t.K0, t.ue, t.K0_alpha, t.ku, t.g, t.lambdaU, t.und_type, t.current_flag, t.drift_flag]
                     for t in tab])
    np.testing.assert_array_equal(rows[0, :7], g["init"])
    seq = g["seq"]
    cols = [0, 1, 2, 3, 5, 6, 7, 8]
    bad = np.nonzero(np.any(rows[1:, cols] != seq[:, cols], axis=1))[0]
    assert bad.size == 0, (bad[:5], rows[1 + bad[:3]], seq[bad[:3]])
    # The quad gradient of a partly covered half step is (overlap)/dz*2*g. The
    # reference is built with -ffast-math (makefile:38-52), under which gcc
    # reassociates that product differently from branch to branch; the product
    # evaluates it as written: 9 of the 2400 entries differ by one ulp.
    gg = rows[1:, 4]
    np.testing.assert_array_less(np.abs(gg - seq[:, 4]), 2 * np.spacing(np.abs(seq[:, 4])) + 1e-300)
    assert np.count_nonzero(gg != seq[:, 4]) <= 12
    # the sequence exercises every branch: both blocks, gaps, air, both quad lists
    assert {1.0, 2.0} <= set(seq[:, 6]) and set(seq[:, 7]) == {0.0, 1.0} and set(seq[:, 8]) == {0.0, 1.0}
    assert (seq[:, 4] > 0).any() and (seq[:, 4] < 0).any() and len(set(seq[:, 0])) > 100


@pytest.mark.parametrize("name", ["deck_lattice.inp", "cfg0.inp", "cfg1.inp", "cfg2.inp", "cfg2e.inp", "cfg2s.inp",
                                  "cfg3.inp"])
def test_deck_derived_scalars(name):
    """ks, lambda0, the grid (dx from the deck's nx, then nx + 1), the radial
    grid, gamR, dz, maxStep and sliceN (with the Static mode's window) as
    parameterSetting.cpp derives them."""
    want = json.loads((GOLD / "deck_desc.json").read_text())[name]
    path = GOLD / name if name == "deck_lattice.inp" else DECKS / name
    deck = fl.Deck(path)
    d = deck.desc(fl.PLAIN_FP64)
    for k in ("nx", "ny", "nr", "num_slice", "slice_n", "abc_n", "bessel_n", "sc_l", "sc_f"):
        assert getattr(d, k) == want[k], (k, getattr(d, k), want[k])
    for k in ("dx", "dy", "min_x", "min_y", "min_z", "max_z", "ks", "lambda0", "gamR", "dr", "dz", "abc_sig"):
        assert getattr(d, k) == pytest.approx(want[k], rel=1e-15, abs=0), (k, getattr(d, k), want[k])
    assert deck.max_step == want["max_step"]
    assert deck.slice_n == want["slice_n"]
