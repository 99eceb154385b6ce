"""The wake-function table of wakeField.cpp:53-159 on the device
(felis_wake_function, felis_deck_wake_function_device).

Pinned to the reference's own tables (tests/golden/wake_function_*.npz,
written by wakeField.cpp compiled from the reference's sources) for the
circular AC, circular DC and flat AC impedances, to the tolerance the oracle
meets (1e-10 relative), and to the host table of the same deck. The flat
pipe's impedance integral is summed per wavenumber in a fixed tree on the
device instead of in order, the source of the remaining differences.
"""
import time
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


@pytest.mark.parametrize("tag", ["circ_ac", "circ_dc", "flat_ac"])
def test_device_wake_function_matches_reference(tag):
    g = np.load(GOLDEN / f"wake_function_{tag}.npz")
    dZ = float(g["num_slice"]) * float(g["lambda0"])
    t0 = time.perf_counter()
    wf, imp = fl.wake_function(int(g["shape"]), int(g["acdc"]), float(g["radius"]), float(g["cond"]),
                               float(g["ctau"]), int(g["slice_n"]), dZ)
    print(f"{tag}: {time.perf_counter() - t0:.4f} s on the device")
    scale = np.max(np.abs(g["wakeF"]))
    np.testing.assert_allclose(wf, g["wakeF"], rtol=1e-10, atol=1e-12 * scale)
    assert imp[0] == 0.0


def test_device_table_matches_the_host_table():
    deck = fl.Deck(ROOT / "felis_amd" / "decks" / "cfg2.inp")
    wf_h, imp_h = deck.wake_function()
    wf_d, imp_d = deck.wake_function_device()
    np.testing.assert_allclose(imp_d, imp_h, rtol=1e-11, atol=1e-13 * np.abs(imp_h).max())
    np.testing.assert_allclose(wf_d, wf_h, rtol=1e-10, atol=1e-12 * np.abs(wf_h).max())


def test_bad_settings_are_refused():
    with pytest.raises(fl.FelisError):
        fl.wake_function(3, 1, 1e-3, 3e7, 2.4e-6, 16, 1e-9)
    with pytest.raises(fl.FelisError):
        fl.wake_function(1, 1, 0.0, 3e7, 2.4e-6, 16, 1e-9)
