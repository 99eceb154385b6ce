"""The 1D oracle (particlePush.cpp:388-524, solveField.cpp:409-549) against
closed forms. No reference build or golden vector covers these functions, so
the checks are the mathematics that fixes them.

CPU only.
"""
import math

import numpy as np

import fel_cases as fc


def test_zero_field_push_is_free_slippage():
    """With no field and no space charge every RK4 stage has the same rate:
    dtheta = (ku - ks/(2 gamma^2) K2) dz, dgamma = -dz wakeE/(mc^2 1e6)."""
    c = fc.make_case_1d(sc_l=1)
    c.state["Ux"][:] = 0
    c.state["Uy"][:] = 0
    c.state["Ez"][:] = 0
    o = c.oracle_state()
    lat = c.lattice()
    assert o.push_theta_gamma(lat) == 0
    p = c.params
    K2 = 1 + lat.K0 ** 2 * 0.5 * (1 + lat.ue ** 2)
    g0 = c.state["gamma"]
    dth = (lat.ku - p.ks / (2 * g0 ** 2) * K2) * p.dz
    np.testing.assert_allclose(o.theta - c.state["theta"], dth, rtol=1e-9, atol=1e-12)
    sl = np.repeat(np.arange(c.slice_n), np.diff(c.state["seg_off"]))
    dg = -p.dz * c.state["wakeE"][sl] / (0.510998918 * 1e6)
    # to the rounding of gamma ~ 8000 (one ulp is 9e-13)
    np.testing.assert_allclose(o.gamma - g0, dg, rtol=0, atol=2e-12)


def test_quiet_start_source_cancels():
    """Four mirrored phases per beamlet (exact, no jitter): the harmonic-1 and
    harmonic-3 sources and the first three space-charge modes sum to zero,
    while mode 4 keeps 4 e^{-i 4 theta0} per beamlet."""
    c = fc.make_case_1d(sc_l=4, harmonics=(1, 3))
    th = c.state["theta"]
    nb = th.size // 4
    th0 = np.random.default_rng(9).uniform(0, 2 * math.pi, nb)
    c.state["theta"] = np.repeat(th0, 4) + np.tile(np.arange(4) * math.pi / 2, nb)
    c.state["gamma"] = np.full_like(c.state["gamma"], 8000.0)
    o = c.oracle_state()
    lat = c.lattice()
    assert o.solve_field(lat) == 0
    src = o.ScUx[:, 1:-1]
    scale = np.abs(c.state["Ux"]).max()
    assert np.abs(src).max() < 1e-9 * scale
    Ez = o.Ez
    assert np.abs(Ez[:3, 1:-1]).max() < 1e-9 * np.abs(Ez[3, 1:-1]).max()
    # mode 4 (l = 3): -i coef/4 * sum macro * 4 e^{-i4 theta0}
    p = c.params
    area = 2 * math.pi * 30e-6 * 25e-6
    coef = 1.60217662e-19 * 299792458.0 ** 2 * 1.2566370614e-6 / p.ks / area / p.lambda0
    so = c.state["seg_off"]
    for sl in range(c.slice_n):
        t0 = th0[so[sl] // 4:so[sl + 1] // 4]
        want = -1j * coef / 4 * c.state["seg_weight"][sl] * np.sum(4 * np.exp(-4j * t0))
        assert abs(Ez[3, sl + 1] - want) < 1e-9 * abs(want)


def test_field_update_adds_the_source():
    """solveField.cpp:409-420: U += Sc on slots 1..slice_n, ghosts untouched."""
    c = fc.make_case_1d()
    o = c.oracle_state()
    U0 = o.Ux.copy()
    assert o.solve_field(c.lattice()) == 0
    np.testing.assert_allclose(o.Ux[:, 1:-1], U0[:, 1:-1] + o.ScUx[:, 1:-1], rtol=1e-15)
    assert np.array_equal(o.Ux[:, 0], U0[:, 0]) and np.array_equal(o.Ux[:, -1], U0[:, -1])
