"""The 1D oracle (particlePush.cpp:388-524, solveField.cpp:409-549) against
closed forms. No reference build or golden vector covers these functions, so
the checks are the mathematics that fixes them:

* the push's field coupling: the Madey low-gain curve of a frozen field
  (second-order perturbation theory, energy loss above resonance and gain
  below, with the coupling's a^2 setting the scale);
* the radiation source against the push: one step's field-energy gain (the
  cross term of the source with the field, as power over the beam area
  2 pi sx sy times the slice duration) equals the beam's energy loss,
  converging linearly in dz;
* the longitudinal space charge against the push: a cold beam at resonance
  with a small phase modulation oscillates at the 1D plasma wavenumber
  k_p^2 = K2 e^2 n_e / (eps0 m c^2 gamma^3) (a wrong factor, power of gamma
  or sign turns the oscillation into another frequency or a growth);
* zero-field free slippage, quiet-start cancellation and the field update.

CPU only.
"""
import ctypes as C
import math

import numpy as np
from scipy import special

import fel_cases as fc
from oracle import pyoracle as po

MU0 = 1.2566370614e-6
C_LIGHT = 299792458.0
E_MASS = 9.10938291e-31
E_CHARGE = 1.60217662e-19
MC2 = 0.510998918
Z0 = 376.730313668
EPS0 = 1.0 / (MU0 * C_LIGHT ** 2)


def _solve_sc_1d(o, lat):
    """solveField.cpp:422-549 alone: the radiation source into ScUx/ScUy and
    the space-charge modes into Ez, the field left as it is."""
    o.ScUx[:] = 0
    o.ScUy[:] = 0
    return po.lib().orc_solve_sc_1d(C.byref(o.st), C.byref(lat))


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


def _one_slice(n, K0=1.87, sc_l=1, harmonics=(1,), weight=1e3):
    c = fc.make_case_1d(slice_n=1, harmonics=harmonics, ppslice=n, sc_l=sc_l, K0=K0, wake=False, in_phases=1)
    c.state["seg_weight"][:] = weight
    c.state["Ux"][:] = 0
    c.state["Uy"][:] = 0
    c.state["Ez"][:] = 0
    return c


def test_push_madey_low_gain_curve_1d():
    """Twenty periods of the 1D push (planar, harmonic 1) in a frozen field
    on a quiet start of 32 phases per detuning u = (ku - ks K2/(2 gamma^2)) L.
    Second-order perturbation theory of dgamma/dz = a cos(theta), dtheta/dz =
    u/L + c dgamma + (a/gamma) sin(theta) with c = ks K2/gamma^3 and
    a = ks K0 (J0(xi) - J1(xi)) |U| / (2 gamma) (particlePush.cpp:470-512)
    gives <dgamma> = -(a^2 c L^3/(2u)) (2(1 - cos u)/u^2 - sin u/u)
    - a^2 L^2 (1 - cos u)/(gamma u^2): loss above resonance, gain below
    (the second term, from the field's phase modulation, checked against a
    direct integration of the same two equations). At Omega L = 0.05 the
    push follows it to 1e-4 of the peak (measured 2e-6); a coupling off by a
    factor k moves the curve by k^2."""
    Nu, nph = 20, 32
    us = np.array([-6.0, -4.0, -2.6, -1.0, 1.0, 2.6, 4.0, 6.0])
    c = _one_slice(len(us) * nph)
    p, lat = c.params, c.lattice()
    ks, ku, K0, dz = p.ks, lat.ku, lat.K0, p.dz
    K2 = 1 + K0 * K0 / 2
    L = Nu * dz
    gam = np.sqrt(ks * K2 / (2 * (ku - us / L)))
    c.state["theta"] = np.tile(np.arange(nph) * 2 * math.pi / nph + 0.3, len(us))
    c.state["gamma"] = np.repeat(gam, nph)
    g0 = float(np.sqrt(ks * K2 / (2 * ku)))
    xi0 = 0.25 * K0 * K0 / (1 + K0 * K0 / 2)
    c0 = ks * K2 / g0 ** 3
    Omega = 0.05 / L  # (Omega L)^2: the size of the orders past the second
    U0 = (Omega ** 2 / c0) * 2 * g0 / (ks * K0 * (special.jv(0, xi0) - special.jv(1, xi0)))
    c.state["Ux"][0, 1] = U0
    o = c.oracle_state()
    g_start = o.gamma.copy()
    for _ in range(Nu):
        assert o.push_theta_gamma(lat) == 0
    dg = (o.gamma - g_start).reshape(len(us), nph).mean(1)
    xi = 0.25 * K0 * K0 / (1 + K0 * K0 / 2)  # the 1D push's xi (its gamma is the resonance's)
    a = ks * K0 * (special.jv(0, xi) - special.jv(1, xi)) * U0 / (2 * gam)
    cc = ks * K2 / gam ** 3
    want = (-(a * a * cc * L ** 3 / (2 * us)) * (2 * (1 - np.cos(us)) / us ** 2 - np.sin(us) / us)
            - a * a * L ** 2 * (1 - np.cos(us)) / (gam * us * us))
    assert np.all(np.sign(dg) == -np.sign(us))
    np.testing.assert_allclose(dg, want, rtol=0, atol=1e-4 * np.max(np.abs(want)))


def _energy_balance_1d(dz, harmonics, ue, K0_alpha):
    c = _one_slice(64, harmonics=harmonics, weight=2e3)
    c.params.dz = dz
    lat = c.lattice(ue=ue, K0_alpha=float(K0_alpha), und_type=2 if ue else 1)
    rng = np.random.default_rng(4)
    c.state["theta"] = rng.uniform(0, 2 * math.pi, 64)
    c.state["gamma"] = 8000.0 * (1 + 1e-4 * rng.normal(size=64))
    for h in range(len(harmonics)):
        c.state["Ux"][h, 1] = 3e-6 * np.exp(1j * (0.4 + h))
        c.state["Uy"][h, 1] = 2e-6 * np.exp(1j * (1.3 - h))
    o = c.oracle_state()
    U0x, U0y = o.Ux[:, 1].copy(), o.Uy[:, 1].copy()
    g0 = o.gamma.copy()
    assert _solve_sc_1d(o, lat) == 0
    Sx, Sy = o.ScUx[:, 1].copy(), o.ScUy[:, 1].copy()
    o.Ez[:] = 0                      # the space-charge field is not radiation
    assert o.push_theta_gamma(lat) == 0
    p = c.params
    area = 2 * math.pi * c.state["beam_sig"][0] * c.state["beam_sig"][1]
    # E = m c^2 ks U / e over the beam area, P = |E|^2 area / (2 Z0), times the
    # slice duration
    coef2 = (E_MASS * C_LIGHT ** 2 * p.ks / E_CHARGE) ** 2 / (2 * Z0) * area * p.lambda0 * p.num_slice / C_LIGHT
    cross = np.sum(2 * np.real(np.conj(U0x) * Sx) + 2 * np.real(np.conj(U0y) * Sy))
    dE = np.sum(o.gamma - g0) * c.state["seg_weight"][0] * MC2 * 1e6 * E_CHARGE
    return cross * coef2, dE


import pytest  # noqa: E402


@pytest.mark.parametrize("harmonics,ue,K0_alpha", [((1,), 0.0, 1), ((1, 3), 0.0, 1), ((1, 3), 0.6, 1),
                                                   ((1, 3), 0.6, -1)])
def test_source_and_push_conserve_energy_1d(harmonics, ue, K0_alpha):
    """The work the field does on the beam in the 1D push
    (particlePush.cpp:470-512) equals the field energy the 1D source adds
    (solveField.cpp:470-500, then U += Sc at 409-420): the cross term
    2 Re(U* S) as power over the beam area times the slice duration. Odd
    harmonics, ue != 0 and both K0_alpha signs exercise the coupling's
    branches (J1 -+ K0_alpha J2 in the source against K0_alpha J1 -+ J2 in
    the push). The mismatch is O(dz) (the phase slips within a step; most
    for the elliptical undulator, whose beam here is off resonance) and falls
    tenfold with dz (measured at dz = 2.6e-5: 2e-6 planar, 1.4e-3
    elliptical)."""
    errs = []
    for dz in (0.0026, 0.00026, 0.000026):
        gain, dE = _energy_balance_1d(dz, harmonics, ue, K0_alpha)
        assert abs(dE) > 0
        errs.append(abs(gain / (-dE) - 1))
    assert errs[2] < 3e-3, errs
    assert errs[0] > 5 * errs[1] > 25 * errs[2], errs


def test_space_charge_plasma_oscillation_1d():
    """A cold beam at resonance (gamma with ks K2/(2 gamma^2) = ku), 256
    evenly spaced phases modulated by eps cos(theta0), no radiation field
    (K0 = 0.2, the field left at zero): each step's space-charge mode
    (solveField.cpp:527-546) drives the push (particlePush.cpp:466-468, 509),
    and the bunching b = <e^{-i theta}> = -i eps/2 cos(k_p z) oscillates at
    k_p^2 = K2 e^2 n_e / (eps0 m c^2 gamma^3), n_e the slice's electrons over
    its volume (beam area 2 pi sx sy times lambda0). The fitted wavenumber
    agrees to 1%."""
    n, eps = 256, 0.02
    c = _one_slice(n, K0=0.2)
    p, lat = c.params, c.lattice()
    ks, ku, K0 = p.ks, lat.ku, lat.K0
    K2 = 1 + K0 * K0 / 2
    g0 = math.sqrt(ks * K2 / (2 * ku))
    area = 2 * math.pi * c.state["beam_sig"][0] * c.state["beam_sig"][1]
    vol = area * p.lambda0 * p.num_slice
    # electrons per macro-particle for k_p = pi / (600 dz): half a period in 600 steps
    kp_want = math.pi / (600 * p.dz)
    ne = kp_want ** 2 * EPS0 * E_MASS * C_LIGHT ** 2 * g0 ** 3 / (K2 * E_CHARGE ** 2)
    c.state["seg_weight"][:] = ne * vol / n
    th0 = np.arange(n) * 2 * math.pi / n
    c.state["theta"] = th0 + eps * np.cos(th0)
    c.state["gamma"] = np.full(n, g0)
    o = c.oracle_state()
    z, b = [], []
    for k in range(900):
        assert _solve_sc_1d(o, lat) == 0
        assert o.push_theta_gamma(lat) == 0
        o.Ux[:] = 0
        o.Uy[:] = 0
        z.append((k + 1) * p.dz)
        b.append(np.mean(np.exp(-1j * o.theta)))
    z, b = np.array(z), np.array(b)
    assert np.max(np.abs(b.real)) < 0.05 * np.max(np.abs(b.imag))
    # fit -eps/2 cos(k z + phi): the wavenumber by least squares over a grid
    ks_grid = kp_want * np.linspace(0.8, 1.2, 4001)
    res = []
    for kk in ks_grid:
        A = np.stack([np.cos(kk * z), np.sin(kk * z)], 1)
        coef, r, *_ = np.linalg.lstsq(A, b.imag, rcond=None)
        res.append(float(np.sum((A @ coef - b.imag) ** 2)))
    k_fit = ks_grid[int(np.argmin(res))]
    assert abs(k_fit / kp_want - 1) < 0.01, k_fit / kp_want
    # the amplitude: eps/2, the oscillation neither grows nor decays
    assert abs(np.max(np.abs(b.imag)) / (eps / 2) - 1) < 0.05
