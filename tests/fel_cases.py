"""Small seeded FEL states shared by the oracle tests and the GPU parity tests.

A case holds the grid description (as both the oracle's Params and the
library's Desc), the setup tables, one lattice state and an initial state:
Gaussian beam with quiet-start phases and a Gaussian seed plus a seeded
perturbation, so every coupling term of the push, deposit and ADI matters.
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

from oracle import pyoracle as po  # noqa: E402

C_LIGHT = 299792458.0
PLANK_H = 4.135667696e-15


@dataclass
class Case:
    nx: int
    ny: int
    slice_n: int
    harmonics: tuple
    n_species: int
    params: po.Params
    lat: po.Lattice
    state: dict          # numpy arrays: particles, fields, Ez, wakeE, seg_off, seg_weight
    tables: dict         # alpha, alphaP, beta, betaP, bessel
    desc_kw: dict        # kwargs for felis_amd._lib.Desc

    def oracle_state(self) -> po.OracleState:
        s = {k: np.array(v, copy=True) for k, v in self.state.items()}
        t = self.tables
        return po.OracleState(params=self.params, alpha=t["alpha"], alphaP=t["alphaP"], beta=t["beta"],
                              betaP=t["betaP"], bessel=t["bessel"], **s)

    def lattice(self, **over) -> po.Lattice:
        lat = self.lat.copy()
        for k, v in over.items():
            setattr(lat, k, v)
        return lat


def resonant_photon_energy(gamma0: float, K0: float, lambdaU: float) -> float:
    lam = lambdaU / (2 * gamma0 ** 2) * (1 + K0 ** 2 / 2)
    return PLANK_H * C_LIGHT / lam


def make_case(nx=17, ny=17, slice_n=4, harmonics=(1,), n_species=1, ppslice=64, num_slice=1,
              K0=1.87, lambdaU=0.026, gamma0=8000.0, ue=0.0, K0_alpha=1, und_type=1, g=12.0,
              box=300e-6, sig=30e-6, field_amp=2e-6, seed=1, nr=None, sc_l=1, sc_f=1, bn=2001,
              abc_n=3, abc_sig=1.0, wake=True, uneven=False, dimension=3, in_phases=4) -> Case:
    rng = np.random.default_rng(seed)
    maxh = max(harmonics)
    photon = resonant_photon_energy(gamma0, K0, lambdaU)
    ks = photon * 2 * math.pi / (PLANK_H * C_LIGHT)
    lambda0 = 2 * math.pi / ks
    ku = 2 * math.pi / lambdaU
    dz = lambdaU
    # parameterSetting.cpp:108-111 (deck n -> array n+1) and 175, 194-195
    dx = 2 * box / (nx - 1)
    dy = 2 * box / (ny - 1)
    gamR = math.sqrt(0.5 * ks / ku * (1 + 0.5 * (1 + ue * ue) * K0 * K0))
    if nr is None:
        nr = int(math.sqrt(nx * nx + ny * ny))
    dr = math.sqrt((2 * box) ** 2 + (2 * box) ** 2) / nx
    p = po.Params()
    p.nx, p.ny, p.slice_n, p.n_harm = nx, ny, slice_n, len(harmonics)
    for i, h in enumerate(harmonics):
        p.harmony[i] = h
    p.num_slice, p.n_species, p.nr, p.sc_l, p.sc_f = num_slice, n_species, nr, sc_l, sc_f
    p.bn, p.b_orders, p.wake_on, p.dimension = bn, maxh // 2 + 2, int(wake), dimension
    p.dx, p.dy, p.dz, p.min_x, p.min_y = dx, dy, dz, -box, -box
    p.ks, p.lambda0, p.gamR, p.dr = ks, lambda0, gamR, dr
    p.dBessel = maxh * 2.0 * math.pi / bn
    al, alP, be, beP = po.abc_setup(nx, ny, abc_n, abc_sig, dx, dy, dz, ks)
    bes = po.bessel_table(bn, maxh)

    # particles: segment-major (species, slice)
    counts = []
    for s in range(n_species):
        for sl in range(slice_n):
            n = ppslice
            if uneven:
                n = max(in_phases, int(ppslice * math.exp(-((sl - slice_n / 2) / (slice_n / 3)) ** 2)))
                n -= n % in_phases
            counts.append(n)
    seg_off = np.zeros(len(counts) + 1, np.int64)
    seg_off[1:] = np.cumsum(counts)
    npart = int(seg_off[-1])
    nb = npart // in_phases
    xs = rng.normal(0, sig, nb)
    ys = rng.normal(0, sig, nb)
    sigp = 0.4e-6 / gamma0 / sig * gamma0  # p = gamma * x', emittance 0.4 mm mrad scale
    pxs = rng.normal(0, sigp, nb)
    pys = rng.normal(0, sigp, nb)
    gs = gamma0 * (1 + 1e-4 * rng.normal(size=nb))
    th0 = rng.uniform(0, 2 * math.pi * num_slice, nb)
    rep = lambda a: np.repeat(a, in_phases)
    x, y, px, py, gamma = rep(xs), rep(ys), rep(pxs), rep(pys), rep(gs)
    theta = rep(th0) + np.tile(np.arange(in_phases) * 2 * math.pi / in_phases, nb)
    theta += 1e-3 * rng.normal(size=npart)          # break the quiet start a little
    theta = np.mod(theta, 2 * math.pi * num_slice)
    seg_weight = 1e3 * (1 + 0.1 * rng.random(len(counts)))

    # fields: Gaussian mode plus a seeded perturbation, all slots incl. ghosts
    S = slice_n + 2
    xx = -box + dx * np.arange(nx)
    yy = -box + dy * np.arange(ny)
    X, Y = np.meshgrid(xx, yy)
    gauss = np.exp(-(X ** 2 + Y ** 2) / (4 * (40e-6) ** 2))
    Ux = np.zeros((len(harmonics), S, ny, nx), np.complex128)
    Uy = np.zeros_like(Ux)
    for h in range(len(harmonics)):
        for slot in range(1, slice_n + 1):
            ph = rng.uniform(0, 2 * math.pi)
            Ux[h, slot] = field_amp * gauss * np.exp(1j * ph) * (1 + 0.05 * rng.normal(size=(ny, nx)))
            Uy[h, slot] = 0.3 * field_amp * gauss * np.exp(1j * (ph + 0.7)) * (1 + 0.05 * rng.normal(size=(ny, nx)))
    Ez = np.zeros((sc_l, S * sc_f * nr), np.complex128)
    for l in range(sc_l):
        for slot in range(1, slice_n + 1):
            base = slot * sc_f * nr
            Ez[l, base:base + sc_f * nr] = 1e3 * (rng.normal(size=sc_f * nr) + 1j * rng.normal(size=sc_f * nr))
    wakeE = 1e4 * rng.normal(size=slice_n)
    lat = po.Lattice()
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU = K0, ue, float(K0_alpha), ku, g, lambdaU
    lat.und_type, lat.current_flag, lat.drift_flag = und_type, 1, 0
    state = dict(seg_off=seg_off, seg_weight=seg_weight, x=x, y=y, px=px, py=py, theta=theta, gamma=gamma,
                 Ux=Ux, Uy=Uy, ScUx=np.zeros_like(Ux), ScUy=np.zeros_like(Ux), Ez=Ez, wakeE=wakeE)
    tables = dict(alpha=al, alphaP=alP, beta=be, betaP=beP, bessel=bes)
    desc_kw = dict(dimension=dimension, mode=2, nx=nx, ny=ny, slice_n=slice_n, n_harm=len(harmonics),
                   harmony=list(harmonics), num_slice=num_slice, n_species=n_species, nr=nr, sc_l=sc_l,
                   sc_f=sc_f, bessel_n=bn, abc_n=abc_n, wake_on=int(wake), abc_sig=abc_sig, dx=dx, dy=dy,
                   dz=dz, min_x=-box, min_y=-box, min_z=0.0, max_z=slice_n * lambda0 * num_slice, ks=ks,
                   lambda0=lambda0, gamR=gamR, dr=dr, gamma_ref=gamma0)
    return Case(nx, ny, slice_n, tuple(harmonics), n_species, p, lat, state, tables, desc_kw)


def make_desc(case: Case, precision: int = 0):
    from felis_amd import _lib as fl
    d = fl.Desc()
    for k, v in case.desc_kw.items():
        if k == "harmony":
            for i, h in enumerate(v):
                d.harmony[i] = h
        else:
            setattr(d, k, v)
    d.precision = precision
    return d


def to_lib_lattice(lat: po.Lattice):
    from felis_amd import _lib as fl
    out = fl.Lattice()
    for name, _ in po.Lattice._fields_:
        setattr(out, name, getattr(lat, name))
    return out


def gpu_context(case: Case, precision: int = 0):
    """A library context loaded with the case's initial state."""
    from felis_amd import _lib as fl
    ctx = fl.Context(make_desc(case, precision))
    s = case.state
    ctx.upload_particles(s["seg_off"], s["seg_weight"], s["x"], s["y"], s["px"], s["py"], s["theta"], s["gamma"])
    ctx.upload_field(fl.UX, s["Ux"])
    ctx.upload_field(fl.UY, s["Uy"])
    ctx.upload_field(fl.EZ, s["Ez"])
    ctx.set_wake(s["wakeE"])
    if case.desc_kw["dimension"] == 1:
        ctx.set_beam_sigma(s["beam_sig"])
    return ctx


def rel_err(a, b) -> float:
    a = np.asarray(a); b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


def make_case_1d(slice_n=6, harmonics=(1, 3), ppslice=64, sc_l=2, n_species=1, seed=5, K0=1.87,
                 lambdaU=0.026, gamma0=8000.0, ue=0.0, K0_alpha=1, field_amp=2e-6, in_phases=4,
                 sigma=(30e-6, 25e-6), wake=True, bn=2001) -> Case:
    """A dimension = 1 state: one field cell per slot (parameterSetting.cpp:112,
    196-200), beam sizes for the source area (solveField.cpp:463-468)."""
    rng = np.random.default_rng(seed)
    maxh = max(harmonics)
    photon = resonant_photon_energy(gamma0, K0, lambdaU)
    ks = photon * 2 * math.pi / (PLANK_H * C_LIGHT)
    lambda0 = 2 * math.pi / ks
    ku = 2 * math.pi / lambdaU
    dz = lambdaU
    gamR = math.sqrt(0.5 * ks / ku * (1 + 0.5 * (1 + ue * ue) * K0 * K0))
    p = po.Params()
    p.nx, p.ny, p.slice_n, p.n_harm = 1, 1, slice_n, len(harmonics)
    for i, h in enumerate(harmonics):
        p.harmony[i] = h
    p.num_slice, p.n_species, p.nr, p.sc_l, p.sc_f = 1, n_species, 1, sc_l, 1
    p.bn, p.b_orders, p.wake_on, p.dimension = bn, maxh // 2 + 2, int(wake), 1
    p.dx, p.dy, p.dz, p.min_x, p.min_y = 0.0, 0.0, dz, -300e-6, -300e-6
    p.ks, p.lambda0, p.gamR, p.dr = ks, lambda0, gamR, 0.0
    p.dBessel = maxh * 2.0 * math.pi / bn
    bes = po.bessel_table(bn, maxh)
    counts = [ppslice] * (n_species * slice_n)
    seg_off = np.zeros(len(counts) + 1, np.int64)
    seg_off[1:] = np.cumsum(counts)
    npart = int(seg_off[-1])
    nb = npart // in_phases
    rep = lambda a: np.repeat(a, in_phases)  # noqa: E731
    gamma = rep(gamma0 * (1 + 1e-4 * rng.normal(size=nb)))
    theta = rep(rng.uniform(0, 2 * math.pi, nb)) + np.tile(np.arange(in_phases) * 2 * math.pi / in_phases, nb)
    theta = np.mod(theta + 1e-3 * rng.normal(size=npart), 2 * math.pi)
    x = rep(rng.normal(0, 30e-6, nb)); y = rep(rng.normal(0, 30e-6, nb))
    px = rep(rng.normal(0, 0.5, nb)); py = rep(rng.normal(0, 0.5, nb))
    seg_weight = 1e3 * (1 + 0.1 * rng.random(len(counts)))
    S = slice_n + 2
    nh = len(harmonics)
    Ux = np.zeros((nh, S, 1, 1), np.complex128)
    Uy = np.zeros_like(Ux)
    for h in range(nh):
        for slot in range(1, slice_n + 1):
            Ux[h, slot] = field_amp * np.exp(1j * rng.uniform(0, 2 * math.pi))
            Uy[h, slot] = 0.3 * field_amp * np.exp(1j * rng.uniform(0, 2 * math.pi))
    Ez = np.zeros((sc_l, S), np.complex128)
    Ez[:, 1:slice_n + 1] = 1e3 * (rng.normal(size=(sc_l, slice_n)) + 1j * rng.normal(size=(sc_l, slice_n)))
    wakeE = 1e4 * rng.normal(size=slice_n)
    lat = po.Lattice()
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU = K0, ue, float(K0_alpha), ku, 0.0, lambdaU
    lat.und_type, lat.current_flag, lat.drift_flag = 1, 1, 0
    state = dict(seg_off=seg_off, seg_weight=seg_weight, x=x, y=y, px=px, py=py, theta=theta, gamma=gamma,
                 Ux=Ux, Uy=Uy, ScUx=np.zeros_like(Ux), ScUy=np.zeros_like(Ux), Ez=Ez, wakeE=wakeE,
                 beam_sig=np.tile(np.array(sigma, np.float64), n_species))
    tables = dict(alpha=np.zeros(1, np.complex128), alphaP=np.zeros(1, np.complex128),
                  beta=np.zeros(1, np.complex128), betaP=np.zeros(1, np.complex128), bessel=bes)
    desc_kw = dict(dimension=1, mode=2, nx=1, ny=1, slice_n=slice_n, n_harm=nh, harmony=list(harmonics),
                   num_slice=1, n_species=n_species, nr=1, sc_l=sc_l, sc_f=1, bessel_n=bn, abc_n=1,
                   wake_on=int(wake), abc_sig=0.0, dx=0.0, dy=0.0, dz=dz, min_x=-300e-6, min_y=-300e-6,
                   min_z=0.0, max_z=slice_n * lambda0, ks=ks, lambda0=lambda0, gamR=gamR, dr=0.0,
                   gamma_ref=gamma0)
    return Case(1, 1, slice_n, tuple(harmonics), n_species, p, lat, state, tables, desc_kw)
