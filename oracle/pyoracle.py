"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never imports it.

The state mirrors the reference layouts (see felis_oracle.h): fields are
complex128 arrays of shape (n_harm, slice_n+2, ny, nx), particles are SoA
float64 arrays in segment-major order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libfelisref.so"

MAX_HARM = 8


def build(force: bool = False) -> Path:
    """Compiles liboracle.so with make (gcc)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-B", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def build_ref() -> Path | None:
    """Compiles oracle/_ref from the reference sources when they are present."""
    if not Path("/root/reference/updateTotalEnergy.cpp").exists():
        return None
    subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)
    return REF_PATH


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("slice_n", C.c_int), ("n_harm", C.c_int),
        ("harmony", C.c_int * MAX_HARM),
        ("num_slice", C.c_int), ("n_species", C.c_int),
        ("nr", C.c_int), ("sc_l", C.c_int), ("sc_f", C.c_int),
        ("bn", C.c_int), ("b_orders", C.c_int), ("wake_on", C.c_int), ("dimension", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("min_x", C.c_double),
        ("min_y", C.c_double), ("ks", C.c_double), ("lambda0", C.c_double), ("gamR", C.c_double),
        ("dr", C.c_double), ("dBessel", C.c_double),
    ]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class State(C.Structure):
    _fields_ = [
        ("p", C.POINTER(Params)),
        ("alpha", _dp), ("alphaP", _dp), ("beta", _dp), ("betaP", _dp),
        ("bessel", _dp), ("seg_off", _i64p), ("seg_weight", _dp),
        ("x", _dp), ("y", _dp), ("px", _dp), ("py", _dp), ("theta", _dp), ("gamma", _dp),
        ("Ux", _dp), ("Uy", _dp), ("ScUx", _dp), ("ScUy", _dp), ("Ez", _dp),
        ("wakeE", _dp), ("sc_scratch", _dp), ("beam_sig", _dp),
    ]


class Lattice(C.Structure):
    _fields_ = [
        ("K0", C.c_double), ("ue", C.c_double), ("K0_alpha", C.c_double), ("ku", C.c_double),
        ("g", C.c_double), ("lambdaU", C.c_double),
        ("und_type", C.c_int), ("current_flag", C.c_int), ("drift_flag", C.c_int), ("pad", C.c_int),
    ]

    def copy(self) -> "Lattice":
        n = Lattice()
        C.memmove(C.byref(n), C.byref(self), C.sizeof(Lattice))
        return n


class UndUnit(C.Structure):
    _fields_ = [
        ("unit_start", C.c_double), ("unit_end", C.c_double), ("und_start", C.c_double),
        ("und_end", C.c_double), ("K0", C.c_double), ("slopeK", C.c_double), ("ue", C.c_double),
        ("lambdaU", C.c_double), ("K0_alpha", C.c_int), ("und_type", C.c_int), ("air", C.c_int),
        ("pad", C.c_int),
    ]


class Quad(C.Structure):
    _fields_ = [("q_start", C.c_double), ("q_end", C.c_double), ("g", C.c_double),
                ("list_id", C.c_int), ("pad", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_abc_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_bessel_table.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_update_k_quadg.argtypes = [C.POINTER(Lattice), C.c_double, C.c_int, C.c_double,
                                         C.POINTER(UndUnit), C.c_int, C.POINTER(Quad), C.c_int]
        for name in ("orc_transverse_push", "orc_drift_theta_gamma", "orc_solve_field", "orc_solve_sc_3d",
                     "orc_solve_sc_1d"):
            getattr(L, name).argtypes = [C.POINTER(State), C.POINTER(Lattice)]
            getattr(L, name).restype = C.c_int
        for name in ("orc_push_theta_gamma_3d", "orc_push_theta_gamma_1d"):
            getattr(L, name).argtypes = [C.POINTER(State), C.POINTER(Lattice), C.c_int]
            getattr(L, name).restype = C.c_int
        L.orc_periodic_particles.argtypes = [C.POINTER(State)]
        L.orc_periodic_particles.restype = C.c_int
        L.orc_solve_field_u_3d.argtypes = [C.POINTER(State), _dp, _dp, C.c_int]
        L.orc_shift_field.argtypes = [C.POINTER(State)]
        L.orc_solve_field_u_1d.argtypes = [C.POINTER(State)]
        L.orc_total_energy.argtypes = [C.POINTER(State), _dp, _dp]
        L.orc_bfactor.argtypes = [C.POINTER(State), _dp, _dp]
        L.orc_twiss.argtypes = [C.POINTER(State), _dp, _dp]
        L.orc_wake_function.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, _dp, _dp]
        L.orc_wake_function.restype = C.c_int
        L.orc_update_wake.argtypes = [C.POINTER(State), _dp, _dp, _dp, _dp]
        L.orc_load_seed_3d.argtypes = [C.POINTER(Params), _dp, _dp, C.c_int] + [C.c_double] * 9
        L.orc_run.argtypes = [C.POINTER(State), C.POINTER(Lattice), C.c_int, C.c_int, C.c_int, C.c_int,
                              _dp, _dp, _dp, _dp, _dp]
        L.orc_run.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.int64:
        return a.ctypes.data_as(_i64p)
    assert a.dtype in (np.float64, np.complex128), a.dtype
    return a.ctypes.data_as(_dp)


def abc_setup(nx, ny, abc_n, abc_sig, dx, dy, dz, ks):
    al = np.zeros(nx, np.complex128); alP = np.zeros(nx, np.complex128)
    be = np.zeros(ny, np.complex128); beP = np.zeros(ny, np.complex128)
    lib().orc_abc_setup(nx, ny, abc_n, abc_sig, dx, dy, dz, ks, ptr(al), ptr(alP), ptr(be), ptr(beP))
    return al, alP, be, beP


def bessel_table(bn: int, max_h: int) -> np.ndarray:
    orders = max_h // 2 + 2
    t = np.zeros((bn, orders), np.float64)
    lib().orc_bessel_table(bn, max_h, ptr(t))
    return t


@dataclass
class OracleState:
    """numpy-owned oracle state; .st is the ctypes view handed to the C code."""
    params: Params
    alpha: np.ndarray
    alphaP: np.ndarray
    beta: np.ndarray
    betaP: np.ndarray
    bessel: np.ndarray
    seg_off: np.ndarray
    seg_weight: np.ndarray
    x: np.ndarray
    y: np.ndarray
    px: np.ndarray
    py: np.ndarray
    theta: np.ndarray
    gamma: np.ndarray
    Ux: np.ndarray
    Uy: np.ndarray
    ScUx: np.ndarray
    ScUy: np.ndarray
    Ez: np.ndarray
    wakeE: np.ndarray
    beam_sig: np.ndarray = field(default_factory=lambda: np.zeros(2))
    sc_scratch: np.ndarray = None
    st: State = None

    def __post_init__(self):
        p = self.params
        self.sc_scratch = np.zeros(p.sc_l * (p.sc_f * p.nr + 1), np.complex128)
        self.refresh()

    def refresh(self):
        s = State()
        s.p = C.pointer(self.params)
        for name in ("alpha", "alphaP", "beta", "betaP", "bessel", "seg_off", "seg_weight", "x", "y", "px",
                     "py", "theta", "gamma", "Ux", "Uy", "ScUx", "ScUy", "Ez", "wakeE", "sc_scratch",
                     "beam_sig"):
            setattr(s, name, ptr(getattr(self, name)))
        self.st = s

    # -- the reference operators ------------------------------------------------
    def transverse_push(self, lat: Lattice) -> int:
        return lib().orc_transverse_push(C.byref(self.st), C.byref(lat))

    def push_theta_gamma(self, lat: Lattice, it: int = 0) -> int:
        f = lib().orc_push_theta_gamma_3d if self.params.dimension == 3 else lib().orc_push_theta_gamma_1d
        return f(C.byref(self.st), C.byref(lat), it)

    def drift_theta_gamma(self, lat: Lattice) -> int:
        return lib().orc_drift_theta_gamma(C.byref(self.st), C.byref(lat))

    def periodic(self) -> int:
        return lib().orc_periodic_particles(C.byref(self.st))

    def solve_field(self, lat: Lattice) -> int:
        return lib().orc_solve_field(C.byref(self.st), C.byref(lat))

    def solve_sc(self, lat: Lattice) -> int:
        return lib().orc_solve_sc_3d(C.byref(self.st), C.byref(lat))

    def solve_u(self, which: str, current_flag: int) -> int:
        U = self.Ux if which == "x" else self.Uy
        Sc = self.ScUx if which == "x" else self.ScUy
        return lib().orc_solve_field_u_3d(C.byref(self.st), ptr(U), ptr(Sc), int(current_flag))

    def shift(self) -> None:
        lib().orc_shift_field(C.byref(self.st))

    def total_energy(self):
        nh = self.params.n_harm
        ex = np.zeros(nh); ey = np.zeros(nh)
        lib().orc_total_energy(C.byref(self.st), ptr(ex), ptr(ey))
        return ex, ey

    def bfactor(self):
        nh = self.params.n_harm
        b = np.zeros(2 * nh); tot = np.zeros(1)
        lib().orc_bfactor(C.byref(self.st), ptr(b), ptr(tot))
        return b[0::2] + 1j * b[1::2], float(tot[0])

    def twiss(self):
        ns = self.params.n_species
        sums = np.zeros(7 * ns); der = np.zeros(6 * ns)
        lib().orc_twiss(C.byref(self.st), ptr(sums), ptr(der))
        return sums.reshape(ns, 7), der.reshape(ns, 6)

    def update_wake(self, wakeF: np.ndarray):
        sn = self.params.slice_n
        den = np.zeros(sn); tot = np.zeros(1)
        lib().orc_update_wake(C.byref(self.st), ptr(np.ascontiguousarray(wakeF, np.float64)), ptr(den),
                              ptr(self.wakeE), ptr(tot))
        return den, float(tot[0])

    def run(self, lat_tab, it0: int, n_steps: int, time_dependent: bool, wake_step: int = 0,
            wakeF: np.ndarray | None = None):
        p = self.params
        nh, ns = p.n_harm, p.n_species
        energy = np.zeros((n_steps, 2 * nh)); bfac = np.zeros((n_steps, nh + 1))
        twiss = np.zeros((n_steps, 7 * ns))
        arr = (Lattice * len(lat_tab))(*lat_tab)
        wf = None if wakeF is None else np.ascontiguousarray(wakeF, np.float64)
        rc = lib().orc_run(C.byref(self.st), arr, it0, n_steps, int(time_dependent), wake_step, ptr(wf),
                           ptr(self.wakeE), ptr(energy), ptr(bfac), ptr(twiss))
        return rc, energy, bfac, twiss


def wake_function(shape: int, acdc: int, radius: float, cond: float, ctau: float, slice_n: int, dZ: float):
    impZ = np.zeros(1001); wakeF = np.zeros(slice_n)
    rc = lib().orc_wake_function(shape, acdc, radius, cond, ctau, slice_n, dZ, ptr(impZ), ptr(wakeF))
    return rc, impZ, wakeF


def update_k_quadg(lat: Lattice, dz: float, it: int, half: float, units, quads) -> Lattice:
    ua = (UndUnit * max(1, len(units)))(*units)
    qa = (Quad * max(1, len(quads)))(*quads)
    lib().orc_update_k_quadg(C.byref(lat), dz, it, half, ua, len(units), qa, len(quads))
    return lat


def load_seed_3d(params: Params, Ux, Uy, h_pick, a0, duration, min_z, zR, focus, spotSigR, polarity,
                 laserAlpha, laserPsi):
    lib().orc_load_seed_3d(C.byref(params), ptr(Ux), ptr(Uy), h_pick, a0, duration, min_z, zR, focus,
                           spotSigR, polarity, laserAlpha, laserPsi)


# ------------------------------------------------------------------ reference
_ref = None


def ref():
    """oracle/_ref/libfelisref.so (the reference's own TUs), or None."""
    global _ref
    if _ref is None and REF_PATH.exists():
        R = C.CDLL(str(REF_PATH))
        _ref = R
    return _ref


def state_from_desc(d, seg_off, seg_w, particles, Ux, Uy) -> "OracleState":
    """An oracle state for a felis_desc-like object `d` (fields of
    include/felis.h) with the given beam and field (copied)."""
    p = Params()
    p.nx, p.ny, p.slice_n, p.n_harm = d.nx, d.ny, d.slice_n, d.n_harm
    for h in range(d.n_harm):
        p.harmony[h] = d.harmony[h]
    maxh = d.harmony[d.n_harm - 1]
    p.num_slice, p.n_species, p.nr, p.sc_l, p.sc_f = d.num_slice, d.n_species, d.nr, d.sc_l, d.sc_f
    p.bn, p.b_orders, p.wake_on, p.dimension = d.bessel_n, maxh // 2 + 2, d.wake_on, d.dimension
    p.dx, p.dy, p.dz, p.min_x, p.min_y = d.dx, d.dy, d.dz, d.min_x, d.min_y
    p.ks, p.lambda0, p.gamR, p.dr = d.ks, d.lambda0, d.gamR, d.dr
    p.dBessel = maxh * 2.0 * np.pi / d.bessel_n
    al, alP, be, beP = abc_setup(d.nx, d.ny, d.abc_n, d.abc_sig, d.dx, d.dy, d.dz, d.ks)
    bes = bessel_table(d.bessel_n, maxh)
    S = d.slice_n + 2
    return OracleState(params=p, alpha=al, alphaP=alP, beta=be, betaP=beP, bessel=bes,
                       seg_off=np.ascontiguousarray(seg_off, np.int64),
                       seg_weight=np.ascontiguousarray(seg_w, np.float64),
                       x=particles["x"].astype(np.float64), y=particles["y"].astype(np.float64),
                       px=particles["px"].astype(np.float64), py=particles["py"].astype(np.float64),
                       theta=particles["theta"].astype(np.float64), gamma=particles["gamma"].astype(np.float64),
                       Ux=np.ascontiguousarray(Ux, np.complex128).copy(),
                       Uy=np.ascontiguousarray(Uy, np.complex128).copy(),
                       ScUx=np.zeros((d.n_harm, S, d.ny, d.nx), np.complex128),
                       ScUy=np.zeros((d.n_harm, S, d.ny, d.nx), np.complex128),
                       Ez=np.zeros((d.sc_l, S * d.sc_f * d.nr), np.complex128), wakeE=np.zeros(d.slice_n))


def lattice_table(tab) -> list:
    """Oracle lattices from any objects with the felis_lattice fields."""
    out = []
    for t in tab:
        lat = Lattice()
        for name, _ in Lattice._fields_:
            setattr(lat, name, getattr(t, name))
        out.append(lat)
    return out
