"""ctypes bindings of libfelis_hip.so (the C-ABI declared in include/felis.h).

The product path: every call goes to the gfx950 library. Loading fails loudly
when the library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libfelis_hip.so"
HDF_PATH = PKG / "libfelis_hdf.so"
MAX_HARM = 8

OK = 0
ERR_ARG, ERR_HIP, ERR_GUARD, ERR_RANGE, ERR_NAN, ERR_UNSUPPORTED, ERR_STATE, ERR_IO = -1, -2, -3, -4, -5, -6, -7, -8
PLAIN_FP64, MIXED_FP32 = 0, 1
DEPOSIT_ATOMIC, DEPOSIT_FIXED = 0, 1
UX, UY, SCUX, SCUY, EZ = 0, 1, 2, 3, 4

# every entry point include/felis.h declares
EXPORTS = [
    "felis_create", "felis_destroy", "felis_last_error", "felis_sync", "felis_get_tables",
    "felis_upload_particles", "felis_download_particles", "felis_upload_field", "felis_download_field",
    "felis_slice_energy", "felis_solve_field", "felis_transverse_push", "felis_push_theta_gamma",
    "felis_drift_theta_gamma", "felis_periodic_particles", "felis_shift_field", "felis_update_wake",
    "felis_set_wake", "felis_update_total_energy", "felis_update_bfactor", "felis_calculate_twiss",
    "felis_run", "felis_run_noalloc", "felis_find_parameter", "felis_deck_open", "felis_deck_close", "felis_deck_error",
    "felis_deck_desc", "felis_deck_steps", "felis_deck_lattice", "felis_deck_wake_function",
    "felis_deck_beam_layout", "felis_deck_load_beam", "felis_deck_load_seed", "felis_profile",
    "felis_profile_read", "felis_device_info", "felis_set_beam_sigma", "felis_deck_beam_sigma",
    "felis_mlp_create", "felis_mlp_destroy", "felis_mlp_last_error", "felis_mlp_get_layer", "felis_mlp_set_layer",
    "felis_mlp_forward", "felis_mlp_train_step", "felis_mlp_get_grad", "felis_gemm_bf16",
    "felis_run_surrogate", "felis_state_save", "felis_state_restore", "felis_slice_moments",
    "felis_jump_residual", "felis_slice_features", "felis_mlp_shape", "felis_mlp_set_input_affine",
    "felis_mlp_get_input_affine", "felis_jump_emission_residual", "felis_live_polarisations",
    "felis_set_tuning", "felis_get_tuning", "felis_set_deposit_mode", "felis_get_deposit_mode", "felis_get_desc", "felis_particle_layout", "felis_deck_save_settings",
    "felis_field_norms", "felis_field_diag", "felis_wake_function", "felis_deck_wake_function_device",
    "felis_fno_create", "felis_fno_destroy", "felis_fno_last_error", "felis_fno_shape", "felis_fno_param_count",
    "felis_fno_get_params", "felis_fno_set_params", "felis_fno_get_grad", "felis_fno_set_input_affine",
    "felis_fno_forward", "felis_fno_train_step", "felis_fno_jump_maps",
]
# every entry point include/felis_hdf.h declares (libfelis_hdf.so)
HDF_EXPORTS = ["felis_save_field_hdf", "felis_restore_field_hdf", "felis_save_particles_hdf",
               "felis_restore_particles_hdf", "felis_hdf_last_error"]


def device_info() -> dict:
    info = (C.c_int * 8)()
    rc = load().felis_device_info(info)
    if rc != OK:
        raise FelisError(rc, "felis_device_info")
    keys = ["cus", "lds_per_block", "lds_per_cu", "wave", "l2_bytes", "clock_khz", "mem_clock_khz", "bus_bits"]
    return dict(zip(keys, list(info)))

# felis_kernel_id order
KERNELS = ["transverse", "push", "drift", "periodic", "deposit", "ez_deposit", "ez_solve", "adi_x", "adi_y",
           "shift", "energy", "bfactor", "twiss", "wake", "zero", "particle_step", "adi_fused", "jump_window"]


class FelisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"felis error {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [
        ("dimension", C.c_int), ("mode", C.c_int), ("precision", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("slice_n", C.c_int), ("n_harm", C.c_int),
        ("harmony", C.c_int * MAX_HARM),
        ("num_slice", C.c_int), ("n_species", C.c_int), ("nr", C.c_int), ("sc_l", C.c_int), ("sc_f", C.c_int),
        ("bessel_n", C.c_int), ("abc_n", C.c_int), ("wake_on", C.c_int), ("pad0", C.c_int),
        ("abc_sig", C.c_double), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("min_x", C.c_double), ("min_y", C.c_double), ("min_z", C.c_double), ("max_z", C.c_double),
        ("ks", C.c_double), ("lambda0", C.c_double), ("gamR", C.c_double), ("dr", C.c_double),
        ("gamma_ref", C.c_double),
    ]


class Lattice(C.Structure):
    _fields_ = [
        ("K0", C.c_double), ("ue", C.c_double), ("K0_alpha", C.c_double), ("ku", C.c_double),
        ("g", C.c_double), ("lambdaU", C.c_double),
        ("und_type", C.c_int), ("current_flag", C.c_int), ("drift_flag", C.c_int), ("pad", C.c_int),
    ]


class Tuning(C.Structure):
    _fields_ = [("step_graph", C.c_int), ("deposit_lds", C.c_int), ("deposit_threads", C.c_int),
                ("adi_threads", C.c_int), ("deposit_run", C.c_int), ("adi_lanes", C.c_int)]


class JumpOpts(C.Structure):
    _fields_ = [("periods", C.c_int), ("max_jumps", C.c_int), ("bunching_max", C.c_double),
                ("spread_max", C.c_double), ("net", C.c_void_p), ("net_bunching_max", C.c_double),
                ("fno", C.c_void_p)]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_lib = None
_hdf = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built: run python -m felis_amd.build")
        L = C.CDLL(str(LIB_PATH))
        L.felis_create.argtypes = [C.POINTER(Desc), C.POINTER(C.c_void_p)]
        L.felis_destroy.argtypes = [C.c_void_p]
        L.felis_last_error.argtypes = [C.c_void_p]
        L.felis_last_error.restype = C.c_char_p
        L.felis_sync.argtypes = [C.c_void_p]
        L.felis_get_tables.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.felis_upload_particles.argtypes = [C.c_void_p, _i64p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.felis_download_particles.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.felis_upload_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.felis_download_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.felis_slice_energy.argtypes = [C.c_void_p, _dp]
        L.felis_solve_field.argtypes = [C.c_void_p, C.POINTER(Lattice)]
        L.felis_transverse_push.argtypes = [C.c_void_p, C.POINTER(Lattice)]
        L.felis_push_theta_gamma.argtypes = [C.c_void_p, C.POINTER(Lattice), C.c_int]
        L.felis_drift_theta_gamma.argtypes = [C.c_void_p, C.POINTER(Lattice), C.c_int]
        L.felis_periodic_particles.argtypes = [C.c_void_p, C.c_int]
        L.felis_shift_field.argtypes = [C.c_void_p, C.c_int]
        L.felis_update_wake.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.felis_set_wake.argtypes = [C.c_void_p, _dp]
        L.felis_update_total_energy.argtypes = [C.c_void_p, _dp, _dp]
        L.felis_update_bfactor.argtypes = [C.c_void_p, _dp, _dp]
        L.felis_calculate_twiss.argtypes = [C.c_void_p, _dp]
        L.felis_run.argtypes = [C.c_void_p, C.POINTER(Lattice), C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.felis_run_noalloc.argtypes = [C.c_void_p, C.POINTER(Lattice), C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                        _dp]
        L.felis_state_save.argtypes = [C.c_void_p, C.c_int]
        L.felis_state_restore.argtypes = [C.c_void_p, C.c_int]
        L.felis_slice_moments.argtypes = [C.c_void_p, _dp]
        L.felis_jump_residual.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.felis_jump_emission_residual.argtypes = [C.c_void_p, C.c_int, _dp]
        L.felis_live_polarisations.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.felis_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.felis_get_desc.argtypes = [C.c_void_p, C.POINTER(Desc)]
        L.felis_particle_layout.argtypes = [C.c_void_p, _i64p, _dp, _i64p]
        L.felis_deck_save_settings.argtypes = [C.c_void_p, _dp, _dp]
        L.felis_field_norms.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.felis_wake_function.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                          _dp, _dp]
        L.felis_deck_wake_function_device.argtypes = [C.c_void_p, _dp, _dp]
        L.felis_field_diag.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.felis_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.felis_set_deposit_mode.argtypes = [C.c_void_p, C.c_int]
        L.felis_get_deposit_mode.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.felis_slice_features.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.felis_mlp_shape.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.felis_mlp_set_input_affine.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.felis_mlp_get_input_affine.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.felis_run_surrogate.argtypes = [C.c_void_p, C.POINTER(Lattice), C.c_int, C.c_int, C.c_int, _dp,
                                          C.POINTER(JumpOpts), _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.felis_find_parameter.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.felis_deck_open.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.felis_deck_close.argtypes = [C.c_void_p]
        L.felis_deck_error.argtypes = [C.c_void_p]
        L.felis_deck_error.restype = C.c_char_p
        L.felis_deck_desc.argtypes = [C.c_void_p, C.c_int, C.POINTER(Desc)]
        L.felis_deck_steps.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.felis_deck_lattice.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Lattice)]
        L.felis_set_beam_sigma.argtypes = [C.c_void_p, _dp]
        _fp = C.POINTER(C.c_float)
        L.felis_mlp_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]
        L.felis_mlp_destroy.argtypes = [C.c_void_p]
        L.felis_mlp_last_error.argtypes = [C.c_void_p]
        L.felis_mlp_last_error.restype = C.c_char_p
        L.felis_mlp_get_layer.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
        L.felis_mlp_set_layer.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
        L.felis_mlp_forward.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
        L.felis_mlp_train_step.argtypes = [C.c_void_p, C.c_int, _fp, _fp, _fp, C.c_float, _dp]
        L.felis_fno_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                       C.POINTER(C.c_void_p)]
        L.felis_fno_destroy.argtypes = [C.c_void_p]
        L.felis_fno_last_error.argtypes = [C.c_void_p]
        L.felis_fno_last_error.restype = C.c_char_p
        L.felis_fno_shape.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.felis_fno_param_count.argtypes = [C.c_void_p]
        L.felis_fno_param_count.restype = C.c_int64
        L.felis_fno_get_params.argtypes = [C.c_void_p, _fp]
        L.felis_fno_set_params.argtypes = [C.c_void_p, _fp]
        L.felis_fno_get_grad.argtypes = [C.c_void_p, _fp]
        L.felis_fno_set_input_affine.argtypes = [C.c_void_p, _fp, _fp]
        L.felis_fno_forward.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
        L.felis_fno_train_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                           C.c_float, _dp]
        L.felis_fno_jump_maps.argtypes = [C.c_void_p, C.c_int, _fp, C.c_int, C.POINTER(C.c_int), _fp, _fp, _fp, _fp]
        L.felis_mlp_get_grad.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
        L.felis_gemm_bf16.argtypes = [C.c_int, C.c_int, C.c_int, _fp, _fp, _fp]
        L.felis_deck_beam_sigma.argtypes = [C.c_void_p, _dp]
        L.felis_deck_wake_function.argtypes = [C.c_void_p, _dp, _dp]
        L.felis_deck_beam_layout.argtypes = [C.c_void_p, _i64p, _dp]
        L.felis_deck_load_beam.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.felis_deck_load_seed.argtypes = [C.c_void_p, C.c_void_p]
        L.felis_profile.argtypes = [C.c_void_p, C.c_int]
        L.felis_profile_read.argtypes = [C.c_void_p, _dp, _i64p]
        L.felis_device_info.argtypes = [C.POINTER(C.c_int)]
        _lib = L
    return _lib


def wake_function(shape: int, acdc: int, radius: float, cond: float, ctau: float, slice_n: int, dZ: float):
    """wakeField.cpp:53-159 on the device for explicit settings: (wakeF, impZ)."""
    wf, imp = np.zeros(slice_n), np.zeros(1001)
    rc = load().felis_wake_function(int(shape), int(acdc), float(radius), float(cond), float(ctau), int(slice_n),
                                    float(dZ), _p(wf), _p(imp))
    if rc != OK:
        raise FelisError(rc, "felis_wake_function")
    return wf, imp


def load_hdf() -> C.CDLL:
    """libfelis_hdf.so: the HDF5 dump/restore of a context (reference layout)."""
    global _hdf
    if _hdf is None:
        load()
        if not HDF_PATH.exists():
            raise FileNotFoundError(f"{HDF_PATH} is not built: run python -m felis_amd.build (needs HDF5)")
        H = C.CDLL(str(HDF_PATH))
        H.felis_save_field_hdf.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        H.felis_restore_field_hdf.argtypes = [C.c_void_p, C.c_char_p]
        H.felis_save_particles_hdf.argtypes = [C.c_void_p, C.c_char_p, _dp]
        H.felis_restore_particles_hdf.argtypes = [C.c_void_p, C.c_char_p]
        H.felis_hdf_last_error.restype = C.c_char_p
        _hdf = H
    return _hdf


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    if a.dtype == np.int64:
        return a.ctypes.data_as(_i64p)
    assert a.dtype in (np.float64, np.complex128), a.dtype
    return a.ctypes.data_as(_dp)


class Context:
    """One device context (the reference Domain on the GPU)."""

    def __init__(self, desc: Desc):
        self.L = load()
        self.desc = desc
        h = C.c_void_p()
        rc = self.L.felis_create(C.byref(desc), C.byref(h))
        if rc != OK:
            raise FelisError(rc, "felis_create failed")
        self.h = h
        d = desc
        self.field_shape = (d.n_harm, d.slice_n + 2, d.ny, d.nx)
        self.ez_shape = (d.sc_l, (d.slice_n + 2) * d.sc_f * d.nr)
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.felis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc != OK:
            raise FelisError(rc, self.L.felis_last_error(self.h).decode())

    # -- state -------------------------------------------------------------------
    def tables(self):
        d = self.desc
        al = np.zeros(d.nx, np.complex128); alP = np.zeros(d.nx, np.complex128)
        be = np.zeros(d.ny, np.complex128); beP = np.zeros(d.ny, np.complex128)
        maxh = d.harmony[d.n_harm - 1]
        bes = np.zeros((d.bessel_n, maxh // 2 + 2))
        self._chk(self.L.felis_get_tables(self.h, _p(al), _p(alP), _p(be), _p(beP), _p(bes)))
        return al, alP, be, beP, bes

    def upload_particles(self, seg_off, seg_weight, x, y, px, py, theta, gamma):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, px, py, theta, gamma)]
        so = np.ascontiguousarray(seg_off, np.int64)
        sw = np.ascontiguousarray(seg_weight, np.float64)
        self.np = int(so[-1])
        self._chk(self.L.felis_upload_particles(self.h, _p(so), _p(sw), *[_p(a) for a in arrs]))

    def download_particles(self):
        out = [np.zeros(self.np) for _ in range(6)]
        self._chk(self.L.felis_download_particles(self.h, *[_p(a) for a in out]))
        x, y, px, py, th, ga = out
        return dict(x=x, y=y, px=px, py=py, theta=th, gamma=ga)

    def upload_field(self, which: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, np.complex128)
        self._chk(self.L.felis_upload_field(self.h, which, _p(a)))

    def download_field(self, which: int) -> np.ndarray:
        shape = self.ez_shape if which == EZ else self.field_shape
        a = np.zeros(shape, np.complex128)
        self._chk(self.L.felis_download_field(self.h, which, _p(a)))
        return a

    def set_wake(self, wakeE):
        self._chk(self.L.felis_set_wake(self.h, _p(np.ascontiguousarray(wakeE, np.float64))))

    def set_beam_sigma(self, sig):
        """dimension = 1: rms beam sizes [n_species][2] of the source area."""
        self._chk(self.L.felis_set_beam_sigma(self.h, _p(np.ascontiguousarray(sig, np.float64).ravel())))

    # -- operators ---------------------------------------------------------------
    def solve_field(self, lat: Lattice):
        self._chk(self.L.felis_solve_field(self.h, C.byref(lat)))

    def transverse_push(self, lat: Lattice):
        self._chk(self.L.felis_transverse_push(self.h, C.byref(lat)))

    def push_theta_gamma(self, lat: Lattice, it: int = 0):
        self._chk(self.L.felis_push_theta_gamma(self.h, C.byref(lat), it))

    def drift_theta_gamma(self, lat: Lattice, it: int = 0):
        self._chk(self.L.felis_drift_theta_gamma(self.h, C.byref(lat), it))

    def periodic(self, it: int = 0):
        self._chk(self.L.felis_periodic_particles(self.h, it))

    def shift(self, it: int = 0):
        self._chk(self.L.felis_shift_field(self.h, it))

    def update_wake(self, wakeF):
        sn = self.desc.slice_n
        wE = np.zeros(sn); den = np.zeros(sn); tot = np.zeros(1)
        self._chk(self.L.felis_update_wake(self.h, _p(np.ascontiguousarray(wakeF, np.float64)), _p(wE), _p(den),
                                           _p(tot)))
        return wE, den, float(tot[0])

    def total_energy(self):
        nh = self.desc.n_harm
        ex = np.zeros(nh); ey = np.zeros(nh)
        self._chk(self.L.felis_update_total_energy(self.h, _p(ex), _p(ey)))
        return ex, ey

    def slice_energy(self):
        d = self.desc
        out = np.zeros((2, d.n_harm, d.slice_n))
        self._chk(self.L.felis_slice_energy(self.h, _p(out)))
        return out

    def bfactor(self):
        nh = self.desc.n_harm
        b = np.zeros(2 * nh); tot = np.zeros(1)
        self._chk(self.L.felis_update_bfactor(self.h, _p(b), _p(tot)))
        return b[0::2] + 1j * b[1::2], float(tot[0])

    def twiss(self):
        ns = self.desc.n_species
        s = np.zeros(7 * ns)
        self._chk(self.L.felis_calculate_twiss(self.h, _p(s)))
        return s.reshape(ns, 7)

    def run(self, lat_tab, it0: int, n_steps: int, wake_step: int = 0, wakeF=None, outputs: bool = True):
        d = self.desc
        arr = (Lattice * len(lat_tab))(*lat_tab)
        e = np.zeros((n_steps, 2 * d.n_harm)) if outputs else None
        b = np.zeros((n_steps, 2 * d.n_harm + 1)) if outputs else None
        t = np.zeros((n_steps, 7 * d.n_species)) if outputs else None
        wf = None if wakeF is None else np.ascontiguousarray(wakeF, np.float64)
        self._chk(self.L.felis_run(self.h, arr, it0, n_steps, wake_step, _p(wf), _p(e), _p(b), _p(t)))
        return e, b, t

    def run_surrogate(self, lat_tab, it0: int, n_steps: int, periods: int, bunching_max: float = 0.05,
                      max_jumps: int = -1, wake_step: int = 0, wakeF=None, net=None, spread_max: float = 0.0,
                      net_bunching_max: float = 0.0, fno=None):
        """felis_run with multi-period jumps while |b1| < bunching_max (and
        the rms relative energy spread < spread_max when that is > 0); net is
        the learned emission correction (an MLP) or None, applied while
        |b1| < net_bunching_max (<= 0: always); fno (an FNO) the
        field-to-field correction in its place.
        Returns energy, bfac, twiss (NaN rows inside jumps) and the step kinds."""
        d = self.desc
        arr = (Lattice * len(lat_tab))(*lat_tab)
        e = np.zeros((n_steps, 2 * d.n_harm))
        b = np.zeros((n_steps, 2 * d.n_harm + 1))
        t = np.zeros((n_steps, 7 * d.n_species))
        kind = np.zeros(max(1, n_steps), np.int32)
        o = JumpOpts(int(periods), int(max_jumps), float(bunching_max), float(spread_max),
                     net.h if net is not None else None, float(net_bunching_max), fno.h if fno is not None else None)
        wf = None if wakeF is None else np.ascontiguousarray(wakeF, np.float64)
        self._chk(self.L.felis_run_surrogate(self.h, arr, it0, n_steps, wake_step, _p(wf), C.byref(o), _p(e), _p(b),
                                             _p(t), kind.ctypes.data_as(C.POINTER(C.c_int))))
        return e, b, t, kind[:n_steps]

    def save_state(self, slot: int = 0):
        self._chk(self.L.felis_state_save(self.h, slot))

    def restore_state(self, slot: int = 0):
        self._chk(self.L.felis_state_restore(self.h, slot))

    def slice_moments(self):
        d = self.desc
        out = np.zeros((d.slice_n, 9 + 8 * d.n_harm))
        self._chk(self.L.felis_slice_moments(self.h, _p(out)))
        return out

    def slice_features(self):
        """The correction network's inputs of every slice (float32)."""
        d = self.desc
        out = np.zeros((d.slice_n, 5 + 5 * d.n_harm), np.float32)
        self._chk(self.L.felis_slice_features(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def jump_residual(self, slot: int):
        """Current state (a) against snapshot `slot` (b): field sums
        [2*n_harm, slice_n, 4] and particle least-squares sums [slice_n, 12]."""
        d = self.desc
        f = np.zeros((2 * d.n_harm, d.slice_n, 4))
        p = np.zeros((d.slice_n, 12))
        self._chk(self.L.felis_jump_residual(self.h, slot, _p(f), _p(p)))
        return f, p

    def fno_jump_maps(self, slot: int, features, rows):
        """After a jump (state a) against snapshot `slot` (b): for the slices
        in rows, the field-to-field network's input maps X [n][4096][cin],
        the emission E and residual R [n][4096][2 n_harm] (in units of the
        emission's rms per harmonic) and info [n][3][n_harm] (felis_fno_jump_maps)."""
        d = self.desc
        nh = d.n_harm
        rows = np.ascontiguousarray(rows, np.int32)
        n = len(rows)
        X = np.zeros((n, 4096, 4 * nh + 12), np.float32)
        E = np.zeros((n, 4096, 2 * nh), np.float32)
        R = np.zeros((n, 4096, 2 * nh), np.float32)
        info = np.zeros((n, 3, nh), np.float32)
        f, pf = _f32(features)
        self._chk(self.L.felis_fno_jump_maps(self.h, slot, pf, n, rows.ctypes.data_as(C.POINTER(C.c_int)),
                                             _f32(X)[1], _f32(E)[1], _f32(R)[1], _f32(info)[1]))
        return X, E, R, info

    def jump_emission_residual(self, slot: int):
        """After a jump (state a) against snapshot `slot` (b): per (pol x
        harmonic, slice) sum conj(E)(b - a) (re, im) and sum |E|^2, E the
        jump's emission into that slice's slot, [2*n_harm, slice_n, 3]."""
        d = self.desc
        out = np.zeros((2 * d.n_harm, d.slice_n, 3))
        self._chk(self.L.felis_jump_emission_residual(self.h, slot, _p(out)))
        return out

    def set_tuning(self, step_graph: int = 0, deposit_lds: int = 0, deposit_threads: int = 0, adi_threads: int = 0,
                   deposit_run: int = 0, adi_lanes: int = 0):
        """Launch shapes of the batched step (felis_tuning)."""
        t = Tuning(int(step_graph), int(deposit_lds), int(deposit_threads), int(adi_threads), int(deposit_run),
                   int(adi_lanes))
        self._chk(self.L.felis_set_tuning(self.h, C.byref(t)))

    def set_deposit_mode(self, mode: int):
        """DEPOSIT_ATOMIC or DEPOSIT_FIXED (felis_set_deposit_mode)."""
        self._chk(self.L.felis_set_deposit_mode(self.h, int(mode)))

    def deposit_mode(self) -> int:
        m = C.c_int()
        self._chk(self.L.felis_get_deposit_mode(self.h, C.byref(m)))
        return m.value

    def tuning(self) -> dict:
        t = Tuning()
        self._chk(self.L.felis_get_tuning(self.h, C.byref(t)))
        return dict(step_graph=t.step_graph, deposit_lds=t.deposit_lds, deposit_threads=t.deposit_threads,
                    adi_threads=t.adi_threads, deposit_run=t.deposit_run, adi_lanes=t.adi_lanes)

    def field_norms(self, spot_sig_r: float = 0.0):
        """(powerNorm, intensityNorm) of the field dump."""
        pn, inn = np.zeros(1), np.zeros(1)
        self._chk(self.L.felis_field_norms(self.h, float(spot_sig_r), _p(pn), _p(inn)))
        return float(pn[0]), float(inn[0])

    def field_diag(self, spot_sig_r: float = 0.0):
        """felField on the live field: power [2][n_harm][slice_n] (W) and
        fluence [2][n_harm][nx*ny] (J/m^2)."""
        d = self.desc
        p = np.zeros((2, d.n_harm, d.slice_n))
        f = np.zeros((2, d.n_harm, d.nx * d.ny))
        self._chk(self.L.felis_field_diag(self.h, float(spot_sig_r), _p(p), _p(f)))
        return p, f

    def particle_layout(self):
        """(seg_off, seg_weight) of the last particle upload."""
        d = self.desc
        so = np.zeros(d.n_species * d.slice_n + 1, np.int64)
        sw = np.zeros(d.n_species * d.slice_n)
        n = np.zeros(1, np.int64)
        self._chk(self.L.felis_particle_layout(self.h, _p(so), _p(sw), _p(n)))
        return so, sw

    def _hdf(self, rc: int):
        if rc != OK:
            raise FelisError(rc, load_hdf().felis_hdf_last_error().decode())

    def save_field_hdf(self, path, spot_sig_r: float = 0.0):
        """saveFieldHDF: field<it>.h5 in the reference's layout."""
        self._hdf(load_hdf().felis_save_field_hdf(self.h, str(path).encode(), float(spot_sig_r)))

    def restore_field_hdf(self, path):
        self._hdf(load_hdf().felis_restore_field_hdf(self.h, str(path).encode()))

    def save_particles_hdf(self, path, gamma0):
        """saveParticleHDF: particle<it>.h5; gamma0 per species (energy/mc2 + 1)."""
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma0, np.float64), (self.desc.n_species,)))
        self._hdf(load_hdf().felis_save_particles_hdf(self.h, str(path).encode(), _p(g)))

    def restore_particles_hdf(self, path):
        self._hdf(load_hdf().felis_restore_particles_hdf(self.h, str(path).encode()))
        so, _ = self.particle_layout()
        self.np = int(so[-1])

    def live_polarisations(self) -> int:
        """1 while the y field and its source are exactly zero (skipped by the
        batched step), else 2."""
        n = C.c_int()
        self._chk(self.L.felis_live_polarisations(self.h, C.byref(n)))
        return n.value

    def sync(self):
        self._chk(self.L.felis_sync(self.h))

    def profile(self, enable: bool):
        self._chk(self.L.felis_profile(self.h, int(enable)))

    def profile_read(self) -> dict:
        ms = np.zeros(len(KERNELS)); n = np.zeros(len(KERNELS), np.int64)
        self._chk(self.L.felis_profile_read(self.h, _p(ms), _p(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS) if n[i]}


def find_parameter(block: str, rank: int, key: str, path: str):
    """findparam.cpp:30-116 lookup: the value string, or None."""
    buf = C.create_string_buffer(512)
    rc = load().felis_find_parameter(block.encode(), rank, key.encode(), str(path).encode(), buf, 512)
    if rc < 0:
        raise FelisError(rc, f"cannot read {path}")
    return buf.value.decode() if rc == 1 else None


class Deck:
    """An input deck parsed by the C++ host (parameterSetting.cpp semantics)."""

    def __init__(self, path):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.felis_deck_open(str(path).encode(), C.byref(h))
        self.h = h
        if rc != OK:
            msg = self.L.felis_deck_error(h).decode() if h else "open failed"
            self.close()
            raise FelisError(rc, msg)
        info = (C.c_int * 8)()
        self.L.felis_deck_steps(self.h, info)
        self.max_step, self.wake_step, self.save_step, self.save_start = info[0], info[1], info[2], info[3]
        self.field_save, self.particle_save, self.mode, self.slice_n = bool(info[4]), bool(info[5]), info[6], info[7]

    def close(self):
        if getattr(self, "h", None):
            self.L.felis_deck_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def desc(self, precision: int = PLAIN_FP64) -> Desc:
        d = Desc()
        rc = self.L.felis_deck_desc(self.h, precision, C.byref(d))
        if rc != OK:
            raise FelisError(rc, "felis_deck_desc")
        return d

    def lattice(self, it0: int, n_steps: int):
        tab = (Lattice * (2 * n_steps + 1))()
        self.L.felis_deck_lattice(self.h, it0, n_steps, tab)
        return list(tab)

    def wake_function(self):
        wf = np.zeros(self.slice_n); imp = np.zeros(1001)
        rc = self.L.felis_deck_wake_function(self.h, _p(wf), _p(imp))
        if rc != OK:
            raise FelisError(rc, "wake function")
        return wf, imp

    def beam_sigma(self, n_species: int):
        sig = np.zeros(2 * n_species)
        self.L.felis_deck_beam_sigma(self.h, _p(sig))
        return sig.reshape(n_species, 2)

    def wake_function_device(self):
        """(wakeF, impZ) of wake_function, computed on the device."""
        wf = np.zeros(self.slice_n); imp = np.zeros(1001)
        rc = self.L.felis_deck_wake_function_device(self.h, _p(wf), _p(imp))
        if rc != OK:
            raise FelisError(rc, "felis_deck_wake_function_device")
        return wf, imp

    def save_settings(self, n_species: int):
        """(gamma0 per species, spot_sig_r) that the savers record."""
        g = np.zeros(n_species)
        r = np.zeros(1)
        rc = self.L.felis_deck_save_settings(self.h, _p(g), _p(r))
        if rc != OK:
            raise FelisError(rc, "felis_deck_save_settings")
        return g, float(r[0])

    def beam_layout(self, n_species: int):
        so = np.zeros(n_species * self.slice_n + 1, np.int64)
        sw = np.zeros(n_species * self.slice_n)
        self.L.felis_deck_beam_layout(self.h, _p(so), _p(sw))
        return so, sw

    def load(self, ctx: Context, seed: int = 20261203):
        """Beam (loadBeam.cpp) and seed field (loadSeed.cpp) generated on the device."""
        rc = self.L.felis_deck_load_beam(self.h, ctx.h, seed)
        if rc != OK:
            raise FelisError(rc, self.L.felis_last_error(ctx.h).decode())
        ctx.np = int(self.beam_layout(ctx.desc.n_species)[0][-1])
        rc = self.L.felis_deck_load_seed(self.h, ctx.h)
        if rc != OK:
            raise FelisError(rc, self.L.felis_last_error(ctx.h).decode())


def _f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def gemm_bf16(A, B):
    """C = A B on the matrix cores with BF16-rounded operands (FP32 result)."""
    A, pa = _f32(A)
    B, pb = _f32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.zeros((M, N), np.float32)
    rc = load().felis_gemm_bf16(M, N, K, pa, pb, Cm.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != OK:
        raise FelisError(rc, "felis_gemm_bf16")
    return Cm


class MLP:
    """The matrix-core perceptron of felis_mlp.hip (BF16 operands, FP32
    accumulation, GELU hidden layers, linear output, Adam)."""

    def __init__(self, dims, capacity: int, seed: int = 1):
        self.L = load()
        self.dims = [int(d) for d in dims]
        arr = (C.c_int * len(self.dims))(*self.dims)
        h = C.c_void_p()
        rc = self.L.felis_mlp_create(len(self.dims) - 1, arr, int(capacity), int(seed), C.byref(h))
        if rc != OK:
            raise FelisError(rc, "felis_mlp_create")
        self.h = h

    def close(self):
        if self.h:
            self.L.felis_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != OK:
            raise FelisError(rc, self.L.felis_mlp_last_error(self.h).decode())

    def get_layer(self, l):
        W = np.zeros((self.dims[l], self.dims[l + 1]), np.float32)
        b = np.zeros(self.dims[l + 1], np.float32)
        self._chk(self.L.felis_mlp_get_layer(self.h, l, _f32(W)[1], _f32(b)[1]))
        return W, b

    def set_layer(self, l, W, b):
        W, pw = _f32(W)
        b, pb = _f32(b)
        self._chk(self.L.felis_mlp_set_layer(self.h, l, pw, pb))

    def set_input_affine(self, mu, inv_sd):
        mu, pm = _f32(mu)
        inv_sd, ps = _f32(inv_sd)
        self._chk(self.L.felis_mlp_set_input_affine(self.h, pm, ps))

    def get_input_affine(self):
        mu = np.zeros(self.dims[0], np.float32)
        isd = np.zeros(self.dims[0], np.float32)
        self._chk(self.L.felis_mlp_get_input_affine(self.h, _f32(mu)[1], _f32(isd)[1]))
        return mu, isd

    def forward(self, X):
        X, px = _f32(X)
        Y = np.zeros((X.shape[0], self.dims[-1]), np.float32)
        self._chk(self.L.felis_mlp_forward(self.h, X.shape[0], px, _f32(Y)[1]))
        return Y

    def train_step(self, X, T, lr: float, wrow=None) -> float:
        X, px = _f32(X)
        T, pt = _f32(T)
        pw = None
        if wrow is not None:
            wrow, pw = _f32(wrow)
        loss = C.c_double()
        self._chk(self.L.felis_mlp_train_step(self.h, X.shape[0], px, pt, pw, float(lr), C.byref(loss)))
        return float(loss.value)

    def get_grad(self, l):
        gW = np.zeros((self.dims[l], self.dims[l + 1]), np.float32)
        gb = np.zeros(self.dims[l + 1], np.float32)
        self._chk(self.L.felis_mlp_get_grad(self.h, l, _f32(gW)[1], _f32(gb)[1]))
        return gW, gb


class FNO:
    """The Fourier neural operator of felis_fno.hip on 64 x 64 channel-last
    maps (BF16 operands on the matrix cores, FP32 accumulation and Adam)."""

    def __init__(self, cin: int, cout: int, width: int = 32, modes: int = 8, layers: int = 4, proj: int = 64,
                 capacity: int = 16, seed: int = 1):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.felis_fno_create(int(cin), int(cout), int(width), int(modes), int(layers), int(proj),
                                     int(capacity), int(seed), C.byref(h))
        if rc != OK:
            raise FelisError(rc, "felis_fno_create")
        self.h = h
        self.cin, self.cout, self.width, self.modes, self.layers, self.proj = cin, cout, width, modes, layers, proj
        self.capacity = capacity
        self.n_params = int(self.L.felis_fno_param_count(h))

    def close(self):
        if self.h:
            self.L.felis_fno_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != OK:
            raise FelisError(rc, self.L.felis_fno_last_error(self.h).decode())

    def params(self):
        P = np.zeros(self.n_params, np.float32)
        self._chk(self.L.felis_fno_get_params(self.h, _f32(P)[1]))
        return P

    def set_params(self, P):
        P, pp = _f32(P)
        assert P.size == self.n_params
        self._chk(self.L.felis_fno_set_params(self.h, pp))

    def grad(self):
        G = np.zeros(self.n_params, np.float32)
        self._chk(self.L.felis_fno_get_grad(self.h, _f32(G)[1]))
        return G

    def set_input_affine(self, mu, inv_sd):
        mu, pm = _f32(mu)
        inv_sd, ps = _f32(inv_sd)
        self._chk(self.L.felis_fno_set_input_affine(self.h, pm, ps))

    def forward(self, X):
        X, px = _f32(X)
        Y = np.zeros((X.shape[0], X.shape[1], self.cout), np.float32)
        self._chk(self.L.felis_fno_forward(self.h, X.shape[0], px, _f32(Y)[1]))
        return Y

    def train_step(self, X, E, R, wb, lr: float) -> float:
        """Host arrays X [B][4096][cin], E, R [B][4096][cout], wb [B]."""
        X, px = _f32(X)
        E, pe = _f32(E)
        R, pr = _f32(R)
        wb, pw = _f32(wb)
        loss = C.c_double()
        self._chk(self.L.felis_fno_train_step(self.h, X.shape[0], C.cast(px, C.c_void_p), C.cast(pe, C.c_void_p),
                                              C.cast(pr, C.c_void_p), C.cast(pw, C.c_void_p), 0, float(lr),
                                              C.byref(loss)))
        return float(loss.value)

    def train_step_dev(self, B: int, X, E, R, wb, lr: float) -> float:
        """Device pointers (ints) X, E, R, wb as in train_step."""
        loss = C.c_double()
        self._chk(self.L.felis_fno_train_step(self.h, int(B), C.c_void_p(X), C.c_void_p(E), C.c_void_p(R),
                                              C.c_void_p(wb), 1, float(lr), C.byref(loss)))
        return float(loss.value)
