"""Writes golden vectors with the reference's own compiled routines.

    make -C oracle ref && python tests/golden/make_golden.py

oracle/_ref/libfelisref.so is built from the reference's translation units
(rearrangeParticles.cpp, updateTotalEnergy.cpp, twiss.cpp, saveFile.cpp,
wakeField.cpp, updateK_quadG.cpp, fieldShareZ.cpp, loadSeed.cpp,
findparam.cpp) plus oracle/ref_harness.cpp, which fills a Domain with the
inputs below and reads the results back. The fixtures hold inputs and
expected outputs only.
"""
from __future__ import annotations

import ctypes as C
import json
import re
import math
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import pyoracle as po  # noqa: E402

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)


def P(a):
    if a.dtype == np.int64:
        return a.ctypes.data_as(i64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(ip)
    return a.ctypes.data_as(dp)


def layout(rng, n_species, slice_n, lo, hi):
    counts = rng.integers(lo, hi, n_species * slice_n)
    so = np.zeros(counts.size + 1, np.int64)
    so[1:] = np.cumsum(counts)
    return so


def main():
    R = po.ref()
    if R is None:
        po.build_ref()
        R = po.ref()
    assert R is not None, "oracle/_ref/libfelisref.so is not built"
    rng = np.random.default_rng(20261203)
    work = tempfile.mkdtemp()
    wd = work.encode()
    out = {}

    # rearrangeParticles.cpp:77-114 -------------------------------------------
    so = layout(rng, 2, 5, 3, 40)
    n = int(so[-1])
    for num_slice in (1, 3):
        dth = 2 * math.pi * num_slice
        th = rng.uniform(-3 * dth, 4 * dth, n)
        th[:6] = [0.0, dth, -dth, 2 * dth, -1e-300, dth * (1 - 1e-16)]
        w = rng.uniform(0.5, 2, so.size - 1)
        res = th.copy()
        R.ref_periodic(5, num_slice, 2, P(so), P(w), P(res))
        out[f"periodic_ns{num_slice}"] = dict(seg_off=so, theta_in=th, theta_out=res, num_slice=num_slice)

    # updateTotalEnergy.cpp:7-87 ----------------------------------------------
    nx, ny, sn, nh = 7, 6, 4, 2
    fs = (sn + 2) * nx * ny
    Ux = (rng.normal(size=(nh, fs)) + 1j * rng.normal(size=(nh, fs))).astype(np.complex128)
    Uy = (rng.normal(size=(nh, fs)) + 1j * rng.normal(size=(nh, fs))).astype(np.complex128)
    raw = np.zeros(2 * nh); line = np.zeros(1 + 2 * nh)
    args = dict(dx=4.7e-6, dy=5.1e-6, ks=1.1e10, lambda0=2 * math.pi / 1.1e10, num_slice=2, min_z=-1e-6,
                max_z=3e-6, dz=0.026, iteration=7)
    for mode in (1, 2):
        R.ref_total_energy.argtypes = [C.c_char_p] + [C.c_int] * 4 + [dp, dp] + [C.c_double] * 4 + [C.c_int] * 2 + \
            [C.c_double] * 3 + [C.c_int, dp, dp]
        R.ref_total_energy(wd, nx, ny, sn, nh, P(Ux), P(Uy), args["dx"], args["dy"], args["ks"], args["lambda0"],
                           args["num_slice"], mode, args["min_z"], args["max_z"], args["dz"], args["iteration"],
                           P(raw), P(line))
        out[f"total_energy_mode{mode}"] = dict(Ux=Ux.reshape(nh, sn + 2, ny, nx), Uy=Uy.reshape(nh, sn + 2, ny, nx),
                                               raw=raw.copy(), line=line.copy(), mode=mode,
                                               **{k: np.float64(v) for k, v in args.items()})

    # saveFile.cpp:109-172 (bFactor) --------------------------------------------
    so = layout(rng, 2, 6, 10, 60)
    n = int(so[-1])
    harm = np.array([1, 2, 3], np.int32)
    w = rng.uniform(0.5, 2, so.size - 1)
    cases = {"random": rng.uniform(0, 2 * math.pi, n),
             "bunched": (0.4 + 0.3 * rng.normal(size=n)) % (2 * math.pi)}
    for name, th in cases.items():
        line = np.zeros(4)
        R.ref_bfactor.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, ip, i64p, dp, dp, C.c_double, C.c_double,
                                  C.c_int, dp]
        R.ref_bfactor(wd, 6, 2, 3, P(harm), P(so), P(w), P(th), 0.026, -2e-6, 11, P(line))
        out[f"bfactor_{name}"] = dict(seg_off=so, weight=w, theta=th, harmony=harm, line=line.copy(), dz=0.026,
                                      min_z=-2e-6, iteration=11)

    # twiss.cpp:7-99 -----------------------------------------------------------
    so = layout(rng, 2, 5, 20, 80)
    n = int(so[-1])
    x = rng.normal(0, 30e-6, n); y = rng.normal(0, 25e-6, n)
    gam = 8000 * (1 + 1e-3 * rng.normal(size=n))
    px = gam * (rng.normal(0, 1.7e-6, n) + 0.01 * x); py = gam * (rng.normal(0, 2e-6, n) - 0.02 * y)
    w = rng.uniform(0.5, 2, so.size - 1)
    line = np.zeros(13)
    R.ref_twiss.argtypes = [C.c_char_p, C.c_int, C.c_int, i64p, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double,
                            C.c_double, C.c_int, dp]
    R.ref_twiss(wd, 5, 2, P(so), P(w), P(x), P(y), P(px), P(py), P(gam), 8000.0, 0.026, -1e-6, 4, P(line))
    out["twiss"] = dict(seg_off=so, weight=w, x=x, y=y, px=px, py=py, gamma=gam, line=line, gamma0=8000.0,
                        dz=0.026, min_z=-1e-6, iteration=4)

    # wakeField.cpp:53-159 (wake function) -----------------------------------------
    R.ref_wake_function.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_double, dp]
    lam0 = 5.58e-10
    for shape, acdc, sn_, tag in ((2, 1, 300, "circ_ac"), (2, 2, 300, "circ_dc"), (1, 1, 40, "flat_ac")):
        wf = np.zeros(sn_)
        R.ref_wake_function(wd, shape, acdc, 3e-3, 5.8e7, 2.4e-6, sn_, 1, lam0, P(wf))
        out[f"wake_function_{tag}"] = dict(shape=shape, acdc=acdc, radius=3e-3, cond=5.8e7, ctau=2.4e-6, slice_n=sn_,
                                           num_slice=1, lambda0=lam0, wakeF=wf)

    # wakeField.cpp:9-50 + 161-223 (update) ------------------------------------------
    so = layout(rng, 2, 30, 0, 50)
    w = rng.uniform(0.5, 2, so.size - 1)
    wf = out["wake_function_circ_ac"]["wakeF"][:30].copy()
    den = np.zeros(30); wE = np.zeros(30); tot = np.zeros(1)
    R.ref_update_wake.argtypes = [C.c_char_p, C.c_int, C.c_int, i64p, dp, C.c_int, C.c_double, C.c_double,
                                  C.c_double, dp, dp, dp, dp]
    R.ref_update_wake(wd, 30, 2, P(so), P(w), 1, lam0, -1e-6, 0.026, P(wf), P(den), P(wE), P(tot))
    out["update_wake"] = dict(seg_off=so, weight=w, wakeF=wf, den=den, wakeE=wE, total_cnt=tot[0],
                              num_slice=1, lambda0=lam0)

    # updateK_quadG.cpp:8-94 ------------------------------------------------------------
    # two undulator blocks (taper, slope, air gaps), two FODO quad lists
    und = []
    for i in range(6):
        us, ue = 5.0 * i, 5.0 * (i + 1)
        K0 = 1.87 + i * (-0.01 / math.sqrt(2)) if i < 3 else 1.80
        blk = 0 if i < 3 else 1
        und.append([us, ue, us + 0.5, ue - 0.5, K0, -0.001 if blk else 0.0, 0.0 if blk == 0 else 0.3,
                    0.026 if blk == 0 else 0.03, 1 if blk == 0 else -1, 1 if blk == 0 else 2, 1 if i % 2 else 0, blk])
    und = np.array(und, np.float64)
    quad = []
    for q in range(8):
        quad.append([q * 3.75 - 0.05, q * 3.75 + 0.05, 30.0, 1])
    for q in range(8):
        quad.append([q * 3.75 + 1.875 - 0.06, q * 3.75 + 1.875 + 0.06, -30.0, 2])
    quad = np.array(quad, np.float64)
    dz = 0.026
    state = np.array([1.87, 0.0, 1.0, 2 * math.pi / 0.026, 0.0, 0.026, 1.0, 0.0])
    flags = np.zeros(2, np.int32)
    seq = []
    R.ref_update_k_quadg.argtypes = [dp, C.c_int, dp, C.c_int, C.c_double, C.c_int, C.c_double, dp, ip]
    for it in range(0, 1200):
        for half in (0.0, 0.5):
            R.ref_update_k_quadg(P(und), len(und), P(quad), len(quad), dz, it, half, P(state), P(flags))
            seq.append(np.concatenate([state[:7], flags.astype(np.float64)]))
    out["update_k_quadg"] = dict(und=und, quad=quad, dz=dz, seq=np.array(seq))

    # loadSeed.cpp:25-80 -------------------------------------------------------------
    nx, ny, sn = 9, 8, 3
    harm = np.array([1, 3], np.int32)
    Ux = np.zeros((2, (sn + 2) * nx * ny), np.complex128); Uy = np.zeros_like(Ux)
    ks = 1.1e10
    seed = dict(dx=60e-6, dy=70e-6, dz=0.026, min_x=-240e-6, min_y=-245e-6, min_z=-3e-6, lambda0=2 * math.pi / ks,
                num_slice=2, ks=ks, a0=3e-8, duration=30e-15, zR=2 * (20e-6) ** 2 * ks * 3, focus=0.7,
                spotSigR=20e-6, loadH=3, laserAlpha=1.0, laserPsi=0.4)
    R.ref_load_seed_3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ip] + [C.c_double] * 7 + [C.c_int] + \
        [C.c_double] * 6 + [C.c_int, C.c_double, C.c_double, dp, dp]
    R.ref_load_seed_3d(nx, ny, sn, 2, P(harm), seed["dx"], seed["dy"], seed["dz"], seed["min_x"], seed["min_y"],
                       seed["min_z"], seed["lambda0"], seed["num_slice"], seed["ks"], seed["a0"], seed["duration"],
                       seed["zR"], seed["focus"], seed["spotSigR"], seed["loadH"], seed["laserAlpha"],
                       seed["laserPsi"], P(Ux), P(Uy))
    out["load_seed"] = dict(nx=nx, ny=ny, slice_n=sn, harmony=harm, Ux=Ux.reshape(2, sn + 2, ny, nx),
                            Uy=Uy.reshape(2, sn + 2, ny, nx), **{k: np.float64(v) for k, v in seed.items()})

    # fieldShareZ.cpp:39-78 on one rank ----------------------------------------------
    f = (rng.normal(size=(2, 5 * 12)) + 1j * rng.normal(size=(2, 5 * 12))).astype(np.complex128)
    g = f.copy()
    R.ref_transfer_zplus.argtypes = [C.c_int] * 5 + [dp]
    R.ref_transfer_zplus(2, 12, 4, 1, 5, P(g))
    out["transfer_zplus"] = dict(f_in=f, f_out=g)

    for name, d in out.items():
        np.savez(HERE / f"{name}.npz", **{k: np.asarray(v) for k, v in d.items()})

    # findparam.cpp:30-116 on the reference's sample deck and a crafted one ----------
    R.ref_find_parameters.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]
    crafted = HERE / "deck_crafted.inp"
    decks = {"test.inp": Path("/root/reference/input/test.inp"), "deck_crafted.inp": crafted}
    queries = [("Domain", 1, k) for k in ("dimension", "mode", "minZ", "maxZ", "nx", "ny", "photon_energy",
                                          "num_harmony", "harmony0", "harmony1", "harmony2", "slices_in_bucket",
                                          "lambdaUs_in_iteration", "ABC_N", "radial_grids", "missing_key")]
    queries += [("Quad", r, k) for r in (1, 2, 3) for k in ("numbers", "unit_start", "g", "quad_end")]
    queries += [("EBeam", 1, k) for k in ("load_type", "beam_energy", "z_n1", "energySpread_z0", "noise_ONOFF")]
    queries += [("Undulator", 1, k) for k in ("undulator_type", "polarity", "K0", "linear_taper", "lambdaU")]
    queries += [("Seed", 1, k) for k in ("power", "focus")] + [("Save", 1, "total_length"), ("Save", 1, "max_step")]
    queries += [("Wake_field", 1, "shape"), ("Phase_shifter", 1, "number"), ("Space_charge", 1, "radial_grids")]
    result = {}
    for dname, path in decks.items():
        rows = []
        for blk, rank, key in queries:
            buf = C.create_string_buffer(512)
            r = R.ref_find_parameters(blk.encode(), rank, key.encode(), str(path).encode(), buf)
            rows.append([blk, rank, key, int(r), buf.value.decode() if r else None])
        result[dname] = rows
    (HERE / "findparam.json").write_text(json.dumps(result, indent=1))

    # deck -> lattice table: the reference's FindParameters reads the blocks of
    # deck_lattice.inp, parameterSetting.cpp:287-429 expands them into units and
    # quads (restated below), parameterSetting.cpp:168-174 sets the state before
    # the loop and the reference's updateK_quadG.cpp walks 1200 periods.
    deck = HERE / "deck_lattice.inp"

    def fp(block, rank, key):
        buf = C.create_string_buffer(512)
        r = R.ref_find_parameters(block.encode(), rank, key.encode(), str(deck).encode(), buf)
        return buf.value.decode() if r else None

    def atof(v):
        m = re.match(r"\s*[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?", v or "")
        return float(m.group(0)) if m else 0.0

    und, rank = [], 1
    while True:
        t = fp("Undulator", rank, "undulator_type")
        if t is None:
            break
        utype = 1 if t.strip().startswith("BiPolar") else 2
        n = int(atof(fp("Undulator", rank, "numbers")))
        air = 1 if (fp("Undulator", rank, "in_air") or "").strip().startswith("ON") else 0
        lam = atof(fp("Undulator", rank, "lambdaU")) * 0.01
        lin = atof(fp("Undulator", rank, "linear_taper")) / math.sqrt(2.0) if fp("Undulator", rank, "linear_taper") else 0.0
        qt = atof(fp("Undulator", rank, "quad_taper")) / math.sqrt(2.0) if fp("Undulator", rank, "quad_taper") else 0.0
        slope = atof(fp("Undulator", rank, "slope_K")) / math.sqrt(2.0) if fp("Undulator", rank, "slope_K") else 0.0
        qid = int(atof(fp("Undulator", rank, "quad_taper_start_index"))) if fp("Undulator", rank, "quad_taper_start_index") else 30000
        pol = atof(fp("Undulator", rank, "polarity"))
        ka = int(atof(fp("Undulator", rank, "K0_alpha"))) if fp("Undulator", rank, "K0_alpha") else 1
        K0 = atof(fp("Undulator", rank, "K0"))
        us, ue_ = atof(fp("Undulator", rank, "unit_start")), atof(fp("Undulator", rank, "unit_end"))
        ds, de = atof(fp("Undulator", rank, "undulator_start")), atof(fp("Undulator", rank, "undulator_end"))
        rows = [[us, ue_, ds, de, K0]]
        unit_len, und_off, und_len = ue_ - us, ds - us, de - ds
        for i in range(1, n):
            a = rows[i - 1][1]
            b = a + unit_len
            c = a + und_off
            Ki = K0 + i * lin + ((i - qid) * (i - qid) * qt if i >= qid else 0.0)
            rows.append([a, b, c, c + und_len, Ki])
        for r in rows:
            und.append(r + [slope, pol, lam, ka, utype, air, rank - 1])
        rank += 1
    quad, rank = [], 1
    while fp("Quad", rank, "numbers") is not None:
        n = int(atof(fp("Quad", rank, "numbers")))
        us, ue_ = atof(fp("Quad", rank, "unit_start")), atof(fp("Quad", rank, "unit_end"))
        qs, qe, g = atof(fp("Quad", rank, "quad_start")), atof(fp("Quad", rank, "quad_end")), atof(fp("Quad", rank, "g"))
        unit_len, q_off, q_len = ue_ - us, qs - us, qe - qs
        rows = [[us, ue_, qs, qe]]
        for i in range(1, n):
            a = rows[i - 1][1]
            rows.append([a, a + unit_len, a + q_off, a + q_off + q_len])
        quad += [[r[2], r[3], g, rank] for r in rows]
        rank += 1
    und = np.array(und, np.float64)
    quad = np.array(quad, np.float64)
    u0 = und[0]
    state = np.array([u0[4], u0[6], u0[8], 2.0 * math.pi / u0[7], 0.0, u0[7], u0[9], 0.0])
    init = state[:7].copy()
    dz = u0[7] * 1
    seq = []
    for it in range(0, 1200):
        for half in (0.0, 0.5):
            R.ref_update_k_quadg(P(und), len(und), P(quad), len(quad), dz, it, half, P(state), P(flags))
            seq.append(np.concatenate([state[:7], flags.astype(np.float64)]))
    np.savez(HERE / "update_k_quadg_deck.npz", und=und, quad=quad, dz=dz, init=init, seq=np.array(seq))

    # deck -> derived Domain scalars (parameterSetting.cpp:71-77, 89-111, 168-175,
    # 186-196, 236-240 restated on the reference's FindParameters values) ----------
    PLANK_H, C_LIGHT, MC2_MEV = 4.135667696e-15, 299792458.0, 0.510998918

    def derived(path):
        def f(block, key, default=None):
            buf = C.create_string_buffer(512)
            r = R.ref_find_parameters(block.encode(), 1, key.encode(), str(path).encode(), buf)
            return buf.value.decode() if r else default
        dim = int(atof(f("Domain", "dimension")))
        static = (f("Domain", "mode") or "").strip().startswith("Static")
        ks = atof(f("Domain", "photon_energy")) * 2 * math.pi / (PLANK_H * C_LIGHT)
        lambda0 = 2 * math.pi / ks
        Lz = atof(f("Save", "total_length"))
        minX, maxX = atof(f("Domain", "minX")) * 1e-6, atof(f("Domain", "maxX")) * 1e-6
        minY, maxY = atof(f("Domain", "minY")) * 1e-6, atof(f("Domain", "maxY")) * 1e-6
        minZ, maxZ = atof(f("Domain", "minZ")) * 1e-6, atof(f("Domain", "maxZ")) * 1e-6
        nx, ny = int(atof(f("Domain", "nx"))), int(atof(f("Domain", "ny")))
        dx, dy = (maxX - minX) / (nx * 1.0), (maxY - minY) / (ny * 1.0)
        nx, ny = nx + 1, ny + 1
        if dim == 1:
            nx, ny, dx, dy = 1, 1, 0.0, 0.0
        num_slice = int(atof(f("Domain", "slices_in_bucket", "1")))
        num_lu = int(atof(f("Domain", "lambdaUs_in_iteration", "1")))
        gamma0 = atof(f("EBeam", "beam_energy")) / MC2_MEV + 1.0
        lam_u = atof(f("Undulator", "lambdaU")) * 0.01
        ku = 2 * math.pi / lam_u
        ue = atof(f("Undulator", "polarity"))
        K0 = atof(f("Undulator", "K0")) if f("Undulator", "K0") else math.sqrt(
            (4 * gamma0 * gamma0 * ku / ks - 2) / (1.0 + ue * ue))
        gamR = math.sqrt(0.5 * ks / ku * (1 + 0.5 * (1 + ue * ue) * K0 * K0))
        nr = int(atof(f("Domain", "radial_grids"))) if f("Domain", "radial_grids") else int(math.sqrt(nx * nx + ny * ny))
        dr = math.sqrt((maxX - minX) ** 2 + (maxY - minY) ** 2) / (1.0 * nx)
        if dim == 1:
            nr, dr = 1, 0.0
        if static:
            minZ, maxZ = 0.0, lambda0 * num_slice
        dz = lam_u * num_lu
        # seed (parameterSetting.cpp:198-229); D.polarity is never parsed (0)
        Z0_OHM, E_CH, E_MASS = 376.730313668, 1.60217662e-19, 9.10938291e-31
        spot = atof(f("Seed", "spot_sigma_R")) * 1e-6
        load_h = int(atof(f("Seed", "loading_harmonics", "1")))
        s_ue = atof(f("Seed", "polarity", "0"))
        area = 2.0 * math.pi * spot * spot
        a0 = math.sqrt(atof(f("Seed", "power")) * 2.0 * Z0_OHM / area) * E_CH / (E_MASS * C_LIGHT ** 2 * ks * load_h)
        seed = dict(a0=a0, zR=2.0 * spot * spot * ks * load_h, spot_sig_r=spot, load_h=load_h,
                    duration=atof(f("Seed", "rms_duration")) * 1e-15, focus=atof(f("Seed", "focus")),
                    laser_alpha=atof(f("Seed", "laser_alpha", "-1")),
                    laser_psi=float(np.angle((1.0 - s_ue * s_ue + 2j * s_ue) / (1.0 + s_ue * s_ue))))
        return dict(seed=seed, nx=nx, ny=ny, dx=dx, dy=dy, min_x=minX, min_y=minY, min_z=minZ, max_z=maxZ, ks=ks,
                    lambda0=lambda0, gamR=gamR, dr=dr, nr=nr, dz=dz, num_slice=num_slice,
                    max_step=int(Lz / dz + 1), slice_n=int((maxZ - minZ) / (lambda0 * num_slice)),
                    abc_n=int(atof(f("Domain", "ABC_N", "10"))), abc_sig=atof(f("Domain", "ABC_coef", "1")),
                    bessel_n=int(atof(f("Bessel_table", "num_grids", "2001"))),
                    sc_l=int(atof(f("Domain", "number_longitudinal_mode", "1"))),
                    sc_f=int(atof(f("Domain", "number_azimuthal_mode", "1"))))

    decks_d = {"deck_lattice.inp": deck}
    for name in ("cfg0.inp", "cfg1.inp", "cfg2.inp", "cfg2e.inp", "cfg2s.inp", "cfg3.inp"):
        decks_d[name] = ROOT / "felis_amd" / "decks" / name
    (HERE / "deck_desc.json").write_text(json.dumps({k: derived(v) for k, v in decks_d.items()}, indent=1))
    print("wrote", len(out), "npz fixtures and findparam.json to", HERE)


if __name__ == "__main__":
    main()
