#!/usr/bin/env python3
"""Benchmark of the FELIS per-step slice kernel on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg1] [--precision plain|mixed]

A step is one undulator period of the whole slice loop of main.cpp:92-156
(diagnostics, field solve, transverse and longitudinal push, theta wrap,
slippage) over every slice of the configuration, with I/O off. The beam and
seed are generated on the device from the configuration's input deck
(felis_amd/decks/), so inputs are resident in HBM when timing starts.

Prints one JSON line: value = slices x periods / s over the timed steps, the
roofline of the dominant kernel (algorithmic bytes per launch over its mean
HIP-event duration), and the FP64 oracle timed on the host cores on a bounded
sample of the same workload (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "slice·periods/s at ≤1e-3 rel power err vs FP64 oracle; batched+surrogate vs plain"
PEAK_HBM_GBS = 8000.0
WORKLOADS = {
    "cfg1": "time-dependent 256 slices x 4096 particles, 101x101 grid, harmonic 1, flat 3 kA, 10 m (385 periods)",
    "cfg2": "time-dependent 2048 slices, 256-16384 particles/slice (11,670,128), 129x129, harmonics 1,3,5, "
            "Gaussian 4 kA, 20 m (770 periods)",
    "cfg2e": "cfg2 in an elliptical undulator (polarity 0.5, K 1.6726 for the same resonance): both polarisations "
             "coupled and solved",
    "cfg3": "time-dependent 4096 slices x 8192 particles (33,554,432), 151x151, harmonic 1, Gaussian 3 kA, six 5 m "
            "units with 1 m drifts, -1%/unit taper from the fourth, FODO quads, circular AC wake, 29 m "
            "(1116 periods)",
}


def kernel_bytes(name: str, d, np_: int, mixed: bool, npol: int = 2) -> float:
    """Algorithmic HBM bytes of one launch of a kernel: each operand it must
    read or write, once (the bilinear field gather is counted as cached).
    npol: polarisations the launch solves (a y field that is exactly zero is
    skipped by the batched step and not counted)."""
    rs = 4 if mixed else 8          # particle coordinate bytes (theta is 8)
    es = 8 if mixed else 16         # complex field element bytes
    cells = d.nx * d.ny
    interior = (d.nx - 2) * (d.ny - 2)
    nsys = npol * d.n_harm * d.slice_n
    table = {
        "transverse": np_ * 9 * rs,                       # x,y,px,py,gamma in; x,y,px,py out
        "push": np_ * (8 + 5 * rs + 8 + rs),              # theta + 5 coords in; theta, gamma out
        "drift": np_ * (8 + 4 * rs + 8),
        "periodic": np_ * 16,
        "deposit": np_ * (8 + 5 * rs),                    # particle reads (scatter counted as cached)
        "ez_deposit": np_ * (8 + 2 * rs),
        "adi_x": nsys * interior * 3 * es,                # U in, Sc in, Uc out
        "adi_y": nsys * interior * 3 * es,                # Uc in, Sc in, U out
        "energy": nsys * cells * es,
        "particle_step": np_ * 2 * (8 + 5 * rs),          # full particle state in and out
        "adi_fused": nsys * interior * 2 * es,            # U in and out
    }
    return float(table.get(name, 0.0))


# kernel-timer name -> rocprofv3 kernel name, for the measured HBM traffic
# (the field solve's name carries its workgroup size, a tuned launch shape)
PROFILED = {("cfg2", "mixed"): {"adi_fused": "felis::k_adi_scan<8, {adi_threads}>",
                                "particle_step": "felis::k_pstep<3, false, false>",
                                "deposit": "felis::k_deposit_slot<3, 1, false, false>"},
            ("cfg2e", "mixed"): {"adi_fused": "felis::k_adi_scan<8, {adi_threads}>",
                                 "particle_step": "felis::k_pstep<3, true, false>",
                                 "deposit": "felis::k_deposit_slot<3, 2, false, false>"}}


def measured_traffic(config: str, precision: str, kernel: str, tuning=None):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE summary of this workload (scaled by a copy of known size), or
    None when that workload was not profiled."""
    adi_threads = (tuning or {}).get("adi_threads") or 512
    names = {k: v.format(adi_threads=adi_threads) for k, v in PROFILED.get((config, precision), {}).items()}
    path = ROOT / "profiles" / f"hbm_traffic_{config}_batched.md"
    if kernel not in names or not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.strip().split(" ")
        if len(parts) >= 3 and parts[0] == "traffic" and " ".join(parts[1:-1]) == names[kernel]:
            return float(parts[-1])
    return None


def run_gpu(args, precision=None, steps=None, warmup=None, keep_init=None):
    from felis_amd import _lib as fl
    deck = fl.Deck(ROOT / "felis_amd" / "decks" / f"{args.config}.inp")
    precision = precision or args.precision
    prec = fl.MIXED_FP32 if precision == "mixed" else fl.PLAIN_FP64
    desc = deck.desc(prec)
    ctx = fl.Context(desc)
    deck.load(ctx, seed=args.seed)
    tuning = None
    if prec == fl.MIXED_FP32:
        from felis_amd import tune
        tuning = tune.apply(ctx, args.tuning)
        if getattr(args, "deposit_mode", "atomic") == "fixed":
            ctx.set_deposit_mode(fl.DEPOSIT_FIXED)
    init = None
    if (not args.no_cpu or not args.no_whole_run) if keep_init is None else keep_init:
        init = dict(particles=ctx.download_particles(), Ux=ctx.download_field(fl.UX), Uy=ctx.download_field(fl.UY))
    W = args.warmup if warmup is None else warmup
    K = args.steps if steps is None else steps
    P = args.prof_steps
    tab = deck.lattice(0, W + K + P)
    wakeF = deck.wake_function()[0] if desc.wake_on else None
    ctx.sync()
    if W:
        ctx.run(tab[0:2 * W + 1], 0, W, deck.wake_step, wakeF)
    ctx.sync()
    # the timed steps run as they would in use (the batched path with the
    # autotuned launch shapes); per-kernel HIP-event times come from the next
    # P steps, launched one kernel at a time
    t0 = time.perf_counter()
    energy, bfac, twiss = ctx.run(tab[2 * W:2 * (W + K) + 1], W, K, deck.wake_step, wakeF)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.profile(True)
    ctx.run(tab[2 * (W + K):2 * (W + K + P) + 1], W + K, P, deck.wake_step, wakeF)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    return dict(deck=deck, desc=desc, ctx=ctx, seconds=t1 - t0, prof=prof, energy=energy, init=init, tab=tab,
                tuning=tuning,
                wakeF=wakeF, np=ctx.np, npol=ctx.live_polarisations() if precision == "mixed" else 2)


def roofline(prof, d, np_, mixed, npol, config, precision, tuning):
    """The dominant kernel of a profiled run against HBM peak."""
    dom = max(prof, key=lambda k: prof[k][0])
    ms, n = prof[dom]
    avg_s = ms / n * 1e-3
    bpl = kernel_bytes(dom, d, np_, mixed, npol)
    achieved = bpl / avg_s / 1e9 if bpl else 0.0
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(config, precision, dom, tuning),
            "bytes_per_launch": bpl, "avg_launch_us": avg_s * 1e6}


def two_polarisation_leg(args, steps: int = 20):
    """The same beam in an elliptical undulator (cfg2e): both polarisations
    are coupled, so the batched step solves, deposits and gathers Ux and Uy.
    Timed like the main workload, over fewer steps."""
    import copy
    a = copy.copy(args)
    a.config = "cfg2e"
    g = run_gpu(a, steps=steps, warmup=3, keep_init=False)
    d = g["desc"]
    out = {"workload": f"cfg2e: {WORKLOADS['cfg2e']}", "steps": steps,
           "ms_per_step": g["seconds"] / steps * 1e3, "value": d.slice_n * steps / g["seconds"],
           "polarisations_solved": g["npol"],
           "roofline": roofline(g["prof"], d, g["np"], True, g["npol"], "cfg2e", "mixed", g["tuning"]),
           "kernels_ms_per_step": {k: v[0] / args.prof_steps
                                   for k, v in sorted(g["prof"].items(), key=lambda kv: -kv[1][0])}}
    g["ctx"].close()
    return out


def host_cores():
    """(physical cores of the host from lscpu, CPUs this process may run on)."""
    import subprocess
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        phys = len(cores) or None
    except (OSError, subprocess.SubprocessError):
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return phys, avail


def run_cpu(args, g):
    """The FP64 oracle on a bounded sample of the same workload: one untimed
    period from the initial state (period 0 of main.cpp:92-156 has no
    deposit), then args.cpu_steps timed steady-state periods. Threads: the
    CPU share this process is given (the physical cores of the host, capped
    by the CPUs it may use and OMP_NUM_THREADS), one OpenMP thread per core
    over the slices, as the reference's ranks split them."""
    from oracle import pyoracle as po
    d = g["desc"]
    phys, avail = host_cores()
    threads = min(x for x in (phys, avail, int(os.environ.get("OMP_NUM_THREADS", "0")) or None) if x)
    threads = int(args.cpu_threads or threads)
    po.set_threads(threads)
    seg_off, seg_w = g["deck"].beam_layout(d.n_species)
    st = po.state_from_desc(d, seg_off, seg_w, g["init"]["particles"], g["init"]["Ux"], g["init"]["Uy"])
    n = 1 + args.cpu_steps
    lat = po.lattice_table(g["tab"][:2 * n + 1])
    rc, e0, _, _ = st.run(lat[:3], 0, 1, d.mode == 2, g["deck"].wake_step, g["wakeF"])
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    t0 = time.perf_counter()
    rc, e1, _, _ = st.run(lat[2:], 1, args.cpu_steps, d.mode == 2, g["deck"].wake_step, g["wakeF"])
    t1 = time.perf_counter()
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    return dict(seconds=t1 - t0, threads=threads, host_physical_cores=phys, host_cpus_available=avail,
                energy=np.concatenate([e0, e1]))


def gpu_parity(args, g, cpu):
    """GPU from the same initial state over the sampled periods vs the oracle."""
    from felis_amd import _lib as fl
    d = g["desc"]
    ctx = fl.Context(d)
    part = g["init"]["particles"]
    seg_off, seg_w = g["deck"].beam_layout(d.n_species)
    ctx.upload_particles(seg_off, seg_w, part["x"], part["y"], part["px"], part["py"], part["theta"],
                         part["gamma"])
    ctx.upload_field(fl.UX, g["init"]["Ux"])
    ctx.upload_field(fl.UY, g["init"]["Uy"])
    n = 1 + args.cpu_steps
    e, _, _ = ctx.run(g["tab"][:2 * n + 1], 0, n, g["deck"].wake_step, g["wakeF"])
    ctx.close()
    nh = d.n_harm
    tot_g = e[:, :nh].sum(1) + e[:, nh:].sum(1)
    tot_o = cpu["energy"][:, :nh].sum(1) + cpu["energy"][:, nh:].sum(1)
    return float(np.max(np.abs(tot_g - tot_o) / np.abs(tot_o)))


def whole_run_parity(args, g):
    """The batched path against the plain FP64 path (itself oracle-matched to
    1e-9) over every period of the deck, both on the GPU from the initial
    state: the largest relative error of the integrated power of each
    harmonic over the run, and of the fundamental's slice power at its end
    (slices above 1e-3 of the peak)."""
    from felis_amd import _lib as fl
    d = g["desc"]
    deck = g["deck"]
    n = deck.max_step
    tab = deck.lattice(0, n)
    out = {}
    curves = {}
    slices = {}
    for prec in (fl.MIXED_FP32, fl.PLAIN_FP64):
        ctx = fl.Context(deck.desc(prec))
        part = g["init"]["particles"]
        seg_off, seg_w = deck.beam_layout(d.n_species)
        ctx.upload_particles(seg_off, seg_w, part["x"], part["y"], part["px"], part["py"], part["theta"],
                             part["gamma"])
        ctx.upload_field(fl.UX, g["init"]["Ux"])
        ctx.upload_field(fl.UY, g["init"]["Uy"])
        if prec == fl.MIXED_FP32:
            from felis_amd import tune
            tune.apply(ctx, args.tuning)
        e, _, _ = ctx.run(tab, 0, n, deck.wake_step, g["wakeF"])
        curves[prec] = e
        slices[prec] = ctx.slice_energy()[0, 0]
        ctx.close()
    nh = d.n_harm
    em, ep = curves[fl.MIXED_FP32], curves[fl.PLAIN_FP64]
    for h in range(nh):
        ref = ep[:, h] + ep[:, nh + h]
        on = ref > 0
        err = np.abs(em[on, h] + em[on, nh + h] - ref[on]) / ref[on]
        out[f"max_rel_err_power_h{d.harmony[h]}"] = float(err.max())
    sp, sm = slices[fl.PLAIN_FP64], slices[fl.MIXED_FP32]
    strong = sp > 1e-3 * sp.max()
    out["max_rel_err_slice_power_h1_end"] = float(np.max(np.abs(sm[strong] - sp[strong]) / sp[strong]))
    out["periods"] = n
    out["against"] = "plain FP64 path on the GPU (oracle-matched to 1e-9 in tests/test_gpu_workloads.py)"
    return out


def surrogate_runs(args, periods: int = 10, gate: float = 0.05, wide_gate: float = 0.1):
    """The surrogate-assisted run against the batched one. Jumps of `periods`
    periods stand in for the fine steps while |b1| < gate (0.05, the range
    the surrogate is specified for), with the committed correction network;
    the batched step runs and scores everything else.

    to_saturation (the figure the 3x target is set on, for the 2048-slice
    workload): the same beam in a 40 m undulator, batched vs jumps without
    and with the network, to saturation (the first maximum of the power):
    wall times of the step loop, saturation length and power errors, and the
    mean |error| of the field energy along the run to saturation (what the
    network gains over the jumps alone). wider_gate repeats the run with the
    network while |b1| < 0.1, outside the specified range.
    whole_run: every period of the deck itself, batched vs jumps with the
    network (the run ends before saturation)."""
    from felis_amd import _lib as fl
    from felis_amd import surrogate as su
    deck = ROOT / "felis_amd" / "decks" / f"{args.config}.inp"
    dk = fl.Deck(deck)
    slices = dk.slice_n
    dk.close()
    model = su.load()
    net = model.network(capacity=slices)
    fno_model = su.load_fno()

    def filled(r):
        ok = ~np.isnan(r["E"])
        idx = np.nonzero(ok)[0]
        e = r["E"].copy()
        e[~ok] = np.interp(np.nonzero(~ok)[0], idx, r["E"][idx])
        return e

    def versus(rf, r, z=None, zs=None, ps=None):
        ok = ~np.isnan(r["E"])
        rel = r["E"][ok] / rf["E"][ok] - 1
        o = {"seconds": r["seconds"], "speedup_of_batched": rf["seconds"] / r["seconds"],
             "jumps": int(np.sum(r["kind"] == 1)), "periods_in_jumps": int(np.sum(r["kind"] > 0)),
             "max_power_rel_err": float(np.max(np.abs(rel)))}
        if zs is None:
            o["final_power_rel_err"] = float(r["E"][-1] / rf["E"][-1] - 1)
        else:
            zj, pj, _ = su.saturation(z, filled(r))
            m = ok & (z <= zs)
            o.update(z_sat_rel_err=float(zj / zs - 1), power_sat_rel_err=float(pj / ps - 1),
                     mean_power_rel_err_to_sat=float(np.mean(np.abs(r["E"][m] / rf["E"][m] - 1))))
        return o

    rf = su.run_curves(deck, seed=args.seed)
    rn = su.run_curves(deck, seed=args.seed, periods=periods, gate=gate, net=net, net_gate=gate)
    out = {"periods_per_jump": periods, "bunching_gate": gate, "model": "felis_amd/models/jump_corr_J10.npz",
           "network": "per-slice 20-64-64-6 perceptron (BF16 MFMA, FP32 accumulate) scaling the jump's emission; "
                      "the specified field-to-field FNO is the to_saturation leg's jump_field_to_field",
           "whole_run": {"periods": int(rf["steps"]), "batched_seconds": rf["seconds"], "jump_learned": versus(rf, rn)},
           "model_meta": model.meta}
    if args.config == "cfg2":
        import tempfile
        from felis_amd.decks import make_decks as md
        with tempfile.TemporaryDirectory() as td:
            p40 = Path(td) / "cfg2_40m.inp"
            p40.write_text(md.deck(**dict(md.DECKS["cfg2.inp"], length=40.0)))
            dk40 = fl.Deck(p40)
            dzs = dk40.desc(fl.MIXED_FP32).dz  # m per step
            dk40.close()
            sf = su.run_curves(p40, seed=args.seed)
            sj = su.run_curves(p40, seed=args.seed, periods=periods, gate=gate)
            sn = su.run_curves(p40, seed=args.seed, periods=periods, gate=gate, net=net, net_gate=gate)
            sw = su.run_curves(p40, seed=args.seed, periods=periods, gate=wide_gate, net=net, net_gate=gate)
            fno = fno_model.network(capacity=slices)
            sfno = su.run_curves(p40, seed=args.seed, periods=periods, gate=gate, fno=fno, net_gate=gate)
            fno.close()
        z = np.arange(sf["steps"], dtype=float)
        zs, ps, found = su.saturation(z, sf["E"])
        learned = versus(sf, sn, z, zs, ps)
        out["to_saturation"] = {
            "undulator_m": 40.0, "periods": int(sf["steps"]), "saturated": bool(found),
            "z_sat_m": float(zs * dzs), "batched_seconds": sf["seconds"],
            "speedup_of_batched": learned["speedup_of_batched"], "target_speedup": 3.0,
            "jump": versus(sf, sj, z, zs, ps), "jump_learned": learned,
            "jump_field_to_field": dict(versus(sf, sfno, z, zs, ps), model="felis_amd/models/jump_fno_J10.npz",
                                        network="4-layer Fourier neural operator, width 32, 8 modes per axis, on "
                                                "64x64 maps of each slice's emission and field plus 8 moments "
                                                "(BF16 MFMA, FP32 accumulate): factor maps on the emission",
                                        model_meta=fno_model.meta),
            "wider_gate": dict(versus(sf, sw, z, zs, ps), bunching_gate=wide_gate)}
    net.close()
    return out


def calibrate_copy(mib: int):
    """A plain device-to-device copy of known size, for scaling the HBM byte
    counters (FETCH_SIZE, WRITE_SIZE) of a profiler run."""
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["plain", "mixed"])
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--prof-steps", type=int, default=10,
                    help="steps after the timed ones, each kernel timed with HIP events (roofline)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-whole-run", action="store_true",
                    help="skip the whole-run comparison of the batched path against the plain FP64 path")
    ap.add_argument("--plain-steps", type=int, default=10,
                    help="periods of the plain FP64 path timed beside a mixed run (0: skip)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the whole-run comparison of multi-period jumps (with and without the learned "
                         "correction) against the batched step")
    ap.add_argument("--deposit-mode", default="atomic", choices=["atomic", "fixed"],
                    help="the batched deposit's accumulation (felis_set_deposit_mode): float atomics, or 64-bit fixed "
                         "point (bitwise reproducible)")
    ap.add_argument("--no-two-pol", action="store_true",
                    help="skip the elliptical-undulator leg (both polarisations solved) of the 2048-slice workload")
    ap.add_argument("--tuning", default=None, metavar="JSON",
                    help="autotuner cache the batched step's launch shapes come from (default: "
                         "felis_amd/tuning.json)")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on this workload first and add its result to the --tuning cache")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB three times (profiler counter scale)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (sharding the slices across "
              "GPUs is a design, not code); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.tune and args.precision == "mixed":
        from felis_amd import tune
        e = tune.tune(ROOT / "felis_amd" / "decks" / f"{args.config}.inp", seed=args.seed, verbose=False)
        tune.write_cache(e, args.tuning)
    g = run_gpu(args)
    d = g["desc"]
    K = args.steps
    value = d.slice_n * K / g["seconds"]
    mixed = args.precision == "mixed"
    prof = g["prof"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "slice·periods/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": g["seconds"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic: quiet-start beam and Gaussian seed generated on the device from the deck",
        "config": {"workload": f"{args.config}: {WORKLOADS[args.config]}", "precision": args.precision,
                   "slices": d.slice_n, "particles": g["np"], "grid": f"{d.nx}x{d.ny}",
                   "harmonics": [d.harmony[i] for i in range(d.n_harm)], "polarisations_solved": g["npol"],
                   "tuning": g["tuning"], "deposit_mode": args.deposit_mode},
        "roofline": roofline(prof, d, g["np"], mixed, g["npol"], args.config, args.precision, g["tuning"]),
        "kernels_ms_per_step": {k: v[0] / args.prof_steps for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])},
    }
    if mixed and args.plain_steps > 0:
        gp = run_gpu(args, precision="plain", steps=args.plain_steps, warmup=1, keep_init=False)
        plain_ms = gp["seconds"] / args.plain_steps * 1e3
        gp["ctx"].close()
        out["plain_fp64"] = {"ms_per_step": plain_ms, "steps": args.plain_steps,
                             "value": d.slice_n * args.plain_steps / gp["seconds"],
                             "speedup_of_batched": plain_ms / out["ms_per_step"]}
    if mixed and args.config == "cfg2" and not args.no_two_pol:
        out["two_polarisation"] = two_polarisation_leg(args)
    if mixed and not args.no_surrogate and g["deck"].mode == 2:
        out["surrogate"] = surrogate_runs(args)
    if not args.no_cpu:
        cpu = run_cpu(args, g)
        out["cpu_baseline"] = {"value": d.slice_n * args.cpu_steps / cpu["seconds"], "unit": "slice·periods/s",
                               "cores": cpu["threads"], "kind": "port",
                               "host_physical_cores": cpu["host_physical_cores"],
                               "host_cpus_available": cpu["host_cpus_available"],
                               "thread_model": "one process, OpenMP threads over the slices (the reference splits "
                                               "slices over MPI ranks, boundary.cpp:33-46)",
                               "sample": f"periods 1-{args.cpu_steps} (steady state, after an untimed period 0) of "
                                         f"the same workload and initial state, FP64 oracle "
                                         f"(oracle/felis_oracle.c), {cpu['threads']} threads"}
        out["parity"] = {"steps": 1 + args.cpu_steps, "max_rel_err_total_energy": gpu_parity(args, g, cpu)}
    if mixed and not args.no_whole_run:
        out.setdefault("parity", {})["whole_run_vs_plain_fp64"] = whole_run_parity(args, g)
    g["ctx"].close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
