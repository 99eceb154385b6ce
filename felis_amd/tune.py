"""Autotuner of the batched step's launch shapes (felis_tuning).

For one workload (an input deck) it runs the beam a few warm-up periods into
the undulator, keeps that state on the device (felis_state_save), and then
times every candidate tuning over the same stretch of periods from that same
state (felis_state_restore before each run), a few times each. The candidate
with the smallest median wins, provided its per-period field energies agree
with the default tuning's (the deposit's float atomics reorder sums: 1e-5
relative). Winners go into a JSON cache keyed by the workload's shape, which
`apply` reads.

The tunables:
  step_graph       replay one captured step (a HIP graph) over uniform
                   undulator stretches, or launch each kernel;
  deposit_threads  threads per deposit workgroup (one workgroup per slice);
  deposit_lds      LDS bytes a deposit workgroup uses for its windows (less
                   LDS: more workgroups per CU, but harmonics deposited in
                   chunks);
  adi_threads      threads per field-solve workgroup (one system per
                   workgroup, its LDS allows one per CU: 1024 threads give
                   each SIMD four waves to hide latency with, at 128
                   registers per lane);
  deposit_run      consecutive particles a deposit thread walks (the
                   phases of a beamlet share a cell and are combined in
                   registers; longer runs combine more, shorter runs have
                   more threads in flight);
  adi_lanes        0: the field solve keeps the whole y right-hand side of
                   a system in LDS; 8: only the first 8 lanes' columns, the
                   rest through scratch in HBM, for less LDS per system.

    python -m felis_amd.tune felis_amd/decks/cfg2.inp [--out felis_amd/tuning.json]
"""
from __future__ import annotations

import argparse
import itertools
import json
import time
from pathlib import Path

import numpy as np

from . import _lib as fl

CACHE = Path(__file__).resolve().parent / "tuning.json"
ARCH = "gfx950"
DEFAULT = dict(step_graph=0, deposit_threads=1024, deposit_lds=0, adi_threads=512, deposit_run=8, adi_lanes=0)
SPACE = dict(step_graph=(0, 1), deposit_threads=(256, 512, 1024), deposit_lds=(0, 65536, 32768),
             adi_threads=(512, 1024), deposit_run=(4, 8, 16), adi_lanes=(0, 8))


def key(desc, n_particles: int, npol: int) -> str:
    """The workload shape a tuning is kept for: grid, slices, harmonics,
    particle count (three significant digits) and live polarisations."""
    harms = ",".join(str(desc.harmony[i]) for i in range(desc.n_harm))
    npart = int(round(n_particles, -max(0, len(str(int(n_particles))) - 3)))
    return f"{ARCH}:{desc.nx}x{desc.ny}:s{desc.slice_n}:h{harms}:np{npart}:pol{npol}"


def candidates():
    names = list(SPACE)
    for vals in itertools.product(*(SPACE[n] for n in names)):
        yield dict(zip(names, vals))


def _rel_energy_err(a, b) -> float:
    """Largest difference of the per-period energies (every polarisation and
    harmonic column) relative to that period's total energy."""
    scale = np.maximum(np.abs(b).sum(axis=1, keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def tune(deck_path, seed: int = 20261203, warmup: int = 5, periods: int = 0, repeats: int = 3,
         tol: float = 1e-5, verbose: bool = True) -> dict:
    """Times every candidate on the device; returns the cache entry
    {"key", "best", "ms_per_step", "table"} for the deck's workload."""
    deck = fl.Deck(deck_path)
    desc = deck.desc(fl.MIXED_FP32)
    ctx = fl.Context(desc)
    deck.load(ctx, seed=seed)
    k = key(desc, ctx.np, ctx.live_polarisations())
    wakeF = deck.wake_function()[0] if desc.wake_on else None
    # periods per timed run: about 30 ms of work (a period takes roughly
    # 0.1 ns per particle), between 10 and 200 periods
    if periods <= 0:
        periods = int(np.clip(3.0e8 / max(ctx.np, 1), 10, 200))
    tab = deck.lattice(0, warmup + periods)
    ctx.run(tab[:2 * warmup + 1], 0, warmup, deck.wake_step, wakeF)
    ctx.sync()
    ctx.save_state(0)
    run_tab = tab[2 * warmup:2 * (warmup + periods) + 1]

    def timed(t):
        ctx.set_tuning(**t)
        times, e = [], None
        for r in range(repeats + 1):  # the first run captures the graph, untimed
            ctx.restore_state(0)
            ctx.sync()
            t0 = time.perf_counter()
            e, _, _ = ctx.run(run_tab, warmup, periods, deck.wake_step, wakeF)
            ctx.sync()
            if r:
                times.append(time.perf_counter() - t0)
        return float(np.median(times)) / periods * 1e3, e

    ref_ms, ref_e = timed(DEFAULT)
    table = []
    best, best_ms = dict(DEFAULT), ref_ms
    for t in candidates():
        ms, e = (ref_ms, ref_e) if t == DEFAULT else timed(t)
        err = _rel_energy_err(e, ref_e)
        ok = err <= tol
        table.append(dict(t, ms_per_step=ms, rel_energy_err=err, ok=ok))
        if verbose:
            print(f"tune {t} {ms:.4f} ms/step err {err:.1e}{'' if ok else ' REJECTED'}", flush=True)
        if ok and ms < best_ms:
            best, best_ms = dict(t), ms
    ctx.close()
    deck.close()
    return dict(key=k, best=best, ms_per_step=best_ms, default_ms_per_step=ref_ms, periods=periods,
                repeats=repeats, table=table)


def read_cache(path=None) -> dict:
    p = Path(path) if path else CACHE
    if not p.exists():
        return {}
    with open(p) as f:
        return json.load(f)


def write_cache(entry: dict, path=None):
    p = Path(path) if path else CACHE
    cache = read_cache(p)
    cache[entry["key"]] = {k: v for k, v in entry.items() if k != "key"}
    p.parent.mkdir(parents=True, exist_ok=True)
    with open(p, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
        f.write("\n")


def lookup(ctx, path=None):
    """The cached tuning of this context's workload, or None."""
    if ctx.desc.precision != fl.MIXED_FP32:
        return None
    e = read_cache(path).get(key(ctx.desc, ctx.np, ctx.live_polarisations()))
    return dict(e["best"]) if e else None


def apply(ctx, path=None) -> dict:
    """Sets the cached tuning on ctx (the defaults when the workload has no
    entry); returns the tuning in force."""
    t = lookup(ctx, path) or dict(DEFAULT)
    ctx.set_tuning(**t)
    return t


def _main():
    ap = argparse.ArgumentParser()
    ap.add_argument("deck", nargs="+")
    ap.add_argument("--out", default=str(CACHE))
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args()
    for dk in a.deck:
        e = tune(dk, seed=a.seed, repeats=a.repeats)
        write_cache(e, a.out)
        print(json.dumps(dict(key=e["key"], best=e["best"], ms_per_step=e["ms_per_step"],
                              default_ms_per_step=e["default_ms_per_step"])), flush=True)


if __name__ == "__main__":
    _main()
