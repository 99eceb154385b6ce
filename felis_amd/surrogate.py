"""Learned correction of the batched step's multi-period jumps.

A jump (felis_run_surrogate) advances every slice J undulator periods with
one field solve and one particle push of step J*dz, keeping the slippage of
the J fine steps through window passes: each field slot takes the source of
the J slices it slips past, and each slice is pushed in the field slots that
slip past it, each weighted in time between the field before and after the
jump's solve by the period in which it meets the slice. Its residual against
J fine steps from the same state is, per (slice, harmonic), mostly in the
field the jump emits: E = 2 Sw / H, Sw the windowed source. The correction
network (felis_mlp: BF16 operands on the matrix cores, FP32 accumulation,
FP32 master weights and Adam, all on the GPU) maps each slice's state
before the jump (felis_slice_features: current, field intensity at the
beam, bunching, coherent fraction and phase per harmonic, energy spread and
modulation) to a complex factor e_h on that emission, as (log|e_h|,
arg e_h) per harmonic; after the push the jump's field gains (e_h - 1) E.

The plain batched step is both the fallback and the measure of the error:
jumps run only while the fundamental bunching |b1| stays below the gate
(the linear-gain regime), every training target is the fine steps' own
result, and evaluate() scores a run with jumps against the plain run from
the same initial state.

Pipeline (all on the GPU through the C-ABI):
  training_decks()  random peak current (2-5 kA), energy spread (0.5e-4 to
                    2e-4) and seed power (1-100 kW) on the geometry of the
                    2048-slice workload (129 x 129, harmonics 1, 3, 5), with
                    fewer slices
  pairs()           along a fine run of each deck, at every J-th step while
                    |b1| < gate: the features of the state, J fine steps and
                    one jump from that state (device snapshots), and the
                    least-squares emission factor per (slice, harmonic) as
                    the target (emission_targets)
  train()           Adam on the weighted mean squared error
  save() / load()   the network as .npz (dims, FP32 weights, input affine)
"""
from __future__ import annotations

import json
import math
import tempfile
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib as fl

PKG = Path(__file__).resolve().parent
MODELS = PKG / "models"
DEFAULT_MODEL = MODELS / "jump_corr_J10.npz"


@dataclass
class Model:
    dims: list
    W: list
    b: list
    mu: np.ndarray
    inv_sd: np.ndarray
    periods: int
    gate: float
    meta: dict = field(default_factory=dict)

    def network(self, capacity: int) -> fl.MLP:
        """The network on the GPU with room for `capacity` slices."""
        net = fl.MLP(self.dims, capacity=capacity, seed=1)
        for l, (W, b) in enumerate(zip(self.W, self.b)):
            net.set_layer(l, W, b)
        net.set_input_affine(self.mu, self.inv_sd)
        return net


def save(model: Model, path) -> None:
    arrs = {f"W{l}": w for l, w in enumerate(model.W)}
    arrs.update({f"b{l}": b for l, b in enumerate(model.b)})
    np.savez(path, dims=np.array(model.dims, np.int32), mu=model.mu, inv_sd=model.inv_sd,
             periods=np.int32(model.periods), gate=np.float64(model.gate),
             meta=np.array(json.dumps(model.meta)), **arrs)


def load(path=DEFAULT_MODEL) -> Model:
    z = np.load(path, allow_pickle=False)
    dims = [int(v) for v in z["dims"]]
    L = len(dims) - 1
    return Model(dims=dims, W=[z[f"W{l}"] for l in range(L)], b=[z[f"b{l}"] for l in range(L)], mu=z["mu"],
                 inv_sd=z["inv_sd"], periods=int(z["periods"]), gate=float(z["gate"]),
                 meta=json.loads(str(z["meta"])))


# ---------------------------------------------------------------- data

def training_decks(n: int, seed: int, out_dir, slices: int = 768, length: float = 20.0):
    """n decks of the 2048-slice workload's geometry with `slices` slices and
    a Gaussian current of proportionate width; per deck a random peak current
    in [2, 5] kA, energy spread in [0.5e-4, 2e-4] and seed power in [1, 100]
    kW (log-uniform). Returns [(path, params)]."""
    from .decks import make_decks as md
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        cur = float(rng.uniform(2000.0, 5000.0))
        spread = float(rng.uniform(0.5e-4, 2e-4))
        power = float(10 ** rng.uniform(3.0, 5.0))
        sigma = max(8, int(round(slices * 400 / 2048)))
        text = md.deck(mode="Time_Dependent", slices=slices, n=129, harmonics=(1, 3, 5), length=length,
                       beam=md.gauss_beam(cur, 2048, 8, slices // 2, sigma, 32), seed_power=power,
                       energy_spread_pct=spread * 100.0)
        path = Path(out_dir) / f"train_{i:03d}.inp"
        path.write_text(text)
        out.append((path, dict(current=cur, spread=spread, power=power)))
    return out


def gain_targets(f: np.ndarray, nh: int, periods: int):
    """Complex field gain per (slice, harmonic) from felis_jump_residual field
    sums (both polarisations), as [slice, 2 nh] = (log|g|, arg g), and the
    fine field energy of the fundamental per slice. |g|^2 is the ratio of the
    fine steps' field energy to the jump's and arg g the phase of their
    overlap sum conj(U_jump) U_fine: the corrected field keeps the fine
    steps' energy and phase. (The least-squares factor <a,b>/<a,a> would
    shrink the field by the part of the fine field the jump's shape misses,
    and that loss compounds over the jumps.) The residual is taken
    after the slippage of `periods` slots, so the field solved at slot s is
    compared at slot s + periods: row s of the result is that slot's (rows
    whose field has left the window are marked not ok)."""
    f = np.asarray(f)
    sn = f.shape[1]
    fa = np.zeros_like(f)
    if periods < sn:
        fa[:, :sn - periods] = f[:, periods:]
    ab = (fa[:nh, :, 0] + fa[nh:, :, 0]) + 1j * (fa[:nh, :, 1] + fa[nh:, :, 1])
    aa = fa[:nh, :, 2] + fa[nh:, :, 2]
    bb = fa[:nh, :, 3] + fa[nh:, :, 3]
    ok = aa > 0
    ok = ok & (bb > 0) & (np.abs(ab) > 0)
    ratio = np.where(ok, bb / np.where(ok, aa, 1.0), 1.0)
    T = np.empty((f.shape[1], 2 * nh))
    T[:, 0::2] = (0.5 * np.log(ratio)).T
    T[:, 1::2] = np.where(ok, np.angle(ab), 0.0).T
    return T, bb[0], ok.all(0)


def emission_targets(fe: np.ndarray, f: np.ndarray, nh: int, periods: int):
    """Emission factor per (slice, harmonic) from felis_jump_emission_residual
    sums fe (both polarisations): the fine steps' field is the jump's plus
    (e - 1) E, E the jump's emission, with (e - 1) = sum conj(E) (b - a) /
    sum |E|^2; as [slice, 2 nh] = (log|e|, arg e), plus the fine field energy
    of the fundamental per slice (f: felis_jump_residual field sums) as the
    row weight. The sums are taken after the slippage of `periods` slots:
    row s is the field solved at slot s + 1 (rows whose field has left the
    window, or that took no emission, are not ok)."""
    fe = np.asarray(fe)
    f = np.asarray(f)
    sn = fe.shape[1]
    fa = np.zeros_like(fe)
    ff = np.zeros_like(f)
    if periods < sn:
        fa[:, :sn - periods] = fe[:, periods:]
        ff[:, :sn - periods] = f[:, periods:]
    num = (fa[:nh, :, 0] + fa[nh:, :, 0]) + 1j * (fa[:nh, :, 1] + fa[nh:, :, 1])
    den = fa[:nh, :, 2] + fa[nh:, :, 2]
    ok = den > 0
    e = np.where(ok, 1.0 + num / np.where(ok, den, 1.0), 1.0)
    ok &= np.abs(e) > 0
    e = np.where(ok, e, 1.0)
    T = np.empty((sn, 2 * nh))
    T[:, 0::2] = np.log(np.abs(e)).T
    T[:, 1::2] = np.angle(e).T
    bb = ff[0, :, 3] + ff[nh, :, 3]
    return T, bb, ok.all(0)


def pairs(deck_path, periods: int, gate: float, seed: int, rows_per_jump: int | None = None, rng=None,
          max_step: int | None = None, policy: "Model | None" = None):
    """Training rows of one deck: (features, targets, row weights) at every
    jump point while |b1| < gate. The states follow the fine run, or with a
    policy network the run of corrected jumps that network makes (the states
    it will meet in use); the targets are always the fine steps' result."""
    rng = rng or np.random.default_rng(seed)
    deck = fl.Deck(deck_path)
    d = deck.desc(fl.MIXED_FP32)
    nh = d.n_harm
    ctx = fl.Context(d)
    deck.load(ctx, seed=seed)
    n_steps = deck.max_step if max_step is None else min(max_step, deck.max_step)
    tab = deck.lattice(0, n_steps)
    Xs, Ts, Ws = [], [], []
    k = 0
    J = periods
    pnet = policy.network(capacity=d.slice_n) if policy is not None else None
    # the first step runs under the lattice handed in before the loop: fine
    ctx.run(tab[0:3], 0, 1, outputs=False)
    k = 1
    while k + J <= n_steps:
        b, tot = ctx.bfactor()
        if not abs(b[0]) / tot < gate:
            break
        X = ctx.slice_features()
        ctx.save_state(0)
        ctx.run(tab[2 * k:2 * (k + J) + 1], k, J, outputs=False)
        ctx.save_state(1)
        ctx.restore_state(0)
        _, _, _, kind = ctx.run_surrogate(tab[2 * k:2 * (k + J) + 1], k, J, periods=J, bunching_max=math.inf,
                                          max_jumps=1)
        if kind[0] != 1:  # no jump here (the lattice changes within the J periods)
            ctx.restore_state(1)
            k += J
            continue
        f, _ = ctx.jump_residual(1)
        fe = ctx.jump_emission_residual(1)
        if pnet is None:
            ctx.restore_state(1)
        else:
            ctx.restore_state(0)
            ctx.run_surrogate(tab[2 * k:2 * (k + J) + 1], k, J, periods=J, bunching_max=math.inf, max_jumps=1,
                              net=pnet)
        T, e_fine, okf = emission_targets(fe, f, nh, J)
        w = np.where(okf & np.isfinite(T).all(1), e_fine, 0.0)
        if w.sum() > 0:
            w = np.clip(w / w[w > 0].mean(), 0.0, 10.0)
            sel = np.nonzero(w > 0)[0]
            if rows_per_jump and len(sel) > rows_per_jump:
                sel = rng.choice(sel, rows_per_jump, replace=False)
            Xs.append(X[sel])
            Ts.append(T[sel])
            Ws.append(w[sel])
        k += J
    if pnet is not None:
        pnet.close()
    ctx.close()
    deck.close()
    if not Xs:
        F = 5 + 5 * nh
        return np.zeros((0, F), np.float32), np.zeros((0, 2 * nh), np.float32), np.zeros(0, np.float32)
    return (np.concatenate(Xs).astype(np.float32), np.concatenate(Ts).astype(np.float32),
            np.concatenate(Ws).astype(np.float32))


def dataset(n_decks: int, periods: int = 10, gate: float = 0.05, seed: int = 2026, slices: int = 768,
            rows_per_jump: int | None = 256, log=print, policy: "Model | None" = None):
    """Training rows from n_decks random decks (see training_decks), on the
    fine run's states or on the policy network's corrected-jump states."""
    Xs, Ts, Ws = [], [], []
    rng = np.random.default_rng(seed)
    with tempfile.TemporaryDirectory() as td:
        for i, (path, p) in enumerate(training_decks(n_decks, seed, td, slices=slices)):
            t0 = time.perf_counter()
            X, T, W = pairs(path, periods, gate, seed + 17 * i, rows_per_jump, rng, policy=policy)
            Xs.append(X)
            Ts.append(T)
            Ws.append(W)
            if log:
                log(f"deck {i}: {p['current'] / 1e3:.2f} kA, spread {p['spread']:.2e}, seed {p['power'] / 1e3:.1f} kW:"
                    f" {len(X)} rows in {time.perf_counter() - t0:.1f} s")
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws)


# ---------------------------------------------------------------- training

def train(X, T, W, periods: int, gate: float, hidden=(64, 64), steps: int = 3000, batch: int = 4096,
          lr: float = 2e-3, seed: int = 0, log=print) -> Model:
    """Adam on the GPU (felis_mlp_train_step) on z-scored targets; the input
    z-score is the network's FP32 input affine, the target scale is folded
    into the output layer."""
    n, F = X.shape
    O = T.shape[1]
    mu = X.mean(0).astype(np.float32)
    sd = (X.std(0) + 1e-6).astype(np.float32)
    tmu = T.mean(0).astype(np.float32)
    tsd = (T.std(0) + 1e-6).astype(np.float32)
    Tn = ((T - tmu) / tsd).astype(np.float32)
    Wn = (W / W.mean()).astype(np.float32)
    dims = [F, *hidden, O]
    batch = min(batch, n)
    net = fl.MLP(dims, capacity=batch, seed=seed + 1)
    net.set_input_affine(mu, (1.0 / sd).astype(np.float32))
    rng = np.random.default_rng(seed)
    losses = []
    for it in range(steps):
        idx = rng.integers(0, n, batch)
        lr_t = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        losses.append(net.train_step(X[idx], Tn[idx], lr_t, Wn[idx]))
        if log and (it % max(1, steps // 10) == 0 or it == steps - 1):
            log(f"step {it}: loss {np.mean(losses[-50:]):.4f}")
    Ws_, bs_ = [], []
    for l in range(len(dims) - 1):
        Wl, bl = net.get_layer(l)
        Ws_.append(Wl)
        bs_.append(bl)
    net.close()
    Ws_[-1] = (Ws_[-1] * tsd[None, :]).astype(np.float32)
    bs_[-1] = (bs_[-1] * tsd + tmu).astype(np.float32)
    return Model(dims=dims, W=Ws_, b=bs_, mu=mu, inv_sd=(1.0 / sd).astype(np.float32), periods=periods, gate=gate,
                 meta=dict(rows=int(n), steps=steps, batch=batch, hidden=list(hidden), final_loss=float(
                     np.mean(losses[-50:]))))


def fit(n_decks: int = 200, rounds: int = 2, periods: int = 10, gate: float = 0.05, seed: int = 2026,
        slices: int = 768, rows_per_jump: int = 256, steps: int = 3000, hidden=(64, 64), log=print) -> Model:
    """The full training: a network on the fine runs' states, then `rounds`-1
    rounds on the states of the previous network's corrected jumps (new
    random decks each round), each time trained on all rows so far."""
    Xs, Ts, Ws = [], [], []
    model = None
    for r in range(rounds):
        t0 = time.perf_counter()
        X, T, W = dataset(n_decks, periods, gate, seed + 1000 * r, slices, rows_per_jump, log=None, policy=model)
        Xs.append(X)
        Ts.append(T)
        Ws.append(W)
        t1 = time.perf_counter()
        model = train(np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws), periods, gate, hidden=hidden,
                      steps=steps, seed=seed + r, log=None)
        if log:
            log(f"round {r}: {len(X)} new rows ({sum(len(x) for x in Xs)} in all) from {n_decks} decks in "
                f"{t1 - t0:.1f} s, trained in {time.perf_counter() - t1:.1f} s, loss {model.meta['final_loss']:.4f}")
    model.meta.update(decks_per_round=n_decks, rounds=rounds, slices=slices, seed=seed)
    return model


def predict(model: Model, X) -> np.ndarray:
    """The network's outputs for feature rows X on the GPU."""
    net = model.network(capacity=max(1, len(X)))
    try:
        return net.forward(X)
    finally:
        net.close()


def fit_quality(model: Model, X, T, W) -> dict:
    """Weighted mean squared error of the predicted (log|g|, arg g) against
    the targets, next to the error of predicting no correction (g = 1)."""
    Y = predict(model, X)
    w = W / W.sum()
    mse = float(np.sum(w[:, None] * (Y - T) ** 2))
    base = float(np.sum(w[:, None] * T ** 2))
    return dict(mse=mse, mse_no_correction=base, ratio=mse / base if base > 0 else float("nan"))


# ------------------------------------------------- field-to-field network

FNO_MODEL = MODELS / "jump_fno_J10.npz"


@dataclass
class FnoModel:
    """The field-to-field correction (felis_fno): parameters, input affine
    and shape."""
    P: np.ndarray
    mu: np.ndarray
    inv_sd: np.ndarray
    dims: dict
    periods: int
    gate: float
    meta: dict = field(default_factory=dict)

    def network(self, capacity: int) -> fl.FNO:
        d = self.dims
        net = fl.FNO(d["cin"], d["cout"], d["width"], d["modes"], d["layers"], d["proj"], capacity=capacity, seed=1)
        net.set_params(self.P)
        net.set_input_affine(self.mu, self.inv_sd)
        return net


def save_fno(model: FnoModel, path) -> None:
    np.savez(path, P=model.P, mu=model.mu, inv_sd=model.inv_sd, dims=np.array(json.dumps(model.dims)),
             periods=np.int32(model.periods), gate=np.float64(model.gate), meta=np.array(json.dumps(model.meta)))


def load_fno(path=FNO_MODEL) -> FnoModel:
    z = np.load(path, allow_pickle=False)
    return FnoModel(P=z["P"], mu=z["mu"], inv_sd=z["inv_sd"], dims=json.loads(str(z["dims"])),
                    periods=int(z["periods"]), gate=float(z["gate"]), meta=json.loads(str(z["meta"])))


def fno_pairs(deck_path, periods: int, gate: float, seed: int, rows_per_jump: int, rng):
    """Training maps of one deck along its fine run: at every jump point
    while |b1| < gate, J fine steps and one jump from the same state, then
    felis_fno_jump_maps for `rows_per_jump` random slices whose field is
    still in the window and took an emission. Returns X, E, R, weights (the
    fine field energy of the fundamental at the slice, over its mean, at
    most 10)."""
    deck = fl.Deck(deck_path)
    d = deck.desc(fl.MIXED_FP32)
    ctx = fl.Context(d)
    deck.load(ctx, seed=seed)
    n_steps = deck.max_step
    tab = deck.lattice(0, n_steps)
    J = periods
    Xs, Es, Rs, Ws = [], [], [], []
    ctx.run(tab[0:3], 0, 1, outputs=False)
    k = 1
    while k + J <= n_steps:
        b, tot = ctx.bfactor()
        if not abs(b[0]) / tot < gate:
            break
        F = ctx.slice_features()
        ctx.save_state(0)
        ctx.run(tab[2 * k:2 * (k + J) + 1], k, J, outputs=False)
        ctx.save_state(1)
        ctx.restore_state(0)
        _, _, _, kind = ctx.run_surrogate(tab[2 * k:2 * (k + J) + 1], k, J, periods=J, bunching_max=math.inf,
                                          max_jumps=1)
        if kind[0] == 1:
            rows = rng.choice(d.slice_n - J, min(rows_per_jump, d.slice_n - J), replace=False)
            X, E, R, info = ctx.fno_jump_maps(1, F, rows)
            ok = info[:, 0, 0] > 0
            Xs.append(X[ok])
            Es.append(E[ok])
            Rs.append(R[ok])
            Ws.append(info[ok, 2, 0])
        ctx.restore_state(1)
        k += J
    ctx.close()
    deck.close()
    if not Xs:
        return None
    W = np.concatenate(Ws).astype(np.float64)
    W = np.clip(W / max(W.mean(), 1e-300), 0.0, 10.0).astype(np.float32)
    return np.concatenate(Xs), np.concatenate(Es), np.concatenate(Rs), W


def fit_fno(n_decks: int = 40, periods: int = 10, gate: float = 0.05, seed: int = 2026, slices: int = 768,
            rows_per_jump: int = 6, steps: int = 3000, batch: int = 16, lr: float = 1e-3, width: int = 32,
            modes: int = 8, layers: int = 4, proj: int = 64, log=print) -> FnoModel:
    """Maps from n_decks random decks (training_decks), then Adam on the
    GPU (felis_fno_train_step) with a cosine schedule. The input affine
    z-scores the scalar channels (the maps are already normalised)."""
    rng = np.random.default_rng(seed)
    Xs, Es, Rs, Ws = [], [], [], []
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as td:
        for i, (path, p) in enumerate(training_decks(n_decks, seed, td, slices=slices)):
            out = fno_pairs(path, periods, gate, seed + 17 * i, rows_per_jump, rng)
            if out is not None:
                Xs.append(out[0])
                Es.append(out[1])
                Rs.append(out[2])
                Ws.append(out[3])
    X, E, R, W = (np.concatenate(a) for a in (Xs, Es, Rs, Ws))
    if log:
        log(f"fno data: {len(X)} maps from {n_decks} decks in {time.perf_counter() - t0:.1f} s")
    nh = E.shape[2] // 2
    # each (map, harmonic) enters the loss at the scale of its own residual:
    # E and R are both divided by the rms of R (the factor that minimises
    # the loss does not change), so maps whose residual the emission cannot
    # explain do not dominate the ones it can
    for i in range(0, len(R), 256):
        r2 = R[i:i + 256, :, 0::2] ** 2 + R[i:i + 256, :, 1::2] ** 2
        c = 1.0 / np.sqrt(r2.mean(1) + 1e-2)  # [maps][nh]
        cc = np.repeat(c, 2, axis=1)[:, None, :].astype(np.float32)
        E[i:i + 256] *= cc
        R[i:i + 256] *= cc
    cin = X.shape[2]
    mu = np.zeros(cin, np.float32)
    sd = np.ones(cin, np.float32)
    sc = slice(4 * nh, cin - 2)  # the slice's scalars; maps and coordinates stay as they are
    mu[sc] = X[:, 0, sc].mean(0)
    sd[sc] = X[:, 0, sc].std(0) + 1e-6
    net = fl.FNO(cin, 2 * nh, width, modes, layers, proj, capacity=batch, seed=seed + 1)
    net.set_input_affine(mu, (1.0 / sd).astype(np.float32))
    # the output layer (Q2 and c2, the last proj * cout + cout parameters)
    # starts at zero: no correction until the data asks for one
    P0 = net.params()
    P0[len(P0) - (proj * 2 * nh + 2 * nh):] = 0.0
    net.set_params(P0)
    n = len(X)
    losses = []
    t1 = time.perf_counter()
    for it in range(steps):
        idx = rng.integers(0, n, batch)
        lr_t = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        losses.append(net.train_step(X[idx], E[idx], R[idx], W[idx], lr_t))
        if log and (it % max(1, steps // 10) == 0 or it == steps - 1):
            log(f"fno step {it}: loss {np.mean(losses[-50:]):.4f}")
    P = net.params()
    net.close()
    # the loss of the best constant factor per (map, harmonic), the per-slice
    # emission factor's model: sum w (|R|^2 - |<E, R>|^2 / |E|^2) / P
    best = 0.0
    for i in range(0, n, 256):
        e = E[i:i + 256, :, 0::2] + 1j * E[i:i + 256, :, 1::2]
        r = R[i:i + 256, :, 0::2] + 1j * R[i:i + 256, :, 1::2]
        ee = np.sum(np.abs(e) ** 2, axis=1)
        er = np.sum(np.conj(e) * r, axis=1)
        rr = np.sum(np.abs(r) ** 2, axis=1)
        best += float(np.sum(W[i:i + 256, None] * (rr - np.abs(er) ** 2 / np.maximum(ee, 1e-30))))
    best /= n * 4096
    # the loss of no correction (o = 0) on the same maps: sum w |R|^2
    base = float(sum(np.sum(W[i:i + 256, None, None] * R[i:i + 256].astype(np.float64) ** 2)
                     for i in range(0, n, 256)) / (n * 4096))
    dims = dict(cin=cin, cout=2 * nh, width=width, modes=modes, layers=layers, proj=proj)
    return FnoModel(P=P, mu=mu, inv_sd=(1.0 / sd).astype(np.float32), dims=dims, periods=periods, gate=gate,
                    meta=dict(maps=int(n), decks=n_decks, slices=slices, steps=steps, batch=batch, lr=lr,
                              final_loss=float(np.mean(losses[-100:])), first_loss=float(np.mean(losses[:20])),
                              loss_no_correction=base, loss_best_constant_factor=best, seed=seed,
                              train_seconds=time.perf_counter() - t1))


# ---------------------------------------------------------------- evaluation

def saturation(z, P):
    """Saturation point of a gain curve: the first local maximum of P whose
    value is not exceeded within the next tenth of the curve, else the end."""
    P = np.asarray(P)
    n = len(P)
    span = max(2, n // 10)
    for k in range(1, n - 1):
        if P[k] >= P[k - 1] and P[k] >= np.max(P[k + 1:min(n, k + 1 + span)]):
            return float(z[k]), float(P[k]), True
    return float(z[-1]), float(P[-1]), False


def run_curves(deck_path, seed: int, periods: int | None = None, gate: float = 0.05, net=None,
               n_steps: int | None = None, net_gate: float = 0.0, fno=None):
    """One run of the deck: plain batched (periods None) or with jumps, with
    the workload's autotuned launch shapes. Returns steps, total field energy
    per step (NaN inside jumps), |b1|, kinds and wall seconds of the step
    loop."""
    from . import tune
    deck = fl.Deck(deck_path)
    d = deck.desc(fl.MIXED_FP32)
    nh = d.n_harm
    ctx = fl.Context(d)
    deck.load(ctx, seed=seed)
    tune.apply(ctx)  # the workload's autotuned launch shapes (defaults when it has none)
    N = deck.max_step if n_steps is None else n_steps
    tab = deck.lattice(0, N)
    ctx.sync()
    t0 = time.perf_counter()
    if periods is None:
        e, b, _ = ctx.run(tab, 0, N)
        kind = np.zeros(N, np.int32)
    else:
        e, b, _, kind = ctx.run_surrogate(tab, 0, N, periods=periods, bunching_max=gate, net=net,
                                          net_bunching_max=net_gate, fno=fno)
    ctx.sync()
    dt = time.perf_counter() - t0
    E = e[:, :nh].sum(1) + e[:, nh:].sum(1)
    b1 = np.abs(b[:, 0] + 1j * b[:, 1]) / b[:, 2 * nh]
    # the state after the last step: slice energies of the final field
    fin = ctx.slice_energy().sum()
    ctx.close()
    deck.close()
    return dict(E=E, b1=b1, kind=kind, seconds=dt, final=float(fin), steps=N)


def _main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train the jump correction network on the GPU")
    ap.add_argument("--decks", type=int, default=100, help="random decks per round")
    ap.add_argument("--rounds", type=int, default=2, help="rounds (the first on the fine runs' states)")
    ap.add_argument("--periods", type=int, default=10)
    ap.add_argument("--gate", type=float, default=0.05)
    ap.add_argument("--slices", type=int, default=768)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=None)
    ap.add_argument("--fno", action="store_true",
                    help="train the field-to-field network (felis_fno) instead: --decks random decks (40 for the "
                         "committed model), --steps Adam steps (4000), 6 maps per jump point")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.fno:
        model = fit_fno(n_decks=a.decks, periods=a.periods, gate=a.gate, seed=2026, slices=a.slices, steps=a.steps,
                        log=lambda m: print(m, flush=True))
        out = a.out or str(FNO_MODEL)
        save_fno(model, out)
        print(f"saved {out} after {time.perf_counter() - t0:.1f} s", flush=True)
        return
    model = fit(n_decks=a.decks, rounds=a.rounds, periods=a.periods, gate=a.gate, seed=a.seed, slices=a.slices,
                steps=a.steps, log=lambda m: print(m, flush=True))
    out = a.out or str(DEFAULT_MODEL)
    Path(out).parent.mkdir(parents=True, exist_ok=True)
    save(model, out)
    print(f"saved {out} after {time.perf_counter() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    _main()
