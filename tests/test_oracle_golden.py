"""The FP64 oracle against the golden vectors written by the reference's own
compiled routines (tests/golden/make_golden.py), and against closed forms.

CPU only. Every oracle function that a golden vector covers is checked here;
the tolerances follow how the reference reports each quantity (raw doubles,
or a printed line with %g / %10.4g / %12.6g precision).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(GOLD / f"{name}.npz", allow_pickle=False)


def make_state(slice_n=1, n_species=1, nx=4, ny=4, harmony=(1,), num_slice=1, seg_off=None, weight=None, **arr):
    """A minimal oracle state; arrays not given are zero."""
    p = po.Params()
    p.nx, p.ny, p.slice_n, p.n_harm = nx, ny, slice_n, len(harmony)
    for h, H in enumerate(harmony):
        p.harmony[h] = int(H)
    p.num_slice, p.n_species, p.nr, p.sc_l, p.sc_f = num_slice, n_species, 4, 1, 1
    p.bn, p.b_orders, p.wake_on, p.dimension = 4, 2, 0, 3
    for k in ("dx", "dy", "dz", "min_x", "min_y", "ks", "lambda0", "gamR", "dr", "dBessel"):
        setattr(p, k, float(arr.pop(k, 1.0)))
    if seg_off is None:
        seg_off = np.zeros(n_species * slice_n + 1, np.int64)
    if weight is None:
        weight = np.ones(n_species * slice_n)
    n = int(seg_off[-1])
    S = slice_n + 2
    nh = len(harmony)
    z = lambda: np.zeros(n)  # noqa: E731
    fields = dict(Ux=np.zeros((nh, S, ny, nx), np.complex128), Uy=np.zeros((nh, S, ny, nx), np.complex128))
    fields.update({k: np.ascontiguousarray(v) for k, v in arr.items() if k in ("Ux", "Uy")})
    parts = {k: np.ascontiguousarray(arr.get(k, z()), np.float64) for k in ("x", "y", "px", "py", "theta", "gamma")}
    if "gamma" not in arr:
        parts["gamma"] = np.ones(n)
    return po.OracleState(
        params=p, alpha=np.zeros(nx, np.complex128), alphaP=np.zeros(nx, np.complex128),
        beta=np.zeros(ny, np.complex128), betaP=np.zeros(ny, np.complex128), bessel=np.zeros((4, 2)),
        seg_off=np.ascontiguousarray(seg_off, np.int64), seg_weight=np.ascontiguousarray(weight, np.float64),
        ScUx=np.zeros((nh, S, ny, nx), np.complex128), ScUy=np.zeros((nh, S, ny, nx), np.complex128),
        Ez=np.zeros((1, S * 4), np.complex128), wakeE=np.zeros(slice_n), **fields, **parts)


@pytest.mark.parametrize("ns", [1, 3])
def test_periodic_bit_exact(ns):
    """rearrangeParticles.cpp:77-114: theta wrap, including values on the
    period boundaries and a negative denormal."""
    g = load(f"periodic_ns{ns}")
    so = g["seg_off"]
    st = make_state(slice_n=5, n_species=2, num_slice=int(g["num_slice"]), seg_off=so,
                    theta=g["theta_in"].copy())
    assert st.periodic() == 0
    assert np.array_equal(st.theta, g["theta_out"])


@pytest.mark.parametrize("mode", [1, 2])
def test_total_energy(mode):
    """updateTotalEnergy.cpp:25-36: the raw per-harmonic sums (exact doubles)."""
    g = load(f"total_energy_mode{mode}")
    nh, S, ny, nx = g["Ux"].shape
    st = make_state(slice_n=S - 2, nx=nx, ny=ny, harmony=tuple(range(1, nh + 1)), Ux=g["Ux"], Uy=g["Uy"])
    ex, ey = st.total_energy()
    np.testing.assert_allclose(np.concatenate([ex, ey]), g["raw"], rtol=1e-13)
    # the printed z of the appended line
    z = int(g["iteration"]) * float(g["dz"]) + (float(g["min_z"]) + float(g["max_z"])) * 0.5
    assert abs(g["line"][0] - z) <= 1e-14 * max(1.0, abs(z))


@pytest.mark.parametrize("case", ["random", "bunched"])
def test_bfactor(case):
    """saveFile.cpp:109-172: |b_h / total weight| as printed with %10.4g."""
    g = load(f"bfactor_{case}")
    st = make_state(slice_n=6, n_species=2, harmony=tuple(int(h) for h in g["harmony"]), seg_off=g["seg_off"],
                    weight=g["weight"], theta=g["theta"])
    b, tot = st.bfactor()
    mag = np.abs(b / tot)
    assert [float(f"{v:10.4g}") for v in mag] == list(g["line"][1:])
    assert abs(g["line"][0] - (int(g["iteration"]) * float(g["dz"]) + float(g["min_z"]))) < 1e-4 * float(g["dz"])
    # the weight sum is the plain sum of segment weight x count
    cnt = np.diff(g["seg_off"])
    assert tot == pytest.approx(float(np.sum(g["weight"] * cnt)), rel=1e-14)


def test_twiss():
    """twiss.cpp:7-99: emittance (times gamma0), beta, alpha per species as
    printed with %12.6g."""
    g = load("twiss")
    st = make_state(slice_n=5, n_species=2, seg_off=g["seg_off"], weight=g["weight"], x=g["x"], y=g["y"],
                    px=g["px"], py=g["py"], gamma=g["gamma"])
    sums, der = st.twiss()
    got = []
    for s in range(2):
        d = der[s]
        got += [d[0] * float(g["gamma0"]), d[1], d[2], d[3] * float(g["gamma0"]), d[4], d[5]]
    # the oracle's values printed as the reference prints them
    assert [float(f"{v:12.6g}") for v in got] == list(g["line"][1:])
    assert sums[0, 6] + sums[1, 6] == g["x"].size


@pytest.mark.parametrize("tag", ["circ_ac", "circ_dc", "flat_ac"])
def test_wake_function(tag):
    """wakeField.cpp:53-159 for the circular AC/DC and flat AC impedances."""
    g = load(f"wake_function_{tag}")
    dZ = float(g["num_slice"]) * float(g["lambda0"])
    rc, _, wf = po.wake_function(int(g["shape"]), int(g["acdc"]), float(g["radius"]), float(g["cond"]),
                                 float(g["ctau"]), int(g["slice_n"]), dZ)
    assert rc == 0
    scale = np.max(np.abs(g["wakeF"]))
    np.testing.assert_allclose(wf, g["wakeF"], rtol=1e-10, atol=1e-12 * scale)


def test_update_wake():
    """wakeField.cpp:9-50 and makeDensity 161-223: density, wake energy loss."""
    g = load("update_wake")
    st = make_state(slice_n=30, n_species=2, seg_off=g["seg_off"], weight=g["weight"],
                    num_slice=int(g["num_slice"]), lambda0=float(g["lambda0"]))
    den, tot = st.update_wake(g["wakeF"])
    np.testing.assert_allclose(den, g["den"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(st.wakeE, g["wakeE"], rtol=1e-12, atol=1e-15 * np.max(np.abs(g["wakeE"])))
    # summation order differs from the reference's by one rounding
    assert tot == pytest.approx(float(g["total_cnt"]), rel=1e-15)


def _units_quads(g):
    # the reference keeps one air flag per undulator block (mesh.h:44-64, the
    # last one written wins), so each unit takes its block's final flag
    block_air = {int(r[11]): int(r[10]) for r in g["und"]}
    units = []
    for r in g["und"]:
        u = po.UndUnit()
        (u.unit_start, u.unit_end, u.und_start, u.und_end, u.K0, u.slopeK, u.ue, u.lambdaU) = map(float, r[:8])
        u.K0_alpha, u.und_type, u.air = int(r[8]), int(r[9]), block_air[int(r[11])]
        units.append(u)
    quads = []
    for r in g["quad"]:
        q = po.Quad()
        q.q_start, q.q_end, q.g = map(float, r[:3])
        q.list_id = int(r[3])
        quads.append(q)
    return units, quads


def test_update_k_quadg_sequence():
    """updateK_quadG.cpp:8-94 over 1200 periods x 2 halves: tapered K, air
    gaps, helicity change between blocks, two FODO quad lists; K, ue, K_alpha,
    ku, g, lambdaU, type and the current / drift flags, exactly."""
    g = load("update_k_quadg")
    units, quads = _units_quads(g)
    lat = po.Lattice()
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU = 1.87, 0.0, 1.0, 2 * math.pi / 0.026, 0.0, 0.026
    lat.und_type = 1
    k = 0
    for it in range(1200):
        for half in (0.0, 0.5):
            po.update_k_quadg(lat, float(g["dz"]), it, half, units, quads)
            want = g["seq"][k]
            got = [lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU, lat.und_type, lat.current_flag,
                   lat.drift_flag]
            assert got == list(want), (it, half, got, list(want))
            k += 1


def test_load_seed_3d():
    """loadSeed.cpp:25-80: Gaussian seed with focus, curvature and Gouy phase,
    loaded into the harmonic equal to loadH."""
    g = load("load_seed")
    harm = tuple(int(h) for h in g["harmony"])
    nx, ny, sn = int(g["nx"]), int(g["ny"]), int(g["slice_n"])
    st = make_state(slice_n=sn, nx=nx, ny=ny, harmony=harm, num_slice=int(g["num_slice"]),
                    dx=float(g["dx"]), dy=float(g["dy"]), dz=float(g["dz"]), min_x=float(g["min_x"]),
                    min_y=float(g["min_y"]), lambda0=float(g["lambda0"]), ks=float(g["ks"]))
    h_pick = harm.index(int(g["loadH"]))
    po.load_seed_3d(st.params, st.Ux, st.Uy, h_pick, float(g["a0"]), float(g["duration"]), float(g["min_z"]),
                    float(g["zR"]), float(g["focus"]), float(g["spotSigR"]), 0.0, float(g["laserAlpha"]),
                    float(g["laserPsi"]))
    scale = np.max(np.abs(g["Ux"]))
    np.testing.assert_allclose(st.Ux, g["Ux"], rtol=1e-12, atol=1e-14 * scale)
    np.testing.assert_allclose(st.Uy, g["Uy"], rtol=1e-12, atol=1e-14 * scale)


def test_transfer_zplus_single_rank_is_identity():
    """fieldShareZ.cpp:39-78 on one rank exchanges nothing: the slippage of a
    single-GPU run is the local slot shift alone."""
    g = load("transfer_zplus")
    assert np.array_equal(g["f_in"], g["f_out"])


def test_shift_field_closed_form():
    """solveField.cpp:19-47: slots slice_n+1 .. 1 take slots slice_n .. 0 and
    the ghost slot 0 keeps its content (one rank receives nothing)."""
    rng = np.random.default_rng(3)
    nh, sn, ny, nx = 2, 4, 3, 5
    Ux = rng.normal(size=(nh, sn + 2, ny, nx)) + 1j * rng.normal(size=(nh, sn + 2, ny, nx))
    Uy = rng.normal(size=(nh, sn + 2, ny, nx)) + 1j * rng.normal(size=(nh, sn + 2, ny, nx))
    st = make_state(slice_n=sn, nx=nx, ny=ny, harmony=(1, 3), Ux=Ux.copy(), Uy=Uy.copy())
    st.shift()
    for U0, U1 in ((Ux, st.Ux), (Uy, st.Uy)):
        assert np.array_equal(U1[:, 1:sn + 2], U0[:, 0:sn + 1])
        assert np.array_equal(U1[:, 0], U0[:, 0])


def test_bessel_table_against_scipy():
    """boundary.cpp:109-123: J_n(x) on the table grid."""
    from scipy.special import jv
    t = po.bessel_table(400, 5)
    x = 5 * 2 * math.pi / 400 * np.arange(400)
    for n in range(t.shape[1]):
        np.testing.assert_allclose(t[:, n], jv(n, x), atol=1e-14)


def test_abc_closed_form():
    """boundary.cpp:73-106: the absorbing-layer stencil coefficients; with
    sigma = 0 they reduce to -i dz / (4 ks dx^2)."""
    nx, ny, dx, dy, dz, ks = 12, 10, 4e-6, 5e-6, 0.026, 1.1e10
    al, alP, be, beP = po.abc_setup(nx, ny, 3, 0.0, dx, dy, dz, ks)
    np.testing.assert_allclose(al[1:-1], -1j * dz / (4 * ks * dx * dx), rtol=1e-15)
    np.testing.assert_allclose(alP[1:-1], -1j * dz / (4 * ks * dx * dx), rtol=1e-15)
    np.testing.assert_allclose(be[1:-1], -1j * dz / (4 * ks * dy * dy), rtol=1e-15)
    # with the layer on, the interior away from it keeps the free-space value
    al, alP, be, beP = po.abc_setup(nx, ny, 3, 1.0, dx, dy, dz, ks)
    np.testing.assert_allclose(al[4:nx - 4], -1j * dz / (4 * ks * dx * dx), rtol=1e-15)
    assert np.all(al[1:3].real != 0)


@pytest.mark.parametrize("deck", ["test.inp", "deck_crafted.inp"])
def test_find_parameter_matches_reference(deck):
    """findparam.cpp:30-116 through the product's C-ABI (no GPU needed): the
    same keys are found with the same value string, trailing tabs and line
    ends included."""
    from felis_amd import _lib as fl
    rows = json.loads((GOLD / "findparam.json").read_text())[deck]
    path = GOLD / ("deck_reference_sample.inp" if deck == "test.inp" else deck)
    for blk, rank, key, found, val in rows:
        got = fl.find_parameter(blk, rank, key, path)
        assert (got is not None) == bool(found), (blk, rank, key, got, val)
        if found:
            assert got == val, (blk, rank, key, got, val)


@pytest.mark.parametrize("und_type,ka,g", [(1, 1, 0.0), (1, 1, 25.0), (1, -1, 8.0), (2, 0, -12.0)])
def test_transverse_rk4_is_the_matrix_taylor_polynomial(und_type, ka, g):
    """particlePush.cpp:32-117 with ue = 0 (no x-y coupling): the planes are
    the oscillators u' = p/gamma, p' = -f u, whose RK4 step of h = dz/2 is
    exactly the fourth-order Taylor polynomial of exp(h A),
        u <- c u + s p/gamma,  p <- c p - s f u,  w = f/gamma,
        c = 1 - h^2 w/2 + h^4 w^2/24,  s = h (1 - h^2 w/6),
    the form the batched push evaluates; and that polynomial is exp(h A)
    (scipy) to its truncation, O((h^2 w)^3). A dropped stage weight or a
    wrong sign in the focusing term fails both."""
    from scipy.linalg import expm
    rng = np.random.default_rng(11)
    n = 32
    x, y = rng.normal(scale=1e-4, size=(2, n))
    px, py = rng.normal(scale=2e-3, size=(2, n))
    gamma = 4000.0 + 200.0 * rng.random(n)
    dz = 0.026
    st = make_state(seg_off=np.array([0, n]), x=x.copy(), y=y.copy(), px=px.copy(), py=py.copy(),
                    gamma=gamma.copy(), dz=dz)
    lat = po.Lattice()
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.und_type = 2.5, 0.0, float(ka), 2 * np.pi / dz, g, und_type
    assert st.transverse_push(lat) == 0
    h = dz / 2
    bb = lat.K0 ** 2 * 0.25 * lat.ku ** 2
    cc = 1.60217662e-19 * g / 9.10938291e-31 / 299792458.0
    xc, yc = (1.0, 1.0) if und_type == 2 else ((1.0, 0.0) if ka == 1 else (0.0, 1.0))

    def plane(u, p, f, gam):
        w = f / gam
        c = 1 - h * h * w / 2 + h ** 4 * w * w / 24
        s = h * (1 - h * h * w / 6)
        return c * u + s * p / gam, c * p - s * f * u

    fx = bb / gamma * xc + cc
    fy = bb / gamma * yc - cc
    xw, pxw = plane(x, px, fx, gamma)
    yw, pyw = plane(y, py, fy, gamma)
    for got, want in ((st.x, xw), (st.y, yw), (st.px, pxw), (st.py, pyw)):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    for k in range(3):
        for u0, p0, f, u1, p1 in ((x, px, fx, st.x, st.px), (y, py, fy, st.y, st.py)):
            A = np.array([[0.0, 1.0 / gamma[k]], [-f[k], 0.0]])
            e = expm(h * A) @ np.array([u0[k], p0[k]])
            tol = (h * h * abs(f[k]) / gamma[k]) ** 3 * max(abs(u0[k]), abs(p0[k]) * h / gamma[k]) + 1e-18
            assert abs(u1[k] - e[0]) <= tol and abs(p1[k] - e[1]) <= tol * gamma[k] / h + 1e-18
