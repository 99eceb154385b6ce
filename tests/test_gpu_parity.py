"""GPU parity: each C-ABI operator against the FP64 oracle on the same seeded state.

Tolerances: the plain FP64 path differs from the oracle only by rounding
(precomputed reciprocals in the Thomas sweeps, atomic summation order in
the deposit), so per-phase results agree to 1e-11 relative or better;
integer/index work (theta wrap, slippage) is bit-exact.
"""
import math

import numpy as np
import pytest

import fel_cases as fc
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")


def _ctx(case, precision=0):
    return fc.gpu_context(case, precision)


CASES = {
    "planar_h1": dict(),
    "planar_h135": dict(harmonics=(1, 3, 5), nx=21, ny=19),
    "helical_h12": dict(harmonics=(1, 2), ue=0.6, und_type=2),
    "ka_minus1_h123": dict(harmonics=(1, 2, 3), ue=0.3, K0_alpha=-1),
    "two_species_uneven": dict(n_species=2, uneven=True, ppslice=96),
    # radial_grids 3 and a wide beam: in-grid particles beyond the radial grid carry Em
    # (particlePush.cpp:248, 283-304), the radial source spills (A.5)
    "em_carry_nr3": dict(nr=3, sig=60e-6, harmonics=(1, 3), sc_l=2),
}


@pytest.fixture(params=list(CASES))
def case(request):
    return fc.make_case(**CASES[request.param])


def test_transverse_push(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    for half in (0, 1):
        lat = case.lattice(g=(12.0 if half else -7.0))
        assert o.transverse_push(lat) == 0
        ctx.transverse_push(fc.to_lib_lattice(lat))
    p = ctx.download_particles()
    for k in ("x", "y", "px", "py"):
        assert fc.rel_err(p[k], getattr(o, k)) < 1e-13, k
    assert np.array_equal(p["theta"], o.theta) and np.array_equal(p["gamma"], o.gamma)


def test_push_theta_gamma(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    lat = case.lattice()
    assert o.push_theta_gamma(lat) == 0
    ctx.push_theta_gamma(fc.to_lib_lattice(lat))
    p = ctx.download_particles()
    th0 = case.state["theta"]
    g0 = case.state["gamma"]
    # compare the increments: they carry the physics
    assert fc.rel_err(p["theta"] - th0, o.theta - th0) < 1e-11
    assert fc.rel_err(p["gamma"] - g0, o.gamma - g0) < 1e-10
    assert np.array_equal(p["x"], o.x)


def test_drift(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    lat = case.lattice(ku=0.0, drift_flag=1, current_flag=0)
    assert o.drift_theta_gamma(lat) == 0
    ctx.drift_theta_gamma(fc.to_lib_lattice(lat))
    p = ctx.download_particles()
    th0 = case.state["theta"]
    assert fc.rel_err(p["theta"] - th0, o.theta - th0) < 1e-12
    assert np.array_equal(p["gamma"], o.gamma)


def test_periodic_bit_exact():
    case = fc.make_case(num_slice=2)
    s = case.state
    rng = np.random.default_rng(5)
    dth = 2 * math.pi * 2
    s["theta"][:] = rng.uniform(-3 * dth, 4 * dth, s["theta"].size)
    s["theta"][:4] = [-dth, 0.0, dth, -1e-17]
    o = case.oracle_state()
    ctx = _ctx(case)
    assert o.periodic() == 0
    ctx.periodic()
    assert np.array_equal(ctx.download_particles()["theta"], o.theta)


def test_solve_field_with_current(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    lat = case.lattice()
    assert o.solve_field(lat) == 0
    ctx.solve_field(fc.to_lib_lattice(lat))
    for which, ref in ((fl.SCUX, o.ScUx), (fl.SCUY, o.ScUy)):
        got = ctx.download_field(which)
        assert fc.rel_err(got, ref) < 1e-12
    ez = ctx.download_field(fl.EZ)
    assert fc.rel_err(ez, o.Ez) < 1e-12
    for which, ref in ((fl.UX, o.Ux), (fl.UY, o.Uy)):
        got = ctx.download_field(which)
        assert fc.rel_err(got, ref) < 1e-12


def test_solve_field_no_current(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    lat = case.lattice(current_flag=0)
    assert o.solve_field(lat) == 0
    ctx.solve_field(fc.to_lib_lattice(lat))
    for which, ref in ((fl.UX, o.Ux), (fl.UY, o.Uy)):
        assert fc.rel_err(ctx.download_field(which), ref) < 1e-12


def test_shift_bit_exact(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    for _ in range(3):
        o.shift()
        ctx.shift()
    for which, ref in ((fl.UX, o.Ux), (fl.UY, o.Uy)):
        assert np.array_equal(ctx.download_field(which), ref)


def test_diagnostics(case):
    o = case.oracle_state()
    ctx = _ctx(case)
    ex, ey = o.total_energy()
    gx, gy = ctx.total_energy()
    assert fc.rel_err(gx, ex) < 1e-13 and fc.rel_err(gy, ey) < 1e-13
    b, tot = o.bfactor()
    gb, gtot = ctx.bfactor()
    assert abs(gtot - tot) <= 1e-12 * tot
    assert np.max(np.abs(gb - b)) <= 1e-11 * tot
    sums, _ = o.twiss()
    gs = ctx.twiss()
    assert fc.rel_err(gs, sums) < 1e-12


def test_update_wake():
    case = fc.make_case(slice_n=40, n_species=2, uneven=True)
    o = case.oracle_state()
    ctx = _ctx(case)
    rng = np.random.default_rng(3)
    wakeF = 1e14 * rng.normal(size=case.slice_n)
    den, tot = o.update_wake(wakeF)
    wE, gden, gtot = ctx.update_wake(wakeF)
    # the line density is host arithmetic in both; contraction differs by an ulp
    assert abs(gtot - tot) <= 1e-15 * tot
    assert fc.rel_err(gden, den) < 1e-15
    assert fc.rel_err(wE, o.wakeE) < 1e-14


@pytest.mark.parametrize("harm,first_cf,sc_l,sc_f", [((1,), 0, 1, 1), ((1, 3, 5), 0, 1, 1), ((1, 2), 0, 1, 1),
                                                     ((1, 3), 1, 1, 1), ((1, 3), 1, 3, 1), ((1,), 1, 2, 2)])
def test_run_mixed_matches_oracle(harm, first_cf, sc_l, sc_f):
    """Batched FP32/FP64 loop (fused kernels) against the FP64 oracle loop,
    with one, two and three longitudinal space-charge modes and two azimuthal
    ones (solveField.cpp:315-398, particlePush.cpp:287-334).

    Tolerance: FP32 coupling, FP64 phase; after 12 periods the total energy
    agrees to 1e-4 relative and the bunching factors to 1e-3 of the beam's
    total weight (they start near the quiet-start noise floor)."""
    case = fc.make_case(slice_n=6, harmonics=harm, ppslice=256, nx=33, ny=31, ue=0.3 if 2 in harm else 0.0,
                        und_type=2 if 2 in harm else 1, sc_l=sc_l, sc_f=sc_f)
    o = case.oracle_state()
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    nsteps = 12
    tab = [case.lattice(current_flag=first_cf)] + [case.lattice() for _ in range(2 * nsteps)]
    rc, e_o, b_o, t_o = o.run(tab, 0, nsteps, True)
    assert rc == 0
    e_g, b_g, t_g = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    nh = len(harm)
    tot_g = e_g[:, :nh].sum(1) + e_g[:, nh:].sum(1)
    tot_o = e_o[:, :nh].sum(1) + e_o[:, nh:].sum(1)
    assert np.max(np.abs(tot_g - tot_o) / tot_o) < 1e-4
    bg = np.abs(b_g[:, 0:2 * nh:2] + 1j * b_g[:, 1:2 * nh:2]) / b_g[:, 2 * nh:2 * nh + 1]
    assert np.max(np.abs(bg - b_o[:, :nh])) < 1e-3
    assert fc.rel_err(t_g, t_o) < 1e-4
    p = ctx.download_particles()
    assert fc.rel_err(p["gamma"] - case.state["gamma"], o.gamma - case.state["gamma"]) < 1e-3


def _carried(o):
    """In-grid particles at idxR >= nr, from the oracle state's positions."""
    p = o.params
    ix = np.trunc((o.x - p.min_x) / p.dx)
    iy = np.trunc((o.y - p.min_y) / p.dy)
    inside = (ix >= 0) & (ix < p.nx - 1) & (iy >= 0) & (iy < p.ny - 1)
    return inside & (np.trunc(np.sqrt(o.x ** 2 + o.y ** 2) / p.dr + 0.5) >= p.nr)


def test_run_mixed_carries_em():
    """particlePush.cpp:248, 283-304 in the batched step: with radial_grids 2
    about a tenth of the in-grid particles lie at idxR >= nr and take the Em of
    the particle pushed before them (flat order: slice, then particle). One
    period from a prebunched beam, so that Ez is large: every particle's gamma
    change agrees with the oracle's to 2e-4 of the largest, and the carried
    particles' Em terms (the oracle's push with and without Ez) are far larger
    than the differences, so a carried particle with a wrong or zero Em fails."""
    case = fc.make_case(slice_n=6, harmonics=(1,), ppslice=256, nx=33, ny=31, nr=3, sc_l=2)
    th = case.state["theta"]
    th[:] = th + 0.8 * np.sin(th)
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2)]
    g0 = case.state["gamma"].copy()
    o = case.oracle_state()
    rc, _, _, _ = o.run(tab, 0, 1, True)
    assert rc == 0
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    ctx.run([fc.to_lib_lattice(l) for l in tab], 0, 1)
    dg_g = ctx.download_particles()["gamma"] - g0
    dg_o = o.gamma - g0
    diff = np.max(np.abs(dg_g - dg_o))
    assert diff < 2e-4 * np.max(np.abs(dg_o))
    # the Em terms: the same period up to the push, with and without Ez
    pushed = []
    for zero_ez in (False, True):
        q = case.oracle_state()
        assert q.solve_field(tab[0]) == 0
        assert q.transverse_push(tab[1]) == 0 and q.transverse_push(tab[2]) == 0
        carried = _carried(q)
        if zero_ez:
            q.Ez[:] = 0
        assert q.push_theta_gamma(tab[2]) == 0
        pushed.append(q.gamma.copy())
    em_term = np.abs(pushed[0] - pushed[1])
    assert carried.mean() > 0.05
    assert np.min(np.sort(em_term[carried])[-20:]) > 30 * diff


def test_run_mixed_radial_source_takes_each_particles_weights():
    """solveField.cpp:322-359 in the batched deposit: every particle adds to
    bins idxR and idxR+1 with its own radial weights. Here every particle has
    its own position (no beamlets), so consecutive particles in one bin have
    different weights; after one period with the deposit on, Ez (solved from
    that source) agrees with the oracle's to 1e-5 of its largest value (the
    positions are FP32 in the batched path). Two longitudinal modes."""
    case = fc.make_case(slice_n=4, harmonics=(1,), ppslice=512, nx=33, ny=31, in_phases=1, sc_l=2)
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2)]
    o = case.oracle_state()
    rc, _, _, _ = o.run(tab, 0, 1, True)
    assert rc == 0
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    ctx.run([fc.to_lib_lattice(l) for l in tab], 0, 1)
    ez = ctx.download_field(fl.EZ).reshape(o.Ez.shape)
    sl = slice(case.desc_kw["nr"] * 1, case.desc_kw["nr"] * 5)  # slots 1..4
    want = o.Ez[:, sl]
    assert np.max(np.abs(want)) > 0
    assert np.max(np.abs(ez[:, sl] - want)) < 1e-5 * np.max(np.abs(want))


@pytest.mark.parametrize("mode", ["atomic", "fixed"])
def test_run_mixed_window_from_the_push_tile_boxes(mode):
    """After the first period the batched deposit takes each slice's window
    from the boxes of in-grid cells the push leaves per tile (the union over
    every species' tiles of the slice) instead of reading the positions again.
    Two species with uneven slices, every particle at its own position,
    current on throughout, both deposit modes: after 3 periods Ez (solved
    from the last deposit's radial source) agrees with the oracle's to 1e-5
    of its largest value and the loop's diagnostics as in
    test_run_mixed_matches_oracle; a window that missed a cell would drop
    that cell's deposit."""
    case = fc.make_case(slice_n=5, harmonics=(1, 3), n_species=2, ppslice=384, nx=33, ny=31, in_phases=1,
                        uneven=True)
    nsteps = 3
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2 * nsteps)]
    o = case.oracle_state()
    rc, e_o, b_o, _ = o.run(tab, 0, nsteps, True)
    assert rc == 0
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    if mode == "fixed":
        ctx.set_deposit_mode(fl.DEPOSIT_FIXED)
    e_g, b_g, _ = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    ez = ctx.download_field(fl.EZ).reshape(o.Ez.shape)
    nr = case.desc_kw["nr"]
    sl = slice(nr * 1, nr * (case.slice_n + 1))
    want = o.Ez[:, sl]
    assert np.max(np.abs(want)) > 0
    assert np.max(np.abs(ez[:, sl] - want)) < 1e-5 * np.max(np.abs(want))
    tot_g, tot_o = e_g[:, :2].sum(1), e_o[:, :2].sum(1)
    assert np.max(np.abs(tot_g - tot_o) / tot_o) < 1e-4
    bg = np.abs(b_g[:, 0:4:2] + 1j * b_g[:, 1:4:2]) / b_g[:, 4:5]
    assert np.max(np.abs(bg - b_o[:, :2])) < 1e-3


def test_run_mixed_beam_partly_outside_the_grid():
    """A beam wider than the field grid (rms 110 um in a +-300 um box): the
    particles outside the grid are pushed without field (particlePush.cpp:
    254-257 skips them) and deposit nothing (solveField.cpp:273), and the
    push's tile boxes leave them out of the deposit windows. 12 periods of
    the batched loop against the oracle, tolerances as in
    test_run_mixed_matches_oracle."""
    case = fc.make_case(slice_n=4, harmonics=(1, 3), ppslice=512, nx=33, ny=31, sig=110e-6, in_phases=1)
    o = case.oracle_state()
    p = o.params
    ix = np.trunc((o.x - p.min_x) / p.dx)
    iy = np.trunc((o.y - p.min_y) / p.dy)
    outside = ~((ix >= 0) & (ix < p.nx - 1) & (iy >= 0) & (iy < p.ny - 1))
    assert 0.005 < outside.mean() < 0.3
    nsteps = 12
    tab = [case.lattice(current_flag=1)] + [case.lattice() for _ in range(2 * nsteps)]
    rc, e_o, b_o, t_o = o.run(tab, 0, nsteps, True)
    assert rc == 0
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    e_g, b_g, t_g = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    tot_g, tot_o = e_g[:, :2].sum(1), e_o[:, :2].sum(1)
    assert np.max(np.abs(tot_g - tot_o) / tot_o) < 1e-4
    bg = np.abs(b_g[:, 0:4:2] + 1j * b_g[:, 1:4:2]) / b_g[:, 4:5]
    assert np.max(np.abs(bg - b_o[:, :2])) < 1e-3
    assert fc.rel_err(t_g, t_o) < 1e-4


@pytest.mark.parametrize("n", [33, 151, 171], ids=["fused_adi", "split_lds_adi", "two_pass_adi"])
def test_run_mixed_lattice_and_wake(n):
    """The batched loop through the features of the multi-undulator workload
    (4096 slices, 151 x 151): a linear K taper, alternating quad gradients, a
    two-period drift between undulator sections (no current, no slippage) and
    wake updates every four periods from the circular AC resistive-wall wake
    (radius 3 mm, 5.8e7 S/m, ctau 2.4 um), against the FP64 oracle loop. A
    151 x 151 system does not fit one workgroup's LDS whole: the fused ADI
    keeps 80 columns in LDS and passes the rest through scratch; 171 x 171
    takes the two-pass solve. Tolerances as in test_run_mixed_matches_oracle."""
    case = fc.make_case(slice_n=6, harmonics=(1,), ppslice=256, nx=n, ny=n)
    o = case.oracle_state()
    ctx = fc.gpu_context(case, precision=fl.MIXED_FP32)
    nsteps = 12
    base = case.lattice()
    tab = [case.lattice(current_flag=0)]
    for k in range(nsteps):
        drift = k in (5, 6)
        for half in (0, 1):
            lat = case.lattice(K0=base.K0 * (1.0 - 0.002 * k), g=(1 if (2 * k + half) % 4 < 2 else -1) * base.g)
            if drift:
                lat.ku, lat.drift_flag, lat.current_flag = 0.0, 1, 0
            tab.append(lat)
    d = case.desc_kw
    rc, _, wakeF = po.wake_function(2, 1, 3e-3, 5.8e7, 2.4e-6, d["slice_n"], d["num_slice"] * d["lambda0"])
    assert rc == 0
    rc, e_o, b_o, t_o = o.run(tab, 0, nsteps, True, 4, wakeF)
    assert rc == 0
    e_g, b_g, t_g = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps, 4, wakeF)
    tot_g = e_g[:, :1].sum(1) + e_g[:, 1:].sum(1)
    tot_o = e_o[:, :1].sum(1) + e_o[:, 1:].sum(1)
    assert np.max(np.abs(tot_g - tot_o) / tot_o) < 1e-4
    bg = np.abs(b_g[:, 0] + 1j * b_g[:, 1]) / b_g[:, 2]
    assert np.max(np.abs(bg - b_o[:, 0])) < 1e-3
    assert fc.rel_err(t_g, t_o) < 1e-4
    p = ctx.download_particles()
    assert fc.rel_err(p["gamma"] - case.state["gamma"], o.gamma - case.state["gamma"]) < 1e-3


def test_run_plain_matches_oracle_loop():
    """The whole step loop (main.cpp:92-156) for 12 steps: gain curve and bunching."""
    case = fc.make_case(slice_n=6, harmonics=(1, 3), ppslice=128)
    o = case.oracle_state()
    ctx = _ctx(case)
    nsteps = 12
    lat0 = case.lattice(current_flag=0)
    tab = [lat0] + [case.lattice() for _ in range(2 * nsteps)]
    rc, e_o, b_o, t_o = o.run(tab, 0, nsteps, True)
    assert rc == 0
    e_g, b_g, t_g = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    assert fc.rel_err(e_g, e_o) < 1e-9
    nh = case.params.n_harm
    bg = np.abs(b_g[:, 0:2 * nh:2] + 1j * b_g[:, 1:2 * nh:2]) / b_g[:, 2 * nh:2 * nh + 1]
    assert fc.rel_err(bg, b_o[:, :nh]) < 1e-8
    assert fc.rel_err(t_g, t_o) < 1e-9


# ---------------------------------------------------------------- 1D path

@pytest.mark.parametrize("harm,L", [((1,), 1), ((1, 3), 2), ((1, 2, 3), 3)])
def test_1d_push(harm, L):
    """push_theta_gamma_1D (particlePush.cpp:388-524) in FP64: theta
    increments to 1e-11 of their size, gamma to a few ulps."""
    c = fc.make_case_1d(harmonics=harm, sc_l=L)
    o = c.oracle_state()
    ctx = fc.gpu_context(c)
    lat = c.lattice()
    assert o.push_theta_gamma(lat) == 0
    ctx.push_theta_gamma(fc.to_lib_lattice(lat))
    p = ctx.download_particles()
    th0, g0 = c.state["theta"], c.state["gamma"]
    assert fc.rel_err(p["theta"] - th0, o.theta - th0) < 1e-11
    # gamma ~ 8000 carries 9e-13 per ulp: a few ulps of the increments' size
    assert np.max(np.abs((p["gamma"] - g0) - (o.gamma - g0))) < 4e-12
    ctx.close()


@pytest.mark.parametrize("harm,L,ns", [((1,), 1, 1), ((1, 3), 2, 2), ((1, 2, 3), 3, 1)])
def test_1d_solve_field(harm, L, ns):
    """solve_Sc_1D + solve_Field_U_1D (solveField.cpp:409-549): sources,
    fields and the space-charge modes."""
    c = fc.make_case_1d(harmonics=harm, sc_l=L, n_species=ns)
    o = c.oracle_state()
    ctx = fc.gpu_context(c)
    lat = c.lattice()
    assert o.solve_field(lat) == 0
    ctx.solve_field(fc.to_lib_lattice(lat))
    # the sources are sums over a jittered quiet start that cancel to ~1e-3 of
    # their terms, so a different summation order shows at ~1e-12 relative
    dU = o.Ux - c.state["Ux"]
    assert fc.rel_err(ctx.download_field(fl.UX) - c.state["Ux"], dU) < 1e-10
    assert fc.rel_err(ctx.download_field(fl.UY) - c.state["Uy"], o.Uy - c.state["Uy"]) < 1e-10
    assert fc.rel_err(ctx.download_field(fl.EZ), o.Ez) < 1e-10
    ctx.close()


def test_1d_run_matches_oracle_loop():
    """The 1D loop (main.cpp:92-156 with no transverse push) for 12 steps with
    slippage, the wake loss and the space-charge modes."""
    c = fc.make_case_1d(harmonics=(1, 3), sc_l=2, n_species=2)
    o = c.oracle_state()
    ctx = fc.gpu_context(c)
    nsteps = 12
    tab = [c.lattice(current_flag=0)] + [c.lattice() for _ in range(2 * nsteps)]
    rc, e_o, b_o, _ = o.run(tab, 0, nsteps, True)
    assert rc == 0
    e_g, b_g, _ = ctx.run([fc.to_lib_lattice(l) for l in tab], 0, nsteps)
    assert fc.rel_err(e_g, e_o) < 1e-10
    nh = 2
    bg = np.abs(b_g[:, 0:2 * nh:2] + 1j * b_g[:, 1:2 * nh:2]) / b_g[:, 2 * nh:2 * nh + 1]
    assert np.max(np.abs(bg - b_o[:, :nh])) < 1e-10
    ctx.close()
