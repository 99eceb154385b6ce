"""Pins the oracle's 3D field solve, bunching deposit, radial solve, push and
drift with checks the mathematics fixes. CPU only.

The reference TUs that hold these routines (solveField.cpp, particlePush.cpp,
boundary.cpp) include a GSL header that this image lacks, so no reference
build produces golden vectors for them. Each check below is chosen so that a
plausible transcription mistake (a dropped term, a wrong sign, index or
factor, a transposed operand) fails it:

* ADI (solveField.cpp:80-185): dense complex solves of the row equations the
  reference writes, with the special last row (solveField.cpp:146-147,
  174-175) and the untouched boundary ring; free-space Gaussian-beam
  diffraction (beam width and Gouy phase) converging at second order in the
  cell size; conservation of sum |U|^2 with the absorbing layer off.
* Absorbing layer (boundary.cpp:86-105): the profile's closed form at
  sigma0 != 0; a tilted beam entering the layer loses energy step by step.
* theta/gamma push (particlePush.cpp:198-386): the zero-field closed form
  with the radius- and momentum-dependent K2 and the wake loss; one RK4 step
  in a uniform field against an adaptive integrator of the same ODE, with
  local error O(dz^5); the Madey low-gain curve of a frozen field.
* Deposit (solveField.cpp:187-313): single-particle corner writes with the
  Bessel coupling from scipy.special.jv, truncation toward zero and the flat
  index spill; quiet-start cancellation for harmonics that the number of
  phases does not divide.
* Deposit against push: one step's field-energy gain equals the beam's energy
  loss, with the field energy normalised as updateTotalEnergy.cpp:41-45 does,
  converging to 1 linearly in dz, for planar and elliptical undulators, both
  K0_alpha signs and harmonics 1, 2, 3.
* Radial Ez solve (solveField.cpp:315-398): single-particle Green's columns
  against a dense solve of the tridiagonal rows, including the back
  substitution that reads slot 0 (the forward-eliminated d'_j / b'_j, with b'_j
  from leading principal minors).
* Transverse push (particlePush.cpp:32-117) with ue != 0: the step map is the
  fourth-order Taylor polynomial of exp(h M) for the linear system M written
  from the reference's equations; symplectic to O(h^5) for K0_alpha = 1.
* Drift (particlePush.cpp:119-195): zero-field and uniform-field closed forms
  with gamR, the wake loss taken from theta, particles outside the grid kept.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate, linalg, special

from oracle import pyoracle as po

# constants.h:1-10
MU0 = 1.2566370614e-6
C_LIGHT = 299792458.0
E_MASS = 9.10938291e-31
E_CHARGE = 1.60217662e-19
MC2 = 0.510998918
Z0 = 376.730313668


def _coupling(ue, K0_alpha):
    """Phi, etaX, etaY of particlePush.cpp:222-238."""
    Phi = float(np.angle((1 - ue * ue + 2j * ue) / (1 + ue * ue)))
    if K0_alpha == 1:
        return Phi, math.cos(Phi / 2), 1j * math.sin(Phi / 2)
    return Phi, 1j * math.sin(Phi / 2), math.cos(Phi / 2)


def make_state(parts=None, nx=9, ny=9, dx=10e-6, harmony=(1,), K0=1.87, lambdaU=0.026, gamma0=8000.0,
               ue=0.0, K0_alpha=1, und_type=1, weight=1e3, nr=12, dr=None, sc_l=1, sc_f=1, bn=2001,
               abc_n=3, abc_sig=0.0, wakeE=None, dz=None, slice_n=1, ks=None, center=True):
    """An oracle state on an (nx, ny) grid centred on the axis, with the
    photon wavenumber resonant at gamma0 unless ks is given."""
    ku = 2 * math.pi / lambdaU
    aw2 = (1 + ue * ue) * K0 * K0 / 2
    if ks is None:
        ks = 2 * math.pi / (lambdaU * (1 + aw2) / (2 * gamma0 ** 2))
    lambda0 = 2 * math.pi / ks
    p = po.Params()
    p.nx, p.ny, p.slice_n, p.n_harm = nx, ny, slice_n, len(harmony)
    for i, h in enumerate(harmony):
        p.harmony[i] = h
    maxh = max(harmony)
    p.num_slice, p.n_species, p.nr, p.sc_l, p.sc_f = 1, 1, nr, sc_l, sc_f
    p.bn, p.b_orders, p.wake_on, p.dimension = bn, maxh // 2 + 2, int(wakeE is not None), 3
    p.dx, p.dy, p.dz = dx, dx, (lambdaU if dz is None else dz)
    p.min_x = -(nx - 1) / 2 * dx if center else 0.0
    p.min_y = -(ny - 1) / 2 * dx if center else 0.0
    p.ks, p.lambda0, p.gamR = ks, lambda0, math.sqrt(0.5 * ks / ku * (1 + aw2))
    p.dr = dx if dr is None else dr
    p.dBessel = maxh * 2.0 * math.pi / bn
    al, alP, be, beP = po.abc_setup(nx, ny, abc_n, abc_sig, dx, dx, p.dz, ks)
    if parts is None:
        parts = {k: np.zeros(0) for k in ("x", "y", "px", "py", "theta", "gamma")}
    n = len(parts["x"])
    seg_off = np.array([0] + [n * (k + 1) // slice_n for k in range(slice_n)], np.int64)
    S, nh = slice_n + 2, len(harmony)
    st = po.OracleState(
        params=p, alpha=al, alphaP=alP, beta=be, betaP=beP, bessel=po.bessel_table(bn, maxh),
        seg_off=seg_off, seg_weight=np.full(slice_n, float(weight)),
        **{k: np.array(parts[k], np.float64) for k in ("x", "y", "px", "py", "theta", "gamma")},
        Ux=np.zeros((nh, S, ny, nx), complex), Uy=np.zeros((nh, S, ny, nx), complex),
        ScUx=np.zeros((nh, S, ny, nx), complex), ScUy=np.zeros((nh, S, ny, nx), complex),
        Ez=np.zeros((sc_l, S * sc_f * nr), complex),
        wakeE=np.zeros(slice_n) if wakeE is None else np.array(wakeE, np.float64))
    lat = po.Lattice()
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU = K0, ue, float(K0_alpha), ku, 0.0, lambdaU
    lat.und_type, lat.current_flag, lat.drift_flag = und_type, 1, 0
    return st, lat


def grid(st):
    p = st.params
    xs = p.min_x + p.dx * np.arange(p.nx)
    ys = p.min_y + p.dy * np.arange(p.ny)
    return np.meshgrid(xs, ys)


def particles(n, **kw):
    out = {k: np.zeros(n) for k in ("x", "y", "px", "py", "theta", "gamma")}
    for k, v in kw.items():
        out[k] = np.broadcast_to(np.asarray(v, np.float64), (n,)).copy()
    return out


# ------------------------------------------------------------------ ADI

def _implicit_matrix(a, aP, H, n):
    """Rows i = 1..n-2 of the implicit sweep as solveField.cpp:138-152 solves
    them: a_i U_{i-1} + (H - a'_i - a_i) U_i + a'_i U_{i+1} = dd_i, with U_0 = 0
    and, in the last row, coefficient 1 on U_{n-3} and no U_{n-1} term
    (the normalised recurrence of lines 146-147)."""
    m = n - 2
    A = np.zeros((m, m), complex)
    for k in range(m):
        i = k + 1
        A[k, k] = H - aP[i] - a[i]
        if i < n - 2:
            if k > 0:
                A[k, k - 1] = a[i]
            A[k, k + 1] = aP[i]
        else:
            A[k, k - 1] = 1.0
    return A


def _adi_dense(U, Sc, al, alP, be, beP, H, cf):
    """One ADI step of solveField.cpp:110-182 by dense solves: x implicit row
    by row into Uc (boundary zero), then y implicit column by column into U;
    U's boundary ring keeps its values."""
    ny, nx = U.shape
    Uc = np.zeros_like(U)
    Ax = _implicit_matrix(al, alP, H, nx)
    for j in range(1, ny - 1):
        dd = (-beP[j] * U[j + 1, 1:nx - 1] + (H + beP[j] + be[j]) * U[j, 1:nx - 1] - be[j] * U[j - 1, 1:nx - 1]
              + cf * Sc[j, 1:nx - 1])
        Uc[j, 1:nx - 1] = np.linalg.solve(Ax, dd)
    out = U.copy()
    Ay = _implicit_matrix(be, beP, H, ny)
    for i in range(1, nx - 1):
        dd = (-alP[i] * Uc[1:ny - 1, i + 1] + (H + alP[i] + al[i]) * Uc[1:ny - 1, i] - al[i] * Uc[1:ny - 1, i - 1]
              + cf * Sc[1:ny - 1, i])
        out[1:ny - 1, i] = np.linalg.solve(Ay, dd)
    return out


@pytest.mark.parametrize("abc_sig,cf", [(0.0, 1), (2.0, 1), (2.0, 0)])
def test_adi_equals_dense_solves_of_the_row_equations(abc_sig, cf):
    """solveField.cpp:80-185 on a non-square grid (11 x 9), harmonics 1 and 3,
    two slices, a random field with a nonzero boundary ring and a random
    source: the oracle's sweeps equal the dense solves of the rows above."""
    rng = np.random.default_rng(11)
    st, _ = make_state(nx=11, ny=9, harmony=(1, 3), slice_n=2, abc_n=3, abc_sig=abc_sig, dx=3e-6,
                       ks=2 * math.pi / 1e-9, dz=0.05)
    shape = st.Ux.shape
    st.Ux[:] = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    st.ScUx[:] = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    U0, Sc = st.Ux.copy(), st.ScUx.copy()
    assert st.solve_u("x", cf) == 0
    for h, H in enumerate((1, 3)):
        for slot in (1, 2):
            want = _adi_dense(U0[h, slot], Sc[h, slot], st.alpha, st.alphaP, st.beta, st.betaP, float(H), cf)
            np.testing.assert_allclose(st.Ux[h, slot], want, rtol=0, atol=1e-12 * np.max(np.abs(want)))
        # the ghost slots are not solved
        np.testing.assert_array_equal(st.Ux[h, 0], U0[h, 0])
        np.testing.assert_array_equal(st.Ux[h, 3], U0[h, 3])


def _adi_run(ncell, steps, w0=12e-6, ks=2 * math.pi / 1e-9, H=1):
    """A Gaussian beam of waist w0 at its focus on a grid of w0/ncell cells
    reaching 6.7 w0 from the axis, propagated one Rayleigh length of the
    harmonic H in `steps` ADI steps with no source and no absorbing layer."""
    dx = w0 / ncell
    n = 2 * int(round(6.7 * ncell)) + 1
    k = H * ks
    zR = k * w0 * w0 / 2
    st, _ = make_state(nx=n, ny=n, dx=dx, harmony=(H,), dz=zR / steps, ks=ks, abc_sig=0.0)
    X, Y = grid(st)
    U0 = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    st.Ux[0, 1] = U0
    for _ in range(steps):
        assert st.solve_u("x", 0) == 0
    return X, Y, U0, st.Ux[0, 1].copy()


def test_adi_gaussian_beam_diffraction():
    """Free space (Sc = 0, sigma0 = 0): the ADI advances dU/dz = i/(2 H ks)
    lap(U), whose focused Gaussian solution is U = exp(-r^2/(w0^2 q))/q with
    q = 1 + i z/zR, zR = H ks w0^2/2: at z = zR the width is sqrt(2) w0 and the
    on-axis (Gouy) phase is -pi/4. The field error falls at second order in
    the cell size; the step size (Crank-Nicolson, O(dz^2)) barely matters.
    Harmonic 3 checks that the wavenumber is H ks."""
    errs = []
    for ncell in (6, 12):
        X, Y, U0, U = _adi_run(ncell, 40)
        q = 1 + 1j
        exact = np.exp(-(X ** 2 + Y ** 2) / (12e-6 ** 2 * q)) / q
        errs.append(np.linalg.norm(U - exact) / np.linalg.norm(exact))
        I = np.abs(U) ** 2
        w = math.sqrt(4 * np.sum(X ** 2 * I) / np.sum(I))
        c = U.shape[0] // 2
        assert abs(w / (12e-6 * math.sqrt(2)) - 1) < (1e-2 if ncell == 6 else 3e-3)
        assert abs(np.angle(U[c, c] / U0[c, c]) + math.pi / 4) < 2e-4
    assert 3.0 < errs[0] / errs[1] < 5.0, errs          # second order in dx
    assert errs[1] < 3e-3
    X, Y, U0, U = _adi_run(6, 40, H=3)
    exact = np.exp(-(X ** 2 + Y ** 2) / (12e-6 ** 2 * (1 + 1j))) / (1 + 1j)
    assert np.linalg.norm(U - exact) / np.linalg.norm(exact) < 1e-2


def test_adi_conserves_energy_without_the_layer():
    """With sigma0 = 0 every half step is a Cayley transform of an
    anti-Hermitian operator, so sum |U|^2 is conserved (for a field that stays
    clear of the grid's last rows, where the special row of
    solveField.cpp:146-147 is not unitary)."""
    X, Y, U0, U = _adi_run(6, 40)
    assert abs(np.sum(np.abs(U) ** 2) / np.sum(np.abs(U0) ** 2) - 1) < 1e-12


def test_absorbing_layer_profile_closed_form():
    """boundary.cpp:86-105 at sigma0 = 1.5, L = 4: sigma_i = sigma0 (L-i)^2/L^2
    on the low edge and sigma0 (i-n+L)^2/L^2 on the high edge, s = 1 + i sigma,
    alpha_i = -i dz / (2 ks dx^2 s_i (s_i + s_{i-1})), alpha'_i with s_{i+1}."""
    nx, ny, L, s0, dx, dy, dz, ks = 12, 10, 4, 1.5, 2e-6, 3e-6, 0.03, 4e9
    al, alP, be, beP = po.abc_setup(nx, ny, L, s0, dx, dy, dz, ks)

    def prof(n):
        sig = np.zeros(n)
        for i in range(n):
            if i < L:
                sig[i] = s0 * (L - i) ** 2 / L ** 2
            if i >= n - L:
                sig[i] = s0 * (i - n + L) ** 2 / L ** 2
        return 1 + 1j * sig

    for n, d, a, aP in ((nx, dx, al, alP), (ny, dy, be, beP)):
        s = prof(n)
        i = np.arange(1, n - 1)
        np.testing.assert_allclose(a[1:-1], -1j * dz / (2 * ks * d * d * s[i] * (s[i] + s[i - 1])), rtol=1e-14)
        np.testing.assert_allclose(aP[1:-1], -1j * dz / (2 * ks * d * d * s[i] * (s[i] + s[i + 1])), rtol=1e-14)
        assert a[0] == 0 and aP[-1] == 0


def test_absorbing_layer_absorbs_a_beam_and_never_amplifies():
    """A Gaussian beam tilted toward +x reaches the layer (the decks' ABC_N =
    10, ABC_coef = 1) and loses energy: sum |U|^2 never rises from one step to
    the next and falls below 30% after twice the time to reach the edge. Without
    the layer the same beam reaches the special last row and grows, so the
    profile is what keeps the reference's scheme stable."""
    w0, ks = 12e-6, 2 * math.pi / 1e-9
    zR = ks * w0 * w0 / 2
    out = {}
    for sig in (1.0, 0.0):
        st, _ = make_state(nx=81, ny=81, dx=w0 / 6, dz=zR / 40, ks=ks, abc_n=10, abc_sig=sig)
        X, Y = grid(st)
        kx = 2.0 / w0
        st.Ux[0, 1] = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2) * np.exp(1j * kx * X)
        e = [np.sum(np.abs(st.Ux[0, 1]) ** 2)]
        steps = int(2 * (40 * w0 / 6) / (kx / ks) / (zR / 40))
        for _ in range(steps):
            st.solve_u("x", 0)
            e.append(np.sum(np.abs(st.Ux[0, 1]) ** 2))
        out[sig] = np.array(e) / e[0]
    assert np.all(np.diff(out[1.0]) <= 1e-12)
    assert out[1.0][-1] < 0.3
    assert out[0.0][-1] > 10.0


# ------------------------------------------------------------- theta/gamma push

def _k2(p, ue, K0, ku, x, y, px, py):
    return 1.0 + px * px + py * py + (1.0 + ue * ue) * K0 * K0 * 0.5 * (1.0 + ku * ku * 0.5 * (x * x + y * y))


@pytest.mark.parametrize("ue,K0_alpha", [(0.0, 1), (0.5, 1), (0.5, -1)])
def test_push_zero_field_closed_form(ue, K0_alpha):
    """particlePush.cpp:260-378 with U = 0 and Ez = 0: every RK4 stage has the
    same rates, so dtheta = dz (ku - ks K2 / (2 gamma^2)) with K2 = 1 + p^2 +
    (1+ue^2) K0^2/2 (1 + ku^2 r^2 / 2), and dgamma = -dz wakeE / (mc2 1e6).
    A particle with -1 < (x-minX)/dx < 0 truncates to cell 0 and is pushed
    (particlePush.cpp:281-285); one at or past the last cell is not."""
    rng = np.random.default_rng(3)
    n = 12
    pr = particles(n, x=rng.normal(0, 30e-6, n), y=rng.normal(0, 30e-6, n), px=rng.normal(0, 0.1, n),
                   py=rng.normal(0, 0.1, n), theta=rng.uniform(0, 2 * math.pi, n),
                   gamma=8000 * (1 + 1e-3 * rng.normal(size=n)))
    st, lat = make_state(pr, nx=21, ny=21, dx=10e-6, ue=ue, K0_alpha=K0_alpha, wakeE=[3.7e4])
    p = st.params
    st.x[0] = p.min_x - 0.4 * p.dx                      # truncates to cell 0: pushed
    st.x[1] = p.min_x + (p.nx - 1) * p.dx + 1e-9        # idxI = nx-1: not pushed
    th0, g0 = st.theta.copy(), st.gamma.copy()
    assert st.push_theta_gamma(lat) == 0
    K2 = _k2(p, ue, lat.K0, lat.ku, pr["x"], pr["y"], pr["px"], pr["py"])
    K2[0] = _k2(p, ue, lat.K0, lat.ku, st.x[0], pr["y"][0], pr["px"][0], pr["py"][0])
    dth = p.dz * (lat.ku - p.ks * K2 / (2 * g0 * g0))
    dg = -p.dz * 3.7e4 / (MC2 * 1e6)
    inside = np.ones(n, bool)
    inside[1] = False
    np.testing.assert_allclose(st.theta[inside] - th0[inside], dth[inside], rtol=0, atol=1e-12 * p.dz * lat.ku)
    np.testing.assert_allclose(st.gamma[inside] - g0[inside], dg, rtol=0, atol=4 * np.spacing(8100.0))
    assert st.theta[1] == th0[1] and st.gamma[1] == g0[1]


def _bessel_interp(st, H, xi):
    """J_order(H xi) by the table's linear interpolation (particlePush.cpp:338-339)."""
    dB = st.params.dBessel
    idx = int(H * xi / dB)
    w1 = H * xi / dB - idx
    return lambda order: st.bessel[idx, order] * (1 - w1) + st.bessel[idx + 1, order] * w1


def _couplings(st, lat, H, xi, xi2, B, Phi):
    """fx, fy of particlePush.cpp:340-360: the odd harmonics' Bessel couplings
    J1 -+ K0_alpha J2 (sign by (H-1)/2), the even ones' e^{+-iB} sums of J1, J2, J3."""
    J = _bessel_interp(st, H, xi)
    a = lat.K0_alpha
    if H % 2:
        sign = 1.0 if ((H - 1) // 2) % 2 == 0 else -1.0
        J1, J2 = J((H - 1) // 2), J((H + 1) // 2)
        return sign * (a * J1 - J2), sign * (a * J1 + J2)
    sign = 1.0 if (H // 2) % 2 == 0 else -1.0
    J1, J2, J3 = J((H - 2) // 2), J(H // 2), J((H + 2) // 2)
    eP = np.exp(1j * (B - Phi / 2))
    eM = np.conj(eP)
    fx = sign * xi2 * H * 0.5 * (eP * (a * J1 - J2) + eM * (a * J2 - J3))
    fy = sign * xi2 * H * 0.5 * (eP * (a * J1 + J2) + eM * (a * J2 + J3))
    return fx, fy


def _push_rhs(st, lat, n, g0, Ux, Uy, Em):
    """The (theta, gamma) ODE of particlePush.cpp:323-367 for particle n, whose
    gamma at the start of the step is g0, with the gathered fields Ux[h], Uy[h]
    and radial field Em[l]."""
    p = st.params
    ue, K0, ku, ks = lat.ue, lat.K0, lat.ku, p.ks
    Phi, etaX, etaY = _coupling(ue, lat.K0_alpha)
    px, py = st.px[n], st.py[n]
    a = lat.K0_alpha
    K2 = _k2(p, ue, K0, ku, st.x[n], st.y[n], px, py)
    xi = ks / ku * K0 * K0 / (8 * g0 * g0) * (1 - ue * ue)
    xi2 = math.sqrt(2) * K0 / (1 + K0 * K0 * 0.5 * (1 + ue * ue)) * math.sqrt(
        (1 + a) * (px * px + ue * ue * py * py) + (1 - a) * (px * px * ue * ue + py * py))
    B = math.atan(2 * a * ue * (px - py) / ((1 + a) * (px + ue * ue * py) + (1 - a) * (px * ue * ue + py)))
    harm = [p.harmony[h] for h in range(p.n_harm)]
    f = [_couplings(st, lat, H, xi, xi2, B, Phi) for H in harm]
    sumU2 = sum((abs(Ux[h]) ** 2 + abs(Uy[h]) ** 2) / (2 * H * H) for h, H in enumerate(harm))
    e_mc2 = E_CHARGE / E_MASS / C_LIGHT / C_LIGHT
    kk = K0 * math.sqrt(1 + ue * ue)

    def rhs(z, s):
        th, g = s
        sumEz = sum(2 * (Em[l] * np.exp(1j * (l + 1) * th)).real for l in range(len(Em)))
        sumTh = sumG = 0.0
        for h, H in enumerate(harm):
            t = np.exp(1j * H * (th + Phi / 2)) * (etaX * f[h][0] * Ux[h] - etaY * f[h][1] * Uy[h])
            sumTh += (1j * t).real / H
            sumG += t.real / (2 * g)
        return [ku - ks / (2 * g * g) * (K2 + sumU2 + kk * sumTh), ks * kk * sumG + e_mc2 * sumEz]

    return rhs, xi, f


AMP = 5e-4   # field of a few synchrotron radians per metre: the RK4 error is well above rounding


@pytest.mark.parametrize("ue,K0_alpha,harm", [(0.0, 1, (1, 3)), (0.6, 1, (1, 2, 3)), (0.4, -1, (1, 2))])
def test_push_rk4_step_against_adaptive_integration(ue, K0_alpha, harm):
    """One push in a uniform field (every cell of slot 1 holds the same
    complex value per harmonic and polarisation, so the bilinear gather is
    exact) and a uniform radial field, against scipy's DOP853 at rtol 1e-13 on
    the ODE of particlePush.cpp:323-367: the error is the RK4 local error,
    falling 2^5 = 32-fold when dz halves. The particle at r = 0 meets the
    unweighted two-point Em of particlePush.cpp:297-304. The Bessel factors
    from the table are checked against scipy.special.jv to the table's
    interpolation error."""
    rng = np.random.default_rng(5)
    n = 4
    pr = particles(n, x=[0.0, 20e-6, -35e-6, 12e-6], y=[0.0, -15e-6, 8e-6, 30e-6],
                   px=rng.normal(0, 0.05, n), py=rng.normal(0, 0.05, n), theta=rng.uniform(0, 6.28, n),
                   gamma=8000 * (1 + 2e-4 * rng.normal(size=n)))
    errs = {}
    for dzf in (1.0, 0.5):
        st, lat = make_state(pr, nx=21, ny=21, harmony=harm, ue=ue, K0_alpha=K0_alpha, nr=12, dr=10e-6,
                             dz=0.026 * dzf)
        Ux = [AMP * np.exp(0.7j * (h + 1)) for h in range(len(harm))]
        Uy = [0.75 * AMP * np.exp(-0.4j * (h + 2)) for h in range(len(harm))]
        for h in range(len(harm)):
            st.Ux[h, 1] = Ux[h]
            st.Uy[h, 1] = Uy[h]
        E0 = 2e11 * AMP * np.exp(0.3j)
        st.Ez[0, 1 * st.params.nr:2 * st.params.nr] = E0
        th0, g0 = st.theta.copy(), st.gamma.copy()
        assert st.push_theta_gamma(lat) == 0
        e = []
        for k in range(n):
            Em = [2 * E0] if k == 0 else [E0]
            rhs, xi, f = _push_rhs(st, lat, k, g0[k], Ux, Uy, Em)
            st_k = [th0[k], g0[k]]
            sol = integrate.solve_ivp(rhs, (0, st.params.dz), st_k, method="DOP853", rtol=1e-13, atol=1e-14)
            want = sol.y[:, -1]
            e.append(max(abs(st.theta[k] - want[0]), abs(st.gamma[k] - want[1]) / 10))
            # the field matters: theta and gamma move well off the first stage's rates
            assert abs((st.theta[k] - th0[k]) - st.params.dz * rhs(0, st_k)[0]) > 1e-5
            if dzf == 1.0 and ue == 0.0:
                for h, H in enumerate(harm):
                    s = 1.0 if ((H - 1) // 2) % 2 == 0 else -1.0
                    fx = s * (K0_alpha * special.jv((H - 1) // 2, H * xi) - special.jv((H + 1) // 2, H * xi))
                    assert abs(f[h][0] - fx) < 2e-5
        errs[dzf] = np.array(e)
    ratio = errs[1.0] / errs[0.5]
    assert np.all(errs[1.0] < 2e-4)
    assert np.all((ratio > 28) & (ratio < 37)), ratio


def test_push_madey_low_gain_curve():
    """Twenty periods of the push in a frozen, weak, uniform field on a quiet
    start of 32 phases per detuning. Second-order perturbation theory of
    dgamma/dz = a cos(theta), dtheta/dz = u/L + c dgamma + (a/gamma) sin(theta)
    (c = ks K2/gamma^3 from particlePush.cpp:366, a = ks K0 f |U| / (2 gamma)
    from :362-367) gives the Madey curve
      <dgamma> = -(a^2 c L^3/(2u)) (2(1 - cos u)/u^2 - sin u/u)
                 - a^2 L^2 (1 - cos u)/(gamma u^2):
    energy loss above resonance (u > 0), gain below, antisymmetric shape. The
    oracle follows it to 1% of the peak (the rest is of relative order
    1/(2 ku L) from the coupling's 1/gamma, which the formula leaves out)."""
    K0, lambdaU, g0, Nu, nph = 1.87, 0.026, 8000.0, 20, 32
    ku = 2 * math.pi / lambdaU
    aw2 = K0 * K0 / 2
    ks = 2 * math.pi / (lambdaU * (1 + aw2) / (2 * g0 ** 2))
    K2 = 1 + aw2
    L = Nu * lambdaU
    us = np.array([-6.0, -4.0, -2.6, -1.0, 1.0, 2.6, 4.0, 6.0])
    gam = np.sqrt(ks * K2 / (2 * (ku - us / L)))
    pr = particles(len(us) * nph, theta=np.tile(np.arange(nph) * 2 * math.pi / nph + 0.3, len(us)),
                   gamma=np.repeat(gam, nph))
    st, lat = make_state(pr, nx=9, ny=9, dx=10e-6)
    c0 = ks * K2 / g0 ** 3
    xi0 = ks / ku * K0 * K0 / (8 * g0 * g0)
    Omega = 0.1 / L
    U0 = (Omega ** 2 / c0) * 2 * g0 / (ks * K0 * (special.jv(0, xi0) - special.jv(1, xi0)))
    st.Ux[0] = U0
    g_start = st.gamma.copy()
    for _ in range(Nu):
        assert st.push_theta_gamma(lat) == 0
    dg = (st.gamma - g_start).reshape(len(us), nph).mean(1)
    xi = ks / ku * K0 * K0 / (8 * gam * gam)
    a = ks * K0 * (special.jv(0, xi) - special.jv(1, xi)) * U0 / (2 * gam)
    c = ks * K2 / gam ** 3
    want = (-(a * a * c * L ** 3 / (2 * us)) * (2 * (1 - np.cos(us)) / us ** 2 - np.sin(us) / us)
            - a * a * L ** 2 * (1 - np.cos(us)) / (gam * us * us))
    assert np.all(np.sign(dg) == -np.sign(us))
    np.testing.assert_allclose(dg, want, rtol=0, atol=0.01 * np.max(np.abs(want)))


# ------------------------------------------------------------------ deposit

def test_deposit_single_particle_corners():
    """solveField.cpp:246-308 for single particles: the four corners get
    wx[ii] wy[jj] of -eta f H w coef_U e^{-iH(theta+Phi/2)}/gamma, with
    coef_U = dz e^2 mu0 K0 sqrt(1+ue^2) / (4 ks m lambda0 numSlice dx dy) and
    the odd-harmonic f = sign (J_{(H-1)/2} - K0_alpha J_{(H+1)/2})(H xi) at the
    global xi = K0^2 (1-ue^2) / (4 (1 + (1+ue^2) K0^2/2)) (scipy's jv, to the
    table's interpolation error). A particle with -1 < (x-minX)/dx < 0
    truncates to column 0 with a negative weight; one in the last column
    spills its ii = 1 corners into the next row's first cell (the flat index,
    solveField.cpp:268-273, 302-307); one past the grid deposits nothing."""
    nx = ny = 9
    ue, K0 = 0.5, 1.87
    for case in ("interior", "negative", "spill", "outside"):
        st, lat = make_state(particles(1, gamma=7990.0, theta=1.234), nx=nx, ny=ny, harmony=(1, 3), ue=ue,
                             K0=K0, weight=2.5e3)
        p = st.params
        fx_i = {"interior": 3.3, "negative": -0.35, "spill": nx - 1 + 0.4, "outside": nx + 0.2}[case]
        st.x[0] = p.min_x + fx_i * p.dx
        st.y[0] = p.min_y + 4.6 * p.dy
        assert st.solve_sc(lat) == 0
        if case == "outside":
            assert not np.any(st.ScUx) and not np.any(st.ScUy)
            continue
        i0 = int(fx_i) if fx_i >= 0 else 0
        wx1 = fx_i - i0
        wy1 = 0.6
        Phi, etaX, etaY = _coupling(ue, 1)
        coef_U = p.dz * E_CHARGE ** 2 * MU0 * K0 * math.sqrt(1 + ue * ue) / (
            4 * p.ks * E_MASS * p.lambda0 * p.num_slice * p.dx * p.dy)
        xi = 0.25 * K0 * K0 * (1 - ue * ue) / (1 + (1 + ue * ue) * K0 * K0 * 0.5)
        for h, H in enumerate((1, 3)):
            s = 1.0 if ((H - 1) // 2) % 2 == 0 else -1.0
            Ja, Jb = special.jv((H - 1) // 2, H * xi), special.jv((H + 1) // 2, H * xi)
            fx, fy = s * (Ja - Jb), s * (Ja + Jb)
            base = H * 2.5e3 * coef_U * np.exp(-1j * H * (1.234 + Phi / 2)) / 7990.0
            want_x = np.zeros(ny * nx + nx + 1, complex)
            want_y = np.zeros_like(want_x)
            for ii, wx in ((0, 1 - wx1), (1, wx1)):
                for jj, wy in ((0, 1 - wy1), (1, wy1)):
                    c = (4 + jj) * nx + i0 + ii
                    want_x[c] -= etaX * wx * wy * fx * base
                    want_y[c] -= etaY * wx * wy * fy * base
            got_x = st.ScUx[h, 1].ravel()
            got_y = st.ScUy[h, 1].ravel()
            scale = abs(base)
            np.testing.assert_allclose(got_x, want_x[:ny * nx], rtol=0, atol=1e-4 * scale)
            np.testing.assert_allclose(got_y, want_y[:ny * nx], rtol=0, atol=1e-4 * scale)
            if case == "spill" and H == 1:
                assert abs(st.ScUx[h, 1, 5, 0]) > 0.05 * scale  # the next row's first cell
            if case == "negative":
                assert wx1 < 0
            # the corner weights sum to one
            assert abs(got_x.sum() + etaX * fx * base) < 1e-4 * scale


def test_deposit_quiet_start_cancels_harmonics_the_phases_do_not_divide():
    """Four particles at one position with phases theta0 + 2 pi k/4 deposit
    sum_k e^{-iH theta_k} = 4 e^{-iH theta0} for H = 4 and exactly nothing for
    H = 1, 2, 3 (solveField.cpp:301): the phase factor is H theta, not theta."""
    th = 0.77 + np.arange(4) * math.pi / 2
    pr = particles(4, x=13e-6, y=-7e-6, px=0.04, py=-0.03, theta=th, gamma=8000.0)
    st, lat = make_state(pr, nx=9, ny=9, harmony=(1, 2, 3, 4), ue=0.3)
    assert st.solve_sc(lat) == 0
    one, lat1 = make_state(particles(1, x=13e-6, y=-7e-6, px=0.04, py=-0.03, theta=0.77, gamma=8000.0),
                           nx=9, ny=9, harmony=(1, 2, 3, 4), ue=0.3)
    assert one.solve_sc(lat1) == 0
    for h in range(3):
        scale = np.max(np.abs(one.ScUx[h])) + np.max(np.abs(one.ScUy[h]))
        assert np.max(np.abs(st.ScUx[h])) < 1e-13 * scale
        assert np.max(np.abs(st.ScUy[h])) < 1e-13 * scale
    np.testing.assert_allclose(st.ScUx[3], 4 * one.ScUx[3], rtol=1e-12, atol=1e-15 * np.max(np.abs(one.ScUx[3])))
    np.testing.assert_allclose(st.ScUy[3], 4 * one.ScUy[3], rtol=1e-12, atol=1e-15 * np.max(np.abs(one.ScUy[3])))


def _energy_balance(ue, K0_alpha, dz, harmony=(1, 2, 3), seed=3, n=4000):
    """(field energy gain from the source, beam energy change) of one step,
    in joules: the ADI is linear, so U_new(U0, Sc) = P U0 + Q Sc and
    |U_new|^2 - |P U0|^2 - |Q Sc|^2 = 2 Re <P U0, Q Sc> is the part of the gain
    that the push's field does work against."""
    rng = np.random.default_rng(seed)
    dx, nx = 10e-6, 41
    pr = particles(n, x=rng.normal(0, 40e-6, n), y=rng.normal(0, 40e-6, n), px=rng.normal(0, 0.08, n),
                   py=rng.normal(0, 0.08, n), theta=rng.uniform(0, 2 * math.pi, n))
    st, lat = make_state(pr, nx=nx, ny=nx, dx=dx, harmony=harmony, ue=ue, K0_alpha=K0_alpha, weight=1e3, nr=30,
                         dz=dz)
    st.gamma[:] = st.params.gamR
    X, Y = grid(st)
    env = np.exp(-(X ** 2 + Y ** 2) / (150e-6) ** 2)
    for h in range(len(harmony)):
        st.Ux[h, 1] = 3e-6 * env * np.exp(1j * (0.4 + h))
        st.Uy[h, 1] = 2e-6 * env * np.exp(1j * (1.3 - h))
    U0x, U0y = st.Ux.copy(), st.Uy.copy()
    g0 = st.gamma.copy()
    p = st.params
    # updateTotalEnergy.cpp:18, 41-45: power per sum |U|^2 times the slice duration
    coef2 = (E_MASS * C_LIGHT ** 2 * p.ks / E_CHARGE) ** 2 / (2 * Z0) * p.dx * p.dy
    coef2 *= p.num_slice * p.lambda0 / C_LIGHT

    def energy():
        return np.sum(np.abs(st.Ux[:, 1]) ** 2 + np.abs(st.Uy[:, 1]) ** 2)

    assert st.solve_sc(lat) == 0
    st.solve_u("x", 1); st.solve_u("y", 1)
    w_both = energy()
    st.Ux[:], st.Uy[:] = U0x, U0y
    st.solve_u("x", 0); st.solve_u("y", 0)
    w_seed = energy()
    st.Ux[:], st.Uy[:] = 0, 0
    st.solve_u("x", 1); st.solve_u("y", 1)
    w_src = energy()
    st.Ux[:], st.Uy[:] = U0x, U0y
    st.Ez[:] = 0                                  # the radial self-field is not in U
    assert st.push_theta_gamma(lat) == 0
    dE = np.sum(st.gamma - g0) * st.seg_weight[0] * MC2 * 1e6 * E_CHARGE
    return (w_both - w_seed - w_src) * coef2, dE


@pytest.mark.parametrize("ue,K0_alpha", [(0.0, 1), (0.6, 1), (0.6, -1), (1.0, 1)])
def test_deposit_and_push_conserve_energy(ue, K0_alpha):
    """The work the field does on the beam in the push (particlePush.cpp:362-
    367) equals the field energy the deposit adds through the ADI
    (solveField.cpp:301-307, 134, 163), with the field energy of
    updateTotalEnergy.cpp:41-45 (E = m c^2 ks U / e, P = |E|^2 dx dy/(2 Z0),
    times the slice duration): harmonics 1, 2, 3, momenta and ue != 0 exercise
    every coupling branch, including the swapped expP/expM and the deposit's
    J1 - a J2 (which are the adjoints of the push's). The mismatch is O(dz)
    (phase slip within a step) and falls tenfold with dz."""
    errs = []
    for dz in (0.0026, 0.00026):
        gain, dE = _energy_balance(ue, K0_alpha, dz)
        assert abs(dE) > 0
        errs.append(abs(gain / (-dE) - 1))
    assert errs[1] < 3e-3, errs
    assert errs[0] > 3 * errs[1], errs


# ---------------------------------------------------------------- radial Ez

def test_radial_solve_against_dense_solve():
    """solveField.cpp:326-396: one particle at r = k dr (so wr = (1, 0)) puts
    i coef_Ez (l+1) e^{-i((l+1) theta + f phi)} w / k into bin k; the rows
    are c_j x_{j-1} + b_j x_j + a_j x_{j+1} = d_j with a_j = j + 1/2,
    b_j = -2j - (l+1)^2 A j - f^2 ln((j+1/2)/(j-1/2)), c_j = j - 1/2, row 0 =
    (b_0 = -(1 + (l+1)^2 A / 8), a_0 = 1) for f = 0 and x_0 = 0 otherwise, and
    x_{nr-1} = 0. The back substitution reads slot 0, which holds zeros, so each
    slot gets the forward-eliminated d'_j / b'_j = x_j + a_j x_{j+1} / b'_j,
    where x solves the dense system and b'_j is the ratio of leading principal
    minors; its top entry is never written."""
    nr, dr, L, F = 14, 8e-6, 2, 2
    for k in (1, 3, 7):
        phi = 0.6
        pr = particles(1, x=k * dr * math.cos(phi), y=k * dr * math.sin(phi), theta=0.9, gamma=8000.0)
        st, lat = make_state(pr, nx=41, ny=41, dx=5e-6, nr=nr, dr=dr, sc_l=L, sc_f=F, ue=0.2, weight=3e3)
        st.Ez[:] = 7.0   # every slot's top entry stays as it was; slot 0's is cleared
        st.Ez[:, 0:F * nr] = 0.0
        assert st.solve_sc(lat) == 0
        p = st.params
        K0, ue = lat.K0, lat.ue
        coef_Ez = (E_CHARGE * C_LIGHT ** 2 * MU0 * p.ks * (1 + 0.5 * (1 + ue * ue) * K0 * K0)
                   / (2 * math.pi * p.lambda0 * p.num_slice * p.gamR ** 2))
        A = p.ks ** 2 * (1 + 0.5 * (1 + ue * ue) * K0 * K0) / p.gamR ** 2 * dr * dr
        for l in range(L):
            for f in range(F):
                M = np.zeros((nr, nr))
                d = np.zeros(nr, complex)
                if f == 0:
                    M[0, 0], M[0, 1] = -(1 + (l + 1) ** 2 * A / 8), 1.0
                else:
                    M[0, 0] = 1.0
                for j in range(1, nr - 1):
                    M[j, j - 1] = j - 0.5
                    M[j, j] = -2 * j - (l + 1) ** 2 * A * j - f * f * math.log((j + 0.5) / (j - 0.5))
                    M[j, j + 1] = j + 0.5
                M[nr - 1, nr - 1] = 1.0
                d[k] = 1j * coef_Ez * (l + 1) * np.exp(-1j * ((l + 1) * 0.9 + f * phi)) * 3e3 / k
                x = np.linalg.solve(M, d)
                minors = [1.0] + [np.linalg.det(M[:j + 1, :j + 1]) for j in range(nr)]
                bp = np.array([minors[j + 1] / minors[j] for j in range(nr)])
                want = x[:-1] + np.diag(M, 1) * x[1:] / bp[:-1]
                got = st.Ez[l, F * nr + f * nr: F * nr + (f + 1) * nr]
                np.testing.assert_allclose(got[:-1], want, rtol=0, atol=1e-11 * np.max(np.abs(want)))
                assert got[-1] == 7.0
                # the check resolves the quirk: d'_j / b'_j differs from x_j by far more than the tolerance
                assert np.max(np.abs(got[:-1] - x[:-1])) > 1e-7 * np.max(np.abs(x))
            assert st.Ez[l, nr - 1] == 0 and st.Ez[l, F * nr - 1] == 0   # slot 0's top entries cleared


# ---------------------------------------------------------------- transverse

def _transverse_matrix(lat, gam, ku, K0, ue, g):
    """The linear system d(x, y, px, py)/dz = M (x, y, px, py) of
    particlePush.cpp:46-48, 53-68, 88, 96-99."""
    a = lat.K0_alpha
    aa = K0 * K0 * 0.25 * ue * ku
    bb = K0 * K0 * 0.25 * (1 + ue * ue) * ku * ku
    cc = E_CHARGE * g / E_MASS / C_LIGHT
    dd = ku * K0 * K0 / (4 * gam * gam) * ue
    if lat.und_type == 1:
        xc, yc = (1.0, 0.0) if a == 1 else (0.0, 1.0)
    else:
        xc, yc = 1.0, 1.0
    return np.array([
        [-aa * a / gam ** 2, -aa / gam ** 2, 1 / gam, 0],
        [aa * a / gam ** 2, -aa / gam ** 2, 0, 1 / gam],
        [-(bb / gam * xc + cc), 0, a * dd, -a * dd],
        [0, -(bb / gam * yc - cc), a * dd, a * dd],
    ])


def _transverse_map(lat, gam, dz, nvec=4):
    """The oracle's one-step map on the unit vectors (the step is linear for a
    fixed gamma), as a 4x4 matrix."""
    pr = particles(nvec, gamma=gam)
    st, _ = make_state(pr, nx=5, ny=5, dz=dz)
    for k, name in enumerate(("x", "y", "px", "py")):
        getattr(st, name)[k] = 1.0
    assert st.transverse_push(lat) == 0
    return np.array([[getattr(st, name)[k] for k in range(4)] for name in ("x", "y", "px", "py")])


@pytest.mark.parametrize("und_type,ka,ue,g", [(1, 1, 0.7, 0.0), (1, 1, 0.4, 18.0), (1, -1, 0.6, -9.0),
                                             (2, 1, 1.0, 12.0)])
def test_coupled_transverse_step_is_the_taylor_polynomial(und_type, ka, ue, g):
    """particlePush.cpp:76-115 with ue != 0 couples x and y. For a fixed gamma
    the system is linear with constant M, and RK4 over h = dz/2 is exactly
    T4(hM) = sum_{k<=4} (hM)^k / k!, within O(h^5) of expm(hM): the gap to the
    exponential falls 32-fold when h halves. For K0_alpha = 1 the flow is
    Hamiltonian, so the map is symplectic to that order (M^T J + J M = 0)."""
    lat = po.Lattice()
    K0, lambdaU, gam = 1.87, 0.026, 300.0   # low gamma: the coupling terms are not negligible
    lat.K0, lat.ue, lat.K0_alpha, lat.ku, lat.g, lat.lambdaU = K0, ue, float(ka), 2 * math.pi / lambdaU, g, lambdaU
    lat.und_type, lat.current_flag, lat.drift_flag = und_type, 1, 0
    M = _transverse_matrix(lat, gam, lat.ku, K0, ue, g)
    J = np.block([[np.zeros((2, 2)), np.eye(2)], [-np.eye(2), np.zeros((2, 2))]])
    gaps = []
    for dz in (0.4, 0.2):
        h = dz * 0.5
        Phi = _transverse_map(lat, gam, dz)
        T4 = sum(np.linalg.matrix_power(h * M, k) / math.factorial(k) for k in range(5))
        np.testing.assert_allclose(Phi, T4, rtol=0, atol=1e-12 * np.max(np.abs(T4)))
        gaps.append(np.max(np.abs(Phi - linalg.expm(h * M))))
        if ka == 1:
            assert np.max(np.abs(M.T @ J + J @ M)) < 1e-12 * np.max(np.abs(M))
            assert np.max(np.abs(Phi.T @ J @ Phi - J)) < 10 * gaps[-1] + 1e-13
    assert 20 < gaps[0] / gaps[1] < 48, gaps


# --------------------------------------------------------------------- drift

def test_drift_closed_forms():
    """particlePush.cpp:140-187: theta += (ku - ks/(2 gamR^2)(1 + px^2 + py^2 +
    sum_h |U_h|^2/(2H^2))) dz - dz wakeE/(mc2 1e6) with the reference gamma gamR
    (not the particle's), gamma unchanged, the wake loss taken from theta; a
    uniform field per harmonic and polarisation makes the bilinear gather
    exact; a particle outside the grid is left alone."""
    rng = np.random.default_rng(9)
    n = 6
    pr = particles(n, x=rng.normal(0, 20e-6, n), y=rng.normal(0, 20e-6, n), px=rng.normal(0, 0.1, n),
                   py=rng.normal(0, 0.1, n), theta=rng.uniform(0, 6, n), gamma=8000 * (1 + 1e-3 * rng.normal(size=n)))
    st, lat = make_state(pr, nx=15, ny=15, harmony=(1, 3), wakeE=[2.2e4])
    st.x[2] = 1.0   # off the grid
    lat.ku = 0.0
    lat.drift_flag = 1
    u = {(0, "x"): 3e-5 + 1e-5j, (0, "y"): -2e-5j, (1, "x"): 4e-5, (1, "y"): 1e-5 + 1e-5j}
    for (h, pol), v in u.items():
        getattr(st, "U" + pol)[h, 1] = v
    th0, g0 = st.theta.copy(), st.gamma.copy()
    assert st.drift_theta_gamma(lat) == 0
    p = st.params
    sumU2 = (abs(u[(0, "x")]) ** 2 + abs(u[(0, "y")]) ** 2) / 2 + (abs(u[(1, "x")]) ** 2 + abs(u[(1, "y")]) ** 2) / 18
    want = (0.0 - p.ks / (2 * p.gamR ** 2) * (1 + pr["px"] ** 2 + pr["py"] ** 2 + sumU2)) * p.dz \
        - p.dz * 2.2e4 / (MC2 * 1e6)
    inside = np.arange(n) != 2
    np.testing.assert_allclose(st.theta[inside] - th0[inside], want[inside], rtol=1e-11)
    assert st.theta[2] == th0[2]
    np.testing.assert_array_equal(st.gamma, g0)


def test_guard_and_nan_in_the_oracle():
    """particlePush.cpp:370-374: a field whose ponderomotive term alone moves
    theta by more than 2 pi numSlice in one period trips the guard (the
    reference exits; the oracle returns ORC_ERR_GUARD). A NaN theta passes the
    reference's push and wrap unnoticed (NaN fails both comparisons of
    rearrangeParticles.cpp:105); the product's error word adds that check."""
    import fel_cases as fc
    case = fc.make_case(slice_n=3, ppslice=64, harmonics=(1, 3))
    case.state["Ux"][:, 2] *= 2e6
    o = case.oracle_state()
    assert o.push_theta_gamma(case.lattice()) == -3
    case = fc.make_case(slice_n=3, ppslice=64, harmonics=(1, 3))
    case.state["theta"][17] = np.nan
    o = case.oracle_state()
    assert o.push_theta_gamma(case.lattice()) == 0
    assert o.periodic() == 0 and np.isnan(o.theta[17])


def test_push_carries_em_past_the_radial_grid():
    """particlePush.cpp:248, 283-304: Em is one vector per push call. An
    in-grid particle with idxR < nr sets it from Ez; one with idxR >= nr uses
    what it holds, i.e. the Em of the last particle before it (slices in
    order) that set it, zero when none did; a particle outside the grid leaves
    it alone. With K0 = 0 and U = 0 a particle's push depends on its own
    (theta, gamma, p) and on Em only, so a carried particle must end exactly
    where a particle with its (theta, gamma, p) placed at its Em source does.
    Two modes, two slices with different Ez; dr = dx, nr = 5: radii from
    4.5 dr on are carried."""
    rng = np.random.default_rng(17)
    dx = 10e-6
    mom = dict(theta=[0.3, 1.1, 2.9, 4.0, 5.2, 0.7], gamma=[8001.0, 7999.0, 8003.0, 7998.5, 8002.0, 8000.5],
               px=[0.01, -0.02, 0.03, 0.0, -0.01, 0.02], py=[0.0, 0.01, -0.02, 0.02, 0.01, -0.01])
    pos = dict(x=[60e-6, 15e-6, 50e-6, 200e-6, 0.0, 20e-6], y=[0.0, 0.0, 50e-6, 0.0, 60e-6, 15e-6])
    # 0 carried, no source yet | 1 sets Em (slot 1) | 2 carried: source 1
    # 3 outside the grid (slot 2) | 4 carried: source 1 in the previous slice | 5 sets Em
    kw = dict(nx=21, ny=21, dx=dx, K0=0.0, nr=5, dr=dx, sc_l=2, slice_n=2)

    def state(parts, seg):
        st, lat = make_state(parts, **kw)
        st.seg_off[:] = seg
        st.Ez[:] = ez
        return st, lat

    S = 4
    ez = (rng.normal(size=(2, S * 5)) + 1j * rng.normal(size=(2, S * 5))) * 1e4
    st, lat = state(particles(6, **pos, **mom), [0, 3, 6])
    th0, g0 = st.theta.copy(), st.gamma.copy()
    assert st.push_theta_gamma(lat) == 0
    assert st.theta[3] == th0[3] and st.gamma[3] == g0[3]
    # particle 0: no Em yet -> the zero-field closed form, gamma unchanged
    dth0 = st.params.dz * (lat.ku - st.params.ks * (1 + 0.01 ** 2) / (2 * 8001.0 ** 2))
    assert st.gamma[0] == g0[0]
    np.testing.assert_allclose(st.theta[0] - th0[0], dth0, rtol=0, atol=4 * np.spacing(th0[0]))
    for n in (2, 4):
        one = {k: [mom[k][n]] for k in mom}
        ref, _ = state(particles(1, x=pos["x"][1], y=pos["y"][1], **one), [0, 1, 1])
        assert ref.push_theta_gamma(lat) == 0
        assert st.theta[n] == ref.theta[0] and st.gamma[n] == ref.gamma[0], n
        assert abs(st.gamma[n] - g0[n]) > 1e-6  # the carried Em moves gamma
    # particles 1 and 5 set Em from their own slots
    for n, seg in ((1, [0, 1, 1]), (5, [0, 0, 1])):
        one = {k: [mom[k][n]] for k in mom}
        ref, _ = state(particles(1, x=pos["x"][n], y=pos["y"][n], **one), seg)
        assert ref.push_theta_gamma(lat) == 0
        assert st.theta[n] == ref.theta[0] and st.gamma[n] == ref.gamma[0], n
