// gfx950 kernels for the FELIS per-step slice path (per-phase forms).
//
// Storage types: R = particle coordinate type (double, or float in the mixed
// path, where gamma is stored as an offset from a reference gamma), T = field
// component type (double or float). theta is always double. The per-phase
// particle kernels evaluate in FP64 whatever the storage, the ADI evaluates in T.
#pragma once

#include "felis_common.h"
#include "felis_emcarry.h"

namespace felis {

// ---------------------------------------------------------- transverse push

// particlePush.cpp:32-117, one RK4 over dz/2.
template <typename R>
__global__ void __launch_bounds__(256) k_transverse(long long np, GridArgs g, LatArgs l,
                                                    R *__restrict__ x, R *__restrict__ y,
                                                    R *__restrict__ px, R *__restrict__ py,
                                                    const R *__restrict__ gamma) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  double dz = g.dz * 0.5, ku = l.ku, K0 = l.K0, ue = l.ue;
  double aa = K0 * K0 * 0.25 * ue * ku;
  double bb = K0 * K0 * 0.25 * (1.0 + ue * ue) * ku * ku;
  double cc = kECharge * l.g / kEMass / kC;
  int ka = (int)l.K0_alpha;
  double xCoef = 0.0, yCoef = 0.0;
  if (l.und_type == 1) {
    if (ka == 1) xCoef = 1.0;
    else if (ka == -1) yCoef = 1.0;
  } else {
    xCoef = 1.0;
    yCoef = 1.0;
  }
  double gam = (double)gamma[n] + g.gref;
  double x0 = x[n], y0 = y[n], px0 = px[n], py0 = py[n];
  double dd = ku * K0 * K0 / (4.0 * gam * gam) * ue;
  double kx = 0, ky = 0, kpx = 0, kpy = 0;
  double sx = 0, sy = 0, spx = 0, spy = 0;
  const double cm[5] = {0, 0, 0.5, 0.5, 1};
  const double wt[5] = {0, 1, 2, 2, 1};
#pragma unroll
  for (int m = 1; m < 5; ++m) {
    double xx = x0 + dz * kx * cm[m];
    double yy = y0 + dz * ky * cm[m];
    double ppx = px0 + dz * kpx * cm[m];
    double ppy = py0 + dz * kpy * cm[m];
    kx = ppx / gam - aa * (ka * xx + yy) / (gam * gam);
    ky = ppy / gam + aa * (ka * xx - yy) / (gam * gam);
    kpx = ka * (ppx - ppy) * dd - (bb / gam * xCoef + cc) * xx;
    kpy = ka * (ppx + ppy) * dd - (bb / gam * yCoef - cc) * yy;
    sx += wt[m] * kx;
    sy += wt[m] * ky;
    spx += wt[m] * kpx;
    spy += wt[m] * kpy;
  }
  x[n] = (R)(x0 + dz / 6.0 * sx);
  y[n] = (R)(y0 + dz / 6.0 * sy);
  px[n] = (R)(px0 + dz / 6.0 * spx);
  py[n] = (R)(py0 + dz / 6.0 * spy);
}

// ---------------------------------------------------------- theta/gamma push

__device__ inline double bessel_interp(const double *__restrict__ tab, int bo, int idx, int order,
                                       double w0, double w1) {
  return tab[idx * bo + order] * w0 + tab[(idx + 1) * bo + order] * w1;
}

// Em of the longitudinal space-charge modes for an in-grid particle at
// radius^2 r2 in slot (particlePush.cpp:283-304): the two radial points
// weighted by wr, summed unweighted on the axis. False when idxR >= nr,
// where the reference keeps the Em of the particle it pushed before.
__device__ inline bool em_gather_f64(const GridArgs &g, const cpx<double> *__restrict__ Ez, int slot, double r2,
                                     cpx<double> (&Em)[4]) {
  const int L = g.L, F = g.F, nr = g.nr;
  const double r = sqrt(r2);
  const int idxR = (int)(r / g.dr + 0.5);
  const long long es = (long long)g.S * F * nr;
  if (idxR >= nr) return false;
  const double wr1 = idxR > 0 ? (r * r / (g.dr * g.dr) - idxR * idxR) / (2.0 * idxR) : 1.0;
  const double wr0 = idxR > 0 ? 1.0 - wr1 : 1.0;
  for (int ll = 0; ll < L; ++ll) {
    cpx<double> e = mk(0.0, 0.0);
    for (int f = 0; f < F; ++f) {
      const cpx<double> *p = Ez + ll * es + (long long)slot * nr * F + nr * f + idxR;
      e = e + p[0] * wr0;
      e = e + p[1] * wr1;
    }
    Em[ll] = e;
  }
  return true;
}

// particlePush.cpp:198-386 for one particle (FP64). Returns false when the
// particle is outside the push grid (it is then left untouched).
template <typename T>
__device__ inline bool push_tg_particle(const GridArgs &g, const LatArgs &l, int slot, double wakeE,
                                        const cpx<T> *__restrict__ Ux, const cpx<T> *__restrict__ Uy,
                                        const cpx<double> *__restrict__ Ez,
                                        const double *__restrict__ bes, double x0, double y0,
                                        double px, double py, double &th, double &gam,
                                        const cpx<double> *__restrict__ emc, long long n, unsigned *err) {
  const double ku = l.ku, ks = g.ks, K0 = l.K0, ue = l.ue, ka = l.K0_alpha;
  const double dz = g.dz;
  double r2 = x0 * x0 + y0 * y0;
  double pr2 = px * px + py * py;
  double th0 = th, gam0 = gam;
  double K2 = 1.0 + pr2 + (1.0 + ue * ue) * K0 * K0 * 0.5 * (1.0 + ku * ku * 0.5 * r2);
  double xi = ks / ku * K0 * K0 / (8.0 * gam0 * gam0) * (1.0 - ue * ue);
  int idxI = (int)((x0 - g.minX) / g.dx);
  int idxJ = (int)((y0 - g.minY) / g.dy);
  double wx1 = (x0 - g.minX) / g.dx - idxI, wx0 = 1.0 - wx1;
  double wy1 = (y0 - g.minY) / g.dy - idxJ, wy0 = 1.0 - wy1;
  if (!(idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1)) return false;

  // longitudinal space-charge field (one mode l for the usual L=F=1); an
  // in-grid particle beyond the radial grid takes the Em the reference
  // carries over from the particle it pushed before (emc, filled by
  // k_em_fill; null: refused)
  cpx<double> Em[4];
  if (!em_gather_f64(g, Ez, slot, r2, Em)) {
    if (emc) {
      for (int ll = 0; ll < g.L; ++ll) Em[ll] = emc[n * g.L + ll];
    } else {
      atomicOr(err, EB_EMCARRY);
      for (int ll = 0; ll < g.L; ++ll) Em[ll] = mk(0.0, 0.0);
    }
  }

  // bilinear field gather per harmonic
  cpx<double> ux[FELIS_MAX_HARM], uy[FELIS_MAX_HARM];
  double sumU2 = 0.0;
  const long long base = (long long)phys_slot(slot, g.ring_base, g.S) * g.N + (long long)idxJ * g.nx + idxI;
  for (int h = 0; h < g.nh; ++h) {
    const cpx<T> *px0 = Ux + (long long)h * g.S * g.N + base;
    const cpx<T> *py0 = Uy + (long long)h * g.S * g.N + base;
    cpx<double> a = mk(0.0, 0.0), b = mk(0.0, 0.0);
    a = a + cvt<double>(px0[0]) * (wx0 * wy0);
    b = b + cvt<double>(py0[0]) * (wx0 * wy0);
    a = a + cvt<double>(px0[g.nx]) * (wx0 * wy1);
    b = b + cvt<double>(py0[g.nx]) * (wx0 * wy1);
    a = a + cvt<double>(px0[1]) * (wx1 * wy0);
    b = b + cvt<double>(py0[1]) * (wx1 * wy0);
    a = a + cvt<double>(px0[g.nx + 1]) * (wx1 * wy1);
    b = b + cvt<double>(py0[g.nx + 1]) * (wx1 * wy1);
    ux[h] = a;
    uy[h] = b;
    int H = g.harm[h];
    sumU2 += (norm2(a) + norm2(b)) / (2.0 * H * H);
  }

  // stage-invariant coupling factors
  double xi2 = 0.0;
  cpx<double> expP = mk(1.0, 0.0), expM = mk(1.0, 0.0);
  bool any_even = false;
  for (int h = 0; h < g.nh; ++h) any_even |= (g.harm[h] % 2 == 0);
  if (any_even) {
    xi2 = sqrt(2.0) * K0 / (1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue)) *
          sqrt((1 + ka) * (px * px + ue * ue * py * py) + (1 - ka) * (px * px * ue * ue + py * py));
    double B = atan(2.0 * ka * ue * (px - py) /
                    ((1.0 + ka) * (px + ue * ue * py) + (1 - ka) * (px * ue * ue + py)));
    double sb, cb;
    sincos(B - l.Phi * 0.5, &sb, &cb);
    expP = mk(cb, sb);
    expM = conjg(expP);
  }
  cpx<double> fx[FELIS_MAX_HARM], fy[FELIS_MAX_HARM];
  for (int h = 0; h < g.nh; ++h) {
    int H = g.harm[h];
    int idx = (int)((1.0 * H * xi) / g.dB);
    double w1 = (H * xi / g.dB) - idx, w0 = 1.0 - w1;
    if (idx < 0 || idx + 1 >= g.bn) {
      atomicOr(err, EB_BESSEL);
      return false;
    }
    if (H % 2 == 1) {
      double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
      double J1 = bessel_interp(bes, g.bo, idx, (H - 1) / 2, w0, w1);
      double J2 = bessel_interp(bes, g.bo, idx, (H + 1) / 2, w0, w1);
      fx[h] = mk(sign * (ka * J1 - J2), 0.0);
      fy[h] = mk(sign * (ka * J1 + J2), 0.0);
    } else {
      double sign = (H / 2 % 2 == 0) ? 1.0 : -1.0;
      double J1 = bessel_interp(bes, g.bo, idx, (H - 2) / 2, w0, w1);
      double J2 = bessel_interp(bes, g.bo, idx, H / 2, w0, w1);
      double J3 = bessel_interp(bes, g.bo, idx, (H + 2) / 2, w0, w1);
      double s = sign * xi2 * H * 0.5;
      fx[h] = (expP * (ka * J1 - J2) + expM * (ka * J2 - J3)) * s;
      fy[h] = (expP * (ka * J1 + J2) + expM * (ka * J2 + J3)) * s;
    }
  }
  const cpx<double> etaX = mk(l.etaXr, l.etaXi), etaY = mk(l.etaYr, l.etaYi);
  const double e_mc2 = kECharge / kEMass / kC / kC;
  const double cm[5] = {0, 0, 0.5, 0.5, 1};
  double kth = 0.0, kg = 0.0, sth = 0.0, sg = 0.0;
  const double wt[5] = {0, 1, 2, 2, 1};
  for (int m = 1; m < 5; ++m) {
    double thm = th0 + dz * kth * cm[m];
    double gm = gam0 + dz * kg * cm[m];
    double sumTh = 0.0, sumG = 0.0, sumEz = 0.0;
    for (int ll = 0; ll < g.L; ++ll) {
      double s, c;
      sincos((ll + 1.0) * thm, &s, &c);
      sumEz += 2.0 * (Em[ll].re * c - Em[ll].im * s);
    }
    for (int h = 0; h < g.nh; ++h) {
      int H = g.harm[h];
      double s, c;
      sincos((1.0 * H) * (thm + l.Phi * 0.5), &s, &c);
      cpx<double> t = mk(c, s) * (etaX * fx[h] * ux[h] - etaY * fy[h] * uy[h]);
      sumTh += -t.im / (1.0 * H);
      sumG += t.re / (2.0 * gm);
    }
    kth = ku - ks / (2.0 * gm * gm) * (K2 + sumU2 + K0 * sqrt(ue * ue + 1.0) * sumTh);
    kg = ks * K0 * sqrt(ue * ue + 1.0) * sumG + e_mc2 * sumEz;
    sth += wt[m] * kth;
    sg += wt[m] * kg;
  }
  double dth = dz / 6.0 * sth;
  if (dth > g.dPhi || dth < -g.dPhi) {
    atomicOr(err, EB_GUARD);
    return false;
  }
  th = th0 + dth;
  gam = gam0 + (dz / 6.0 * sg - dz * wakeE);
  return true;
}

template <typename R, typename T>
__global__ void __launch_bounds__(kTile)
    k_push_tg(GridArgs g, LatArgs l, Tiles tl, double *__restrict__ theta, R *__restrict__ gamma,
              const R *__restrict__ x, const R *__restrict__ y, const R *__restrict__ px,
              const R *__restrict__ py, const cpx<T> *__restrict__ Ux,
              const cpx<T> *__restrict__ Uy, const cpx<double> *__restrict__ Ez,
              const double *__restrict__ bes, const double *__restrict__ wakeE,
              const cpx<double> *__restrict__ emc, unsigned *err) {
  int t = blockIdx.x;
  int k = threadIdx.x;
  if (k >= tl.cnt[t]) return;
  int sl = tl.seg[t] % g.slice_n;
  long long n = tl.start[t] + k;
  double wk = g.wake_on ? wakeE[sl] / (kMc2 * 1.0e6) : 0.0;
  double th = theta[n];
  double gm = (double)gamma[n] + g.gref;
  if (push_tg_particle<T>(g, l, sl + 1, wk, Ux, Uy, Ez, bes, (double)x[n], (double)y[n],
                          (double)px[n], (double)py[n], th, gm, emc, n, err)) {
    theta[n] = th;
    gamma[n] = (R)(gm - g.gref);
  }
}

// Em carry of the per-phase push: the positions are those the transverse
// kernels left, Em is formed in FP64 (see felis_emcarry.h)
template <typename R>
struct EmCarryPlain {
  using Pos = double2;
  using E = cpx<double>;
  GridArgs g;
  const R *x, *y;
  const cpx<double> *Ez;
  __device__ int where(long long n, Pos &q) const {
    const double x0 = (double)x[n], y0 = (double)y[n];
    const int idxI = (int)((x0 - g.minX) / g.dx), idxJ = (int)((y0 - g.minY) / g.dy);
    q = make_double2(x0, y0);
    if (!(idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1)) return 0;
    return (int)(sqrt(x0 * x0 + y0 * y0) / g.dr + 0.5) < g.nr ? 1 : 2;
  }
  __device__ int modes() const { return g.L; }
  __device__ void em(int seg, const Pos &q, E *out) const {
    cpx<double> Em[4];
    em_gather_f64(g, Ez, seg % g.slice_n + 1, q.x * q.x + q.y * q.y, Em);
    for (int l = 0; l < g.L; ++l) out[l] = Em[l];
  }
};

// particlePush.cpp:119-195: gaps; wake loss is applied to theta.
template <typename R, typename T>
__global__ void __launch_bounds__(kTile)
    k_drift(GridArgs g, LatArgs l, Tiles tl, double *__restrict__ theta, const R *__restrict__ x,
            const R *__restrict__ y, const R *__restrict__ px, const R *__restrict__ py,
            const cpx<T> *__restrict__ Ux, const cpx<T> *__restrict__ Uy,
            const double *__restrict__ wakeE, double gamR) {
  int t = blockIdx.x;
  int k = threadIdx.x;
  if (k >= tl.cnt[t]) return;
  int sl = tl.seg[t] % g.slice_n;
  long long n = tl.start[t] + k;
  double wk = g.wake_on ? wakeE[sl] / (kMc2 * 1.0e6) : 0.0;
  double x0 = x[n], y0 = y[n], ppx = px[n], ppy = py[n];
  double K2 = 1.0 + (ppx * ppx + ppy * ppy);
  int idxI = (int)((x0 - g.minX) / g.dx);
  int idxJ = (int)((y0 - g.minY) / g.dy);
  double wx1 = (x0 - g.minX) / g.dx - idxI, wx0 = 1.0 - wx1;
  double wy1 = (y0 - g.minY) / g.dy - idxJ, wy0 = 1.0 - wy1;
  if (!(idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1)) return;
  const long long base =
      (long long)phys_slot(sl + 1, g.ring_base, g.S) * g.N + (long long)idxJ * g.nx + idxI;
  double sumU2 = 0.0;
  for (int h = 0; h < g.nh; ++h) {
    const cpx<T> *a0 = Ux + (long long)h * g.S * g.N + base;
    const cpx<T> *b0 = Uy + (long long)h * g.S * g.N + base;
    cpx<double> a = cvt<double>(a0[0]) * (wx0 * wy0) + cvt<double>(a0[g.nx]) * (wx0 * wy1) +
                    cvt<double>(a0[1]) * (wx1 * wy0) + cvt<double>(a0[g.nx + 1]) * (wx1 * wy1);
    cpx<double> b = cvt<double>(b0[0]) * (wx0 * wy0) + cvt<double>(b0[g.nx]) * (wx0 * wy1) +
                    cvt<double>(b0[1]) * (wx1 * wy0) + cvt<double>(b0[g.nx + 1]) * (wx1 * wy1);
    int H = g.harm[h];
    sumU2 += (norm2(a) + norm2(b)) / (2.0 * H * H);
  }
  double tmp = l.ku - g.ks / (2.0 * gamR * gamR) * (K2 + sumU2);
  double th = theta[n];
  th += tmp * g.dz;
  th -= g.dz * wk;
  theta[n] = th;
}

// rearrangeParticles.cpp:77-114
static __global__ void __launch_bounds__(256) k_periodic(long long np, double dTh,
                                                  double *__restrict__ theta, unsigned *err) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  double th = theta[n], del = 0.0;
  if (th >= dTh) {
    int it = (int)(th / dTh);
    del = dTh * it;
  } else if (th < 0) {
    int it = (int)(th / dTh - 1.0);
    del = dTh * it;
  }
  th -= del;
  theta[n] = th;
  if (th != th) atomicOr(err, EB_NAN);
  else if (th > dTh || th < 0) atomicOr(err, EB_RANGE);
}

// ---------------------------------------------------------------- deposit



template <typename T>
__device__ inline void atomic_add_c(cpx<T> *p, cpx<double> v) {
  unsafeAtomicAdd(&p->re, (T)v.re);
  unsafeAtomicAdd(&p->im, (T)v.im);
}

// solveField.cpp:246-312: bunching deposit of every particle, subtracting
// from the four surrounding cells per harmonic and polarisation.
template <typename R, typename T>
__global__ void __launch_bounds__(kTile)
    k_deposit(GridArgs g, DepArgs d, Tiles tl, const double *__restrict__ theta,
              const R *__restrict__ gamma, const R *__restrict__ x, const R *__restrict__ y,
              const R *__restrict__ px, const R *__restrict__ py,
              const double *__restrict__ seg_w, const double *__restrict__ bes,
              cpx<T> *__restrict__ Scx, cpx<T> *__restrict__ Scy, unsigned *err) {
  int t = blockIdx.x;
  int k = threadIdx.x;
  if (k >= tl.cnt[t]) return;
  int seg = tl.seg[t];
  int slot = seg % g.slice_n + 1;
  long long n = tl.start[t] + k;
  double macro = seg_w[seg];
  double xx = x[n], yy = y[n], th = theta[n], gam = (double)gamma[n] + g.gref;
  double ppx = px[n], ppy = py[n];
  int idxI = (int)((xx - g.minX) / g.dx);
  int idxJ = (int)((yy - g.minY) / g.dy);
  double wx[2], wy[2];
  wx[1] = (xx - g.minX) / g.dx - idxI;
  wx[0] = 1.0 - wx[1];
  wy[1] = (yy - g.minY) / g.dy - idxJ;
  wy[0] = 1.0 - wy[1];
  if (!(idxI >= 0 && idxI < g.nx && idxJ >= 0 && idxJ < g.ny)) return;
  const double ue = d.ue, ka = d.ka;
  cpx<double> expP = mk(1.0, 0.0), expM = mk(1.0, 0.0);
  double xi2 = 0.0;
  bool any_even = false;
  for (int h = 0; h < g.nh; ++h) any_even |= (g.harm[h] % 2 == 0);
  if (any_even) {
    xi2 = d.xi2c * sqrt((1 + ka) * (ppx * ppx + ue * ue * ppy * ppy) +
                        (1 - ka) * (ppx * ppx * ue * ue + ppy * ppy));
    double B = atan(2.0 * ka * ue * (ppx - ppy) /
                    ((1.0 + ka) * (ppx + ue * ue * ppy) + (1 - ka) * (ppx * ue * ue + ppy)));
    double sb, cb;
    sincos(B - d.Phi * 0.5, &sb, &cb);
    expP = mk(cb, sb);
    expM = conjg(expP);
  }
  const long long idx = (long long)slot * g.N + (long long)idxJ * g.nx + idxI;
  const cpx<double> etaX = mk(d.etaXr, d.etaXi), etaY = mk(d.etaYr, d.etaYi);
  for (int h = 0; h < g.nh; ++h) {
    int H = g.harm[h];
    double dbH = (double)H;
    int ib = (int)((dbH * d.xi) / g.dB);
    double w1 = (dbH * d.xi / g.dB) - ib, w0 = 1.0 - w1;
    if (ib < 0 || ib + 1 >= g.bn) {
      atomicOr(err, EB_BESSEL);
      return;
    }
    cpx<double> fx, fy;
    if (H % 2 == 1) {
      double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
      double J1 = bessel_interp(bes, g.bo, ib, (H - 1) / 2, w0, w1);
      double J2 = bessel_interp(bes, g.bo, ib, (H + 1) / 2, w0, w1);
      fx = mk(sign * (J1 - ka * J2), 0.0);
      fy = mk(sign * (J1 + ka * J2), 0.0);
    } else {
      double sign = (H / 2 % 2 == 0) ? 1.0 : -1.0;
      double J1 = bessel_interp(bes, g.bo, ib, (H - 2) / 2, w0, w1);
      double J2 = bessel_interp(bes, g.bo, ib, H / 2, w0, w1);
      double J3 = bessel_interp(bes, g.bo, ib, (H + 2) / 2, w0, w1);
      double s = sign * xi2 * dbH * 0.5;
      fx = (expM * (J1 - ka * J2) + expP * (J2 - ka * J3)) * s;
      fy = (expM * (J1 + ka * J2) + expP * (J2 + ka * J3)) * s;
    }
    double sp, cp;
    sincos(-dbH * (th + d.Phi * 0.5), &sp, &cp);
    cpx<double> mc = mk(cp, sp) * (dbH * macro * d.coef_U / gam);
    cpx<double> cx = etaX * fx * mc, cy = etaY * fy * mc;
    cpx<T> *sx = Scx + (long long)h * g.S * g.N + idx;
    cpx<T> *sy = Scy + (long long)h * g.S * g.N + idx;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        double w = wx[ii] * wy[jj];
        atomic_add_c(sx + jj * g.nx + ii, cx * (-w));
        atomic_add_c(sy + jj * g.nx + ii, cy * (-w));
      }
  }
}

// Per-particle deposit factors of one harmonic (solveField.cpp:274-301):
// returns etaX*fx*mc and etaY*fy*mc (the corner weights are applied later).
__device__ inline void deposit_factors(const GridArgs &g, const DepArgs &d, const double *__restrict__ bes, int H,
                                       double macro, double th, double gam, double xi2, cpx<double> expP,
                                       cpx<double> &cx, cpx<double> &cy, bool &bad) {
  const double ka = d.ka;
  double dbH = (double)H;
  int ib = (int)((dbH * d.xi) / g.dB);
  double w1 = (dbH * d.xi / g.dB) - ib, w0 = 1.0 - w1;
  if (ib < 0 || ib + 1 >= g.bn) {
    bad = true;
    cx = cy = mk(0.0, 0.0);
    return;
  }
  cpx<double> fx, fy;
  if (H % 2 == 1) {
    double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
    double J1 = bessel_interp(bes, g.bo, ib, (H - 1) / 2, w0, w1);
    double J2 = bessel_interp(bes, g.bo, ib, (H + 1) / 2, w0, w1);
    fx = mk(sign * (J1 - ka * J2), 0.0);
    fy = mk(sign * (J1 + ka * J2), 0.0);
  } else {
    double sign = (H / 2 % 2 == 0) ? 1.0 : -1.0;
    double J1 = bessel_interp(bes, g.bo, ib, (H - 2) / 2, w0, w1);
    double J2 = bessel_interp(bes, g.bo, ib, H / 2, w0, w1);
    double J3 = bessel_interp(bes, g.bo, ib, (H + 2) / 2, w0, w1);
    double s = sign * xi2 * dbH * 0.5;
    cpx<double> expM = conjg(expP);
    fx = (expM * (J1 - ka * J2) + expP * (J2 - ka * J3)) * s;
    fy = (expM * (J1 + ka * J2) + expP * (J2 + ka * J3)) * s;
  }
  double sp, cp;
  sincos(-dbH * (th + d.Phi * 0.5), &sp, &cp);
  cpx<double> mc = mk(cp, sp) * (dbH * macro * d.coef_U / gam);
  cx = mk(d.etaXr, d.etaXi) * fx * mc;
  cy = mk(d.etaYr, d.etaYi) * fy * mc;
}

template <typename T>
__device__ inline void lds_add(cpx<T> *p, cpx<double> v) {
  atomicAdd(&p->re, (T)v.re);
  atomicAdd(&p->im, (T)v.im);
}

// Deposit with an LDS-privatised window: one workgroup per chunk of a
// segment (particles of one species in one slice). The window is the
// bounding box of the chunk's corners; harmonics are taken `hp` at a time so
// the window fits the LDS. Particles whose corners leave the slice's last
// row or column (the reference's flat-index spill, A.5) add straight to HBM.
template <typename R, typename T>
__global__ void __launch_bounds__(512)
    k_deposit_win(GridArgs g, DepArgs d, Tiles ch, const double *__restrict__ theta, const R *__restrict__ gamma,
                  const R *__restrict__ x, const R *__restrict__ y, const R *__restrict__ px,
                  const R *__restrict__ py, const double *__restrict__ seg_w, const double *__restrict__ bes,
                  cpx<T> *__restrict__ Scx, cpx<T> *__restrict__ Scy, int lds_bytes, unsigned *err) {
  extern __shared__ double lds_d[];
  cpx<T> *win = reinterpret_cast<cpx<T> *>(lds_d);
  __shared__ int bb[4];
  const int c = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int seg = ch.seg[c];
  const int slot = seg % g.slice_n + 1;
  const long long start = ch.start[c];
  const int cnt = ch.cnt[c];
  const double macro = seg_w[seg];
  if (tid == 0) {
    bb[0] = 1 << 30; bb[1] = -1; bb[2] = 1 << 30; bb[3] = -1;
  }
  __syncthreads();
  int mi = 1 << 30, Mi = -1, mj = 1 << 30, Mj = -1;
  for (int k = tid; k < cnt; k += nt) {
    double xx = x[start + k], yy = y[start + k];
    int idxI = (int)((xx - g.minX) / g.dx), idxJ = (int)((yy - g.minY) / g.dy);
    if (idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1) {
      mi = min(mi, idxI); Mi = max(Mi, idxI); mj = min(mj, idxJ); Mj = max(Mj, idxJ);
    }
  }
  atomicMin(&bb[0], mi); atomicMax(&bb[1], Mi); atomicMin(&bb[2], mj); atomicMax(&bb[3], Mj);
  __syncthreads();
  const int wi0 = bb[0], wj0 = bb[2];
  const bool empty = bb[1] < 0;
  const int ww = empty ? 0 : bb[1] + 2 - wi0, wh = empty ? 0 : bb[3] + 2 - wj0;
  const int W = ww * wh;
  int hp = W > 0 ? lds_bytes / (W * 2 * (int)sizeof(cpx<T>)) : g.nh;
  if (hp > g.nh) hp = g.nh;
  const cpx<double> etaX = mk(d.etaXr, d.etaXi), etaY = mk(d.etaYr, d.etaYi);
  (void)etaX; (void)etaY;
  bool any_even = false;
  for (int h = 0; h < g.nh; ++h) any_even |= (g.harm[h] % 2 == 0);
  const int step = hp > 0 ? hp : 1;
  for (int h0 = 0; h0 < g.nh; h0 += step) {
    const int hc = min(step, g.nh - h0);
    if (hp > 0) {
      for (int q = tid; q < hc * 2 * W; q += nt) win[q] = mk<T>(0, 0);
      __syncthreads();
    }
    bool bad = false;
    for (int k = tid; k < cnt; k += nt) {
      long long n = start + k;
      double xx = x[n], yy = y[n];
      int idxI = (int)((xx - g.minX) / g.dx), idxJ = (int)((yy - g.minY) / g.dy);
      if (!(idxI >= 0 && idxI < g.nx && idxJ >= 0 && idxJ < g.ny)) continue;
      double wx[2], wy[2];
      wx[1] = (xx - g.minX) / g.dx - idxI; wx[0] = 1.0 - wx[1];
      wy[1] = (yy - g.minY) / g.dy - idxJ; wy[0] = 1.0 - wy[1];
      double th = theta[n], gam = (double)gamma[n] + g.gref;
      double xi2 = 0.0;
      cpx<double> expP = mk(1.0, 0.0);
      if (any_even) {
        double ppx = px[n], ppy = py[n], ue = d.ue, ka = d.ka;
        xi2 = d.xi2c * sqrt((1 + ka) * (ppx * ppx + ue * ue * ppy * ppy) + (1 - ka) * (ppx * ppx * ue * ue + ppy * ppy));
        double B = atan(2.0 * ka * ue * (ppx - ppy) / ((1.0 + ka) * (ppx + ue * ue * ppy) + (1 - ka) * (ppx * ue * ue + ppy)));
        double sb, cb;
        sincos(B - d.Phi * 0.5, &sb, &cb);
        expP = mk(cb, sb);
      }
      const bool inwin = hp > 0 && idxI < g.nx - 1 && idxJ < g.ny - 1;
      for (int hh = 0; hh < hc; ++hh) {
        const int h = h0 + hh;
        cpx<double> cx, cy;
        deposit_factors(g, d, bes, g.harm[h], macro, th, gam, xi2, expP, cx, cy, bad);
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
          for (int jj = 0; jj < 2; ++jj) {
            double w = wx[ii] * wy[jj];
            if (inwin) {
              int q = (idxJ + jj - wj0) * ww + (idxI + ii - wi0);
              lds_add(win + (hh * 2 + 0) * W + q, cx * (-w));
              lds_add(win + (hh * 2 + 1) * W + q, cy * (-w));
            } else {
              long long o = (long long)h * g.S * g.N + (long long)slot * g.N + (long long)idxJ * g.nx + idxI +
                            jj * g.nx + ii;
              atomic_add_c(Scx + o, cx * (-w));
              atomic_add_c(Scy + o, cy * (-w));
            }
          }
      }
    }
    if (bad) atomicOr(err, EB_BESSEL);
    if (hp > 0) {
      __syncthreads();
      for (int q = tid; q < hc * 2 * W; q += nt) {
        cpx<T> v = win[q];
        if (v.re == (T)0 && v.im == (T)0) continue;
        int hh = q / (2 * W), pol = (q / W) & 1, r = q % W;
        int jy = wj0 + r / ww, ix = wi0 + r % ww;
        cpx<T> *dst = (pol ? Scy : Scx) + (long long)(h0 + hh) * g.S * g.N + (long long)slot * g.N +
                      (long long)jy * g.nx + ix;
        unsafeAtomicAdd(&dst->re, v.re);
        unsafeAtomicAdd(&dst->im, v.im);
      }
      __syncthreads();
    }
  }
}

// solveField.cpp:322-359: radial deposit of the last species into the
// per-slot source vector [slot][l][F*nr + 1] (the spare entry takes the
// idxR = nr-1 write that spills past the array in the reference).
template <typename R>
__global__ void __launch_bounds__(kTile)
    k_ez_deposit(GridArgs g, DepArgs d, Tiles tl, int tile0, const double *__restrict__ theta,
                 const R *__restrict__ x, const R *__restrict__ y, const double *__restrict__ seg_w,
                 cpx<double> *__restrict__ ScR) {
  extern __shared__ double lds[];
  const int L = g.L, F = g.F, nr = g.nr;
  const int row = F * nr + 1;
  const int nb = L * row;
  cpx<double> *hist = reinterpret_cast<cpx<double> *>(lds);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = mk(0.0, 0.0);
  __syncthreads();
  int t = tile0 + blockIdx.x;
  int seg = tl.seg[t];
  int slot = seg % g.slice_n + 1;
  int k = threadIdx.x;
  if (k < tl.cnt[t]) {
    long long n = tl.start[t] + k;
    double macro = seg_w[seg];
    double xx = x[n], yy = y[n], th = theta[n];
    double phi = (xx == 0.0) ? 0.0 : atan2(yy, xx);
    double r = sqrt(xx * xx + yy * yy);
    int idxR = (int)(r / g.dr + 0.5);
    if (idxR > 0 && idxR < nr) {
      double wr1 = (r * r / (g.dr * g.dr) - idxR * idxR) / (2.0 * idxR), wr0 = 1.0 - wr1;
      for (int ll = 0; ll < L; ++ll)
        for (int f = 0; f < F; ++f) {
          double s, c;
          sincos(-((ll + 1) * th + f * phi), &s, &c);
          // I * coef_Ez * (l+1) * exp(-i(...)) * macro / idxR
          double amp = d.coef_Ez * (ll + 1.0) * macro / (1.0 * idxR);
          cpx<double> cc = mk(-s * amp, c * amp);
          cpx<double> *hp = hist + ll * row + f * nr + idxR;
          atomicAdd(&hp[0].re, cc.re * wr0);
          atomicAdd(&hp[0].im, cc.im * wr0);
          atomicAdd(&hp[1].re, cc.re * wr1);
          atomicAdd(&hp[1].im, cc.im * wr1);
        }
    } else if (idxR == 0) {
      for (int ll = 0; ll < L; ++ll) {
        double s, c;
        sincos(-((ll + 1) * th), &s, &c);
        double amp = d.coef_Ez * (ll + 1.0) * macro;
        atomicAdd(&hist[ll * row].re, -s * amp);
        atomicAdd(&hist[ll * row].im, c * amp);
      }
    }
  }
  __syncthreads();
  cpx<double> *dst = ScR + (long long)slot * nb;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    cpx<double> v = hist[b];
    if (v.re != 0.0 || v.im != 0.0) {
      unsafeAtomicAdd(&dst[b].re, v.re);
      unsafeAtomicAdd(&dst[b].im, v.im);
    }
  }
}

// solveField.cpp:361-391, the field-independent half: forward elimination of
// the real tridiagonal matrix per (l, f), shared by every slice.
// coef[(l*F + f)*2*nr + j] = w_j (multiplier), [.. + nr + j] = b'_j.
static __global__ void k_ez_coef(GridArgs g, double A, double *__restrict__ coef) {
  const int L = g.L, F = g.F, nr = g.nr;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L * F) return;
  int f = t % F, ll = t / F;
  double *w = coef + (long long)t * 2 * nr, *b = w + nr;
  double aprev = (f == 0) ? 1.0 : 0.0;
  b[0] = (f == 0) ? -(1.0 + (ll + 1.0) * (ll + 1.0) * A / 8.0) : 1.0;
  w[0] = 0.0;
  for (int j = 1; j < nr; ++j) {
    double aj, bj, cj;
    if (j < nr - 1) {
      aj = j + 0.5;
      bj = -2.0 * j - (ll + 1.0) * (ll + 1.0) * A * j - f * f * log((j + 0.5) / (j - 0.5));
      cj = j - 0.5;
    } else {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
    }
    double wj = cj / b[j - 1];
    b[j] = bj - wj * aprev;
    w[j] = wj;
    aprev = aj;
  }
}

// solveField.cpp:387-396 per (slot, l, f): the source recurrence and the back
// substitution, which reads slot 0 of Ez (A.3).
// One wavefront per (slice, l, f). The forward elimination of the source,
// d_j = S_j - w_j d_{j-1}, is a first-order linear recurrence solved with a
// wave-wide scan of affine maps (lane owns rows lane*E .. lane*E+E-1); the
// back substitution only reads slot 0 of Ez (A.3), so it is elementwise.
template <int E>
__global__ void __launch_bounds__(64) k_ez_solve(GridArgs g, const double *__restrict__ coef,
                                                 const cpx<double> *__restrict__ ScR,
                                                 cpx<double> *__restrict__ Ez) {
  const int L = g.L, F = g.F, nr = g.nr;
  const int sid = blockIdx.x, lane = threadIdx.x;
  const int f = sid % F, ll = (sid / F) % L, sl = sid / (F * L), slot = sl + 1;
  const int row = F * nr + 1;
  const cpx<double> *S = ScR + ((long long)slot * L + ll) * row;
  const double *w = coef + (long long)(ll * F + f) * 2 * nr, *bp = w + nr;
  const long long es = (long long)g.S * F * nr;
  cpx<double> *ez = Ez + ll * es;
  // slot 0's last entry, which the reference clears after the solve: no
  // block reads it (the elimination below takes it as zero) or writes slot 0
  if (sid == 0)
    for (int t = lane; t < L * F; t += 64) Ez[(long long)(t / F) * es + (t % F) * nr + nr - 1] = mk(0.0, 0.0);
  double av[E];
  cpx<double> bv[E];
  double A = 1.0;
  cpx<double> B = mk(0.0, 0.0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int j = lane * E + e;
    if (j < nr) {
      av[e] = j == 0 ? 0.0 : -w[j];
      bv[e] = j == 0 ? (f == 0 ? S[0] : mk(0.0, 0.0)) : (j < nr - 1 ? S[f * nr + j] : mk(0.0, 0.0));
    } else {
      av[e] = 1.0;
      bv[e] = mk(0.0, 0.0);
    }
    B = B * av[e] + bv[e];
    A = A * av[e];
  }
#pragma unroll
  for (int dd = 1; dd < 64; dd <<= 1) {
    const double Ap = __shfl_up(A, dd);
    const double Br = __shfl_up(B.re, dd), Bi = __shfl_up(B.im, dd);
    if (lane >= dd) {
      B = mk(B.re + A * Br, B.im + A * Bi);
      A = A * Ap;
    }
  }
  const double pr = __shfl_up(B.re, 1), pi = __shfl_up(B.im, 1);
  cpx<double> dprev = lane == 0 ? mk(0.0, 0.0) : mk(pr, pi);
  const double a0 = (f == 0) ? 1.0 : 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int j = lane * E + e;
    dprev = dprev * av[e] + bv[e];
    if (j <= nr - 2) {
      const double aj = (j == 0) ? a0 : j + 0.5;
      // slot 0 of Ez, whose last entry the reference clears before reading
      const cpx<double> z = j + 1 == nr - 1 ? mk(0.0, 0.0) : ez[f * nr + j + 1];
      const cpx<double> v = dprev - z * aj;
      ez[(long long)slot * F * nr + f * nr + j] = mk(v.re / bp[j], v.im / bp[j]);
    }
  }
}

// ---------------------------------------------------------------- ADI

// Per-harmonic, per-direction recurrence coefficients (complex double):
// DD_i = r_i dd_i + a_i DD_{i-1};  U_i = q_i DD_i + p_i U_{i+1}
// (solveField.cpp:137-152 and 165-180 with the field-independent CC hoisted).


// Step 1 (x implicit, solveField.cpp:127-153): one wavefront per row; the
// two first-order recurrences are solved with a wave-wide scan of affine maps.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_adi_x(GridArgs g, AdiCoef c, int cf,
                                               const cpx<T> *__restrict__ U,
                                               const cpx<T> *__restrict__ Sc,
                                               cpx<T> *__restrict__ Uc) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int j = 1 + blockIdx.x * 4 + wv;
  const int sys = blockIdx.y;  // (pol*nh + h)*slice_n + sl
  if (j > g.ny - 2) return;
  const int sl = sys % g.slice_n;
  const int ph = sys / g.slice_n;  // pol*nh + h
  const int h = ph % g.nh;
  const int slot = sl + 1;
  const int nx = g.nx;
  const T H = (T)g.harm[h];
  const cpx<T> *Un = U + ((long long)ph * g.S + phys_slot(slot, g.ring_base, g.S)) * g.N;
  const cpx<T> *Sn = Sc + ((long long)ph * g.S + slot) * g.N;
  const cpx<double> *cx = c.x + (long long)h * 4 * nx;
  const cpx<T> bP = cvt<T>(c.abc[2 * nx + g.ny + j]);
  const cpx<T> bM = cvt<T>(c.abc[2 * nx + j]);
  const cpx<T> diag = mk<T>(H, 0) + bP + bM;
  const int n = nx - 2;
  // lane owns i = 1 + lane*E .. 1 + lane*E + E-1
  cpx<T> av[E], bv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    int i = 1 + lane * E + e;
    if (i <= n) {
      long long idx = (long long)j * nx + i;
      cpx<T> dd = diag * Un[idx] - bP * Un[idx + nx] - bM * Un[idx - nx];
      if (cf) dd = dd + Sn[idx];
      bv[e] = cvt<T>(cx[i]) * dd;       // r_i * dd_i
      av[e] = cvt<T>(cx[nx + i]);       // a_i
    } else {
      av[e] = mk<T>(1, 0);
      bv[e] = mk<T>(0, 0);
    }
  }
  // local composition x -> A x + B over the lane's chunk
  cpx<T> A = mk<T>(1, 0), B = mk<T>(0, 0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    B = av[e] * B + bv[e];
    A = av[e] * A;
  }
  // inclusive scan across lanes (lane l composes lanes 0..l)
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T Ar = __shfl_up(A.re, d), Ai = __shfl_up(A.im, d);
    T Br = __shfl_up(B.re, d), Bi = __shfl_up(B.im, d);
    if (lane >= d) {
      cpx<T> Ap = mk(Ar, Ai), Bp = mk(Br, Bi);
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  T pr = __shfl_up(B.re, 1), pi = __shfl_up(B.im, 1);
  cpx<T> DD = lane == 0 ? mk<T>(0, 0) : mk(pr, pi);
  cpx<T> ddv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    DD = av[e] * DD + bv[e];
    ddv[e] = DD;
  }
  // back substitution U_i = q_i DD_i + p_i U_{i+1}, scanned from the right
  A = mk<T>(1, 0);
  B = mk<T>(0, 0);
#pragma unroll
  for (int e = E - 1; e >= 0; --e) {
    int i = 1 + lane * E + e;
    cpx<T> q, p;
    if (i <= n) {
      q = cvt<T>(cx[2 * nx + i]);
      p = cvt<T>(cx[3 * nx + i]);
    } else {
      q = mk<T>(0, 0);
      p = mk<T>(1, 0);
    }
    av[e] = p;
    bv[e] = q * ddv[e];
    B = p * B + bv[e];
    A = p * A;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T Ar = __shfl_down(A.re, d), Ai = __shfl_down(A.im, d);
    T Br = __shfl_down(B.re, d), Bi = __shfl_down(B.im, d);
    if (lane + d < 64) {
      cpx<T> Ap = mk(Ar, Ai), Bp = mk(Br, Bi);
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  T nr_ = __shfl_down(B.re, 1), ni_ = __shfl_down(B.im, 1);
  cpx<T> Unext = lane == 63 ? mk<T>(0, 0) : mk(nr_, ni_);
  cpx<T> *Uo = Uc + (long long)sys * g.N + (long long)j * nx;
#pragma unroll
  for (int e = E - 1; e >= 0; --e) {
    int i = 1 + lane * E + e;
    Unext = av[e] * Unext + bv[e];
    if (i <= n) Uo[i] = Unext;
  }
}

// Step 2 (y implicit, solveField.cpp:155-181): one thread per column, the
// rows walked in order; DD is parked in the output cell and overwritten by
// the back substitution.
template <typename T>
__global__ void __launch_bounds__(128) k_adi_y(GridArgs g, AdiCoef c, int cf, cpx<T> *__restrict__ U,
                                               const cpx<T> *__restrict__ Sc,
                                               const cpx<T> *__restrict__ Uc) {
  const int sys = blockIdx.x;
  const int i = 1 + blockIdx.y * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (i > nx - 2) return;
  const int sl = sys % g.slice_n;
  const int ph = sys / g.slice_n;
  const int h = ph % g.nh;
  const int slot = sl + 1;
  const T H = (T)g.harm[h];
  cpx<T> *Un = U + ((long long)ph * g.S + phys_slot(slot, g.ring_base, g.S)) * g.N;
  const cpx<T> *Sn = Sc + ((long long)ph * g.S + slot) * g.N;
  const cpx<T> *Ucs = Uc + (long long)sys * g.N;
  const cpx<double> *cy = c.y + (long long)h * 4 * ny;
  const cpx<T> aM = cvt<T>(c.abc[i]);
  const cpx<T> aP = cvt<T>(c.abc[nx + i]);
  const cpx<T> diag = mk<T>(H, 0) + aP + aM;
  cpx<T> DD = mk<T>(0, 0);
  for (int j = 1; j <= ny - 2; ++j) {
    long long idx = (long long)j * nx + i;
    cpx<T> dd = diag * Ucs[idx] - aP * Ucs[idx + 1] - aM * Ucs[idx - 1];
    if (cf) dd = dd + Sn[idx];
    DD = cvt<T>(cy[ny + j]) * DD + cvt<T>(cy[j]) * dd;
    Un[idx] = DD;
  }
  cpx<T> Unext = mk<T>(0, 0);
  for (int j = ny - 2; j >= 1; --j) {
    long long idx = (long long)j * nx + i;
    Unext = cvt<T>(cy[2 * ny + j]) * Un[idx] + cvt<T>(cy[3 * ny + j]) * Unext;
    Un[idx] = Unext;
  }
}

// ---------------------------------------------------------------- slippage

static __global__ void k_fill_u32(unsigned *p, long long n, unsigned v) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}

// zero one physical slot of every (pol, h) field after a ring rotation
template <typename T>
__global__ void k_zero_slot(cpx<T> *U, int phys, int S, long long N, int nph) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int ph = blockIdx.y;
  if (k < N && ph < nph) U[((long long)ph * S + phys) * N + k] = mk<T>(0, 0);
}

// ---------------------------------------------------------------- reductions

// Block-wide FP64 sum in a fixed order (deterministic for a fixed launch shape).
template <int NT>
__device__ inline double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < NT / 64; ++k) r += sh[k];
  }
  __syncthreads();
  return r;
}

// |U|^2 per (pol, h, slot) over all N cells: out[(pol*nh + h)*slice_n + sl]
template <typename T>
__global__ void __launch_bounds__(256) k_energy_slices(GridArgs g, const cpx<T> *__restrict__ U,
                                                       double *__restrict__ out) {
  __shared__ double sh[4];
  const int sys = blockIdx.x;
  const int sl = sys % g.slice_n, ph = sys / g.slice_n;
  const cpx<T> *p = U + ((long long)ph * g.S + phys_slot(sl + 1, g.ring_base, g.S)) * g.N;
  double acc = 0.0;
  for (long long k = threadIdx.x; k < g.N; k += 256) {
    cpx<T> v = p[k];
    acc += (double)v.re * v.re + (double)v.im * v.im;
  }
  double s = block_sum<256>(acc, sh);
  if (threadIdx.x == 0) out[sys] = s;
}

// rows of a [nrow][ncol] matrix, one workgroup of kSumRowThreads per row in a
// fixed order: strided partial sums in four independent chains per thread
// (the loads of a chain do not wait for each other's sums), then a tree
constexpr int kSumRowThreads = 256;
static __global__ void __launch_bounds__(kSumRowThreads) k_sum_rows(const double *__restrict__ in, int ncol,
                                                                    int stride_row, int stride_col,
                                                                    double *__restrict__ out) {
  constexpr int NT = kSumRowThreads;
  __shared__ double sh[NT];
  const int tid = threadIdx.x;
  const double *row = in + (long long)blockIdx.x * stride_row;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int c = tid;
  for (; c + 3 * NT < ncol; c += 4 * NT) {
    s0 += row[(long long)c * stride_col];
    s1 += row[(long long)(c + NT) * stride_col];
    s2 += row[(long long)(c + 2 * NT) * stride_col];
    s3 += row[(long long)(c + 3 * NT) * stride_col];
  }
  for (; c < ncol; c += NT) s0 += row[(long long)c * stride_col];
  sh[tid] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  if (tid == 0) out[blockIdx.x] = sh[0];
}

// One wave per tile for the per-tile diagnostics below: lane l takes the
// tile's particles l, l+64, ..., and the wave sums each value by a fixed
// shuffle tree (no LDS, no barrier; a quarter of the shuffles of a
// 256-thread block per tile).
constexpr int kTilesPerBlock = 4;
__device__ inline double wave_sum(double a) {
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o);
  return a;
}

// b^n for a unit complex b by binary powers (n < 0: the conjugate's)
__device__ inline cpx<double> cpow_int_d(cpx<double> b, int n) {
  if (n < 0) {
    b = conjg(b);
    n = -n;
  }
  cpx<double> r = mk(1.0, 0.0);
  while (n) {
    if (n & 1) r = r * b;
    b = b * b;
    n >>= 1;
  }
  return r;
}

// bFactor partials per tile: [tile][2*nh + 1]. e^{iH theta} from one FP64
// sincos of theta by products (binary powers, then the steps between the
// harmonics), a few ulp from sincos(H theta) directly.
template <typename R>
__global__ void __launch_bounds__(64 * kTilesPerBlock) k_bfactor_tiles(GridArgs g, Tiles tl, int ntiles,
                                                                       const double *__restrict__ theta,
                                                                       const double *__restrict__ seg_w,
                                                                       double *__restrict__ out) {
  const int t = blockIdx.x * kTilesPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= ntiles) return;  // whole waves
  const int nh = g.nh, nv = 2 * nh + 1, cnt = tl.cnt[t];
  const long long st = tl.start[t];
  const double w = seg_w[tl.seg[t]];
  double v[2 * FELIS_MAX_HARM];
#pragma unroll
  for (int q = 0; q < 2 * FELIS_MAX_HARM; ++q) v[q] = 0.0;
  for (int k = lane; k < cnt; k += 64) {
    double s, c;
    sincos(theta[st + k], &s, &c);
    const cpx<double> e1 = mk(c, s);
    cpx<double> pw = cpow_int_d(e1, g.harm[0]);
#pragma unroll
    for (int h = 0; h < FELIS_MAX_HARM; ++h) {
      if (h >= nh) break;
      if (h > 0) pw = pw * cpow_int_d(e1, g.harm[h] - g.harm[h - 1]);
      v[2 * h] += pw.re * w;
      v[2 * h + 1] += pw.im * w;
    }
  }
  double* o = out + (long long)t * nv;
#pragma unroll
  for (int q = 0; q < 2 * FELIS_MAX_HARM; ++q) {
    if (q >= 2 * nh) break;
    const double r = wave_sum(v[q]);
    if (lane == 0) o[q] = r;
  }
  if (lane == 0) o[2 * nh] = w * cnt;
}

// twiss partials per tile: [tile][7]
template <typename R>
__global__ void __launch_bounds__(64 * kTilesPerBlock) k_twiss_tiles(GridArgs g, Tiles tl, int ntiles,
                                                                     const R *__restrict__ x,
                                                                     const R *__restrict__ y, const R *__restrict__ px,
                                                                     const R *__restrict__ py,
                                                                     const R *__restrict__ gamma,
                                                                     double *__restrict__ out) {
  const int t = blockIdx.x * kTilesPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= ntiles) return;  // whole waves
  const int cnt = tl.cnt[t];
  const long long st = tl.start[t];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int k = lane; k < cnt; k += 64) {
    const long long n = st + k;
    const double xx = x[n], yy = y[n], inv = 1.0 / ((double)gamma[n] + g.gref);
    const double xp = (double)px[n] * inv, yp = (double)py[n] * inv;
    v[0] += xx * xx; v[1] += yy * yy; v[2] += xp * xp; v[3] += yp * yp;
    v[4] += xx * xp; v[5] += yy * yp;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double r = wave_sum(v[q]);
    if (lane == 0) out[(long long)t * 7 + q] = r;
  }
  if (lane == 0) out[(long long)t * 7 + 6] = (double)cnt;
}

// wakeField.cpp:28-35: wakeE[ii] = sum_{i>=ii} den[i] wakeF[i-ii] dZ coef,
// accumulated in the reference's order (i descending) without contraction.
static __global__ void k_wake(int sn, double dZ, double coef, const double *__restrict__ den,
                       const double *__restrict__ wakeF, double *__restrict__ wakeE) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= sn) return;
  double acc = 0.0;
  for (int i = sn - 1; i >= ii; --i) {
    double t = __dmul_rn(__dmul_rn(__dmul_rn(den[i], wakeF[i - ii]), dZ), coef);
    acc = __dadd_rn(acc, t);
  }
  wakeE[ii] = acc;
}

// ---------------------------------------------------------------- 1D (nx = ny = 1)

// push_theta_gamma_1D (particlePush.cpp:388-524) for one particle per thread:
// the slice's field and its L longitudinal space-charge modes drive an RK4 of
// theta and gamma; even harmonics do not couple (494-497).
static __global__ void __launch_bounds__(kTile)
    k_push_tg_1d(GridArgs g, LatArgs l, Tiles tl, double *__restrict__ theta, double *__restrict__ gamma,
                 const cpx<double> *__restrict__ Ux, const cpx<double> *__restrict__ Uy,
                 const cpx<double> *__restrict__ Ez, const double *__restrict__ bes,
                 const double *__restrict__ wakeE, unsigned *err) {
  const int t = blockIdx.x, k = threadIdx.x;
  if (k >= tl.cnt[t]) return;
  const int seg = tl.seg[t], sl = seg % g.slice_n, slot = sl + 1;
  const long long n = tl.start[t] + k;
  const int phys = phys_slot(slot, g.ring_base, g.S);
  const double ue = l.ue, ks = g.ks, ku = l.ku, K0 = l.K0, dz = g.dz, ka = l.K0_alpha;
  const double dPhi = g.dPhi;
  const double e_mc2 = kECharge / kEMass / kC / kC;
  const double K2 = 1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue);
  const double xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  const double wk = g.wake_on ? wakeE[sl] / (kMc2 * 1.0e6) : 0.0;
  const cpx<double> etaX = mk(l.etaXr, l.etaXi), etaY = mk(l.etaYr, l.etaYi);
  double sumU2 = 0.0;
  for (int h = 0; h < g.nh; ++h) {
    const cpx<double> ux = Ux[(long long)h * g.S + phys], uy = Uy[(long long)h * g.S + phys];
    const int H = g.harm[h];
    sumU2 += (norm2(ux) + norm2(uy)) / (2.0 * H * H);
  }
  const double th0 = theta[n], gam0 = gamma[n];
  const double cm[5] = {0, 0, 0.5, 0.5, 1};
  double kth[5] = {0, 0, 0, 0, 0}, kg[5] = {0, 0, 0, 0, 0};
  for (int m = 1; m < 5; ++m) {
    const double th = th0 + dz * kth[m - 1] * cm[m];
    const double gam = gam0 + dz * kg[m - 1] * cm[m];
    double sumTh = 0.0, sumG = 0.0, sumEz = 0.0;
    for (int q = 0; q < g.L; ++q) {
      double sn, cs;
      sincos((q + 1.0) * th, &sn, &cs);
      const cpx<double> em = Ez[(long long)q * g.S + slot];
      sumEz += 2.0 * (em.re * cs - em.im * sn);
    }
    for (int h = 0; h < g.nh; ++h) {
      const int H = g.harm[h];
      if (H % 2 == 0) continue;
      const double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
      const int idx = (int)((1.0 * H * xi) / g.dB);
      const double w1 = (H * xi / g.dB) - idx, w0 = 1.0 - w1;
      if (idx < 0 || idx + 1 >= g.bn) {
        atomicOr(err, EB_BESSEL);
        return;
      }
      const double J1 = bessel_interp(bes, g.bo, idx, (H - 1) / 2, w0, w1);
      const double J2 = bessel_interp(bes, g.bo, idx, (H + 1) / 2, w0, w1);
      const double fx = sign * (ka * J1 - J2), fy = sign * (ka * J1 + J2);
      double sn, cs;
      sincos((1.0 * H) * (th + l.Phi * 0.5), &sn, &cs);
      const cpx<double> ux = Ux[(long long)h * g.S + phys], uy = Uy[(long long)h * g.S + phys];
      const cpx<double> c = etaX * ux * fx - etaY * uy * fy;
      const cpx<double> tt = mk(cs, sn) * c;
      sumTh += -tt.im / (1.0 * H);
      sumG += tt.re / (2.0 * gam);
    }
    kth[m] = ku - ks / (2.0 * gam * gam) * (K2 + sumU2 + K0 * sqrt(ue * ue + 1.0) * sumTh);
    kg[m] = ks * K0 * sqrt(ue * ue + 1.0) * sumG + e_mc2 * sumEz;
  }
  const double dth = dz / 6.0 * (kth[1] + 2.0 * kth[2] + 2.0 * kth[3] + kth[4]);
  if (dth > dPhi || dth <= -dPhi) {
    atomicOr(err, EB_GUARD);
    return;
  }
  theta[n] = th0 + dth;
  gamma[n] = gam0 + dz / 6.0 * (kg[1] + 2 * kg[2] + 2 * kg[3] + kg[4]) - dz * wk;
}

// solve_Sc_1D (solveField.cpp:422-549) for one slot per workgroup: the
// bunching source of every harmonic and the L longitudinal modes of Ez, each
// a sum over the slot's particles of every species. sig[2s], sig[2s+1] are
// the species' rms beam sizes that set the source area (463-468).
static __global__ void __launch_bounds__(256)
    k_sc_1d(GridArgs g, LatArgs l, int n_species, const long long *__restrict__ seg_off,
            const double *__restrict__ seg_w, const double *__restrict__ sig, const double *__restrict__ theta,
            const double *__restrict__ gamma, const double *__restrict__ bes, double bucket,
            cpx<double> *__restrict__ Scx, cpx<double> *__restrict__ Scy, cpx<double> *__restrict__ Ez,
            unsigned *err) {
  __shared__ double sh[256];
  const int sl = blockIdx.x, slot = sl + 1, tid = threadIdx.x;
  const double ue = l.ue, K0 = l.K0, ka = l.K0_alpha;
  const double xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  // per harmonic: sum of macro coef e^{-iH(theta+Phi/2)}/gamma; per mode: sum
  // of macro coef2 e^{-i(l+1) theta}
  for (int q = 0; q < g.nh + g.L; ++q) {
    double re = 0.0, im = 0.0;
    for (int s = 0; s < n_species; ++s) {
      const double area = 2.0 * M_PI * sig[2 * s] * sig[2 * s + 1];
      const int seg = s * g.slice_n + sl;
      const double macro = seg_w[seg];
      double c;
      if (q < g.nh)
        c = g.dz * kECharge * kECharge * kMu0 * K0 * sqrt(1.0 + ue * ue) / (2.0 * g.ks * kEMass * bucket * area);
      else
        c = kECharge * kC * kC * kMu0 / g.ks / area / bucket / (1.0 + (q - g.nh));
      for (long long n = seg_off[seg] + tid; n < seg_off[seg + 1]; n += blockDim.x) {
        double sn, cs;
        if (q < g.nh) {
          sincos(-(1.0 * g.harm[q]) * (theta[n] + l.Phi * 0.5), &sn, &cs);
          const double a = macro * c / gamma[n];
          re += cs * a;
          im += sn * a;
        } else {
          sincos(-((q - g.nh) + 1.0) * theta[n], &sn, &cs);
          re += cs * macro * c;
          im += sn * macro * c;
        }
      }
    }
    re = block_sum<256>(re, sh);
    im = block_sum<256>(im, sh);
    if (tid == 0) {
      if (q < g.nh) {
        const int H = g.harm[q];
        double fx = 0.0, fy = 0.0;
        if (H % 2 == 1) {
          const double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
          const int idx = (int)(H * xi / g.dB);
          const double w1 = (H * xi / g.dB) - idx, w0 = 1.0 - w1;
          if (idx < 0 || idx + 1 >= g.bn) {
            atomicOr(err, EB_BESSEL);
          } else {
            const double J1 = bessel_interp(bes, g.bo, idx, (H - 1) / 2, w0, w1);
            const double J2 = bessel_interp(bes, g.bo, idx, (H + 1) / 2, w0, w1);
            fx = sign * (J1 - ka * J2);
            fy = sign * (J1 + ka * J2);
          }
        }
        const cpx<double> m = mk(re, im);
        Scx[(long long)q * g.S + slot] = (mk(l.etaXr, l.etaXi) * m) * (-fx);
        Scy[(long long)q * g.S + slot] = (mk(l.etaYr, l.etaYi) * m) * (-fy);
      } else {
        // -i * sum
        Ez[(long long)(q - g.nh) * g.S + slot] = mk(im, -re);
      }
    }
  }
  if (sl == 0 && tid < g.L) {
    Ez[(long long)tid * g.S] = mk(0.0, 0.0);
    Ez[(long long)tid * g.S + g.S - 1] = mk(0.0, 0.0);
  }
}

// solve_Field_U_1D (solveField.cpp:409-420): U += Sc on slots 1..slice_n.
static __global__ void k_field_1d(GridArgs g, cpx<double> *__restrict__ U, const cpx<double> *__restrict__ Sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // over pol, h, slice
  if (t >= 2 * g.nh * g.slice_n) return;
  const int ph = t / g.slice_n, slot = t % g.slice_n + 1;
  cpx<double> &u = U[(long long)ph * g.S + phys_slot(slot, g.ring_base, g.S)];
  u = u + Sc[(long long)ph * g.S + slot];
}

}  // namespace felis
