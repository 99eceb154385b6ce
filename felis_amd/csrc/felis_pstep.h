// Particle half of the batched step (felis_run, FP32 with FP64 phase).
//
//   k_pstep        one workgroup per tile (<= 256 particles of one segment):
//                  transversePush x2 (particlePush.cpp:32-117), push_theta_gamma
//                  (198-386) or drift (119-195), the theta wrap
//                  (rearrangeParticles.cpp:77-114) and the moments of the next
//                  step's bFactor and twiss (saveFile.cpp:124-154,
//                  twiss.cpp:34-57) as per-tile FP64 partial sums.
//   k_deposit_slot one workgroup per slice: the next solveField's bunching
//                  deposit (solveField.cpp:246-312) into an LDS window over the
//                  slice's beam, stored once, and the radial space-charge
//                  source of the last species (solveField.cpp:322-359).
//   k_reduce_parts, k_reduce_final
//                  the step's diagnostics from the partial sums, in a fixed
//                  order.
#pragma once

#include "felis_common.h"
#include "felis_kernels.h"

namespace felis {

constexpr int kRun = 8;       // consecutive particles one deposit thread combines (DepArgs::run, tuned)
constexpr int kMaxModes = 4;  // longitudinal space-charge modes of the batched path

// transversePush constants of one lattice row (particlePush.cpp:76-90),
// formed on the host: the same for every particle of a launch
struct TransArgs {
  float aa, bb, cc, ddc, ue, xCoef, yCoef;  // dd = ddc ig^2 ue
  int ka;
};

struct PStepArgs {
  LatArgs l0, l1;  // transversePush at (it,0); everything else at (it,0.5)
  TransArgs t0, t1;
  int trans_same;  // t0 and t1 are the same map (a uniform section): formed once, applied twice
  // push_theta_gamma constants of l1 (particlePush.cpp:260-300), formed on
  // the host: K2 = 1 + p^2 + K2a (1 + K2b r^2), xi = xic / gamma0^2,
  // xi2 = xi2c |p|_ue, sq = K0 sqrt(1 + ue^2), rotm = e^{-i Phi/2}
  double K2a, K2b, xic, halfPhi;
  float xi2c, sq, kaf;
  cpx<float> rotm, etaX, etaY;
  int drift;       // driftFlag of (it,0.5)
  int nred;        // partial-sum row length: 2*nh + 1 + 7
  double gamR;
  double inv_dx, inv_dy, inv_dr;  // grid reciprocals
  float inv_drf;
  // per harmonic H: 1/H, H/dBessel, 1/(2 H^2) and H - H_prev (H_prev of the
  // first harmonic is 0), the step of the power chain e^{iH psi}
  float hinv[FELIS_MAX_HARM], hdb[FELIS_MAX_HARM], hu2[FELIS_MAX_HARM];
  int hstep[FELIS_MAX_HARM];
};

// a / b and sqrt(a) in FP32 from the hardware reciprocal and square root
// (1 ulp), where the IEEE sequences would cost a dozen instructions each
__device__ inline float fdiv_fast(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline float fsqrt_fast(float a) { return __builtin_amdgcn_sqrtf(a); }

// 1/x in FP64 for a positive normal x: the hardware reciprocal and two
// Newton steps (each squares the relative error), half the instructions of
// the IEEE division sequence
__device__ inline double rcp_f64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// the Bessel table (FP32 copy) at order `order`, linear in the argument
__device__ inline float bessel_interp_f(const float *__restrict__ tab, int bo, int idx, int order, float w0, float w1) {
  return tab[idx * bo + order] * w0 + tab[(idx + 1) * bo + order] * w1;
}

// word 3 of a raw buffer resource on gfx9-family GPUs (untyped, range checked)
constexpr int kBufW3 = 0x00020000;

// a buffer resource over `bytes` bytes from p: loads take a 32-bit lane
// offset plus a uniform one, with no 64-bit address arithmetic per lane
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, 0, (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL), kBufW3);
}
__device__ inline cpx<float> buf_c(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline float buf_f(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// e^{i H_h psi} for every harmonic from e^{i psi}, by multiplying along the
// ascending harmonic list (steps of 1 and 2 use e and e^2 directly)
template <int NH>
__device__ inline void harm_powers(cpx<float> e1, const int *hstep, cpx<float> (&out)[NH]) {
  const cpx<float> e2 = e1 * e1;
  cpx<float> cur = mk(1.f, 0.f);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int d = hstep[h];
    cpx<float> m;
    if (d == 1) m = e1;
    else if (d == 2) m = e2;
    else {
      m = mk(1.f, 0.f);
      cpx<float> b = e1;
      for (int n = d; n; n >>= 1) {
        if (n & 1) m = m * b;
        b = b * b;
      }
    }
    cur = h == 0 && d == 1 ? e1 : cur * m;
    out[h] = cur;
  }
}

template <typename T>
__device__ inline cpx<T> cpow_int(cpx<T> b, int n) {
  cpx<T> r = mk<T>(1, 0);
  while (n) {
    if (n & 1) r = r * b;
    b = b * b;
    n >>= 1;
  }
  return r;
}

// e^{i psi} in FP32 with the reduction in FP64: psi = q pi/2 + r with
// |r| <= pi/4, then the Taylor series of sin r (to r^9) and cos r (to r^10),
// whose truncation errors there are below 2e-9, and the quadrant's swap.
__host__ __device__ inline cpx<float> expi_f32(double psi) {
  const double q = rint(psi * 0.6366197723675814);
  const float r = (float)fma(-q, 1.5707963267948966, psi);
  const float r2 = r * r;
  const float sr = r + r * r2 * (-1.f / 6 + r2 * (1.f / 120 + r2 * (-1.f / 5040 + r2 * (1.f / 362880))));
  const float cr = 1.f + r2 * (-0.5f + r2 * (1.f / 24 + r2 * (-1.f / 720 + r2 * (1.f / 40320 + r2 * (-1.f / 3628800)))));
  const int qi = (int)(long long)q & 3;
  const float s = (qi & 1) ? cr : sr, c = (qi & 1) ? sr : cr;
  return mk((qi == 1 || qi == 2) ? -c : c, (qi & 2) ? -s : s);
}

// e^{i d} for a small angle in FP32: the Taylor series to d^5 (sin) and d^6
// (cos), whose truncation errors for |d| <= 0.25 are below 1.2e-8
__device__ inline cpx<float> expi_small(float d) {
  const float d2 = d * d;
  const float sn = d * (1.f + d2 * (-1.f / 6 + d2 * (1.f / 120)));
  const float cs = 1.f + d2 * (-0.5f + d2 * (1.f / 24 + d2 * (-1.f / 720)));
  return mk(cs, sn);
}

// transversePush of one particle in FP32 (particlePush.cpp:76-111)
__host__ inline TransArgs trans_args(const LatArgs &l) {
  TransArgs t;
  const float ku = (float)l.ku, K0 = (float)l.K0, ue = (float)l.ue;
  t.aa = K0 * K0 * 0.25f * ue * ku;
  t.bb = K0 * K0 * 0.25f * (1.0f + ue * ue) * ku * ku;
  t.cc = (float)(kECharge * l.g / kEMass / kC);
  t.ddc = ku * K0 * K0 * 0.25f;
  t.ue = ue;
  t.ka = (int)l.K0_alpha;
  t.xCoef = 0.f;
  t.yCoef = 0.f;
  if (l.und_type == 1) {
    if (t.ka == 1) t.xCoef = 1.f;
    else if (t.ka == -1) t.yCoef = 1.f;
  } else {
    t.xCoef = 1.f;
    t.yCoef = 1.f;
  }
  return t;
}

// times: 1, or 2 for the same half step applied twice (the map's
// coefficients are formed once)
__device__ inline void transverse_f32(const TransArgs &t, float dz, float gam, float &x, float &y, float &px,
                                      float &py, int times = 1) {
  const float aa = t.aa, bb = t.bb, cc = t.cc;
  const int ka = t.ka;
  const float ig = __builtin_amdgcn_rcpf(gam), ig2 = ig * ig;
  const float dd = t.ddc * ig2 * t.ue;
  const float fx = bb * ig * t.xCoef + cc, fy = bb * ig * t.yCoef - cc;
  if (aa == 0.f && dd == 0.f) {
    // ue K0 = 0: the planes decouple into u' = ig p, p' = -f u. For a linear
    // system RK4 is the Taylor polynomial of exp(dz A) to fourth order, and
    // here A^2 = -w I (w = ig f), so the step is
    //   u <- c u + s ig p,  p <- c p - s f u,
    //   c = 1 - dz^2 w/2 + dz^4 w^2/24,  s = dz (1 - dz^2 w/6).
    auto plane = [&](float &u, float &p, float f) {
      const float hw = dz * dz * (ig * f);
      const float c = fmaf(hw, fmaf(hw, 1.0f / 24.0f, -0.5f), 1.0f);
      const float s = dz * fmaf(hw, -1.0f / 6.0f, 1.0f);
      const float si = s * ig, sf = s * f;
      for (int k = 0; k < times; ++k) {
        const float u0 = u;
        u = fmaf(c, u0, si * p);
        p = fmaf(c, p, -sf * u0);
      }
    };
    plane(x, px, fx);
    plane(y, py, fy);
    return;
  }
  const float cm[5] = {0, 0, 0.5f, 0.5f, 1};
  const float wt[5] = {0, 1, 2, 2, 1};
  for (int rep = 0; rep < times; ++rep) {
    float kx = 0, ky = 0, kpx = 0, kpy = 0, sx = 0, sy = 0, spx = 0, spy = 0;
#pragma unroll
    for (int m = 1; m < 5; ++m) {
      float xx = x + dz * kx * cm[m], yy = y + dz * ky * cm[m];
      float ppx = px + dz * kpx * cm[m], ppy = py + dz * kpy * cm[m];
      kx = ppx * ig - aa * (ka * xx + yy) * ig2;
      ky = ppy * ig + aa * (ka * xx - yy) * ig2;
      kpx = ka * (ppx - ppy) * dd - fx * xx;
      kpy = ka * (ppx + ppy) * dd - fy * yy;
      sx += wt[m] * kx;
      sy += wt[m] * ky;
      spx += wt[m] * kpx;
      spy += wt[m] * kpy;
    }
    const float dz6 = dz * (1.0f / 6.0f);
    x += dz6 * sx;
    y += dz6 * sy;
    px += dz6 * spx;
    py += dz6 * spy;
  }
}

// Em of the longitudinal space-charge modes (the first L <= ML) for an
// in-grid particle at radius^2 r2 in slot (particlePush.cpp:283-304): the two
// radial points weighted by wr, summed unweighted on the axis. Zero and false
// when idxR >= nr, where the reference keeps the Em of the particle it pushed
// before.
// the radial bin idxR = (int)(r/dr + 0.5) of radius^2 r2, in FP32; rr = r/dr
__device__ inline int radial_bin_f32(float r2, float inv_dr, float &rr) {
  rr = fsqrt_fast(r2) * inv_dr;
  return (int)(rr + 0.5f);
}

template <int ML>
__device__ inline bool em_gather_f32(const GridArgs &g, float inv_dr, int L, const cpx<double> *__restrict__ Ez,
                                     int slot, float r2, cpx<float> (&Em)[ML]) {
  const int F = g.F, nr = g.nr;
  const long long es = (long long)g.S * F * nr;
  float rr;
  const int idxR = radial_bin_f32(r2, inv_dr, rr);
#pragma unroll
  for (int l = 0; l < ML; ++l) Em[l] = mk(0.f, 0.f);
  if (!(idxR >= 0 && idxR < nr)) return false;
  const float wr1 = idxR > 0 ? fdiv_fast(fmaf(rr, rr, -(float)(idxR * idxR)), 2.f * idxR) : 1.f;
  const float wr0 = idxR > 0 ? 1.f - wr1 : 1.f;
#pragma unroll
  for (int l = 0; l < ML; ++l) {
    if (l >= L) break;
    for (int f = 0; f < F; ++f) {
      const cpx<double> *p = Ez + l * es + (long long)slot * nr * F + nr * f + idxR;
      Em[l] = Em[l] + cvt<float>(p[0]) * wr0 + cvt<float>(p[1]) * wr1;
    }
  }
  return true;
}

// push_theta_gamma_3D of one particle (particlePush.cpp:260-379): the field
// coupling in FP32, the slippage rate ku - ks/(2 gamma0^2) K2 in FP64 with the
// RK stages' changes of it in FP32, gamma and theta in FP64. One sincos per particle: the RK stages rotate it by their
// small phase increments; harmonics by complex powers. On return eth_out
// holds e^{i theta} of the new theta (when the particle was pushed: true).
// (fx_, fy_): the position in cells, (idxI, idxJ) its cell, in_grid: the cell
// is inside the field grid
template <int NH, bool UY, bool LM>
__device__ inline bool push_tg_f32(const GridArgs &g, const PStepArgs &a, const LatArgs &l, int slot, double wakeE,
                                   const cpx<float> *__restrict__ Ux, const cpx<float> *__restrict__ Uy,
                                   const cpx<double> *__restrict__ Ez, const float *__restrict__ bes, float x0,
                                   float y0, double fx_, double fy_, int idxI, int idxJ, bool in_grid, float px,
                                   float py, double &th, double &gam, cpx<float> &eth_out,
                                   const cpx<float> *__restrict__ emc, long long n, unsigned *err) {
  const double ku = l.ku, ks = g.ks;
  const float ka = a.kaf;
  const double dz = g.dz;
  if (!in_grid) return false;
  const float wx1 = (float)(fx_ - idxI), wx0 = 1.f - wx1;
  const float wy1 = (float)(fy_ - idxJ), wy0 = 1.f - wy1;
  const float r2 = x0 * x0 + y0 * y0, pr2 = px * px + py * py;
  const double th0 = th, gam0 = gam;
  const double K2 = 1.0 + (double)pr2 + a.K2a * (1.0 + a.K2b * (double)r2);
  const double ig0 = rcp_f64(gam0);
  const float xi = (float)(a.xic * ig0 * ig0);
  // longitudinal space charge, modes l = 0 .. L-1 (particlePush.cpp:287-304);
  // on the axis the two radial points are summed unweighted. LM: more than
  // one mode (otherwise the first only, in fewer registers)
  // an in-grid particle beyond the radial grid takes the Em the reference
  // carries over from the particle it pushed before (emc, filled by
  // k_em_fill; null: refused)
  constexpr int ML = LM ? kMaxModes : 1;
  cpx<float> Em[ML];
  const int L = LM ? g.L : 1;
  if (!em_gather_f32<ML>(g, a.inv_drf, L, Ez, slot, r2, Em)) {
    if (emc) {
#pragma unroll
      for (int l = 0; l < ML; ++l)
        if (l < L) Em[l] = emc[n * L + l];
    } else {
      atomicOr(err, EB_EMCARRY);
    }
  }
  // this slot's field (the same for the whole tile) and the particle's cell
  const long long sbase = (long long)phys_slot(slot, g.ring_base, g.S) * g.N;
  const int cell = idxJ * g.nx + idxI;
  bool any_even = false;
#pragma unroll
  for (int h = 0; h < NH; ++h) any_even |= (g.harm[h] % 2 == 0);
  float xi2 = 0.f;
  cpx<float> expP = mk(1.f, 0.f);
  if (any_even) {
    const float uef = (float)l.ue;
    xi2 = a.xi2c *
          sqrtf((1 + ka) * (px * px + uef * uef * py * py) + (1 - ka) * (px * px * uef * uef + py * py));
    const float B = atanf(2.0f * ka * uef * (px - py) /
                          ((1.0f + ka) * (px + uef * uef * py) + (1 - ka) * (px * uef * uef + py)));
    expP = expi_f32((double)B - a.halfPhi);
  }
  const cpx<float> etaX = a.etaX, etaY = a.etaY;
  cpx<float> cu[NH];
  float sumU2 = 0.f;
  // the four corners of every harmonic's cell through buffer resources: one
  // lane offset for all, the row step and the harmonic blocks uniform
  const long long hblk = g.S * g.N;  // cells per harmonic block
  const long long sbytes = g.N * (long long)sizeof(cpx<float>);
  const int voff = cell * (int)sizeof(cpx<float>);
  const int rowb = g.nx * (int)sizeof(cpx<float>);
  const __amdgpu_buffer_rsrc_t rbes = buf_rsrc(bes, (long long)g.bn * g.bo * (long long)sizeof(float));
  const int bob = g.bo * (int)sizeof(float);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(Ux + h * hblk + sbase, sbytes);
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(Uy + h * hblk + sbase, sbytes);
    const cpx<float> ux = buf_c(ra, voff, 0) * (wx0 * wy0) + buf_c(ra, voff, rowb) * (wx0 * wy1) +
                          buf_c(ra, voff + 8, 0) * (wx1 * wy0) + buf_c(ra, voff + 8, rowb) * (wx1 * wy1);
    const cpx<float> uy = UY ? buf_c(rb, voff, 0) * (wx0 * wy0) + buf_c(rb, voff, rowb) * (wx0 * wy1) +
                                   buf_c(rb, voff + 8, 0) * (wx1 * wy0) + buf_c(rb, voff + 8, rowb) * (wx1 * wy1)
                             : mk(0.f, 0.f);
    const int H = g.harm[h];
    sumU2 += (norm2(ux) + norm2(uy)) * a.hu2[h];
    const float xb = xi * a.hdb[h];
    const int idx = (int)xb;
    const float w1 = xb - idx, w0 = 1.f - w1;
    cpx<float> fx = mk(0.f, 0.f), fy = mk(0.f, 0.f);
    // table rows idx and idx + 1 at an order: lane offset idx * bo, uniform order
    const int boff = idx * bob;
    auto J = [&](int order) {
      return buf_f(rbes, boff, order * (int)sizeof(float)) * w0 + buf_f(rbes, boff, order * (int)sizeof(float) + bob) * w1;
    };
    if (idx < 0 || idx + 1 >= g.bn) {
      atomicOr(err, EB_BESSEL);
    } else if (H % 2 == 1) {
      const float sign = ((H - 1) / 2 % 2 == 0) ? 1.f : -1.f;
      const float J1 = J((H - 1) / 2);
      const float J2 = J((H + 1) / 2);
      fx = mk(sign * (ka * J1 - J2), 0.f);
      fy = mk(sign * (ka * J1 + J2), 0.f);
    } else {
      const float sign = (H / 2 % 2 == 0) ? 1.f : -1.f;
      const float J1 = J((H - 2) / 2);
      const float J2 = J(H / 2);
      const float J3 = J((H + 2) / 2);
      const float s = sign * xi2 * H * 0.5f;
      const cpx<float> expM = conjg(expP);
      fx = (expP * (ka * J1 - J2) + expM * (ka * J2 - J3)) * s;
      fy = (expP * (ka * J1 + J2) + expM * (ka * J2 + J3)) * s;
    }
    cu[h] = UY ? etaX * fx * ux - etaY * fy * uy : etaX * fx * ux;
  }
  // the coupling's imaginary part enters theta over H: cuH = cu / H, once
  cpx<float> cuH[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) cuH[h] = cu[h] * a.hinv[h];
  const float sq = a.sq;
  const float ks_f = (float)ks;
  const float e_mc2 = (float)(kECharge / kEMass / kC / kC);
  const double halfPhi = a.halfPhi;
  const cpx<float> rotm = a.rotm;  // e^{-i Phi/2}
  const float cmv[5] = {0, 0, 0.5f, 0.5f, 1};
  const float wt[5] = {0, 1, 2, 2, 1};
  // The slippage rate ku - ks K2 / (2 gamma^2) is a difference of nearly
  // equal terms and is formed in FP64 once, at gamma0: rate0 = ku - Ak with
  // Ak = ks K2 / (2 gamma0^2). At a stage's gamma = gamma0 (1 + eps) the rate
  // is rate0 + Ak (1 - r^2), r = 1 / (1 + eps), and Ak (1 - r^2) =
  // Ak eps (2 + eps) r^2 is a small term that FP32 carries to its own
  // relative precision, like the field terms. The stages' small terms and
  // gamma increments are summed in FP32; theta and gamma are updated in FP64.
  // e^{i(theta0 + Phi/2)} once; stage m rotates it by e^{i del}, del = dz kth
  // c_m (increments above 0.25 rad take the full evaluation).
  const double Ak = ks * 0.5 * ig0 * ig0 * K2;
  const double rate0 = ku - Ak;
  const float Akf = (float)Ak, rate0f = (float)rate0, ig0f = (float)ig0, dzf = (float)dz;
  const float c20 = (float)(ks * 0.5 * ig0 * ig0);
  const cpx<float> e0 = expi_f32(th0 + halfPhi);
  // the first mode's Em times e^{-i Phi/2}: 2 Re(Em e^{i theta}) = 2 Re(Em' e1)
  const cpx<float> Em0r = Em[0] * rotm;
  float kth = 0.f, kg = 0.f, sth = 0.f, sg = 0.f;
#pragma unroll
  for (int m = 1; m < 5; ++m) {
    const float del = dzf * kth * cmv[m];
    cpx<float> e1 = e0;
    float r = 1.f, eps = 0.f;
    if (m > 1) {
      e1 = fabsf(del) <= 0.25f ? e0 * expi_small(del) : expi_f32(th0 + (double)del + halfPhi);
      eps = dzf * kg * cmv[m] * ig0f;
      r = __builtin_amdgcn_rcpf(1.f + eps);
    }
    const float igm = 0.5f * ig0f * r;
    cpx<float> eh[NH];
    harm_powers<NH>(e1, a.hstep, eh);
    // sumTh = -sum Im(e_h cu_h) / H, sumG = igm sum Re(e_h cu_h), by FMAs
    float sumTh = 0.f, sumG0 = 0.f;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      sumTh = fmaf(-eh[h].re, cuH[h].im, fmaf(-eh[h].im, cuH[h].re, sumTh));
      sumG0 = fmaf(eh[h].re, cu[h].re, fmaf(-eh[h].im, cu[h].im, sumG0));
    }
    const float sumG = sumG0 * igm;
    float sumEz = 2.0f * (Em0r.re * e1.re - Em0r.im * e1.im);
    if (ML > 1) {
      const cpx<float> eth = e1 * rotm;  // e^{i theta}
      cpx<float> ethl = eth;             // e^{i (l+1) theta}
#pragma unroll
      for (int l = 1; l < ML; ++l) {
        if (l >= L) break;
        ethl = ethl * eth;
        sumEz += 2.0f * (Em[l].re * ethl.re - Em[l].im * ethl.im);
      }
    }
    const float r2 = r * r;
    // kth - rate0
    const float small = Akf * eps * (2.f + eps) * r2 - (c20 * r2) * (sumU2 + sq * sumTh);
    kth = rate0f + small;
    kg = ks_f * sq * sumG + e_mc2 * sumEz;
    sth = fmaf(wt[m], small, sth);
    sg = fmaf(wt[m], kg, sg);
  }
  const double dz6 = dz * (1.0 / 6.0);
  const double dth = fma(dz, rate0, dz6 * (double)sth);
  if (dth > g.dPhi || dth < -g.dPhi) {
    atomicOr(err, EB_GUARD);
    return false;
  }
  th = th0 + dth;
  gam = gam0 + (dz6 * (double)sg - dz * wakeE);
  eth_out = fabs(dth) <= 0.25 ? (e0 * rotm) * expi_small((float)dth) : expi_f32(th);
  return true;
}

// lane l of each 16-lane row reads lane l - D (lanes without a source read 0)
template <int D>
__device__ inline float row_shr_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + D, 0xF, 0xF, true));
}
// the same shift, lanes without a source reading their own value (for
// minima and maxima)
template <int D>
__device__ inline int row_shr_keep(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x110 + D, 0xF, 0xF, false);
}

// the step's two transversePush half steps at the particle's gamma
__device__ inline void transverse_step(const GridArgs &g, const PStepArgs &a, double gm, float &xx, float &yy,
                                       float &ppx, float &ppy) {
  const float dzh = (float)(g.dz * 0.5);
  const float gf = (float)gm;
  if (a.trans_same) {
    transverse_f32(a.t0, dzh, gf, xx, yy, ppx, ppy, 2);
    return;
  }
  transverse_f32(a.t0, dzh, gf, xx, yy, ppx, ppy);
  transverse_f32(a.t1, dzh, gf, xx, yy, ppx, ppy);
}

// Em carry of the batched step (felis_emcarry.h): the positions are those
// k_pstep reaches after its transverse half steps, Em is formed as k_pstep
// forms it
template <bool LM>
struct EmCarryMixed {
  using Pos = float2;
  using E = cpx<float>;
  GridArgs g;
  PStepArgs a;
  const float *x, *y, *px, *py, *gamma;
  const cpx<double> *Ez;
  __device__ int where(long long n, Pos &q) const {
    float xx = x[n], yy = y[n], ppx = px[n], ppy = py[n];
    transverse_step(g, a, g.gref + (double)gamma[n], xx, yy, ppx, ppy);
    q = make_float2(xx, yy);
    const double fx_ = ((double)xx - g.minX) * a.inv_dx, fy_ = ((double)yy - g.minY) * a.inv_dy;
    const int idxI = (int)fx_, idxJ = (int)fy_;
    if (!(idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1)) return 0;
    float rr;
    const int idxR = radial_bin_f32(xx * xx + yy * yy, a.inv_drf, rr);
    return idxR >= 0 && idxR < g.nr ? 1 : 2;
  }
  __device__ int modes() const { return LM ? g.L : 1; }
  __device__ void em(int seg, const Pos &q, E *out) const {
    constexpr int ML = LM ? kMaxModes : 1;
    cpx<float> Em[ML];
    em_gather_f32<ML>(g, a.inv_drf, modes(), Ez, seg % g.slice_n + 1, q.x * q.x + q.y * q.y, Em);
    for (int l = 0; l < modes(); ++l) out[l] = Em[l];
  }
};

// One workgroup of kTile threads per tile, one particle per thread. UY =
// false: Uy is exactly zero (see k_deposit_slot) and is not read. LM: more
// than one longitudinal space-charge mode.
// With tile_bb the tile's box of in-grid cells (min i, max i, min j, max j of
// the pushed positions) goes to tile_bb[tile], from which k_deposit_slot takes
// its window without reading the particles' positions for it.
template <int NH, bool UY, bool LM>
__global__ void __launch_bounds__(kTile)
    k_pstep(GridArgs g0, PStepArgs a, Tiles tl, double *__restrict__ theta, float *__restrict__ gamma,
            float *__restrict__ x, float *__restrict__ y, float *__restrict__ px, float *__restrict__ py,
            const double *__restrict__ seg_w, const cpx<float> *__restrict__ Ux, const cpx<float> *__restrict__ Uy,
            const cpx<double> *__restrict__ Ez, const float *__restrict__ bes, const double *__restrict__ wakeE,
            const cpx<float> *__restrict__ emc, double *__restrict__ tile_red, int4 *__restrict__ tile_bb,
            unsigned *err) {
  // per-particle moments (FP32): e^{iH theta} re, im per harmonic and the six
  // twiss products; their row sums, one set per 16-lane row (the tile's
  // particle count comes from the tile table)
  constexpr int NF = 2 * NH + 6;
  __shared__ float rows[kTile / 16][NF];
  __shared__ int rbb[kTile / 16][4];
  __shared__ int scell[kTile];  // each particle's cell (j << 16 | i), -1 off the grid
  GridArgs g = g0;
  g.ring_base = ring_of(g0);
  // workgroups go to the 8 XCDs in turn: each XCD takes a contiguous run of
  // tiles instead, so that the tiles of a slice, which gather from one field
  // window, share one XCD's L2 (the push's HBM reads 0.64 -> 0.40 GB on the
  // 2048-slice step)
  const int nt = gridDim.x, xq = nt >> 3, xr = nt & 7, xcd = blockIdx.x & 7;
  const int t = xcd * xq + min(xcd, xr) + (blockIdx.x >> 3), k = threadIdx.x;
  const int seg = tl.seg[t];
  const int sl = seg % g.slice_n, slot = sl + 1;
  const bool on = k < tl.cnt[t];
  const double w = seg_w[seg];
  float acc[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) acc[q] = 0.f;
  if (on) {
    const long long n = tl.start[t] + k;
    const double wk = g.wake_on ? wakeE[sl] / (kMc2 * 1.0e6) : 0.0;
    double th = theta[n];
    double gm = g.gref + (double)gamma[n];
    float xx = x[n], yy = y[n], ppx = px[n], ppy = py[n];
    transverse_step(g, a, gm, xx, yy, ppx, ppy);
    // the pushed position's cell (the deposit's index: truncation toward zero)
    const double fx_ = ((double)xx - g.minX) * a.inv_dx, fy_ = ((double)yy - g.minY) * a.inv_dy;
    const int idxI = (int)fx_, idxJ = (int)fy_;
    const bool in_grid = idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1;
    // kept in LDS for the tile's box, which is formed after the push
    if (tile_bb) scell[k] = in_grid ? (idxJ << 16) | idxI : -1;
    cpx<float> eth = mk(1.f, 0.f);
    bool have_eth = false;
    if (!a.drift) {
      have_eth = push_tg_f32<NH, UY, LM>(g, a, a.l1, slot, wk, Ux, Uy, Ez, bes, xx, yy, fx_, fy_, idxI, idxJ,
                                         in_grid, ppx, ppy, th, gm, eth, emc, n, err);
    } else {
      if (in_grid) {
        const float wx1 = (float)(fx_ - idxI), wx0 = 1.f - wx1;
        const float wy1 = (float)(fy_ - idxJ), wy0 = 1.f - wy1;
        const long long base =
            (long long)phys_slot(slot, g.ring_base, g.S) * g.N + (long long)idxJ * g.nx + idxI;
        float sumU2 = 0.f;
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const cpx<float> *a0 = Ux + (long long)h * g.S * g.N + base;
          const cpx<float> *b0 = Uy + (long long)h * g.S * g.N + base;
          const cpx<float> ux =
              a0[0] * (wx0 * wy0) + a0[g.nx] * (wx0 * wy1) + a0[1] * (wx1 * wy0) + a0[g.nx + 1] * (wx1 * wy1);
          const cpx<float> uy = UY ? b0[0] * (wx0 * wy0) + b0[g.nx] * (wx0 * wy1) + b0[1] * (wx1 * wy0) +
                                         b0[g.nx + 1] * (wx1 * wy1)
                                   : mk(0.f, 0.f);
          sumU2 += (norm2(ux) + norm2(uy)) * a.hu2[h];
        }
        const double K2 = 1.0 + (double)(ppx * ppx + ppy * ppy);
        const double tt = a.l1.ku - g.ks / (2.0 * a.gamR * a.gamR) * (K2 + (double)sumU2);
        th += tt * g.dz;
        th -= g.dz * wk;
      }
    }
    // theta wrap (rearrangeParticles.cpp:84-108)
    const double dTh = g.dPhi;
    double del = 0.0;
    if (th >= dTh) del = dTh * (int)(th / dTh);
    else if (th < 0) del = dTh * (int)(th / dTh - 1.0);
    th -= del;
    if (th != th) atomicOr(err, EB_NAN);
    else if (th > dTh || th < 0) atomicOr(err, EB_RANGE);
    theta[n] = th;
    gamma[n] = (float)(gm - g.gref);
    x[n] = xx;
    y[n] = yy;
    px[n] = ppx;
    py[n] = ppy;
    // moments of the next step's bFactor and twiss (the wrap moves theta
    // by whole periods of e^{i theta})
    cpx<float> eh[NH];
    harm_powers<NH>(have_eth ? eth : expi_f32(th), a.hstep, eh);
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      acc[2 * h] = eh[h].re;
      acc[2 * h + 1] = eh[h].im;
    }
    const float inv = __builtin_amdgcn_rcpf((float)gm);
    const float xp = ppx * inv, yp = ppy * inv;
    acc[2 * NH] = xx * xx;
    acc[2 * NH + 1] = yy * yy;
    acc[2 * NH + 2] = xp * xp;
    acc[2 * NH + 3] = yp * yp;
    acc[2 * NH + 4] = xx * xp;
    acc[2 * NH + 5] = yy * yp;
  }
  // tile sums in a fixed order: each 16-lane row adds its particles by a
  // DPP tree in FP32 (16 terms), the 16 row sums are added in FP64; the
  // bunching sums take the segment weight, the tile's weight is w times its
  // particle count
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    float v = acc[q];
    v += row_shr_f<1>(v);
    v += row_shr_f<2>(v);
    v += row_shr_f<4>(v);
    v += row_shr_f<8>(v);
    if ((k & 15) == 15) rows[k >> 4][q] = v;
  }
  if (tile_bb) {
    // the cell box by the same 16-lane trees (minima and maxima); the cell is
    // read back from LDS (volatile: not held in a register through the push)
    int bi0 = 1 << 30, bi1 = -1, bj0 = 1 << 30, bj1 = -1;
    const int c = on ? reinterpret_cast<volatile int *>(scell)[k] : -1;
    if (c >= 0) {
      bi0 = bi1 = c & 0xffff;
      bj0 = bj1 = c >> 16;
    }
    bi0 = min(bi0, row_shr_keep<1>(bi0)); bi1 = max(bi1, row_shr_keep<1>(bi1));
    bj0 = min(bj0, row_shr_keep<1>(bj0)); bj1 = max(bj1, row_shr_keep<1>(bj1));
    bi0 = min(bi0, row_shr_keep<2>(bi0)); bi1 = max(bi1, row_shr_keep<2>(bi1));
    bj0 = min(bj0, row_shr_keep<2>(bj0)); bj1 = max(bj1, row_shr_keep<2>(bj1));
    bi0 = min(bi0, row_shr_keep<4>(bi0)); bi1 = max(bi1, row_shr_keep<4>(bi1));
    bj0 = min(bj0, row_shr_keep<4>(bj0)); bj1 = max(bj1, row_shr_keep<4>(bj1));
    bi0 = min(bi0, row_shr_keep<8>(bi0)); bi1 = max(bi1, row_shr_keep<8>(bi1));
    bj0 = min(bj0, row_shr_keep<8>(bj0)); bj1 = max(bj1, row_shr_keep<8>(bj1));
    if ((k & 15) == 15) {
      rbb[k >> 4][0] = bi0;
      rbb[k >> 4][1] = bi1;
      rbb[k >> 4][2] = bj0;
      rbb[k >> 4][3] = bj1;
    }
  }
  __syncthreads();
  if (tile_bb && k == 0) {
    int4 b = make_int4(rbb[0][0], rbb[0][1], rbb[0][2], rbb[0][3]);
    for (int r = 1; r < kTile / 16; ++r) {
      b.x = min(b.x, rbb[r][0]);
      b.y = max(b.y, rbb[r][1]);
      b.z = min(b.z, rbb[r][2]);
      b.w = max(b.w, rbb[r][3]);
    }
    tile_bb[t] = b;
  }
  const long long ntl = gridDim.x;  // tiles: rows of tile_red are this long
  for (int q = k; q <= NF; q += kTile) {
    double v = 0.0;
    if (q < NF) {
#pragma unroll
      for (int r = 0; r < kTile / 16; ++r) v += (double)rows[r][q];
    } else {
      v = (double)tl.cnt[t];
    }
    // output order: bunching re, im per harmonic, weight, six twiss sums, count
    if (q < 2 * NH) {
      tile_red[(long long)q * ntl + t] = v * w;
    } else {
      tile_red[(long long)(q + 1) * ntl + t] = v;
      if (q == NF) tile_red[(long long)(2 * NH) * ntl + t] = v * w;
    }
  }
}

// One workgroup per slice: every species' particles of the slot deposit into
// an LDS window over their bounding box (harmonics hp at a time), which is then
// stored to the interior of Sc with plain stores: those cells are written by
// this workgroup only, and every interior cell of the window is rewritten, so
// Sc needs no clearing between steps (the field solve reads the window only).
// Window cells on the boundary ring and the reference's flat-index spill
// (A.5) use atomics; ring cells are never read by the batched field solve.
// A window larger than the LDS budget for its harmonics goes through LDS in
// strips of rows, one harmonic per pass. The host keeps the budget at one
// node row of the widest window or more, so the branch below for a window of
// which not even a row fits (clear in Sc, then global atomics) is a guard
// that the loop does not take.
// NPOL = 1 deposits the x polarisation only: the y source is exactly zero
// (eta_Y = 0: a planar undulator with K0_alpha = 1) and Uy is zero.
// DET (felis_set_deposit_mode FELIS_DEPOSIT_FIXED): the window and the radial
// row accumulate in 64-bit fixed point instead of float and double atomics.
// Integer addition is exact and commutative, so the result no longer depends
// on the order in which particles reach a cell: the deposit is bitwise the
// same from run to run and for every thread count and LDS budget. The
// scales are powers of two from a bound on the slice's total deposit (the
// particle count times the largest single contribution, computed from the
// slice's largest 1/gamma and xi2 / gamma), which keeps every sum below 2^61;
// the resolution is 2^-61 of that bound, far below the FP32 rounding of the
// source and the FP64 rounding of the radial row.
template <bool DET>
__device__ inline void lds_add(void *base, int k, float v, double sc) {
  if (DET) atomicAdd(reinterpret_cast<unsigned long long *>(base) + k, (unsigned long long)__double2ll_rn((double)v * sc));
  else atomicAdd(reinterpret_cast<float *>(base) + k, v);
}
template <bool DET>
__device__ inline void lds_add(cpx<double> *base, int k, double v, double sc) {
  if (DET) atomicAdd(reinterpret_cast<unsigned long long *>(base) + k, (unsigned long long)__double2ll_rn(v * sc));
  else atomicAdd(reinterpret_cast<double *>(base) + k, v);
}
// power-of-two scale with bound * scale < 2^61 (1 for a zero bound)
__device__ inline double fixed_scale(double bound) {
  return bound > 0.0 ? ldexp(1.0, 60 - ilogb(bound)) : 1.0;
}

template <int NH, int NPOL, bool LM, bool DET>
__global__ void __launch_bounds__(1024)
    k_deposit_slot(GridArgs g, DepArgs d, int n_species, const long long *__restrict__ seg_off,
                   const double *__restrict__ seg_w, const double *__restrict__ theta,
                   const float *__restrict__ gamma, const float *__restrict__ x, const float *__restrict__ y,
                   const float *__restrict__ px, const float *__restrict__ py, const double *__restrict__ bes,
                   cpx<float> *__restrict__ Scx, cpx<float> *__restrict__ Scy, cpx<double> *__restrict__ ScR,
                   int4 *__restrict__ win_out, int lds_bytes, const int4 *__restrict__ tile_bb,
                   const int *__restrict__ seg_tile, int sl0, unsigned *err) {
  extern __shared__ double lds_d[];
  __shared__ int bb[4];
  const int sl = blockIdx.x + sl0, slot = sl + 1, tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) {
    bb[0] = 1 << 30; bb[1] = -1; bb[2] = 1 << 30; bb[3] = -1;
  }
  __syncthreads();
  int mi = 1 << 30, Mi = -1, mj = 1 << 30, Mj = -1;
  if (tile_bb) {
    // the box of in-grid cells from the boxes the push left per tile
    for (int sp = 0; sp < n_species; ++sp) {
      const int seg = sp * g.slice_n + sl;
      for (int tt = seg_tile[seg] + tid; tt < seg_tile[seg + 1]; tt += nt) {
        const int4 q = tile_bb[tt];
        mi = min(mi, q.x); Mi = max(Mi, q.y); mj = min(mj, q.z); Mj = max(Mj, q.w);
      }
    }
  } else {
    for (int sp = 0; sp < n_species; ++sp) {
      const int seg = sp * g.slice_n + sl;
      const long long b = seg_off[seg], e = seg_off[seg + 1];
      // the next iteration's loads in flight while this one is binned
      float xq = b + tid < e ? x[b + tid] : 0.f, yq = b + tid < e ? y[b + tid] : 0.f;
      for (long long n = b + tid; n < e; n += nt) {
        const float xn = xq, yn = yq;
        if (n + nt < e) {
          xq = x[n + nt];
          yq = y[n + nt];
        }
        // the same cell index as the deposit below
        const int idxI = (int)(((double)xn - g.minX) * d.inv_dx), idxJ = (int)(((double)yn - g.minY) * d.inv_dy);
        if (idxI >= 0 && idxI < g.nx - 1 && idxJ >= 0 && idxJ < g.ny - 1) {
          mi = min(mi, idxI); Mi = max(Mi, idxI); mj = min(mj, idxJ); Mj = max(Mj, idxJ);
        }
      }
    }
  }
  // wave-wide min/max first: one LDS atomic per wave instead of per thread
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mi = min(mi, __shfl_xor(mi, o));
    Mi = max(Mi, __shfl_xor(Mi, o));
    mj = min(mj, __shfl_xor(mj, o));
    Mj = max(Mj, __shfl_xor(Mj, o));
  }
  if ((tid & 63) == 0) {
    atomicMin(&bb[0], mi); atomicMax(&bb[1], Mi); atomicMin(&bb[2], mj); atomicMax(&bb[3], Mj);
  }
  __syncthreads();
  bool any_even = false;
#pragma unroll
  for (int h = 0; h < NH; ++h) any_even |= (g.harm[h] % 2 == 0);
  const float ka = (float)d.ka;
  // DET: the slice's largest |macro_u| / gamma and |macro_u| xi2 / gamma (as
  // the deposit below forms them) and its particle count
  __shared__ unsigned dmax[2];
  long long cnt_all = 0;
  if (DET) {
    if (tid < 2) dmax[tid] = 0u;
    __syncthreads();
    float am = 0.f, xm = 0.f;
    for (int sp = 0; sp < n_species; ++sp) {
      const int seg = sp * g.slice_n + sl;
      const long long b = seg_off[seg], e = seg_off[seg + 1];
      cnt_all += e - b;
      const float macro_u = fabsf((float)(seg_w[seg] * d.coef_U));
      for (long long n = b + tid; n < e; n += nt) {
        const float a0 = fdiv_fast(macro_u, (float)(g.gref + (double)gamma[n]));
        am = fmaxf(am, a0);
        if (any_even) {
          const float ppx = px[n], ppy = py[n], ue = (float)d.ue;
          const float xi2 = (float)d.xi2c * sqrtf((1 + ka) * (ppx * ppx + ue * ue * ppy * ppy) +
                                                  (1 - ka) * (ppx * ppx * ue * ue + ppy * ppy));
          xm = fmaxf(xm, a0 * xi2);
        }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      am = fmaxf(am, __shfl_xor(am, o));
      xm = fmaxf(xm, __shfl_xor(xm, o));
    }
    // non-negative floats order as their bit patterns
    if ((tid & 63) == 0) {
      atomicMax(&dmax[0], __float_as_uint(am));
      atomicMax(&dmax[1], __float_as_uint(xm));
    }
  }
  const int wi0 = bb[0], wj0 = bb[2];
  const bool empty = bb[1] < 0;
  const int ww = empty ? 0 : bb[1] + 2 - wi0, wh = empty ? 0 : bb[3] + 2 - wj0;
  const int W = ww * wh;
  constexpr int CB = DET ? 2 * (int)sizeof(long long) : (int)sizeof(cpx<float>);  // bytes per window cell
  int hp = W > 0 ? lds_bytes / (W * NPOL * CB) : NH;
  if (hp > NH) hp = NH;
  const int step = hp > 0 ? hp : 1;
  // A window larger than the budget for one harmonic goes through LDS in
  // strips of node rows that fit, one harmonic per pass over the particles;
  // only when not even one row fits are the window's interior cells cleared
  // in Sc and added to with global atomics.
  const int rows_fit = ww > 0 ? lds_bytes / (ww * NPOL * CB) : 0;
  const bool in_lds = hp > 0 || rows_fit > 0;
  const int sh = hp == 0 && rows_fit > 0 ? rows_fit : max(wh, 1);  // node rows per strip
  // the only interior Sc cells this deposit can make nonzero
  if (tid == 0) win_out[slot] = empty ? make_int4(1, 0, 1, 0) : make_int4(wi0, wi0 + ww - 1, wj0, wj0 + wh - 1);
  // per harmonic, from the lattice only: odd H keeps eta (J1 -+ ka J2) H as
  // two complex factors; even H keeps J1, J2, J3 (solveField.cpp:270-300)
  __shared__ float hfac[FELIS_MAX_HARM][4];
  __shared__ int hbad;
  if (tid == 0) hbad = 0;
  __syncthreads();
  if (tid < NH) {
    const int H = g.harm[tid];
    const int ib = (int)(H * d.xi * d.inv_dB);
    const double w1 = H * d.xi * d.inv_dB - ib, w0 = 1.0 - w1;
    if (ib < 0 || ib + 1 >= g.bn) {
      hbad = 1;
      hfac[tid][0] = hfac[tid][1] = hfac[tid][2] = hfac[tid][3] = 0.f;
    } else if (H % 2 == 1) {
      const double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
      const double J1 = bessel_interp(bes, g.bo, ib, (H - 1) / 2, w0, w1);
      const double J2 = bessel_interp(bes, g.bo, ib, (H + 1) / 2, w0, w1);
      const double fx = sign * (J1 - d.ka * J2) * H, fy = sign * (J1 + d.ka * J2) * H;
      hfac[tid][0] = (float)(d.etaXr * fx);
      hfac[tid][1] = (float)(d.etaXi * fx);
      hfac[tid][2] = (float)(d.etaYr * fy);
      hfac[tid][3] = (float)(d.etaYi * fy);
    } else {
      hfac[tid][0] = (float)bessel_interp(bes, g.bo, ib, (H - 2) / 2, w0, w1);
      hfac[tid][1] = (float)bessel_interp(bes, g.bo, ib, H / 2, w0, w1);
      hfac[tid][2] = (float)bessel_interp(bes, g.bo, ib, (H + 2) / 2, w0, w1);
      hfac[tid][3] = 0.f;
    }
  }
  __syncthreads();
  if (hbad) {
    if (tid == 0) atomicOr(err, EB_BESSEL);
    return;
  }
  // DET: fixed-point scales of the window (from the largest contribution of
  // any harmonic: |eta f_H| |macro_u| / gamma, bilinear weights <= 1) and of
  // the radial row (|coef_Ez macro| (l+1), radial weights below 1.5)
  __shared__ double dsc[2];
  if (DET && tid == 0) {
    const double am = __uint_as_float(dmax[0]), xm = __uint_as_float(dmax[1]);
    const double eta = fmax(hypot(d.etaXr, d.etaXi), hypot(d.etaYr, d.etaYi)), kb = fabs(d.ka);
    double fmaxc = 0.0;
    for (int h = 0; h < NH; ++h) {
      const int H = g.harm[h];
      const double f = H % 2 == 1 ? am * fmax(hypot((double)hfac[h][0], (double)hfac[h][1]),
                                              hypot((double)hfac[h][2], (double)hfac[h][3]))
                                  : xm * eta * 0.5 * H * H *
                                        (fabs(hfac[h][0]) + (1.0 + kb) * fabs(hfac[h][1]) + kb * fabs(hfac[h][2]));
      fmaxc = fmax(fmaxc, f);
    }
    dsc[0] = fixed_scale(2.0 * (double)cnt_all * fmaxc);
    const int segl = (n_species - 1) * g.slice_n + sl;
    const double cl = (double)(seg_off[segl + 1] - seg_off[segl]);
    dsc[1] = fixed_scale(4.0 * cl * fabs(d.coef_Ez * seg_w[segl]) * (LM ? g.L : 1));
  }
  __syncthreads();
  const double fsc = DET ? dsc[0] : 1.0, rsc = DET ? dsc[1] : 1.0;
  const cpx<float> etaX = mk((float)d.etaXr, (float)d.etaXi), etaY = mk((float)d.etaYr, (float)d.etaYi);
  cpx<float> *win = reinterpret_cast<cpx<float> *>(lds_d);
  if (!in_lds && W > 0) {
    // the window's interior cells start from zero for the atomics below
    for (int q = tid; q < NH * NPOL * W; q += nt) {
      const int hh = q / (NPOL * W), pol = (q / W) % NPOL, r = q % W;
      const int jy = wj0 + r / ww, ix = wi0 + r % ww;
      if (jy > 0 && jy < g.ny - 1 && ix > 0 && ix < g.nx - 1)
        ((pol ? Scy : Scx) + (long long)hh * g.S * g.N + (long long)slot * g.N)[(long long)jy * g.nx + ix] =
            mk(0.f, 0.f);
    }
    __syncthreads();
  }
  for (int h0 = 0; h0 < NH; h0 += step) {
    const int hc = min(step, NH - h0);
    for (int sr0 = 0; sr0 < max(wh, 1); sr0 += sh) {
    // this strip's node rows sr0 .. sr0+srh-1 of the window (all of them
    // when it fits); cells outside the window take global atomics, once
    const int srh = min(sh, wh - sr0), Ws = ww * max(srh, 0);
    if (in_lds) {
      for (int q = tid; q < hc * NPOL * Ws; q += nt) {
        if (DET) reinterpret_cast<double2 *>(lds_d)[q] = make_double2(0.0, 0.0);
        else win[q] = mk(0.f, 0.f);
      }
      __syncthreads();
    }
    for (int sp = 0; sp < n_species; ++sp) {
      const int seg = sp * g.slice_n + sl;
      const long long b = seg_off[seg], e = seg_off[seg + 1];
      const float macro_u = (float)(seg_w[seg] * d.coef_U);
      // each thread takes d.run consecutive particles; consecutive particles
      // in the same cell (the phases of a beamlet) are summed in registers
      // and reach the window with one set of atomics
      for (long long r0 = b + (long long)tid * d.run; r0 < e; r0 += (long long)nt * d.run) {
      const long long r1 = min(r0 + (long long)d.run, e);
      int key = -1, kI = 0, kJ = 0;
      bool kin = false;
      cpx<float> acc[4][NH][2];
#define FELIS_DEPOSIT_FLUSH()                                                                                   \
  if (key >= 0) {                                                                                               \
    _Pragma("unroll") for (int c = 0; c < 4; ++c) {                                                             \
      const int ii = c >> 1, jj = c & 1;                                                                        \
      _Pragma("unroll") for (int hh = 0; hh < NH; ++hh) {                                                       \
        if (hh >= hc) break;                                                                                    \
        const cpx<float> vx = acc[c][hh][0], vy = acc[c][hh][1];                                                \
        const int row = kJ + jj - wj0 - sr0;                                                                    \
        if (kin && row >= 0 && row < srh) {                                                                     \
          const int q = row * ww + (kI + ii - wi0);                                                             \
          const int k0 = 2 * ((hh * NPOL + 0) * Ws + q);                                                        \
          lds_add<DET>((void *)lds_d, k0, vx.re, fsc);                                                          \
          lds_add<DET>((void *)lds_d, k0 + 1, vx.im, fsc);                                                      \
          if (NPOL == 2) {                                                                                      \
            const int k1 = 2 * ((hh * NPOL + 1) * Ws + q);                                                      \
            lds_add<DET>((void *)lds_d, k1, vy.re, fsc);                                                        \
            lds_add<DET>((void *)lds_d, k1 + 1, vy.im, fsc);                                                    \
          }                                                                                                     \
        } else if (!kin && sr0 == 0) {                                                                          \
          const long long o =                                                                                   \
              (long long)(h0 + hh) * g.S * g.N + (long long)slot * g.N + (long long)kJ * g.nx + kI + jj * g.nx + ii; \
          unsafeAtomicAdd(&Scx[o].re, vx.re);                                                                   \
          unsafeAtomicAdd(&Scx[o].im, vx.im);                                                                   \
          if (NPOL == 2) {                                                                                      \
            unsafeAtomicAdd(&Scy[o].re, vy.re);                                                                 \
            unsafeAtomicAdd(&Scy[o].im, vy.im);                                                                 \
          }                                                                                                     \
        }                                                                                                       \
      }                                                                                                         \
    }                                                                                                           \
  }
      // the next particle's loads are issued before this one is deposited, so
      // each thread has a particle in flight while it works (measured: the
      // deposit 9% faster; two ahead is slower, at 128 registers)
      float xq = x[r0], yq = y[r0], gq = gamma[r0];
      double thq = theta[r0];
      for (long long n = r0; n < r1; ++n) {
        const float xn = xq, yn = yq, gn = gq;
        const double thn = thq;
        if (n + 1 < r1) {
          xq = x[n + 1];
          yq = y[n + 1];
          gq = gamma[n + 1];
          thq = theta[n + 1];
        }
        const double fx_ = ((double)xn - g.minX) * d.inv_dx, fy_ = ((double)yn - g.minY) * d.inv_dy;
        const int idxI = (int)fx_, idxJ = (int)fy_;
        if (!(idxI >= 0 && idxI < g.nx && idxJ >= 0 && idxJ < g.ny)) continue;
        const int k2 = idxJ * g.nx + idxI;
        if (k2 != key) {
          FELIS_DEPOSIT_FLUSH();
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int hh = 0; hh < NH; ++hh) acc[c][hh][0] = acc[c][hh][1] = mk(0.f, 0.f);
          key = k2;
          kI = idxI;
          kJ = idxJ;
          kin = in_lds && idxI < g.nx - 1 && idxJ < g.ny - 1;
        }
        float wx[2], wy[2];
        wx[1] = (float)(fx_ - idxI); wx[0] = 1.f - wx[1];
        wy[1] = (float)(fy_ - idxJ); wy[0] = 1.f - wy[1];
        const double th = thn;
        const float amp0 = fdiv_fast(macro_u, (float)(g.gref + (double)gn));
        float xi2 = 0.f;
        cpx<float> expP = mk(1.f, 0.f);
        if (any_even) {
          const float ppx = px[n], ppy = py[n], ue = (float)d.ue;
          xi2 = (float)d.xi2c *
                sqrtf((1 + ka) * (ppx * ppx + ue * ue * ppy * ppy) + (1 - ka) * (ppx * ppx * ue * ue + ppy * ppy));
          const float B = atanf(2.0f * ka * ue * (ppx - ppy) /
                                ((1.0f + ka) * (ppx + ue * ue * ppy) + (1 - ka) * (ppx * ue * ue + ppy)));
          expP = expi_f32((double)B - d.Phi * 0.5);
        }
        const cpx<float> em1 = conjg(expi_f32(th + d.Phi * 0.5));  // e^{-i psi}
        const cpx<float> em2 = em1 * em1;
        // e^{-iH psi} by the same chain of products from the first harmonic
        // whichever harmonics this pass takes, so that a pass's values do not
        // depend on how the harmonics are split over passes
        cpx<float> pw = cpow_int(em1, g.harm[0]);
        for (int h = 1; h <= h0; ++h) {
          const int dH = g.harm[h] - g.harm[h - 1];
          pw = pw * (dH == 1 ? em1 : dH == 2 ? em2 : cpow_int(em1, dH));
        }
#pragma unroll
        for (int hh = 0; hh < NH; ++hh) {
          if (hh >= hc) break;
          const int h = h0 + hh, H = g.harm[h];
          if (hh > 0) {
            const int dH = H - g.harm[h - 1];
            pw = pw * (dH == 1 ? em1 : dH == 2 ? em2 : cpow_int(em1, dH));
          }
          const cpx<float> mc = pw * amp0;
          cpx<float> cxv, cyv;
          if (H % 2 == 1) {
            cxv = mk(hfac[h][0], hfac[h][1]) * mc;
            cyv = mk(hfac[h][2], hfac[h][3]) * mc;
          } else {
            const float J1 = hfac[h][0], J2 = hfac[h][1], J3 = hfac[h][2];
            const float s = ((H / 2 % 2 == 0) ? 1.f : -1.f) * xi2 * H * 0.5f * H;
            const cpx<float> expM = conjg(expP);
            const cpx<float> fx = (expM * (J1 - ka * J2) + expP * (J2 - ka * J3)) * s;
            const cpx<float> fy = (expM * (J1 + ka * J2) + expP * (J2 + ka * J3)) * s;
            cxv = etaX * fx * mc;
            cyv = etaY * fy * mc;
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float wgt = -(wx[c >> 1] * wy[c & 1]);
            acc[c][hh][0] = acc[c][hh][0] + cxv * wgt;
            if (NPOL == 2) acc[c][hh][1] = acc[c][hh][1] + cyv * wgt;
          }
        }
      }
      FELIS_DEPOSIT_FLUSH();
#undef FELIS_DEPOSIT_FLUSH
      }
    }
    if (in_lds) {
      __syncthreads();
      for (int q = tid; q < hc * NPOL * Ws; q += nt) {
        cpx<float> v;
        if (DET) {
          const long long *wi = reinterpret_cast<const long long *>(lds_d) + 2 * q;
          v = mk((float)((double)wi[0] / fsc), (float)((double)wi[1] / fsc));
        } else {
          v = win[q];
        }
        const int hh = q / (NPOL * Ws), pol = (q / Ws) % NPOL, r = q % Ws;
        const int jy = wj0 + sr0 + r / ww, ix = wi0 + r % ww;
        cpx<float> *dst = (pol ? Scy : Scx) + (long long)(h0 + hh) * g.S * g.N + (long long)slot * g.N +
                          (long long)jy * g.nx + ix;
        if (jy > 0 && jy < g.ny - 1 && ix > 0 && ix < g.nx - 1) {
          *dst = v;
        } else if (v.re != 0.f || v.im != 0.f) {
          unsafeAtomicAdd(&dst->re, v.re);
          unsafeAtomicAdd(&dst->im, v.im);
        }
      }
      __syncthreads();
    }
    }
  }
  // radial source of the last species, stored whole for this slot. Each wave
  // accumulates into its own copy of the row (the beam core fills a handful
  // of radial bins, which one shared copy would serialise); consecutive
  // particles at one radius (a beamlet) are summed in registers first. The
  // copies are added in a fixed order.
  // Modes l = 0 .. L-1 are rows l * row of the slot's source (L > 1 or F > 1
  // deposit particle by particle).
  const int F = g.F, nr = g.nr, L = LM ? g.L : 1, row = F * nr + 1, rows = L * row;
  const int nwv = nt >> 6, wid = tid >> 6;
  cpx<double> *hist = reinterpret_cast<cpx<double> *>(lds_d);
  const bool priv = (long long)nwv * rows * (long long)sizeof(cpx<double>) <= lds_bytes;
  cpx<double> *mine = priv ? hist + (long long)wid * rows : hist;
  for (int q = tid; q < (priv ? nwv : 1) * rows; q += nt) hist[q] = mk(0.0, 0.0);  // zero bits in DET too
  __syncthreads();
  {
    const int seg = (n_species - 1) * g.slice_n + sl;
    const long long b = seg_off[seg], e = seg_off[seg + 1];
    const double macro = seg_w[seg];
    const float inv_drf = (float)(1.0 / g.dr);
    for (long long r0 = b + (long long)tid * d.run; r0 < e; r0 += (long long)nt * d.run) {
      const long long r1 = min(r0 + (long long)d.run, e);
      int key = -1;
      double amp = 0.0;
      // a run's sums of i e^{-i theta} weighted by each particle's own radial
      // weights (w0 into sum[0], w1 into sum[2]), and the axis sum (sum[1])
      cpx<double> sum[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
#define FELIS_RADIAL_FLUSH()                                                                 \
  if (key > 0) {                                                                             \
    lds_add<DET>(mine, 2 * key, sum[0].re * amp, rsc);                                       \
    lds_add<DET>(mine, 2 * key + 1, sum[0].im * amp, rsc);                                   \
    lds_add<DET>(mine, 2 * key + 2, sum[2].re * amp, rsc);                                   \
    lds_add<DET>(mine, 2 * key + 3, sum[2].im * amp, rsc);                                   \
  } else if (key == 0) {                                                                     \
    lds_add<DET>(mine, 0, sum[1].re, rsc);                                                   \
    lds_add<DET>(mine, 1, sum[1].im, rsc);                                                   \
  }
      for (long long n = r0; n < r1; ++n) {
        const float xf = x[n], yf = y[n];
        // the radial bin and weights in FP32 (1e-7 of a bin), as the push
        // forms them; the sums in FP64
        float rrf;
        const int idxR = radial_bin_f32(xf * xf + yf * yf, inv_drf, rrf);
        const cpx<float> em = conjg(expi_f32(theta[n]));  // e^{-i theta}
        if (idxR > 0 && idxR < nr) {
          const double w1 = (double)fdiv_fast(fmaf(rrf, rrf, -(float)(idxR * idxR)), 2.f * idxR), w0 = 1.0 - w1;
          if (F > 1 || (LM && L > 1)) {
            // azimuthal modes need the angle of every particle, longitudinal
            // modes (l+1) theta: no combining. Mode l adds
            // i (l+1) coef_Ez macro / idxR e^{-i((l+1) theta + f phi)}
            // (solveField.cpp:340-347)
            FELIS_RADIAL_FLUSH();
            key = -1;
            const double ampl = d.coef_Ez * macro / (1.0 * idxR);
            const double xx = xf, yy = yf;
            const double ph = (xx == 0.0) ? 0.0 : atan2(yy, xx);
            for (int l = 0; l < L; ++l)
              for (int f = 0; f < F; ++f) {
                double sn, cs;
                sincos(-((l + 1.0) * theta[n] + f * ph), &sn, &cs);
                const double a = ampl * (l + 1.0);
                const int k0 = 2 * (l * row + f * nr + idxR);
                lds_add<DET>(mine, k0, -sn * a * w0, rsc);
                lds_add<DET>(mine, k0 + 1, cs * a * w0, rsc);
                lds_add<DET>(mine, k0 + 2, -sn * a * w1, rsc);
                lds_add<DET>(mine, k0 + 3, cs * a * w1, rsc);
              }
            continue;
          }
          if (idxR != key) {
            FELIS_RADIAL_FLUSH();
            key = idxR;
            sum[0] = sum[2] = mk(0.0, 0.0);
            // the bin's amplitude coef_Ez macro / idxR, the same for the run
            amp = d.coef_Ez * macro / (1.0 * idxR);
          }
          // i e^{-i theta} = (sin theta, cos theta) = (-em.im, em.re)
          const cpx<double> ie = mk(-(double)em.im, (double)em.re);
          sum[0] = sum[0] + ie * w0;
          sum[2] = sum[2] + ie * w1;
        } else if (idxR == 0) {
          if (LM && L > 1) {
            // the axis row of every mode, without 1/idxR (solveField.cpp:348-356)
            FELIS_RADIAL_FLUSH();
            key = -1;
            for (int l = 0; l < L; ++l) {
              double sn, cs;
              sincos(-(l + 1.0) * theta[n], &sn, &cs);
              const double a = d.coef_Ez * macro * (l + 1.0);
              lds_add<DET>(mine, 2 * l * row, -sn * a, rsc);
              lds_add<DET>(mine, 2 * l * row + 1, cs * a, rsc);
            }
            continue;
          }
          if (key != 0) {
            FELIS_RADIAL_FLUSH();
            key = 0;
            sum[1] = mk(0.0, 0.0);
          }
          sum[1] = sum[1] + mk(-(double)em.im, (double)em.re) * (d.coef_Ez * macro);
        }
      }
      FELIS_RADIAL_FLUSH();
#undef FELIS_RADIAL_FLUSH
    }
  }
  __syncthreads();
  for (int q = tid; q < rows; q += nt) {
    cpx<double> v;
    if (DET) {
      const long long *hi = reinterpret_cast<const long long *>(hist);
      long long re = hi[2 * q], im = hi[2 * q + 1];
      if (priv)
        for (int w = 1; w < nwv; ++w) {
          re += hi[2 * ((long long)w * rows + q)];
          im += hi[2 * ((long long)w * rows + q) + 1];
        }
      v = mk((double)re / rsc, (double)im / rsc);
    } else {
      v = hist[q];
      if (priv)
        for (int w = 1; w < nwv; ++w) v = v + hist[(long long)w * rows + q];
    }
    ScR[(long long)slot * rows + q] = v;
  }
}

// out[o] for o < 2*nh: energy of pol o/nh, harmonic o%nh over the slices that
// stay after the slippage; next 2*nh+1: bFactor sums; then 7 twiss sums per
// species. tile_red holds moment q of tile t at [q * ntiles + t]. In two
// passes with a fixed order: np (at most kRedParts) workgroups per output
// each sum a contiguous part of its row (strided partial sums, four
// independent chains per thread, then a fixed tree) into parts[o][p]; one
// thread per output then adds its parts in order (one part: the first pass
// writes the output).
constexpr int kRedParts = 16;

static __global__ void __launch_bounds__(256)
    k_reduce_parts(GridArgs g, const double *__restrict__ slice_e, int skip, const double *__restrict__ tile_red,
                   const int *__restrict__ species_tile, int n_species, double *__restrict__ parts,
                   double *__restrict__ out, const int *__restrict__ step_dev, int per) {
  __shared__ double sh[256];
  const int o = blockIdx.x, p = blockIdx.y, np = gridDim.y, tid = threadIdx.x, nh = g.nh;
  const long long ntiles = species_tile[n_species];
  const double *row;
  long long k0, k1;
  if (o < 2 * nh) {
    row = slice_e + (long long)o * g.slice_n;
    k0 = 0;
    k1 = g.slice_n - skip;  // slices shifted out since the energies were summed
  } else if (o < 4 * nh + 1) {
    row = tile_red + (long long)(o - 2 * nh) * ntiles;
    k0 = 0;
    k1 = ntiles;
  } else {
    const int q = o - (4 * nh + 1), sp = q / 7, c = q % 7;
    row = tile_red + (long long)(2 * nh + 1 + c) * ntiles;
    k0 = species_tile[sp];
    k1 = species_tile[sp + 1];
  }
  const long long len = k1 > k0 ? k1 - k0 : 0;
  const long long a = k0 + len * p / np, b = k0 + len * (p + 1) / np;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long long k = a + tid;
  for (; k + 768 < b; k += 1024) {
    s0 += row[k];
    s1 += row[k + 256];
    s2 += row[k + 512];
    s3 += row[k + 768];
  }
  for (; k < b; k += 256) s0 += row[k];
  sh[tid] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    if (np > 1) {
      parts[(long long)o * kRedParts + p] = sh[0];
    } else {  // one part: the output itself
      if (step_dev) out += (long long)(*step_dev) * per;  // a graph-captured step: the row of the current step
      out[o] = sh[0];
    }
  }
}

static __global__ void k_reduce_final(int nout, int np, const double *__restrict__ parts, double *__restrict__ out,
                                      const int *__restrict__ step_dev, int per) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  if (step_dev) out += (long long)(*step_dev) * per;  // a graph-captured step: the row of the current step
  double s = 0.0;
  for (int p = 0; p < np; ++p) s += parts[(long long)o * kRedParts + p];
  out[o] = s;
}

}  // namespace felis
