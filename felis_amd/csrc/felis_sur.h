// Device kernels of the learned jump correction (felis_sur.hip): per-slice
// moments of the beam and of the field at the particles (the network's
// inputs), and the residual of a jump against the fine steps from the same
// state (its training targets).
#pragma once

#include "felis_common.h"
#include "felis_jump.h"
#include "felis_pstep.h"

namespace felis {

// per-slice moment layout of k_slice_moments
//   0 sum w, 1 sum w dg, 2 sum w dg^2, 3-4 sum w dg e^{i theta},
//   then per harmonic h (8 each): sum w e^{iH theta}, sum w Ux_h e^{iH theta},
//   sum w Uy_h e^{iH theta} (re, im each), sum w |Ux_h|^2, sum w |Uy_h|^2
//   and last sum w Ux_1, sum w Uy_1 of the first harmonic (re, im each)
//   (dg = gamma - gamma_ref, U bilinear at the particle)
__host__ __device__ constexpr int moments_per_slice(int nh) { return 9 + 8 * nh; }
// network inputs per slice (k_slice_features): log10 of the slice current in
// kA, per harmonic log10 of the mean |U|^2 at the particles, log10 |b_h|,
// the coherent fraction |sum w U e^{iH theta}| / sqrt(sum w sum w |U|^2) and
// the cos, sin of its phase; then log10 of the rms relative energy spread,
// the energy modulation |sum w dg e^{i theta}| / (sum w rms dg) and the cos,
// sin of its phase against the bunching b_1
__host__ __device__ constexpr int features_per_slice(int nh) { return 1 + 5 * nh + 4; }
// per-slice layout of k_part_cmp: sums over the particles of w times
//   1, c, s, c^2, s^2, c s, dth, dth c, dth s, dgm, dgm c, dgm s
// with c, s = cos, sin of theta_a, dth = wrap(theta_b - theta_a), dgm = gamma_b - gamma_a
constexpr int kPartCmp = 12;

template <int NQ>
__device__ inline void block_sums_store(double (&acc)[NQ], double *out) {
  __shared__ double sh[4][NQ];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane == 0) sh[wv][q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NQ; q += blockDim.x) out[q] = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
}

// One 256-thread workgroup per slice over the particles of every species.
// UY = false: Uy is exactly zero (see k_deposit_slot), its sums are zero and
// it is not read. e^{iH theta} is a power of e^{i theta} in FP32.
template <int NH, bool UY>
static __global__ void __launch_bounds__(256)
    k_slice_moments(GridArgs g, int n_species, const long long *__restrict__ seg_off,
                    const double *__restrict__ seg_w, const double *__restrict__ theta,
                    const float *__restrict__ gamma, const float *__restrict__ x, const float *__restrict__ y,
                    const cpx<float> *__restrict__ Ux, const cpx<float> *__restrict__ Uy, double inv_dx,
                    double inv_dy, double *__restrict__ out) {
  constexpr int NQ = 9 + 8 * NH;
  const int sl = blockIdx.x, slot = sl + 1;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  const long long pbase = (long long)phys_slot(slot, g.ring_base, g.S) * g.N;
  for (int sp = 0; sp < n_species; ++sp) {
    const int seg = sp * g.slice_n + sl;
    const long long b = seg_off[seg], e = seg_off[seg + 1];
    const double w = seg_w[seg];
    // the next iteration's loads in flight while this one is summed
    const long long n0 = b + threadIdx.x;
    double thq = n0 < e ? theta[n0] : 0.0;
    float gq = n0 < e ? gamma[n0] : 0.f, xq = n0 < e ? x[n0] : 0.f, yq = n0 < e ? y[n0] : 0.f;
    for (long long n = n0; n < e; n += blockDim.x) {
      const double th = thq, dg = (double)gq;
      const float xn = xq, yn = yq;
      if (n + blockDim.x < e) {
        thq = theta[n + blockDim.x];
        gq = gamma[n + blockDim.x];
        xq = x[n + blockDim.x];
        yq = y[n + blockDim.x];
      }
      const cpx<float> e1 = expi_f32(th);
      acc[0] += w;
      acc[1] += w * dg;
      acc[2] += w * dg * dg;
      acc[3] += w * dg * e1.re;
      acc[4] += w * dg * e1.im;
      const double fx_ = ((double)xn - g.minX) * inv_dx, fy_ = ((double)yn - g.minY) * inv_dy;
      const int i = (int)fx_, j = (int)fy_;
      const bool in = i >= 0 && i < g.nx - 1 && j >= 0 && j < g.ny - 1;
      const float wx1 = in ? (float)(fx_ - i) : 0.f, wy1 = in ? (float)(fy_ - j) : 0.f;
      const float wx0 = 1.f - wx1, wy0 = 1.f - wy1;
      const long long c0 = in ? pbase + (long long)j * g.nx + i : 0;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const cpx<float> eh = g.harm[h] == 1 ? e1 : cpow_int(e1, g.harm[h]);
        acc[5 + 8 * h] += w * eh.re;
        acc[6 + 8 * h] += w * eh.im;
        if (in) {
          const cpx<float> *a0 = Ux + (long long)h * g.S * g.N + c0;
          const cpx<float> *b0 = Uy + (long long)h * g.S * g.N + c0;
          const cpx<float> ux =
              a0[0] * (wx0 * wy0) + a0[g.nx] * (wx0 * wy1) + a0[1] * (wx1 * wy0) + a0[g.nx + 1] * (wx1 * wy1);
          const cpx<float> uy = UY ? b0[0] * (wx0 * wy0) + b0[g.nx] * (wx0 * wy1) + b0[1] * (wx1 * wy0) +
                                         b0[g.nx + 1] * (wx1 * wy1)
                                   : mk(0.f, 0.f);
          const cpx<float> px = ux * eh, py = uy * eh;
          acc[7 + 8 * h] += w * px.re;
          acc[8 + 8 * h] += w * px.im;
          acc[9 + 8 * h] += w * py.re;
          acc[10 + 8 * h] += w * py.im;
          acc[11 + 8 * h] += w * (double)norm2(ux);
          acc[12 + 8 * h] += w * (double)norm2(uy);
          if (h == 0) {
            acc[5 + 8 * NH] += w * ux.re;
            acc[6 + 8 * NH] += w * ux.im;
            acc[7 + 8 * NH] += w * uy.re;
            acc[8 + 8 * NH] += w * uy.im;
          }
        }
      }
    }
  }
  block_sums_store<NQ>(acc, out + (long long)sl * NQ);
}

// Network inputs from the moments, one thread per slice (layout of
// features_per_slice). amp_kA: electrons per slice -> kA (e c / slice length).
static __global__ void k_slice_features(int slice_n, int nh, double amp_kA, double gref,
                                 const double *__restrict__ M, float *__restrict__ X) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= slice_n) return;
  const int NQ = moments_per_slice(nh), F = features_per_slice(nh);
  const double *m = M + (long long)sl * NQ;
  float *x = X + (long long)sl * F;
  const double w = m[0], iw = w > 0 ? 1.0 / w : 0.0;
  const double tiny = 1e-30;
  x[0] = (float)log10(w * amp_kA + 1e-9);
  for (int h = 0; h < nh; ++h) {
    const double *q = m + 5 + 8 * h;
    const double I = (q[6] + q[7]) * iw;
    const double br = q[0] * iw, bi = q[1] * iw;
    const double fr = q[2] + q[4], fi = q[3] + q[5];
    const double fa = sqrt(fr * fr + fi * fi);
    x[1 + 5 * h] = (float)log10(I + tiny);
    x[2 + 5 * h] = (float)log10(sqrt(br * br + bi * bi) + 1e-12);
    x[3 + 5 * h] = (float)(fa / sqrt(w * (q[6] + q[7]) + tiny));
    x[4 + 5 * h] = fa > 0 ? (float)(fr / fa) : 1.f;
    x[5 + 5 * h] = fa > 0 ? (float)(fi / fa) : 0.f;
  }
  const double mu = m[1] * iw, var = fmax(m[2] * iw - mu * mu, 0.0), rms = sqrt(var);
  const double pr = m[3] - mu * m[5], pi = m[4] - mu * m[6];  // about the mean energy
  const double pa = sqrt(pr * pr + pi * pi);
  // relative phase of the energy modulation against the bunching b_1
  const double br = m[5], bi = m[6], ba = sqrt(br * br + bi * bi);
  const double cr = pr * br + pi * bi, ci = pi * br - pr * bi, ca = pa * ba;
  const int o = 1 + 5 * nh;
  x[o] = (float)log10(rms / gref + 1e-12);
  x[o + 1] = (float)(pa / (w * rms + tiny));
  x[o + 2] = ca > 0 ? (float)(cr / ca) : 1.f;
  x[o + 3] = ca > 0 ? (float)(ci / ca) : 0.f;
}

// network outputs per slice: (log|e_h|, arg e_h) for each harmonic, the
// factor on the jump's emission
__host__ __device__ constexpr int outputs_per_slice(int nh) { return 2 * nh; }

// The network's outputs into the emission factor of every (pol x harmonic,
// slice) system; log|e| is held to [-0.5, 0.5] and arg e to [-1, 1] rad.
static __global__ void k_decode_corr(int slice_n, int nh, const float *__restrict__ Y,
                                     cpx<float> *__restrict__ gain) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= slice_n) return;
  const float *y = Y + (long long)sl * outputs_per_slice(nh);
  for (int h = 0; h < nh; ++h) {
    const float la = fminf(fmaxf(y[2 * h], -0.5f), 0.5f);
    const float ph = fminf(fmaxf(y[2 * h + 1], -1.f), 1.f);
    const float a = expf(la);
    const cpx<float> g = mk(a * cosf(ph), a * sinf(ph));
    gain[(long long)h * slice_n + sl] = g;
    gain[(long long)(nh + h) * slice_n + sl] = g;
  }
}

// Per (pol x harmonic, slice) system: sum conj(a) b, sum |a|^2, sum |b|^2
// over the slot's cells; a and b share the ring mapping.
static __global__ void __launch_bounds__(256) k_field_cmp(GridArgs g, const cpx<float> *__restrict__ A,
                                                   const cpx<float> *__restrict__ B, double *__restrict__ out) {
  const int sys = blockIdx.x, sl = sys % g.slice_n, ph = sys / g.slice_n;
  const long long off = ((long long)ph * g.S + phys_slot(sl + 1, g.ring_base, g.S)) * g.N;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long k = threadIdx.x; k < g.N; k += blockDim.x) {
    const cpx<float> a = A[off + k], b = B[off + k];
    acc[0] += (double)a.re * b.re + (double)a.im * b.im;
    acc[1] += (double)a.re * b.im - (double)a.im * b.re;
    acc[2] += (double)a.re * a.re + (double)a.im * a.im;
    acc[3] += (double)b.re * b.re + (double)b.im * b.im;
  }
  block_sums_store<4>(acc, out + (long long)sys * 4);
}

// Per (pol x harmonic, slice) system after a jump of J periods: the jump's
// emission E = 2 Sw[s - J] / H of the field now at slot s = slice + 1 (Sw is
// indexed by the slot the field was solved at, and holds it only inside
// that slot's window win2) against the difference of state b and state a
// there: sum conj(E) (b - a) (re, im) and sum |E|^2.
static __global__ void __launch_bounds__(256) k_emit_cmp(GridArgs g, int J, const cpx<float> *__restrict__ A,
                                                         const cpx<float> *__restrict__ B,
                                                         const cpx<float> *__restrict__ Sw,
                                                         const int4 *__restrict__ win2, double *__restrict__ out) {
  const int sys = blockIdx.x, sl = sys % g.slice_n, ph = sys / g.slice_n, h = ph % g.nh;
  const int s = sl + 1, s0 = s - J;
  double acc[3] = {0.0, 0.0, 0.0};
  if (s0 >= 1) {
    const long long off = ((long long)ph * g.S + phys_slot(s, g.ring_base, g.S)) * g.N;
    const cpx<float> *src = Sw + ((long long)ph * g.S + s0) * g.N;
    const double k2 = 2.0 / g.harm[h];
    const int4 w = win2[s0];
    for (long long k = threadIdx.x; k < g.N; k += blockDim.x) {
      const int i = (int)(k % g.nx), j = (int)(k / g.nx);
      const cpx<float> a = A[off + k], b = B[off + k], e = in_box(w, i, j) ? src[k] : mk(0.f, 0.f);
      const double er = k2 * e.re, ei = k2 * e.im;
      const double dr = (double)b.re - a.re, di = (double)b.im - a.im;
      acc[0] += er * dr + ei * di;
      acc[1] += er * di - ei * dr;
      acc[2] += er * er + ei * ei;
    }
  }
  block_sums_store<3>(acc, out + (long long)sys * 3);
}

// Per slice: the least-squares sums of the particle residual (state b minus
// state a) on the basis 1, cos theta_a, sin theta_a.
static __global__ void __launch_bounds__(256)
    k_part_cmp(GridArgs g, int n_species, const long long *__restrict__ seg_off, const double *__restrict__ seg_w,
               const double *__restrict__ tha, const float *__restrict__ ga, const double *__restrict__ thb,
               const float *__restrict__ gb, double *__restrict__ out) {
  const int sl = blockIdx.x;
  double acc[kPartCmp];
#pragma unroll
  for (int q = 0; q < kPartCmp; ++q) acc[q] = 0.0;
  const double P = g.dPhi, hP = 0.5 * g.dPhi;
  for (int sp = 0; sp < n_species; ++sp) {
    const int seg = sp * g.slice_n + sl;
    const long long b = seg_off[seg], e = seg_off[seg + 1];
    const double w = seg_w[seg];
    for (long long n = b + threadIdx.x; n < e; n += blockDim.x) {
      const double ta = tha[n];
      double dth = thb[n] - ta;
      dth -= P * floor((dth + hP) / P);  // to [-P/2, P/2)
      const double dgm = (double)gb[n] - (double)ga[n];
      const cpx<float> e1 = expi_f32(ta);
      const double c = e1.re, s = e1.im;
      acc[0] += w;
      acc[1] += w * c;
      acc[2] += w * s;
      acc[3] += w * c * c;
      acc[4] += w * s * s;
      acc[5] += w * c * s;
      acc[6] += w * dth;
      acc[7] += w * dth * c;
      acc[8] += w * dth * s;
      acc[9] += w * dgm;
      acc[10] += w * dgm * c;
      acc[11] += w * dgm * s;
    }
  }
  block_sums_store<kPartCmp>(acc, out + (long long)sl * kPartCmp);
}

}  // namespace felis
