// Batched ADI of solveField.cpp:80-185 in FP32 for every (slice, harmonic,
// pol) system of a step, one workgroup each.
//
// Per system the two half steps are tridiagonal solves along rows (x) and
// columns (y) whose Thomas recurrences have field-independent coefficients:
//   forward   DD_i = r_i dd_i + a_i DD_{i-1}
//   backward  U_i  = q_i DD_i + p_i U_{i+1}
// Scaling DD' = q DD folds q into the forward map,
//   DD'_i = (q_i r_i) dd_i + (q_i a_i / q_{i-1}) DD'_{i-1},   U_i = DD'_i + p_i U_{i+1},
// so both sweeps are first-order affine recurrences y_i = c_i y_{i-1} + b_i
// whose multipliers c_i depend on the harmonic only. A line of n-2 unknowns is
// held by a 16-lane DPP row, E consecutive unknowns per lane. Each lane runs
// its segment from zero (the particular part), a 4-level Hillis-Steele scan
// over the lanes combines segment ends, and the homogeneous part is added
// back. The products of multipliers that the scan and the add-back need are
// the same for every line, slice and step, so they are tabulated on the host
// (adi_scan_tables) and each level of the scan is one complex FMA.
//
// The y-implicit right-hand side of the whole system lives in LDS between the
// sweeps, in a skewed layout whose row and column strides per lane segment are
// odd in 8-byte words, so the row-wise writes of the x sweep and the
// column-wise reads of the y sweep both spread over the banks.
#pragma once

#include <vector>

#include "felis_common.h"

namespace felis {

// unknowns per lane for a 16-lane line
__host__ __device__ inline int adi_scan_E(int nx, int ny) {
  const int n = (nx > ny ? nx : ny) - 2;
  return (n + 15) / 16;
}

// skewed LDS layout of `width` columns of the (ny-2)-row y right-hand side
struct DyLayout {
  int E, P, sc, sr;
  __host__ __device__ DyLayout(int width, int E_) : E(E_) {
    sc = (E % 2 == 0) ? 1 : 0;
    P = width + sc * ((width - 1) / E);
    sr = ((E * P) % 2 == 0) ? 1 : 0;
  }
  __host__ __device__ int at(int r, int c) const { return r * P + c + sc * (c / E) + sr * (r / E); }
  __host__ __device__ int size(int rows) const { return at(rows - 1, 0) + P; }
};

// Per harmonic table block (complex FP32), e-major entries [e*16 + lane]:
//   x sweep: ap, rp, pp, alphaP, alpha, Af, Gb
//   y sweep: ap, pp, Af, Gb
//   rows j:  betaP[ny], beta[ny], y-rp[ny]
//   scan:    Cf_x[4][16], Cb_x[4][16], Cf_y[4][16], Cb_y[4][16]
struct AdiTabOff {
  int xap, xrp, xpp, xaP, xaM, xAf, xGb, yap, ypp, yAf, yGb, rbP, rbM, ryrp, scan, block, staged;
  __host__ __device__ AdiTabOff(int E, int ny) {
    const int s = 16 * E;
    xap = 0; xrp = s; xpp = 2 * s; xaP = 3 * s; xaM = 4 * s; xAf = 5 * s; xGb = 6 * s;
    yap = 7 * s; ypp = 8 * s; yAf = 9 * s; yGb = 10 * s;
    rbP = 11 * s; rbM = rbP + ny; ryrp = rbM + ny;
    scan = ryrp + ny;
    block = scan + 4 * 4 * 16;
    staged = block;  // the whole block is staged in LDS
  }
};

// LDS bytes of a workgroup that holds `cols` interior columns of the system
inline size_t adi_scan_lds(int nx, int ny, int cols) {
  const int E = adi_scan_E(nx, ny);
  DyLayout L(cols, E);
  AdiTabOff o(E, ny);
  return ((size_t)L.size(ny - 2) + o.staged) * sizeof(cpx<float>);
}

// Interior columns whose y right-hand side stays in LDS: all of them when
// they fit, else the first 8 lanes' 8E columns (the rest goes through the
// system's scratch in HBM and is solved second); 0 when neither fits.
inline int adi_scan_cols(int nx, int ny, size_t lds_max) {
  const int E = adi_scan_E(nx, ny), nxi = nx - 2;
  if (E > 10) return 0;
  if (adi_scan_lds(nx, ny, nxi) + 256 <= lds_max + 1024) return nxi;
  if (8 * E < nxi && adi_scan_lds(nx, ny, 8 * E) + 256 <= lds_max + 1024) return 8 * E;
  return 0;
}

// Host: the tables above from the hoisted Thomas coefficients (r, a, q, p per
// index, as in the plain path) and the absorbing-layer stencil.
inline void adi_scan_tables(int nx, int ny, int nh, const std::vector<cpx<double>> &adix,
                            const std::vector<cpx<double>> &adiy, const std::vector<cpx<double>> &abc,
                            std::vector<cpx<float>> &out) {
  const int E = adi_scan_E(nx, ny);
  AdiTabOff o(E, ny);
  out.assign((size_t)nh * o.block, mk(0.f, 0.f));
  auto cm = [](cpx<double> a, cpx<double> b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); };
  auto cinv = [](cpx<double> a) {
    const double d = a.re * a.re + a.im * a.im;
    return mk(a.re / d, -a.im / d);
  };
  auto f32 = [](cpx<double> a) { return mk((float)a.re, (float)a.im); };
  for (int h = 0; h < nh; ++h) {
    cpx<float> *T = out.data() + (size_t)h * o.block;
    for (int dir = 0; dir < 2; ++dir) {
      const int n = dir ? ny : nx;
      const cpx<double> *c = (dir ? adiy.data() : adix.data()) + (size_t)h * 4 * n;
      const cpx<double> *r = c, *a = c + n, *q = c + 2 * n, *p = c + 3 * n;
      std::vector<cpx<double>> ap(16 * E), rp(16 * E), pp(16 * E), Af(16 * E), Gb(16 * E);
      std::vector<cpx<double>> At(16), Gt(16);
      for (int sub = 0; sub < 16; ++sub) {
        for (int e = 0; e < E; ++e) {
          const int i = 1 + sub * E + e, k = e * 16 + sub;
          const bool valid = i <= n - 2;
          ap[k] = valid && i > 1 ? cm(cm(q[i], a[i]), cinv(q[i - 1])) : mk(0.0, 0.0);
          rp[k] = valid ? cm(q[i], r[i]) : mk(0.0, 0.0);
          pp[k] = valid ? p[i] : mk(0.0, 0.0);
        }
        cpx<double> acc = mk(1.0, 0.0);
        for (int e = 0; e < E; ++e) {
          acc = cm(acc, ap[e * 16 + sub]);
          Af[e * 16 + sub] = acc;
        }
        At[sub] = acc;
        acc = mk(1.0, 0.0);
        for (int e = E - 1; e >= 0; --e) {
          acc = cm(acc, pp[e * 16 + sub]);
          Gb[e * 16 + sub] = acc;
        }
        Gt[sub] = acc;
      }
      cpx<float> *Cf = T + o.scan + (dir ? 2 : 0) * 64, *Cb = Cf + 64;
      for (int lv = 0; lv < 4; ++lv) {
        const int dd = 1 << lv;
        for (int sub = 0; sub < 16; ++sub) {
          cpx<double> f = mk(1.0, 0.0), b = mk(1.0, 0.0);
          for (int m = (sub - dd + 1 > 0 ? sub - dd + 1 : 0); m <= sub; ++m) f = cm(f, At[m]);
          for (int m = sub; m <= (sub + dd - 1 < 15 ? sub + dd - 1 : 15); ++m) b = cm(b, Gt[m]);
          Cf[lv * 16 + sub] = f32(f);
          Cb[lv * 16 + sub] = f32(b);
        }
      }
      for (int k = 0; k < 16 * E; ++k) {
        if (dir == 0) {
          T[o.xap + k] = f32(ap[k]);
          T[o.xrp + k] = f32(rp[k]);
          T[o.xpp + k] = f32(pp[k]);
          T[o.xAf + k] = f32(Af[k]);
          T[o.xGb + k] = f32(Gb[k]);
          const int i = 1 + (k % 16) * E + k / 16;
          if (i <= nx - 2) {
            T[o.xaP + k] = f32(abc[nx + i]);
            T[o.xaM + k] = f32(abc[i]);
          }
        } else {
          T[o.yap + k] = f32(ap[k]);
          T[o.ypp + k] = f32(pp[k]);
          T[o.yAf + k] = f32(Af[k]);
          T[o.yGb + k] = f32(Gb[k]);
        }
      }
      if (dir == 1)
        for (int j = 1; j <= ny - 2; ++j) T[o.ryrp + j] = f32(cm(q[j], r[j]));
    }
    for (int j = 0; j < ny; ++j) {
      T[o.rbP + j] = f32(abc[2 * nx + ny + j]);
      T[o.rbM + j] = f32(abc[2 * nx + j]);
    }
  }
}

#ifdef __HIPCC__

// lane l of each 16-lane row reads lane l-D (l+D); lanes without a source read 0
template <int D>
__device__ inline float dpp_shr(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + D, 0xF, 0xF, true));
}
template <int D>
__device__ inline float dpp_shl(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + D, 0xF, 0xF, true));
}

__device__ inline cpx<float> cfma(cpx<float> c, cpx<float> x, cpx<float> b) {  // c*x + b
  return mk(fmaf(c.re, x.re, fmaf(-c.im, x.im, b.re)), fmaf(c.re, x.im, fmaf(c.im, x.re, b.im)));
}

// forward scan of lane-segment ends: after it, B of lane l is the recurrence
// value at the end of lane l's segment. C points at this lane's constant of
// level 0; level lv is 16 entries further.
__device__ inline cpx<float> scan_fwd(cpx<float> B, const cpx<float> *C) {
  B = cfma(C[0], mk(dpp_shr<1>(B.re), dpp_shr<1>(B.im)), B);
  B = cfma(C[16], mk(dpp_shr<2>(B.re), dpp_shr<2>(B.im)), B);
  B = cfma(C[32], mk(dpp_shr<4>(B.re), dpp_shr<4>(B.im)), B);
  B = cfma(C[48], mk(dpp_shr<8>(B.re), dpp_shr<8>(B.im)), B);
  return mk(dpp_shr<1>(B.re), dpp_shr<1>(B.im));  // value entering the lane
}
__device__ inline cpx<float> scan_bwd(cpx<float> B, const cpx<float> *C) {
  B = cfma(C[0], mk(dpp_shl<1>(B.re), dpp_shl<1>(B.im)), B);
  B = cfma(C[16], mk(dpp_shl<2>(B.re), dpp_shl<2>(B.im)), B);
  B = cfma(C[32], mk(dpp_shl<4>(B.re), dpp_shl<4>(B.im)), B);
  B = cfma(C[48], mk(dpp_shl<8>(B.re), dpp_shl<8>(B.im)), B);
  return mk(dpp_shl<1>(B.re), dpp_shl<1>(B.im));
}

// word 3 of a raw buffer resource on gfx9-family GPUs (untyped, range checked:
// loads past num_records return zero)
constexpr int kBufWord3 = 0x00020000;

__device__ inline cpx<float> buf_load_c(__amdgpu_buffer_rsrc_t r, int byte_off) {
  return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}
template <int E>
__device__ inline void buf_load_row(__amdgpu_buffer_rsrc_t r, int byte_off, cpx<float> (&q)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) q[e] = buf_load_c(r, byte_off + e * (int)sizeof(cpx<float>));
}

// One workgroup (NT threads: 512 or 1024) per system sys = (pol*nh + h)*slice_n + sl.
// cw: interior columns whose y right-hand side the x sweep keeps in LDS (a
// multiple of E, or all); the remaining nxi - cw columns go to the system's
// part of `scratch` and are brought into LDS after the first cw are solved.
template <int E, int NT>
__global__ void __launch_bounds__(NT)
    k_adi_scan(GridArgs g, const cpx<float> *__restrict__ tab, int cf, cpx<float> *__restrict__ U,
               const cpx<float> *__restrict__ Sc, const int4 *__restrict__ win, double *__restrict__ slice_e, int cw,
               cpx<float> *__restrict__ scratch) {
  extern __shared__ double lds_d[];
  __shared__ double red[NT / 64];
  const int nx = g.nx, ny = g.ny, nxi = nx - 2, nyi = ny - 2;
  const DyLayout L(cw, E);
  const AdiTabOff o(E, ny);
  cpx<float> *DY = reinterpret_cast<cpx<float> *>(lds_d);
  cpx<float> *T = DY + L.size(nyi);
  const int sys = blockIdx.x;
  const int sl = sys % g.slice_n, ph = sys / g.slice_n, h = ph % g.nh, slot = sl + 1;
  const float H = (float)g.harm[h];
  // the ring position is the same for every lane (read from memory when a
  // captured step replays): keep it scalar
  const int rb = __builtin_amdgcn_readfirstlane(ring_of(g));
  cpx<float> *Un = U + ((long long)ph * g.S + phys_slot(slot, rb, g.S)) * g.N;
  const cpx<float> *Sn = Sc + ((long long)ph * g.S + slot) * g.N;
  const cpx<float> *Th = tab + (long long)h * o.block;
  const int w1 = nxi - cw;  // columns solved second, through scratch
  cpx<float> *Gs = scratch + (long long)sys * nyi * w1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int k = tid; k < o.staged; k += blockDim.x) T[k] = Th[k];
  const int grp = lane >> 4, sub = lane & 15;
  __syncthreads();
  const int i0 = 1 + sub * E;
  // The source is read only inside this slot's window: the deposit kernel
  // rewrote every interior cell of it, and cells outside it hold older
  // deposits. cf = 0: no source.
  const int4 wn = cf ? win[slot] : make_int4(1, 0, 1, 0);
  // ---- x-implicit half step ----
  // Each 16-lane group walks a band of consecutive rows, so a row's y
  // stencil needs one new row of U (the other two are the previous row's,
  // kept in registers) and the next row's loads are in flight while the
  // rest of this row is solved. Loads go through buffer resources over this system's
  // slot: elements past the row's last unknown read the next cells (their
  // multipliers in the tables are zero) and rows past the slot read zero.
  // The source counts only inside the deposit window.
  const int ngr = nw * 4, gi = wv * 4 + grp;
  const int R = (nyi + ngr - 1) / ngr;  // rows per band (at most kR)
  constexpr int kR = (16 * E + NT / 16 - 1) / (NT / 16);  // 16E >= ny - 2 rows over NT/16 groups
  const int jb0 = 1 + gi * R;
  const int slot_bytes = (int)(g.N * (long long)sizeof(cpx<float>));
  const __amdgpu_buffer_rsrc_t rU = __builtin_amdgcn_make_buffer_rsrc((void *)Un, 0, slot_bytes, kBufWord3);
  const __amdgpu_buffer_rsrc_t rS = __builtin_amdgcn_make_buffer_rsrc((void *)Sn, 0, slot_bytes, kBufWord3);
  const int elo = wn.x - i0, ehi = wn.y - i0;  // this lane's unknowns in the window's columns
  auto row_off = [&](int j) { return (j * nx + i0) * (int)sizeof(cpx<float>); };
  cpx<float> qm[E], q0[E], qp[E], qs[E];
  // outside the window the offset points past the slot: the load reads zero
  auto load_src = [&](int j, cpx<float>(&q)[E]) {
    const bool jw = j >= wn.z && j <= wn.w;
    const int o = row_off(j);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const bool in = jw && e >= elo && e <= ehi;
      q[e] = buf_load_c(rS, in ? o + e * (int)sizeof(cpx<float>) : slot_bytes);
    }
  };
  buf_load_row<E>(rU, row_off(jb0 - 1), qm);
  buf_load_row<E>(rU, row_off(jb0), q0);
  buf_load_row<E>(rU, row_off(jb0 + 1), qp);
  load_src(jb0, qs);
#pragma unroll
  for (int t = 0; t < kR; ++t) {
    const int j = jb0 + t;
    const bool rowok = t < R && j <= ny - 2;
    const int jj = rowok ? j : 1;
    const cpx<float> bP = T[o.rbP + jj], bM = T[o.rbM + jj], yrp = T[o.ryrp + jj];
    const cpx<float> diag = mk(H + bP.re + bM.re, bP.im + bM.im);
    cpx<float> y[E];
    cpx<float> B = mk(0.f, 0.f);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = e * 16 + sub;
      // x-implicit right-hand side: y-explicit stencil of U plus the source
      cpx<float> dd = diag * q0[e] - bP * qp[e] - bM * qm[e] + qs[e];
      B = cfma(T[o.xap + k], B, T[o.xrp + k] * dd);
      y[e] = B;
    }
    // the next row's loads, into the registers of the row above (dead now);
    // with 1024 threads the source follows at the end of the row, when this
    // row's source is dead (registers: 128 per lane at 16 waves per CU)
    cpx<float> qn[E], qsn[E];
    if (t + 1 < kR) {
      buf_load_row<E>(rU, row_off(j + 2), qn);
      if (NT < 1024) load_src(j + 1, qsn);
    }
    const cpx<float> din = scan_fwd(B, T + o.scan + sub);
    B = mk(0.f, 0.f);
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
      const int k = e * 16 + sub;
      const cpx<float> dp = cfma(T[o.xAf + k], din, y[e]);  // DD'
      B = cfma(T[o.xpp + k], B, dp);
      y[e] = B;
    }
    const cpx<float> uin = scan_bwd(B, T + o.scan + 64 + sub);
#pragma unroll
    for (int e = 0; e < E; ++e) y[e] = cfma(T[o.xGb + e * 16 + sub], uin, y[e]);  // Uc
    // y-implicit right-hand side: x-explicit stencil of Uc (zero on the ring),
    // times the y sweep's first multiplier for this row
    const cpx<float> left = mk(dpp_shr<1>(y[E - 1].re), dpp_shr<1>(y[E - 1].im));
    const cpx<float> right = mk(dpp_shl<1>(y[0].re), dpp_shl<1>(y[0].im));
    // row j-1, columns i0-1 .. i0-1+E-1: consecutive in the LDS layout (a
    // lane's columns are all below cw or all above) or in the scratch rows
    const bool in_lds = i0 - 1 < cw;
    cpx<float> *dst = in_lds ? DY + L.at(jj - 1, i0 - 1) : Gs + (long long)(jj - 1) * w1 + (i0 - 1 - cw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = i0 + e, k = e * 16 + sub;
      const cpx<float> um = e == 0 ? left : y[e - 1];
      const cpx<float> up = e == E - 1 ? right : y[e + 1];
      const cpx<float> v = y[e] * H + T[o.xaP + k] * (y[e] - up) + T[o.xaM + k] * (y[e] - um) + qs[e];
      if (rowok && i <= nx - 2) dst[e] = yrp * v;
    }
    if (t + 1 < kR) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        qm[e] = q0[e];
        q0[e] = qp[e];
        qp[e] = qn[e];
        if (NT < 1024) qs[e] = qsn[e];
      }
      if (NT >= 1024) load_src(j + 1, qs);
    }
  }
  __syncthreads();
  // the boundary ring is untouched by the ADI but counts in the energy: its
  // loads go out before the y sweep and are summed after it
  const int nring = 2 * nx + 2 * (ny - 2);
  constexpr int kRing = (2 * 162 + 2 * 160 + NT - 1) / NT;  // ring cells per thread, grids up to 162 wide
  cpx<float> rv[kRing];
#pragma unroll
  for (int q = 0; q < kRing; ++q) {
    const int k0 = tid + q * NT, k = k0 < nring ? k0 : 0;  // a valid cell for every lane
    int idx;
    if (k < nx) idx = k;
    else if (k < 2 * nx) idx = (ny - 1) * nx + (k - nx);
    else {
      const int r = k - 2 * nx;
      idx = (1 + (r >> 1)) * nx + ((r & 1) ? nx - 1 : 0);
    }
    const cpx<float> v = Un[idx];
    rv[q] = k0 < nring ? v : mk(0.f, 0.f);
  }
  // ---- y-implicit half step down the columns, stored straight to U with
  // the slice energy accumulated in FP64: the columns held in LDS, then
  // (when the system does not fit) the rest from scratch ----
  double es = 0.0;
  auto ysweep = [&](const DyLayout &Ly, int c0, int width) {
    for (int cb = 0; cb < width; cb += nw * 4) {
      const int col = cb + wv * 4 + grp;  // column within the range
      const bool colok = col < width;
      const int cc = colok ? col : width - 1;
      // rows sub*E .. sub*E+E-1 of column cc are P apart in the layout
      const cpx<float> *src = DY + Ly.at(sub * E, cc);
      cpx<float> y[E];
      cpx<float> B = mk(0.f, 0.f);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = sub * E + e;
        const cpx<float> b = r < nyi ? src[e * Ly.P] : mk(0.f, 0.f);
        B = cfma(T[o.yap + e * 16 + sub], B, b);
        y[e] = B;
      }
      const cpx<float> din = scan_fwd(B, T + o.scan + 128 + sub);
      B = mk(0.f, 0.f);
#pragma unroll
      for (int e = E - 1; e >= 0; --e) {
        const int k = e * 16 + sub;
        B = cfma(T[o.ypp + k], B, cfma(T[o.yAf + k], din, y[e]));
        y[e] = B;
      }
      const cpx<float> uin = scan_bwd(B, T + o.scan + 192 + sub);
      cpx<float> *out = Un + (1 + sub * E) * nx + c0 + cc + 1;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = sub * E + e;
        if (colok && r < nyi) {
          const cpx<float> v = cfma(T[o.yGb + e * 16 + sub], uin, y[e]);
          out[e * nx] = v;
          es += (double)v.re * v.re + (double)v.im * v.im;
        }
      }
    }
  };
  ysweep(L, 0, cw);
  if (w1 > 0) {
    // the scratch rows (written above by this workgroup) into LDS
    __syncthreads();
    const DyLayout L1(w1, E);
    for (int q = tid; q < nyi * w1; q += NT) {
      const int r = q / w1, c = q - r * w1;
      DY[L1.at(r, c)] = Gs[q];
    }
    __syncthreads();
    ysweep(L1, cw, w1);
  }
#pragma unroll
  for (int q = 0; q < kRing; ++q) es += (double)rv[q].re * rv[q].re + (double)rv[q].im * rv[q].im;
  for (int off = 32; off > 0; off >>= 1) es += __shfl_down(es, off);
  if (lane == 0) red[wv] = es;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += red[k];
    slice_e[sys] = s;
  }
}

#endif  // __HIPCC__

}  // namespace felis
