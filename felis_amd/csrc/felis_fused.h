// Fused gfx950 kernels of the batched step (felis_run).
//
//   k_adi_fused    solveField.cpp:80-185 for one (slice, harmonic, pol)
//                  system per workgroup: x-implicit sweep, y-explicit
//                  right-hand side kept in LDS, y-implicit sweep, slice
//                  energy (updateTotalEnergy.cpp:25-36) and Sc clearing.
#pragma once

#include "felis_common.h"
#include "felis_kernels.h"

namespace felis {

// ------------------------------------------------------------------ ADI

// x sweep of one row held by a 16-lane group: lane `sub` owns columns
// 1 + sub*E .. 1 + sub*E + E-1. The two first-order recurrences
// DD_i = r_i dd_i + a_i DD_{i-1} and U_i = q_i DD_i + p_i U_{i+1} are solved
// with a group-wide scan of affine maps.
template <typename T, int E>
__device__ inline void row_scan16(int sub, cpx<T> (&av)[E], cpx<T> (&bv)[E], const cpx<T> *cq, const cpx<T> *cp,
                                  bool (&valid)[E], int i0, cpx<T> (&out)[E]) {
  cpx<T> A = mk<T>(1, 0), B = mk<T>(0, 0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    B = av[e] * B + bv[e];
    A = av[e] * A;
  }
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    T Ar = __shfl_up(A.re, d, 16), Ai = __shfl_up(A.im, d, 16);
    T Br = __shfl_up(B.re, d, 16), Bi = __shfl_up(B.im, d, 16);
    if (sub >= d) {
      cpx<T> Ap = mk(Ar, Ai), Bp = mk(Br, Bi);
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  T pr = __shfl_up(B.re, 1, 16), pi = __shfl_up(B.im, 1, 16);
  cpx<T> DD = sub == 0 ? mk<T>(0, 0) : mk(pr, pi);
  cpx<T> ddv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    DD = av[e] * DD + bv[e];
    ddv[e] = DD;
  }
  A = mk<T>(1, 0);
  B = mk<T>(0, 0);
#pragma unroll
  for (int e = E - 1; e >= 0; --e) {
    cpx<T> q = valid[e] ? cq[i0 + e] : mk<T>(0, 0);
    cpx<T> p = valid[e] ? cp[i0 + e] : mk<T>(1, 0);
    av[e] = p;
    bv[e] = q * ddv[e];
    B = p * B + bv[e];
    A = p * A;
  }
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    T Ar = __shfl_down(A.re, d, 16), Ai = __shfl_down(A.im, d, 16);
    T Br = __shfl_down(B.re, d, 16), Bi = __shfl_down(B.im, d, 16);
    if (sub + d < 16) {
      cpx<T> Ap = mk(Ar, Ai), Bp = mk(Br, Bi);
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  T nr_ = __shfl_down(B.re, 1, 16), ni_ = __shfl_down(B.im, 1, 16);
  cpx<T> Un = sub == 15 ? mk<T>(0, 0) : mk(nr_, ni_);
#pragma unroll
  for (int e = E - 1; e >= 0; --e) {
    Un = av[e] * Un + bv[e];
    out[e] = Un;
  }
}

// One workgroup (512 threads) per system sys = (pol*nh + h)*slice_n + sl.
// LDS holds the (ny-2)x(nx-2) y right-hand side (overwritten by DD) and the
// system's recurrence and stencil coefficients.
template <typename T>
__host__ __device__ inline size_t adi_fused_lds(int nx, int ny) {
  return ((size_t)(nx - 2) * (ny - 2) + 6 * (size_t)nx + 6 * (size_t)ny) * sizeof(cpx<T>);
}

template <typename T, int E>
__global__ void __launch_bounds__(512)
    k_adi_fused(GridArgs g, AdiCoef c, int cf, cpx<T> *__restrict__ U, cpx<T> *__restrict__ Sc,
                const int4 *__restrict__ win, double *__restrict__ slice_e) {
  extern __shared__ double lds_d[];
  __shared__ double red[16];
  const int nx = g.nx, ny = g.ny, nxi = nx - 2, nyi = ny - 2;
  cpx<T> *DY = reinterpret_cast<cpx<T> *>(lds_d);
  cpx<T> *xr = DY + nxi * nyi, *xa = xr + nx, *xq = xa + nx, *xp = xq + nx;
  cpx<T> *yr = xp + nx, *ya = yr + ny, *yq = ya + ny, *yp = yq + ny;
  cpx<T> *al = yp + ny, *alP = al + nx, *be = alP + nx, *beP = be + ny;
  const int sys = blockIdx.x;
  const int sl = sys % g.slice_n, ph = sys / g.slice_n, h = ph % g.nh, slot = sl + 1;
  const T H = (T)g.harm[h];
  cpx<T> *Un = U + ((long long)ph * g.S + phys_slot(slot, g.ring_base, g.S)) * g.N;
  cpx<T> *Sn = Sc + ((long long)ph * g.S + slot) * g.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  {
    const cpx<double> *cx = c.x + (long long)h * 4 * nx;
    const cpx<double> *cy = c.y + (long long)h * 4 * ny;
    for (int k = tid; k < 4 * nx; k += blockDim.x) xr[k] = cvt<T>(cx[k]);
    for (int k = tid; k < 4 * ny; k += blockDim.x) yr[k] = cvt<T>(cy[k]);
    for (int k = tid; k < 2 * nx + 2 * ny; k += blockDim.x) al[k] = cvt<T>(c.abc[k]);
  }
  __syncthreads();
  const int grp = lane >> 4, sub = lane & 15;
  const int i0 = 1 + sub * E;
  // ---- x-implicit half step, rows in groups of 4 per wave ----
  // Sc is nonzero only inside this slot's deposit window (written by the
  // deposit kernel); cells outside it are neither read nor cleared.
  const int4 wn = cf ? win[slot] : make_int4(1, 0, 1, 0);
  // rows j-1, j, j+1 and Sc of the group's next row, loaded one pass ahead
  cpx<T> qm[E], q0[E], qp[E], qs[E];
#define FELIS_ADI_LOADX(JB)                                                                           \
  {                                                                                                   \
    const int jl = 1 + (JB) + wv * 4 + grp;                                                           \
    _Pragma("unroll") for (int e = 0; e < E; ++e) {                                                   \
      const int i = i0 + e;                                                                           \
      if (jl <= ny - 2 && i <= nx - 2) {                                                              \
        const long long idx = (long long)jl * nx + i;                                                 \
        qm[e] = Un[idx - nx];                                                                         \
        q0[e] = Un[idx];                                                                              \
        qp[e] = Un[idx + nx];                                                                         \
        qs[e] = (i >= wn.x && i <= wn.y && jl >= wn.z && jl <= wn.w) ? Sn[idx] : mk<T>(0, 0);         \
      } else {                                                                                        \
        qm[e] = q0[e] = qp[e] = qs[e] = mk<T>(0, 0);                                                  \
      }                                                                                               \
    }                                                                                                 \
  }
  FELIS_ADI_LOADX(0);
  for (int jb = 0; jb < nyi; jb += nw * 4) {
    const int j = 1 + jb + wv * 4 + grp;
    const bool rowok = j <= ny - 2;
    const int jj = rowok ? j : 1;
    const cpx<T> bP = beP[jj];
    const cpx<T> bM = be[jj];
    const cpx<T> diag = mk<T>(H, 0) + bP + bM;
    cpx<T> av[E], bv[E], sc[E], uc[E];
    bool valid[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = i0 + e;
      valid[e] = rowok && i <= nx - 2;
      if (valid[e]) {
        sc[e] = qs[e];
        cpx<T> dd = diag * q0[e] - bP * qp[e] - bM * qm[e] + sc[e];
        bv[e] = xr[i] * dd;
        av[e] = xa[i];
      } else {
        sc[e] = mk<T>(0, 0);
        av[e] = mk<T>(1, 0);
        bv[e] = mk<T>(0, 0);
      }
    }
    if (jb + nw * 4 < nyi) FELIS_ADI_LOADX(jb + nw * 4);
    row_scan16<T, E>(sub, av, bv, xq, xp, valid, i0, uc);
    // y right-hand side: x-explicit stencil of Uc on this row (Uc's ring is 0)
    T lr = __shfl_up(uc[E - 1].re, 1, 16), li = __shfl_up(uc[E - 1].im, 1, 16);
    T rr = __shfl_down(uc[0].re, 1, 16), ri = __shfl_down(uc[0].im, 1, 16);
    cpx<T> left = sub == 0 ? mk<T>(0, 0) : mk(lr, li);
    cpx<T> right = sub == 15 ? mk<T>(0, 0) : mk(rr, ri);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = i0 + e;
      if (!valid[e]) continue;
      cpx<T> um = e == 0 ? left : uc[e - 1];
      cpx<T> up = (e == E - 1 || i == nx - 2) ? (i == nx - 2 ? mk<T>(0, 0) : right) : uc[e + 1];
      const cpx<T> aM = al[i], aP = alP[i];
      const cpx<T> dg = mk<T>(H, 0) + aP + aM;
      DY[(j - 1) * nxi + (i - 1)] = dg * uc[e] - aP * up - aM * um + sc[e];
      if (i >= wn.x && i <= wn.y && j >= wn.z && j <= wn.w) Sn[(long long)j * nx + i] = mk<T>(0, 0);
    }
  }
#undef FELIS_ADI_LOADX
  __syncthreads();
  // ---- y-implicit half step: 16-lane scans down the columns, in place ----
  for (int cb = 0; cb < nxi; cb += nw * 4) {
    const int col = cb + wv * 4 + grp;  // 0-based interior column
    const bool colok = col < nxi;
    const int cc = colok ? col : 0;
    cpx<T> av[E], bv[E], uo[E];
    bool valid[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = i0 + e;  // row 1 + sub*E + e
      valid[e] = colok && j <= ny - 2;
      if (valid[e]) {
        bv[e] = yr[j] * DY[(j - 1) * nxi + cc];
        av[e] = ya[j];
      } else {
        av[e] = mk<T>(1, 0);
        bv[e] = mk<T>(0, 0);
      }
    }
    row_scan16<T, E>(sub, av, bv, yq, yp, valid, i0, uo);
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (valid[e]) DY[(i0 + e - 1) * nxi + cc] = uo[e];
  }
  __syncthreads();
  // ---- coalesced store of the interior and the slice energy (FP64) ----
  double es = 0.0;
  for (int q = tid; q < nxi * nyi; q += blockDim.x) {
    const int jr = q / nxi, ic = q - jr * nxi;
    const cpx<T> v = DY[q];
    Un[(long long)(jr + 1) * nx + ic + 1] = v;
    es += (double)v.re * v.re + (double)v.im * v.im;
  }
  // the boundary ring is untouched by the ADI but counts in the energy
  const int nring = 2 * nx + 2 * (ny - 2);
  for (int k = tid; k < nring; k += blockDim.x) {
    long long idx;
    if (k < nx) idx = k;
    else if (k < 2 * nx) idx = (long long)(ny - 1) * nx + (k - nx);
    else {
      int r = k - 2 * nx;
      idx = (long long)(1 + (r >> 1)) * nx + ((r & 1) ? nx - 1 : 0);
    }
    cpx<T> v = Un[idx];
    es += (double)v.re * v.re + (double)v.im * v.im;
  }
  for (int o = 32; o > 0; o >>= 1) es += __shfl_down(es, o);
  if (lane == 0) red[wv] = es;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += red[k];
    slice_e[sys] = s;
  }
}

}  // namespace felis
