// Kernels of the multi-period jump of the batched step (felis_run_surrogate).
//
// A jump advances every slice J undulator periods with one field solve and
// one particle push of step J*dz, and slips the field J slots. One coarse
// step of the reference's equations would let each field slot take J times
// the source of a single slice, and push each slice in a single field slot:
// the shot-noise part of the bunching then adds coherently over the J
// periods, which overstates the spontaneous power J-fold. The two window
// passes here keep the slippage of the J fine steps instead:
//
//   source   Sw[s] = sum_{i<J} Sc[s+i]     the slices field slot s passes
//                                           while it slips to slot s+J
//   field    Ub[s] = 1/J sum_{i<J} U_i[s-i] the field slots that pass slice s,
//                                           each at its time in the jump
//                                           (logical slots < 1 hold no field)
//
// Both run over slots in chunks of kJumpChunk with a halo, one thread per
// (cell, pol x harmonic, chunk), with the running sum in FP64. Only the
// bounding box of the beam's deposit windows is visited, its cells taken by
// consecutive threads, and of a chunk only the cells inside the union of the
// windows its slots are read in (the source of a slot is its deposit window,
// and the push reads the field only around its particles).
#pragma once

#include "felis_common.h"

namespace felis {

constexpr int kJumpChunk = 32;  // slots per thread of the window passes
constexpr int kJumpMargin = 2;  // cells the push-field box extends the beam box by

// bbox = union of win[1..slice_n] (non-empty windows), widened by margin and
// clamped to the grid; an empty union gives an empty box (x0 > x1).
static __global__ void __launch_bounds__(256) k_win_bbox(const int4 *__restrict__ win, int slice_n, int nx, int ny,
                                                  int margin, int4 *__restrict__ bbox) {
  __shared__ int sh[4][256];
  int x0 = 1 << 30, x1 = -1, y0 = 1 << 30, y1 = -1;
  for (int s = 1 + threadIdx.x; s <= slice_n; s += blockDim.x) {
    const int4 w = win[s];
    if (w.x <= w.y && w.z <= w.w) {
      x0 = min(x0, w.x); x1 = max(x1, w.y); y0 = min(y0, w.z); y1 = max(y1, w.w);
    }
  }
  const int t = threadIdx.x;
  sh[0][t] = x0; sh[1][t] = x1; sh[2][t] = y0; sh[3][t] = y1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      sh[0][t] = min(sh[0][t], sh[0][t + o]);
      sh[1][t] = max(sh[1][t], sh[1][t + o]);
      sh[2][t] = min(sh[2][t], sh[2][t + o]);
      sh[3][t] = max(sh[3][t], sh[3][t + o]);
    }
    __syncthreads();
  }
  if (t == 0) {
    if (sh[1][0] < 0) {
      *bbox = make_int4(1, 0, 1, 0);
    } else {
      *bbox = make_int4(max(0, sh[0][0] - margin), min(nx - 1, sh[1][0] + margin), max(0, sh[2][0] - margin),
                        min(ny - 1, sh[3][0] + margin));
    }
  }
}

// win2[s] = union of win[s..min(s+J-1, slice_n)], the cells Sw[s] can be
// nonzero in; slots 0 and slice_n+1 get an empty window.
static __global__ void k_win_union(const int4 *__restrict__ win, int slice_n, int J, int4 *__restrict__ win2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > slice_n + 1) return;
  int4 u = make_int4(1 << 30, -1, 1 << 30, -1);
  if (s >= 1 && s <= slice_n) {
    for (int i = 0; i < J && s + i <= slice_n; ++i) {
      const int4 w = win[s + i];
      if (w.x <= w.y && w.z <= w.w) {
        u.x = min(u.x, w.x); u.y = max(u.y, w.y); u.z = min(u.z, w.z); u.w = max(u.w, w.w);
      }
    }
  }
  win2[s] = u.y < 0 ? make_int4(1, 0, 1, 0) : u;
}

__device__ inline bool in_box(int4 b, int i, int j) { return i >= b.x && i <= b.y && j >= b.z && j <= b.w; }

// out[z] = union of w[s] over the slots s of chunk z (kJumpChunk slots from
// slot 1), widened by margin and clamped to the grid; empty if every window
// of the chunk is. The window passes skip the (cell, chunk) pairs outside it.
static __global__ void k_chunk_windows(const int4 *__restrict__ w, int slice_n, int margin, int nx, int ny,
                                       int4 *__restrict__ out) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = 1 + z * kJumpChunk;
  if (s0 > slice_n) return;
  int4 u = make_int4(1 << 30, -1, 1 << 30, -1);
  for (int s = s0; s < s0 + kJumpChunk && s <= slice_n; ++s) {
    const int4 v = w[s];
    if (v.x <= v.y && v.z <= v.w) {
      u.x = min(u.x, v.x); u.y = max(u.y, v.y); u.z = min(u.z, v.z); u.w = max(u.w, v.w);
    }
  }
  out[z] = u.y < 0 ? make_int4(1, 0, 1, 0)
                   : make_int4(max(0, u.x - margin), min(nx - 1, u.y + margin), max(0, u.z - margin),
                               min(ny - 1, u.w + margin));
}

// Sw[ph][s] = sum_{i<J, s+i<=slice_n} Sc[ph][s+i] for the chunk's slots, at
// one interior cell of the beam box, where Sc[s] counts only inside the
// deposit window win[s] (cells outside it hold older deposits); ring cells
// of Sw are zero. Sc and Sw are indexed by logical slot.
// grid: x = box cells / 256 (at most the grid's cells), y = pol x harmonic, z = slot chunks.
static __global__ void __launch_bounds__(256) k_src_window(int nx, int ny, int S, int slice_n, int J,
                                                    const int4 *__restrict__ bbox,
                                                    const int4 *__restrict__ cwin,
                                                    const int4 *__restrict__ win,
                                                    const cpx<float> *__restrict__ Sc,
                                                    cpx<float> *__restrict__ Sw) {
  const long long N = (long long)nx * ny;
  // threads over the box's cells, row by row
  const int4 b = *bbox;
  const int bw = b.y - b.x + 1, bh = b.w - b.z + 1;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (bw <= 0 || bh <= 0 || q >= bw * bh) return;
  const int i = b.x + q % bw, j = b.z + q / bw;
  const long long c = (long long)j * nx + i;
  const int s0 = 1 + blockIdx.z * kJumpChunk;
  if (s0 > slice_n) return;
  // Sw[s] is read only inside win2[s]: the chunk's union of them
  if (!in_box(cwin[blockIdx.z], i, j)) return;
  const int s1 = min(slice_n, s0 + kJumpChunk - 1);
  const cpx<float> *src = Sc + (long long)blockIdx.y * S * N + c;
  cpx<float> *dst = Sw + (long long)blockIdx.y * S * N + c;
  const bool interior = i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
  // Sc[s] inside its window, zero outside it and past the last slot
  auto at = [&](int s) {
    const int sc = min(s, slice_n);
    return interior && s <= slice_n && in_box(win[sc], i, j) ? src[(long long)sc * N] : mk(0.f, 0.f);
  };
  double sr = 0.0, si = 0.0;
  for (int s = s0; s < s0 + J - 1 && s <= slice_n; ++s) {
    const cpx<float> v = at(s);
    sr += v.re;
    si += v.im;
  }
  constexpr int kB = 4;  // slots whose loads are issued together
  for (int sb = s0; sb <= s1; sb += kB) {
    cpx<float> va[kB], vs[kB];
#pragma unroll
    for (int t = 0; t < kB; ++t) {
      const int s = min(sb + t, s1);
      va[t] = at(s + J - 1);
      vs[t] = at(s);
    }
#pragma unroll
    for (int t = 0; t < kB; ++t) {
      const int s = sb + t;
      if (s > s1) break;  // the same for the whole workgroup
      sr += va[t].re;
      si += va[t].im;
      dst[(long long)s * N] = mk((float)sr, (float)si);
      sr -= vs[t].re;
      si -= vs[t].im;
    }
  }
}

// The push field of a jump: slice s meets field slot s-i in period i of the
// jump (i < J), and in the fine steps that slot's field there is the field
// after i+1 of the J solves. With the field before the jump's solve (Uo) and
// after it (Un), linear in between,
//   Ub[s] = 1/J sum_{i<J, s-i>=1} [(1 - w_i) Uo(s-i) + w_i Un(s-i)], w_i = (i+1)/J.
// With A(s) = sum_i U(s-i) and B(s) = sum_i (i+1) U(s-i), run along the slots
// in FP64 (A(s+1) = A + U(s+1) - U(s-J+1), B(s+1) = B + A + U(s+1) -
// (J+1) U(s-J+1)), the pass over Uo (before the solve, ADD = 0) writes
// (A - B/J)/J and the pass over Un (after it, ADD = 1) adds B/J^2. J = 1
// gives Un: the fine step. U and Ub share the ring; one thread per (cell of
// the widened beam box, pol x harmonic, chunk of slots).
template <int ADD>
static __global__ void __launch_bounds__(256) k_field_window(int nx, int ny, int S, int slice_n, int ring_base, int J,
                                                      const int4 *__restrict__ bbox,
                                                      const int4 *__restrict__ cwin,
                                                      const cpx<float> *__restrict__ U,
                                                      cpx<float> *__restrict__ Ub) {
  const long long N = (long long)nx * ny;
  // threads over the box's cells, row by row
  const int4 b = *bbox;
  const int bw = b.y - b.x + 1, bh = b.w - b.z + 1;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (bw <= 0 || bh <= 0 || q >= bw * bh) return;
  const int i = b.x + q % bw, j = b.z + q / bw;
  const long long c = (long long)j * nx + i;
  const int s0 = 1 + blockIdx.z * kJumpChunk;
  if (s0 > slice_n) return;
  // the push reads Ub[s] only around slice s's particles: the chunk's union
  // of their windows, widened
  if (!in_box(cwin[blockIdx.z], i, j)) return;
  const int s1 = min(slice_n, s0 + kJumpChunk - 1);
  const cpx<float> *src = U + (long long)blockIdx.y * S * N + c;
  cpx<float> *dst = Ub + (long long)blockIdx.y * S * N + c;
  // U at logical slot s, zero below slot 1 (the load's address is clamped)
  auto at = [&](int s) {
    const cpx<float> v = src[(long long)phys_slot(max(s, 1), ring_base, S) * N];
    return s >= 1 ? v : mk(0.f, 0.f);
  };
  const double inv = 1.0 / J;
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
  for (int q = 0; q < J; ++q) {
    const cpx<float> v = at(s0 - q);
    ar += v.re;
    ai += v.im;
    br += (q + 1) * (double)v.re;
    bi += (q + 1) * (double)v.im;
  }
  constexpr int kB = 4;  // slots whose loads are issued together
  for (int sb = s0; sb <= s1; sb += kB) {
    cpx<float> un[kB], uo[kB], ob[kB];
#pragma unroll
    for (int t = 0; t < kB; ++t) {
      const int s = min(sb + t, s1);
      un[t] = at(min(s + 1, s1));  // used when s < s1
      uo[t] = at(s - J + 1);
      if (ADD) ob[t] = dst[(long long)phys_slot(s, ring_base, S) * N];
    }
#pragma unroll
    for (int t = 0; t < kB; ++t) {
      const int s = sb + t;
      if (s > s1) break;  // the same for the whole workgroup
      cpx<float> *o = dst + (long long)phys_slot(s, ring_base, S) * N;
      if (ADD)
        *o = mk((float)(ob[t].re + br * inv * inv), (float)(ob[t].im + bi * inv * inv));
      else
        *o = mk((float)((ar - br * inv) * inv), (float)((ai - bi * inv) * inv));
      if (s < s1) {
        br += ar + un[t].re - (J + 1) * (double)uo[t].re;
        bi += ai + un[t].im - (J + 1) * (double)uo[t].im;
        ar += un[t].re - uo[t].re;
        ai += un[t].im - uo[t].im;
      }
    }
  }
}

// The learned correction of a jump's emission: the field solve adds the
// windowed source Sw in both half sweeps, so the jump's emitted field at slot
// s is E = 2 Sw[s] / H (the source varies slowly on the scale diffraction
// reaches in one jump). U[ph][phys(s)] += (e - 1) E with e = gain[ph][s]
// the network's factor, on Sw[s]'s window win2[s] (Sw is zero outside it);
// the field that came into the jump is left as it is. One workgroup per
// (pol x harmonic, slot): the field's energy change is summed in a fixed
// order and added to the slot's slice energy, which the field solve summed
// before the correction.
static __global__ void __launch_bounds__(256) k_emit_correct(GridArgs g, const int4 *__restrict__ win2,
                                                      const cpx<float> *__restrict__ gain,
                                                      const cpx<float> *__restrict__ Sw,
                                                      cpx<float> *__restrict__ U, double *__restrict__ slice_e) {
  __shared__ double red[4];
  const int s = 1 + blockIdx.x, ph = blockIdx.y, h = ph % g.nh;
  const int4 b = win2[s];
  const int bw = b.y - b.x + 1, bh = b.w - b.z + 1;
  const int ncell = bw > 0 && bh > 0 ? bw * bh : 0;
  const float k2 = 2.f / (float)g.harm[h];
  const cpx<float> e = gain[(long long)ph * g.slice_n + s - 1];
  const cpx<float> f = mk((e.re - 1.f) * k2, e.im * k2);
  const cpx<float> *src = Sw + ((long long)ph * g.S + s) * g.N;
  cpx<float> *u = U + ((long long)ph * g.S + phys_slot(s, g.ring_base, g.S)) * g.N;
  double de = 0.0;
  for (int q = threadIdx.x; q < ncell; q += blockDim.x) {
    const long long c = (long long)(b.z + q / bw) * g.nx + b.x + q % bw;
    const cpx<float> o = u[c];
    const cpx<float> n = o + f * src[c];
    u[c] = n;
    de += ((double)n.re * n.re + (double)n.im * n.im) - ((double)o.re * o.re + (double)o.im * o.im);
  }
  for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = de;
  __syncthreads();
  if (threadIdx.x == 0) slice_e[(long long)ph * g.slice_n + s - 1] += (red[0] + red[1]) + (red[2] + red[3]);
}

// zero the physical slots of logical slots 0..J-1 after a J-slot rotation
// (slippage of J slices: the reference's slots 1..J take its empty slot 0)
static __global__ void k_zero_slots(cpx<float> *U, int ring_base, int S, long long N, int J) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ph = blockIdx.y, l = blockIdx.z;
  if (k < N && l < J) U[((long long)ph * S + phys_slot(l, ring_base, S)) * N + k] = mk(0.f, 0.f);
}

}  // namespace felis
