// The reference's Em carry (particlePush.cpp:248, 283-304). push_theta_gamma
// keeps one Em vector per call: an in-grid particle with idxR < nr sets it
// from Ez, an in-grid particle with idxR >= nr uses whatever it holds, i.e.
// the Em of the last particle before it (species, then slice, then particle
// order) that set it, or zero when none did since the call began. Particles
// outside the grid leave it alone. The push kernels run all particles at
// once, so before a push that can meet such a particle three launches give
// every carried particle the Em it would see in that order:
//   k_em_last  per tile: the last particle that sets Em (-1: none);
//   k_em_scan  per tile: the last earlier tile that has one (-1: none);
//   k_em_fill  per carried particle: the Em of its source, recomputed from
//              the source's position and slice, into emc[n * L + l].
// The walk order is the flat particle index: tiles cover the segments in
// order (species-major), particles keep their order inside a segment. This is
// a single-process reference run; under MPI each rank starts its own walk at
// zero.
//
// Policy P: Pos (position type), E (Em element type),
//   int where(long long n, Pos &q)  0 outside the grid, 1 sets Em, 2 carried;
//   int modes()                      L;
//   void em(int seg, const Pos &q, E *out)   Em of a particle at q in seg.
#pragma once

#include "felis_common.h"

namespace felis {

template <class P>
__global__ void __launch_bounds__(kTile) k_em_last(P p, Tiles tl, int *__restrict__ last) {
  __shared__ int m;
  const int t = blockIdx.x, k = threadIdx.x;
  if (k == 0) m = -1;
  __syncthreads();
  if (k < tl.cnt[t]) {
    typename P::Pos q;
    if (p.where(tl.start[t] + k, q) == 1) atomicMax(&m, k);
  }
  __syncthreads();
  if (k == 0) last[t] = m;
}

// src[t] = the last tile before t whose last[] is set, -1 if none: each of
// 1024 threads walks a contiguous run of tiles, the runs are combined by an
// LDS prefix maximum (tile numbers increase along the walk)
static __global__ void __launch_bounds__(1024) k_em_scan(int nt, const int *__restrict__ last, int *__restrict__ src) {
  __shared__ int part[1024];
  const int k = threadIdx.x;
  const int per = (nt + 1023) / 1024;
  const int b = min(nt, k * per), e = min(nt, b + per);
  int lt = -1;
  for (int t = b; t < e; ++t)
    if (last[t] >= 0) lt = t;
  part[k] = lt;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = part[k];
    if (k >= off) v = max(v, part[k - off]);
    __syncthreads();
    part[k] = v;
    __syncthreads();
  }
  int cur = k ? part[k - 1] : -1;
  for (int t = b; t < e; ++t) {
    src[t] = cur;
    if (last[t] >= 0) cur = t;
  }
}

template <class P>
__global__ void __launch_bounds__(kTile)
    k_em_fill(P p, Tiles tl, const int *__restrict__ last, const int *__restrict__ src,
              typename P::E *__restrict__ emc) {
  __shared__ int pre[kTile];
  __shared__ typename P::Pos pos[kTile];
  const int t = blockIdx.x, k = threadIdx.x;
  int cls = 0;
  typename P::Pos q{};
  if (k < tl.cnt[t]) cls = p.where(tl.start[t] + k, q);
  pre[k] = cls == 1 ? k : -1;
  pos[k] = q;
  __syncthreads();
  // inclusive prefix maximum: the last particle at or before k that sets Em
  for (int off = 1; off < kTile; off <<= 1) {
    int v = pre[k];
    if (k >= off) v = max(v, pre[k - off]);
    __syncthreads();
    pre[k] = v;
    __syncthreads();
  }
  if (cls != 2) return;
  const int L = p.modes();
  typename P::E *out = emc + (tl.start[t] + k) * L;
  const int sk = pre[k];
  if (sk >= 0) {
    p.em(tl.seg[t], pos[sk], out);
    return;
  }
  const int st = src[t];
  if (st < 0) {
    for (int l = 0; l < L; ++l) out[l] = typename P::E{};
    return;
  }
  typename P::Pos qs;
  p.where(tl.start[st] + last[st], qs);
  p.em(tl.seg[st], qs, out);
}

}  // namespace felis
