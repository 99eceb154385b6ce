// Shared device/host definitions for the gfx950 FELIS slice kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "felis.h"

namespace felis {

// Physical constants with the reference's exact values (constants.h:1-10).
constexpr double kMu0 = 1.2566370614e-6;
constexpr double kC = 299792458.0;
constexpr double kEMass = 9.10938291e-31;
constexpr double kECharge = 1.60217662e-19;
constexpr double kMc2 = 0.510998918;
constexpr double kZ0 = 376.730313668;
constexpr double kPlankH = 4.135667696E-15;

// Device error word bits, folded into felis_status by the host.
enum : unsigned {
  EB_GUARD = 1u,    // |dTheta| > dPhi (particlePush.cpp:370-374)
  EB_RANGE = 2u,    // theta out of [0, dTh] after the wrap
  EB_NAN = 4u,      // NaN theta
  EB_BESSEL = 8u,   // Bessel-table index beyond the table
  EB_EMCARRY = 16u  // in-grid particle with idxR >= nr (reference carries Em, A.4)
};

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cpx<T> mk(T r, T i) {
  cpx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ inline cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ inline cpx<T> operator*(cpx<T> a, T s) {
  return mk<T>(a.re * s, a.im * s);
}
template <typename T>
__host__ __device__ inline cpx<T> operator*(T s, cpx<T> a) {
  return mk<T>(a.re * s, a.im * s);
}
template <typename T>
__host__ __device__ inline T norm2(cpx<T> a) {
  return a.re * a.re + a.im * a.im;
}
template <typename T>
__host__ __device__ inline cpx<T> conjg(cpx<T> a) {
  return mk<T>(a.re, -a.im);
}
template <typename T, typename U>
__host__ __device__ inline cpx<T> cvt(cpx<U> a) {
  return mk<T>((T)a.re, (T)a.im);
}
// complex division a / b (plain formula, as -ffast-math std::complex does)
template <typename T>
__host__ __device__ inline cpx<T> cdiv(cpx<T> a, cpx<T> b) {
  T d = b.re * b.re + b.im * b.im;
  return mk<T>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

constexpr int kTile = 256;    // particles per tile (one workgroup)
constexpr int kChunk = 4096;  // particles per deposit workgroup

// ---------------------------------------------------------------- arguments

struct GridArgs {
  int nx, ny, S, slice_n, nh, nr, L, F, bo, bn, ring_base, wake_on;
  int harm[FELIS_MAX_HARM];
  long long N;  // nx*ny
  double dx, dy, dr, minX, minY, dz, ks, dB, dPhi, gref;
  // when set, the ring position is read from device memory instead of
  // ring_base: a step captured once as a graph then serves every period
  const int *ring_dev;
};

__device__ inline int ring_of(const GridArgs &g) { return g.ring_dev ? *g.ring_dev : g.ring_base; }

struct LatArgs {
  double K0, ue, K0_alpha, ku, g, Phi;
  double etaXr, etaXi, etaYr, etaYi;
  int und_type, current_flag;
};

struct Tiles {
  const int *seg;        // segment of each tile
  const long long *start;
  const int *cnt;
};

__device__ inline int phys_slot(int slot, int base, int S) {
  int p = slot + base;
  return p >= S ? p - S : p;
}

struct DepArgs {
  double coef_U, xi, coef_Ez, Phi;
  double etaXr, etaXi, etaYr, etaYi;
  double xi2c;  // sqrt(2) K0 / (1 + K0^2/2 (1+ue^2))
  double ue, ka;
  double inv_dx, inv_dy, inv_dB;  // reciprocals for the batched deposit
  int last_species_first_seg;
  int run;  // consecutive particles one batched deposit thread combines
};

struct AdiCoef {
  const cpx<double> *x;  // [nh][4][nx]: r, a, q, p
  const cpx<double> *y;  // [nh][4][ny]
  const cpx<double> *abc; // alpha[nx], alphaP[nx], beta[ny], betaP[ny]
};

}  // namespace felis
