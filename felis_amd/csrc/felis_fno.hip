// A Fourier neural operator on 64 x 64 per-slice maps, on gfx950 matrix
// cores: BF16 operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16), FP32
// storage, FP32 master weights and Adam. It is the field-to-field network of
// the multi-period jumps (felis_run_surrogate): per slice it maps the jump's
// field and emission on a 64 x 64 grid, downsampled from the field grid, with
// the slice's bunching and energy moments, to a complex factor map on the
// jump's emission (one per harmonic).
//
// Layers (channel-last maps [slice][y][x][c], N = 64, width w, m modes):
//   lift      h0 = x Wl + bl                                   (cin -> w)
//   Fourier   a = h W + b + IDFTx(IDFTy(mix(DFTy(DFTx(h)))))   (w -> w), four
//             layers, GELU after each but the last
//   project   q = gelu(h4 Q1 + c1), out = q Q2 + c2           (w -> p -> cout)
// DFTx keeps kx = 0 .. m-1 of the real transform along x, DFTy ky = 0 .. m-1
// and -m .. -1 along y; mix is a real linear map of the (re, im) x channel
// vector of each kept mode (64 x 64 weights per mode); the inverse x
// transform completes the Hermitian half (factor 2 for kx > 0) and divides
// by N^2. Every transform and channel map is a product of strided operands:
// one batched product kernel forms them all, forward and backward.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "felis.h"
#include "felis_ctx.h"
#include "felis_fno.h"
#include "felis_sur.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kN = 64;        // map side
constexpr int kP = kN * kN;   // positions per map

// FP32 -> BF16 bits, round to nearest even: the hardware conversion
// (v_cvt_pk_bf16_f32 on gfx950)
__device__ inline short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// offset of index i: (i >> sh) * hi + (i & (2^sh - 1)) * lo, or i * lo for sh >= 31
struct Ix {
  int sh;
  long long hi, lo;
};
__host__ __device__ inline long long ix(const Ix &x, long long i) {
  return x.sh >= 31 ? i * x.lo : (i >> x.sh) * x.hi + (i & ((1LL << x.sh) - 1)) * x.lo;
}
// an operand: element (batch, row, col) at p + ix(b, batch) + ix(r, row) + ix(c, col)
struct Op {
  const float *p;
  Ix b, r, c;
};
struct OpW {
  float *p;
  Ix b, r, c;
};

// C(b, m, n) = sum_k A(b, m, k) B(b, k, n) [+ C(b, m, n)] [+ bias(n)], BF16
// operands, FP32 accumulation. One wavefront per 32 x 32 tile of one batch
// entry (2 x 2 blocks of 16 x 16), four wavefronts per workgroup. In a 16 x
// 16 x 16 product lane l holds A[l%16][4(l/16) .. +3], B[4(l/16) .. +3][l%16]
// and returns D[4(l/16) .. +3][l%16].
// operand element as BF16 bits: FP32 storage is rounded, BF16 storage (the
// fused inference path's intermediates, already rounded) is taken as it is
__device__ inline short ld_bf(const float *p, long long o) { return f2bf(p[o]); }
__device__ inline short ld_bf(const uint16_t *p, long long o) { return (short)p[o]; }
__device__ inline void st_out(float *p, float v) { *p = v; }
__device__ inline void st_out(uint16_t *p, float v) { *p = (uint16_t)f2bf(v); }

template <typename TA, typename TB, typename TC>
__global__ void __launch_bounds__(256) k_bgemm(int batch, int M, int N, int K, Op A, Op B, OpW C, int acc,
                                               const float *__restrict__ bias) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tm_n = (M + 31) >> 5, tn_n = (N + 31) >> 5;
  const long long tile = (long long)blockIdx.x * 4 + wave, per = (long long)tm_n * tn_n;
  if (tile >= per * batch) return;
  const long long bi = tile / per;
  const int t = (int)(tile % per), m0 = (t / tn_n) * 32, n0 = (t % tn_n) * 32;
  const int r = l & 15, kq = (l >> 4) * 4;
  const TA *pa = reinterpret_cast<const TA *>(A.p) + ix(A.b, bi);
  const TB *pb = reinterpret_cast<const TB *>(B.p) + ix(B.b, bi);
  long long ra[2], cb[2];
  bool va[2], vb[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int m = m0 + 16 * u + r, n = n0 + 16 * u + r;
    va[u] = m < M;
    vb[u] = n < N;
    ra[u] = va[u] ? ix(A.r, m) : 0;
    cb[u] = vb[u] ? ix(B.c, n) : 0;
  }
  v4f d[2][2] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a[2], b[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kq + q;
      const bool vk = k < K;
      const long long ka = vk ? ix(A.c, k) : 0, kb = vk ? ix(B.r, k) : 0;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a[u][q] = (vk && va[u]) ? ld_bf(pa, ra[u] + ka) : (short)0;
        b[u][q] = (vk && vb[u]) ? ld_bf(pb, kb + cb[u]) : (short)0;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
  }
  TC *pc = reinterpret_cast<TC *>(C.p) + ix(C.b, bi);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + 16 * i + kq + q, n = n0 + 16 * j + r;
        if (m < M && n < N) {
          TC *o = pc + ix(C.r, m) + ix(C.c, n);
          float v = d[i][j][q];
          if (bias) v += bias[n];
          if (acc) v += (float)*reinterpret_cast<float *>(o);  // FP32 storage only
          st_out(o, v);
        }
      }
}

// The y transforms of the fused inference path (width 32, 8 modes), one
// workgroup per map, with the operands' layouts fixed at compile time:
//   forward (INV false): Y = Gy (32 x 128) Xf (128 x 256), Xf [k][n]
//     contiguous, Y written in the mode mixing's layout
//     [(i>>1)*512 + (i&1)*32 + (n>>5)*64 + (n&31)];
//   inverse (INV true): Xb = Giy (128 x 32) Zv (32 x 256), Zv read in that
//     layout, Xb [i][n] contiguous.
// The table goes to LDS as BF16 once per workgroup; each wave takes 32 x 32
// output tiles in turn with the same 16 x 16 x 16 MFMA sequence per tile as
// k_bgemm (k ascending, the same BF16 operands), so the results are k_bgemm's
// bit for bit, without its general index arithmetic.
template <bool INV>
__global__ void __launch_bounds__(256) k_fno_ytr(int batch, const float *__restrict__ T,
                                                 const uint16_t *__restrict__ Bm, uint16_t *__restrict__ Cm) {
  constexpr int M = INV ? 128 : 32, K = INV ? 32 : 128, N = 256;
  constexpr long long SB = INV ? 8192 : 32768, SC = INV ? 32768 : 8192;  // per map
  __shared__ short At[M * K];
  const int map = blockIdx.x;
  if (map >= batch) return;
  for (int q = threadIdx.x; q < M * K; q += 256) At[q] = f2bf(T[q]);
  __syncthreads();
  const uint16_t *B = Bm + map * SB;
  uint16_t *Cp = Cm + map * SC;
  auto boff = [](int k, int n) -> int {
    return INV ? (k >> 1) * 512 + (k & 1) * 32 + (n >> 5) * 64 + (n & 31) : k * N + n;
  };
  auto coff = [](int i, int n) -> int {
    return INV ? i * N + n : (i >> 1) * 512 + (i & 1) * 32 + (n >> 5) * 64 + (n & 31);
  };
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  constexpr int TN = N / 32, TILES = (M / 32) * TN;
  for (int t = wave; t < TILES; t += 4) {
    const int m0 = (t / TN) * 32, n0 = (t % TN) * 32;
    v4f d[2][2] = {};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      v4s a[2], b[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a[u] = *reinterpret_cast<const v4s *>(&At[(m0 + 16 * u + r) * K + k0 + kq]);
#pragma unroll
        for (int q = 0; q < 4; ++q) b[u][q] = (short)B[boff(k0 + kq + q, n0 + 16 * u + r)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) Cp[coff(m0 + 16 * i + kq + q, n0 + 16 * j + r)] = (uint16_t)f2bf(d[i][j][q]);
  }
}

// The mode mixing of the fused inference path (width 32, 8 modes: 128
// modes, 64 x 64 real matrices): Z[map][md*64 + n] = sum_k Y[map][md*64 + k]
// Wm[md][k][n], maps sY = 8192 apart. One workgroup per (mode, 64 maps): the
// mode's matrix goes to LDS as BF16, transposed so that a lane's four k of
// one column are one 8-byte read; each wave forms 16 maps x 64 columns with
// k ascending per 16 x 16 block, as k_bgemm does, so the results are its
// bit for bit.
__global__ void __launch_bounds__(256) k_fno_mix(int batch, const uint16_t *__restrict__ Y,
                                                 const float *__restrict__ Wm, uint16_t *__restrict__ Z) {
  constexpr int W2 = 64, SY = 8192, LDW = W2 + 4;  // padded rows: a 16-lane read spreads over the banks
  __shared__ short Wt[W2 * LDW];  // [n][k]
  const int md = blockIdx.y, map0 = blockIdx.x * 64;
  const float *wm = Wm + (long long)md * W2 * W2;
  for (int q = threadIdx.x; q < W2 * W2; q += 256) {
    const int k = q / W2, n = q % W2;
    Wt[n * LDW + k] = f2bf(wm[q]);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  const int mp = map0 + 16 * wave + r;  // this lane's map for the A operand
  const bool va = mp < batch;
  const uint16_t *ya = Y + (long long)(va ? mp : 0) * SY + md * W2;
  v4f d[4] = {};
#pragma unroll
  for (int k0 = 0; k0 < W2; k0 += 16) {
    const v4s a = va ? *reinterpret_cast<const v4s *>(ya + k0 + kq) : v4s{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const v4s b = *reinterpret_cast<const v4s *>(&Wt[(16 * j + r) * LDW + k0 + kq]);
      d[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = map0 + 16 * wave + kq + q;
      if (m < batch) Z[(long long)m * SY + md * W2 + 16 * j + r] = (uint16_t)f2bf(d[j][q]);
    }
}

// GELU in its tanh form, with tanh(u) = 1 - 2 / (1 + e^{2u}) from the
// hardware exponential and reciprocal (saturating to +-1 for large |u|)
__device__ inline float tanh_fast(float u) { return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * u)); }
__device__ inline float gelu(float z) {
  const float c = 0.7978845608f;
  return 0.5f * z * (1.f + tanh_fast(c * (z + 0.044715f * z * z * z)));
}
__device__ inline float gelu_d(float z) {
  const float c = 0.7978845608f;
  const float u = c * (z + 0.044715f * z * z * z), t = tanh_fast(u);
  return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * z * z);
}

__global__ void k_gelu(long long n, const float *__restrict__ a, float *__restrict__ h) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = gelu(a[i]);
}
__global__ void k_gelu_back(long long n, const float *__restrict__ dh, const float *__restrict__ a,
                            float *__restrict__ da) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) da[i] = dh[i] * gelu_d(a[i]);
}

// out[c] = sum over rows of X[row][c] (rows x cols, row-major), one workgroup
// per column, a fixed order
__global__ void __launch_bounds__(256) k_colsum(long long rows, int cols, const float *__restrict__ X,
                                                float *__restrict__ out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (long long i = threadIdx.x; i < rows; i += 256) s += X[i * cols + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = (float)sh[0];
}

// the input affine per channel
__global__ void k_affine(long long n, int cols, const float *__restrict__ x, const float *__restrict__ mu,
                         const float *__restrict__ isd, float *__restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int c = (int)(i % cols);
    y[i] = (x[i] - mu[c]) * isd[c];
  }
}

// loss of the factor maps: out[b][p][2h..2h+1] = o_h, E and R alike (complex
// per harmonic); L = sum_b w_b sum_{p,h} |o E - R|^2 / (B P), dout = dL/do
__global__ void __launch_bounds__(256) k_fno_loss(int B, int nh, const float *__restrict__ out,
                                                  const float *__restrict__ E, const float *__restrict__ R,
                                                  const float *__restrict__ wb, float *__restrict__ dout,
                                                  double *__restrict__ loss) {
  __shared__ double sh[256];
  const long long n = (long long)B * kP * nh;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const long long b = i / ((long long)kP * nh);
    const float w = wb[b];
    const float orr = out[2 * i], oi = out[2 * i + 1], er = E[2 * i], ei = E[2 * i + 1];
    const float dr = orr * er - oi * ei - R[2 * i], di = orr * ei + oi * er - R[2 * i + 1];
    e = (double)w * ((double)dr * dr + (double)di * di);
    const float g = 2.f * w / (float)((long long)B * kP);
    // d|d|^2 / do = 2 d conj(E): (dr er + di ei, di er - dr ei)
    dout[2 * i] = g * (dr * er + di * ei);
    dout[2 * i + 1] = g * (di * er - dr * ei);
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, sh[0] / ((double)B * kP));
}

// out[i] = sum over s of parts[s][i], in order
__global__ void k_sum_parts(int ns, long long cnt, const float *__restrict__ parts, float *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  float s = 0.f;
  for (int k = 0; k < ns; ++k) s += parts[(long long)k * cnt + i];
  out[i] = s;
}

__global__ void k_adam(long long n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// The fused forward of inference, one wavefront per map row (64 positions):
// out_row (64 x W) = in_row (64 x kin) Wt (kin x W) + bias [+ Fxi (64 x M2)
// Xb_row (M2 x W)], GELU when act, stored; with Xf_next also Xf_next_row (M2
// x W) = Fx (M2 x 64) out_row, whose B operand is the product's own output
// (a 16 x 16 block's output layout, D[4(l/16) + q][l%16], is the B operand
// layout of the K step over those 16 positions). in_row may take the input
// affine (mu, isd) on load. The same products and BF16 rounding points as
// the batched-product path of training.
// Intermediates (the input when TI is uint16_t, Xb, out, Xf_next) are BF16
// bits: each is only ever a product operand, so storing it rounded equals
// rounding it where it is loaded.
template <int W, int M2, typename TI>
__global__ void __launch_bounds__(256) k_fno_row(long long rows, int kin, const TI *__restrict__ in,
                                                 const float *__restrict__ mu, const float *__restrict__ isd,
                                                 const float *__restrict__ Wt, const float *__restrict__ bias,
                                                 const float *__restrict__ Fxi, const uint16_t *__restrict__ Xb,
                                                 const float *__restrict__ Fx, int act, uint16_t *__restrict__ out,
                                                 uint16_t *__restrict__ Xf_next) {
  static_assert(W % 16 == 0 && M2 == 16, "blocks of 16");
  constexpr int NB = W / 16;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long long rr = (long long)blockIdx.x * 4 + wave;
  if (rr >= rows) return;
  const int r = l & 15, kq = (l >> 4) * 4;
  v4f d[4][NB] = {};
  const TI *inr = in + rr * kN * kin;
  for (int k0 = 0; k0 < kin; k0 += 16) {
    v4s a[4], b[NB];
    if constexpr (sizeof(TI) == 2) {
      // four consecutive BF16 channels of one position: one 8-byte load
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = k0 + kq < kin ? *reinterpret_cast<const v4s *>(inr + (16 * i + r) * kin + k0 + kq) : v4s{0, 0, 0, 0};
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kq + q;
      const bool vk = k < kin;
      if constexpr (sizeof(TI) == 4) {
        const float m0 = (vk && mu) ? mu[k] : 0.f, s0 = (vk && isd) ? isd[k] : 1.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          a[i][q] = vk ? f2bf(((float)inr[(16 * i + r) * kin + k] - m0) * s0) : (short)0;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j][q] = vk ? f2bf(Wt[k * W + 16 * j + r]) : (short)0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
  }
  if (Xb) {
    v4s a[4], b[NB];
    const uint16_t *xb = Xb + rr * M2 * W;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = kq + q;
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i][q] = f2bf(Fxi[(16 * i + r) * M2 + k]);
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j][q] = (short)xb[k * W + 16 * j + r];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
  }
  // the row's outputs go through this wave's LDS tile and leave in 16-byte
  // stores (a lane's values are 2-byte pieces of 16 rows)
  __shared__ uint4 tile_o[4][kN * W * 2 / 16], tile_x[4][M2 * W * 2 / 16];
  uint16_t *so = reinterpret_cast<uint16_t *>(tile_o[wave]);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int c = 16 * j + r;
      const float bc = bias[c];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float v = d[i][j][q] + bc;
        if (act) v = gelu(v);
        d[i][j][q] = v;
        so[(16 * i + kq + q) * W + c] = (uint16_t)f2bf(v);
      }
    }
  // the wave's LDS writes complete before its lanes read each other's
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  {
    uint4 *og = reinterpret_cast<uint4 *>(out + rr * kN * W);
    constexpr int NV = kN * W * 2 / 16;  // 16-byte pieces of the row
#pragma unroll
    for (int v = l; v < NV; v += 64) og[v] = tile_o[wave][v];
  }
  if (Xf_next) {
    v4f e[NB] = {};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v4s a, b[NB];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = f2bf(Fx[r * kN + 16 * i + kq + q]);
#pragma unroll
        for (int j = 0; j < NB; ++j) b[j][q] = f2bf(d[i][j][q]);
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) e[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[j], e[j], 0, 0, 0);
    }
    uint16_t *sx = reinterpret_cast<uint16_t *>(tile_x[wave]);
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) sx[(kq + q) * W + 16 * j + r] = (uint16_t)f2bf(e[j][q]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    uint4 *xg = reinterpret_cast<uint4 *>(Xf_next + rr * M2 * W);
    constexpr int NX = M2 * W * 2 / 16;
#pragma unroll
    for (int v = l; v < NX; v += 64) xg[v] = tile_x[wave][v];
  }
}

// The projection of one map row, transposed so that the first product's
// output is the second's B operand: zq^T (PH x 64) = Q1^T h^T + c1, q =
// gelu(zq), out^T (CO x 64) = Q2^T q^T + c2 (CO <= 16).
template <int W, int PH>
__global__ void __launch_bounds__(256) k_fno_proj_row(long long rows, int co, const uint16_t *__restrict__ h,
                                                      const float *__restrict__ Q1, const float *__restrict__ c1,
                                                      const float *__restrict__ Q2, const float *__restrict__ c2,
                                                      float *__restrict__ out) {
  static_assert(W % 16 == 0 && PH % 16 == 0, "blocks of 16");
  constexpr int MB = PH / 16;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long long rr = (long long)blockIdx.x * 4 + wave;
  if (rr >= rows) return;
  const int r = l & 15, kq = (l >> 4) * 4;
  const uint16_t *hr = h + rr * kN * W;
  v4f z[MB][4] = {};
  for (int k0 = 0; k0 < W; k0 += 16) {
    v4s a[MB], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) b[n] = *reinterpret_cast<const v4s *>(hr + (16 * n + r) * W + k0 + kq);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kq + q;
#pragma unroll
      for (int m = 0; m < MB; ++m) a[m][q] = f2bf(Q1[k * PH + 16 * m + r]);
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) z[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b[n], z[m][n], 0, 0, 0);
  }
  v4f y[4] = {};
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    v4s a, b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ph = 16 * m + kq + q;
      a[q] = r < co ? f2bf(Q2[(16 * m + kq + q) * co + r]) : (short)0;
      const float cb = c1[ph];
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n][q] = f2bf(gelu(z[m][n][q] + cb));
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) y[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[n], y[n], 0, 0, 0);
  }
  float *o = out + rr * kN * co;
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = kq + q;
      if (c < co) o[(16 * n + r) * co + c] = y[n][q] + c2[c];
    }
}


Ix U1(long long s) { return Ix{31, 0, s}; }
Ix U2(int sh, long long hi, long long lo) { return Ix{sh, hi, lo}; }
int lg2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}

// the batched product; TA, TB, TC the operands' and result's storage
// (float, or uint16_t BF16 bits; accumulation into FP32 storage only)
template <typename TA = float, typename TB = float, typename TC = float>
void bgemm(hipStream_t st, int batch, int M, int N, int K, Op A, Op B, OpW C, int acc, const float *bias) {
  const long long tiles = (long long)batch * ((M + 31) / 32) * ((N + 31) / 32);
  if (tiles <= 0) return;
  hipLaunchKernelGGL((k_bgemm<TA, TB, TC>), dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, batch, M, N, K, A,
                     B, C, acc, bias);
}

}  // namespace

// ------------------------------------------------------------------ network

struct felis_fno {
  int cin = 0, cout = 0, w = 32, m = 8, L = 4, ph = 64, cap = 0;
  long long np = 0, step = 0;
  // parameters, gradients and Adam moments: one FP32 buffer each
  float *P = nullptr, *G = nullptr, *Am = nullptr, *Av = nullptr;
  long long o_lw = 0, o_lb = 0, o_W[4] = {}, o_b[4] = {}, o_Wm[4] = {}, o_q1 = 0, o_c1 = 0, o_q2 = 0, o_c2 = 0;
  // transform tables: Fx [2m][N], Fxi [N][2m], Gy [4m][2N], Giy [2N][4m]
  float *Fx = nullptr, *Fxi = nullptr, *Gy = nullptr, *Giy = nullptr;
  float *mu = nullptr, *isd = nullptr;
  // work, for cap maps
  float *x = nullptr;            // affine input [cap][P][cin]
  float *a[4] = {}, *h[5] = {};  // pre-activations and layer inputs [cap][P][w]
  float *Xf = nullptr;           // [cap][N][2][m][w]
  float *Y[4] = {};              // [cap][2m][m][2][w] per layer
  float *Zs = nullptr;           // mixed modes, same layout
  float *Xb = nullptr;           // [cap][N][2][m][w]
  float *zq = nullptr, *q = nullptr, *out = nullptr;  // [cap][P][ph], [cap][P][ph], [cap][P][cout]
  float *dout = nullptr, *dq = nullptr, *dh = nullptr, *da = nullptr, *dXf = nullptr, *dY = nullptr,
        *dZ = nullptr, *dXb = nullptr, *E = nullptr, *R = nullptr, *wb = nullptr;
  float *part = nullptr;  // per-map partial weight gradients [cap][largest weight matrix]
  double *loss = nullptr;
  bool grads = false;
  std::vector<void *> allocs;
  std::string err;
  hipStream_t stream = nullptr;
};

namespace {

template <typename T>
int fno_alloc(felis_fno *n, T **p, size_t elems) {
  if (hipMalloc((void **)p, std::max<size_t>(elems, 1) * sizeof(T)) != hipSuccess) {
    n->err = "hipMalloc";
    return FELIS_ERR_HIP;
  }
  n->allocs.push_back((void *)*p);
  return FELIS_OK;
}

int fno_grad_buffers(felis_fno *n) {
  if (n->grads) return FELIS_OK;
  const long long B = n->cap, P = kP, w = n->w, m = n->m;
  const long long spec = B * kN * 2 * m * w, modes = B * 2 * m * m * 2 * w;
  int e;
  if ((e = fno_alloc(n, &n->G, n->np)) || (e = fno_alloc(n, &n->Am, n->np)) || (e = fno_alloc(n, &n->Av, n->np)) ||
      (e = fno_alloc(n, &n->dout, B * P * n->cout)) || (e = fno_alloc(n, &n->dq, B * P * n->ph)) ||
      (e = fno_alloc(n, &n->dh, B * P * w)) || (e = fno_alloc(n, &n->da, B * P * w)) ||
      (e = fno_alloc(n, &n->dXf, spec)) || (e = fno_alloc(n, &n->dY, modes)) || (e = fno_alloc(n, &n->dZ, modes)) ||
      (e = fno_alloc(n, &n->dXb, spec)) || (e = fno_alloc(n, &n->E, B * P * n->cout)) ||
      (e = fno_alloc(n, &n->R, B * P * n->cout)) || (e = fno_alloc(n, &n->wb, B)) || (e = fno_alloc(n, &n->loss, 1)) ||
      (e = fno_alloc(n, &n->part, B * std::max({(long long)n->cin * w, w * w, w * n->ph, (long long)n->ph * n->cout}))))
    return e;
  if (hipMemsetAsync(n->Am, 0, n->np * sizeof(float), n->stream) != hipSuccess ||
      hipMemsetAsync(n->Av, 0, n->np * sizeof(float), n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  n->grads = true;
  return FELIS_OK;
}

// operand helpers over the layouts above (B maps of P positions)
// a channel-last map tensor as (row = b * P + p, col = c)
Op maps(const float *p, int c) { return Op{p, U1(0), U1(c), U1(1)}; }
OpW mapsW(float *p, int c) { return OpW{p, U1(0), U1(c), U1(1)}; }
// a weight matrix (rows x cols, row-major) shared by every batch entry, and its transpose
Op mat(const float *p, int cols) { return Op{p, U1(0), U1(cols), U1(1)}; }
Op matT(const float *p, int cols) { return Op{p, U1(0), U1(1), U1(cols)}; }

// the gradient of a weight shared by every position: out (M x N) = A^T Bm
// over rows = B maps of P positions, A [rows][lda], Bm [rows][ldb]; per map
// into n->part, then the maps' parts summed in order
void wgrad(felis_fno *n, int M, int N, int B, const float *A, int lda, const float *Bm, int ldb, float *out) {
  bgemm(n->stream, B, M, N, kP, Op{A, U1((long long)kP * lda), U1(1), U1(lda)},
        Op{Bm, U1((long long)kP * ldb), U1(ldb), U1(1)}, OpW{n->part, U1((long long)M * N), U1(N), U1(1)}, 0, nullptr);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk((long long)M * N)), dim3(256), 0, n->stream, B, (long long)M * N, n->part,
                     out);
}

// forward of B maps (x already affine-mapped into n->x)
void fno_forward(felis_fno *n, int B) {
  const hipStream_t st = n->stream;
  const int w = n->w, m = n->m, ph = n->ph, lw = lg2(w), lm = lg2(m);
  const long long rows = (long long)B * kP, rowsN = (long long)B * kN;
  const long long sX = 2LL * m * w;                // Xf / Xb: per (b, y) row
  const long long sY = 2LL * m * m * 2 * w;        // Y / Z: per map
  // lift
  bgemm(st, 1, (int)rows, w, n->cin, maps(n->x, n->cin), mat(n->P + n->o_lw, w), mapsW(n->h[0], w), 0,
        n->P + n->o_lb);
  for (int l = 0; l < n->L; ++l) {
    const float *h = n->h[l];
    // x transform per (b, y): Xf[b,y] (2m x w) = Fx (2m x N) h[b,y] (N x w)
    bgemm(st, (int)rowsN, 2 * m, w, kN, Op{n->Fx, U1(0), U1(kN), U1(1)}, Op{h, U1((long long)kN * w), U1(w), U1(1)},
          OpW{n->Xf, U1(sX), U1(w), U1(1)}, 0, nullptr);
    // y transform per map: Y[b] ((ky,r'), (kx,c)) = Gy (4m x 2N) Xf[b] ((y,r), (kx,c)); stored [ky][kx][r'][c]
    bgemm(st, B, 4 * m, m * w, 2 * kN, Op{n->Gy, U1(0), U1(2 * kN), U1(1)},
          Op{n->Xf, U1((long long)kN * sX), U1((long long)m * w), U1(1)},
          OpW{n->Y[l], U1(sY), U2(1, (long long)m * 2 * w, w), U2(lw, 2LL * w, 1)}, 0, nullptr);
    // mode mixing per mode: Z[b, mode] (1 x 2w) = Y[b, mode] (1 x 2w) Wm[mode] (2w x 2w)
    bgemm(st, 2 * m * m, B, 2 * w, 2 * w, Op{n->Y[l], U1(2LL * w), U1(sY), U1(1)},
          Op{n->P + n->o_Wm[l], U1(4LL * w * w), U1(2LL * w), U1(1)}, OpW{n->Zs, U1(2LL * w), U1(sY), U1(1)}, 0,
          nullptr);
    // inverse y per map: Xb[b] ((y,r), (kx,c)) = Giy (2N x 4m) Z[b] ((ky,r'), (kx,c))
    bgemm(st, B, 2 * kN, m * w, 4 * m, Op{n->Giy, U1(0), U1(4 * m), U1(1)},
          Op{n->Zs, U1(sY), U2(1, (long long)m * 2 * w, w), U2(lw, 2LL * w, 1)},
          OpW{n->Xb, U1((long long)kN * sX), U1((long long)m * w), U1(1)}, 0, nullptr);
    // a = h W + b, then a[b,y] += Fxi (N x 2m) Xb[b,y] (2m x w)
    float *a = n->a[l];
    bgemm(st, 1, (int)rows, w, w, maps(h, w), mat(n->P + n->o_W[l], w), mapsW(a, w), 0, n->P + n->o_b[l]);
    bgemm(st, (int)rowsN, kN, w, 2 * m, Op{n->Fxi, U1(0), U1(2 * m), U1(1)}, Op{n->Xb, U1(sX), U1(w), U1(1)},
          OpW{a, U1((long long)kN * w), U1(w), U1(1)}, 1, nullptr);
    if (l < n->L - 1)
      hipLaunchKernelGGL(k_gelu, dim3(nblk(rows * w)), dim3(256), 0, st, rows * w, a, n->h[l + 1]);
    (void)lm;
  }
  // projection
  const float *hl = n->L > 0 ? n->a[n->L - 1] : n->h[0];
  bgemm(st, 1, (int)rows, ph, w, maps(hl, w), mat(n->P + n->o_q1, ph), mapsW(n->zq, ph), 0, n->P + n->o_c1);
  hipLaunchKernelGGL(k_gelu, dim3(nblk(rows * ph)), dim3(256), 0, st, rows * ph, n->zq, n->q);
  bgemm(st, 1, (int)rows, n->cout, ph, maps(n->q, ph), mat(n->P + n->o_q2, n->cout), mapsW(n->out, n->cout), 0,
        n->P + n->o_c2);
}

// gradients of the loss for B maps (after fno_forward and the loss kernel
// filled n->dout): into n->G
void fno_backward(felis_fno *n, int B) {
  const hipStream_t st = n->stream;
  const int w = n->w, m = n->m, ph = n->ph, lw = lg2(w);
  const long long rows = (long long)B * kP, rowsN = (long long)B * kN;
  const long long sX = 2LL * m * w, sY = 2LL * m * m * 2 * w;
  float *G = n->G;
  // out = q Q2 + c2
  wgrad(n, ph, n->cout, B, n->q, ph, n->dout, n->cout, G + n->o_q2);
  hipLaunchKernelGGL(k_colsum, dim3(n->cout), dim3(256), 0, st, rows, n->cout, n->dout, G + n->o_c2);
  bgemm(st, 1, (int)rows, ph, n->cout, maps(n->dout, n->cout), matT(n->P + n->o_q2, n->cout), mapsW(n->dq, ph), 0,
        nullptr);
  // q = gelu(zq), zq = h Q1 + c1 (dq becomes dzq in place)
  hipLaunchKernelGGL(k_gelu_back, dim3(nblk(rows * ph)), dim3(256), 0, st, rows * ph, n->dq, n->zq, n->dq);
  const float *hl = n->L > 0 ? n->a[n->L - 1] : n->h[0];
  wgrad(n, w, ph, B, hl, w, n->dq, ph, G + n->o_q1);
  hipLaunchKernelGGL(k_colsum, dim3(ph), dim3(256), 0, st, rows, ph, n->dq, G + n->o_c1);
  // dh = d(output of the last Fourier layer)
  bgemm(st, 1, (int)rows, w, ph, maps(n->dq, ph), matT(n->P + n->o_q1, ph), mapsW(n->dh, w), 0, nullptr);
  for (int l = n->L - 1; l >= 0; --l) {
    const float *h = n->h[l];
    // da: through the GELU after every layer but the last
    if (l < n->L - 1)
      hipLaunchKernelGGL(k_gelu_back, dim3(nblk(rows * w)), dim3(256), 0, st, rows * w, n->dh, n->a[l], n->da);
    else
      (void)hipMemcpyAsync(n->da, n->dh, rows * w * sizeof(float), hipMemcpyDeviceToDevice, st);
    // a = h W + b
    wgrad(n, w, w, B, h, w, n->da, w, G + n->o_W[l]);
    hipLaunchKernelGGL(k_colsum, dim3(w), dim3(256), 0, st, rows, w, n->da, G + n->o_b[l]);
    bgemm(st, 1, (int)rows, w, w, maps(n->da, w), matT(n->P + n->o_W[l], w), mapsW(n->dh, w), 0, nullptr);
    // a[b,y] += Fxi Xb[b,y]: dXb[b,y] = Fxi^T da[b,y]
    bgemm(st, (int)rowsN, 2 * m, w, kN, Op{n->Fxi, U1(0), U1(1), U1(2 * m)},
          Op{n->da, U1((long long)kN * w), U1(w), U1(1)}, OpW{n->dXb, U1(sX), U1(w), U1(1)}, 0, nullptr);
    // Xb[b] = Giy Z[b]: dZ[b] = Giy^T dXb[b]
    bgemm(st, B, 4 * m, m * w, 2 * kN, Op{n->Giy, U1(0), U1(1), U1(4 * m)},
          Op{n->dXb, U1((long long)kN * sX), U1((long long)m * w), U1(1)},
          OpW{n->dZ, U1(sY), U2(1, (long long)m * 2 * w, w), U2(lw, 2LL * w, 1)}, 0, nullptr);
    // Z[b, mode] = Y[b, mode] Wm[mode]: dWm[mode] = Y[:, mode]^T dZ[:, mode], dY = dZ Wm^T
    bgemm(st, 2 * m * m, 2 * w, 2 * w, B, Op{n->Y[l], U1(2LL * w), U1(1), U1(sY)}, Op{n->dZ, U1(2LL * w), U1(sY), U1(1)},
          OpW{G + n->o_Wm[l], U1(4LL * w * w), U1(2LL * w), U1(1)}, 0, nullptr);
    bgemm(st, 2 * m * m, B, 2 * w, 2 * w, Op{n->dZ, U1(2LL * w), U1(sY), U1(1)},
          Op{n->P + n->o_Wm[l], U1(4LL * w * w), U1(1), U1(2LL * w)}, OpW{n->dY, U1(2LL * w), U1(sY), U1(1)}, 0,
          nullptr);
    // Y[b] = Gy Xf[b]: dXf[b] = Gy^T dY[b]
    bgemm(st, B, 2 * kN, m * w, 4 * m, Op{n->Gy, U1(0), U1(1), U1(2 * kN)},
          Op{n->dY, U1(sY), U2(1, (long long)m * 2 * w, w), U2(lw, 2LL * w, 1)},
          OpW{n->dXf, U1((long long)kN * sX), U1((long long)m * w), U1(1)}, 0, nullptr);
    // Xf[b,y] = Fx h[b,y]: dh[b,y] += Fx^T dXf[b,y]
    bgemm(st, (int)rowsN, kN, w, 2 * m, Op{n->Fx, U1(0), U1(1), U1(kN)}, Op{n->dXf, U1(sX), U1(w), U1(1)},
          OpW{n->dh, U1((long long)kN * w), U1(w), U1(1)}, 1, nullptr);
  }
  // lift: h0 = x Wl + bl
  wgrad(n, n->cin, w, B, n->x, n->cin, n->dh, w, G + n->o_lw);
  hipLaunchKernelGGL(k_colsum, dim3(w), dim3(256), 0, st, rows, w, n->dh, G + n->o_lb);
}

// the fused inference forward (k_fno_row, k_fno_proj_row) for the shape it
// is built for: width 32, 8 modes, projection 64, at most 16 outputs
bool fno_fusable(const felis_fno *n) { return n->w == 32 && n->m == 8 && n->ph == 64 && n->cout <= 16; }

// X: FP32 input maps (the input affine applied on load), or XB: the BF16
// bits of the affine-applied inputs
void fno_forward_fused(felis_fno *n, int B, const float *X, const uint16_t *XB = nullptr) {
  const hipStream_t st = n->stream;
  constexpr int w = 32, m = 8;
  const int lw = lg2(w);
  const long long rows = (long long)B * kN;
  const long long sX = 2LL * m * w, sY = 2LL * m * m * 2 * w;
  const unsigned nb = (unsigned)((rows + 3) / 4);
  // BF16 intermediates in the training path's FP32 buffers
  uint16_t *hA = reinterpret_cast<uint16_t *>(n->h[0]), *hB = reinterpret_cast<uint16_t *>(n->a[0]);
  uint16_t *Xf = reinterpret_cast<uint16_t *>(n->Xf), *Yb = reinterpret_cast<uint16_t *>(n->Y[0]),
           *Zb = reinterpret_cast<uint16_t *>(n->Zs), *Xb = reinterpret_cast<uint16_t *>(n->Xb);
  if (XB)
    hipLaunchKernelGGL((k_fno_row<32, 16, uint16_t>), dim3(nb), dim3(256), 0, st, rows, n->cin, XB, nullptr, nullptr,
                       n->P + n->o_lw, n->P + n->o_lb, nullptr, nullptr, n->Fx, 0, hA, Xf);
  else
    hipLaunchKernelGGL((k_fno_row<32, 16, float>), dim3(nb), dim3(256), 0, st, rows, n->cin, X, n->mu, n->isd,
                       n->P + n->o_lw, n->P + n->o_lb, nullptr, nullptr, n->Fx, 0, hA, Xf);
  for (int l = 0; l < n->L; ++l) {
    hipLaunchKernelGGL((k_fno_ytr<false>), dim3(B), dim3(256), 0, st, B, n->Gy, Xf, Yb);
    hipLaunchKernelGGL(k_fno_mix, dim3((B + 63) / 64, 2 * m * m), dim3(256), 0, st, B, Yb, n->P + n->o_Wm[l], Zb);
    hipLaunchKernelGGL((k_fno_ytr<true>), dim3(B), dim3(256), 0, st, B, n->Giy, Zb, Xb);
    const bool last = l == n->L - 1;
    hipLaunchKernelGGL((k_fno_row<32, 16, uint16_t>), dim3(nb), dim3(256), 0, st, rows, w, hA, nullptr, nullptr,
                       n->P + n->o_W[l], n->P + n->o_b[l], n->Fxi, Xb, n->Fx, last ? 0 : 1, hB,
                       last ? nullptr : Xf);
    std::swap(hA, hB);
  }
  hipLaunchKernelGGL((k_fno_proj_row<32, 64>), dim3(nb), dim3(256), 0, st, rows, n->cout, hA, n->P + n->o_q1,
                     n->P + n->o_c1, n->P + n->o_q2, n->P + n->o_c2, n->out);
}

int fno_input(felis_fno *n, int B, const float *X) {
  const long long cnt = (long long)B * kP * n->cin;
  hipLaunchKernelGGL(k_affine, dim3(nblk(cnt)), dim3(256), 0, n->stream, cnt, n->cin, X, n->mu, n->isd, n->x);
  return hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP;
}

}  // namespace

// forward of B device maps for the jump path: X [B][P][cin] (device), the
// outputs in the network's own buffer (returned)
int felis_fno_forward_dev(felis_fno *n, int B, const float *X, const float **out, hipStream_t st) {
  if (!n || B < 1 || B > n->cap || !X) return FELIS_ERR_ARG;
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return FELIS_ERR_HIP;
  // the network's stream follows the caller's and the caller's waits for it
  (void)hipEventRecord(ev, st);
  (void)hipStreamWaitEvent(n->stream, ev, 0);
  int e = FELIS_OK;
  if (fno_fusable(n)) {
    fno_forward_fused(n, B, X);
  } else {
    e = fno_input(n, B, X);
    if (!e) fno_forward(n, B);
  }
  (void)hipEventRecord(ev, n->stream);
  (void)hipStreamWaitEvent(st, ev, 0);
  (void)hipEventDestroy(ev);
  if (out) *out = n->out;
  return e ? e : (hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP);
}

// forward of B device maps given as the BF16 bits of the affine-applied
// inputs (the fused shape only), as felis_fno_forward_dev
int felis_fno_forward_bf16_dev(felis_fno *n, int B, const uint16_t *XB, const float **out, hipStream_t st) {
  if (!n || B < 1 || B > n->cap || !XB || !fno_fusable(n)) return FELIS_ERR_ARG;
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return FELIS_ERR_HIP;
  (void)hipEventRecord(ev, st);
  (void)hipStreamWaitEvent(n->stream, ev, 0);
  fno_forward_fused(n, B, nullptr, XB);
  (void)hipEventRecord(ev, n->stream);
  (void)hipStreamWaitEvent(st, ev, 0);
  (void)hipEventDestroy(ev);
  if (out) *out = n->out;
  return hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP;
}

int felis_fno_input_affine_dev(const felis_fno *n, const float **mu, const float **isd) {
  if (!n) return FELIS_ERR_ARG;
  if (mu) *mu = n->mu;
  if (isd) *isd = n->isd;
  return fno_fusable(n) ? FELIS_OK : FELIS_ERR_UNSUPPORTED;
}

int felis_fno_channels(const felis_fno *n, int *cin, int *cout) {
  if (!n) return FELIS_ERR_ARG;
  if (cin) *cin = n->cin;
  if (cout) *cout = n->cout;
  return FELIS_OK;
}

extern "C" {

int felis_fno_create(int cin, int cout, int width, int modes, int layers, int proj, int capacity, uint64_t seed,
                     felis_fno **out) {
  if (!out || cin < 1 || cout < 1 || cout > 64 || capacity < 1 || layers < 1 || layers > 4 || proj < 1 ||
      proj > 256 || width < 2 || width > 128 || (width & (width - 1)) || modes < 1 || modes > kN / 2 ||
      (modes & (modes - 1)))
    return FELIS_ERR_ARG;
  felis_fno *n = new felis_fno();
  n->cin = cin;
  n->cout = cout;
  n->w = width;
  n->m = modes;
  n->L = layers;
  n->ph = proj;
  n->cap = capacity;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
    delete n;
    return FELIS_ERR_HIP;
  }
  const long long w = width, m = modes;
  long long o = 0;
  n->o_lw = o; o += (long long)cin * w;
  n->o_lb = o; o += w;
  for (int l = 0; l < layers; ++l) {
    n->o_W[l] = o; o += w * w;
    n->o_b[l] = o; o += w;
    n->o_Wm[l] = o; o += 2 * m * m * 4 * w * w;
  }
  n->o_q1 = o; o += w * proj;
  n->o_c1 = o; o += proj;
  n->o_q2 = o; o += (long long)proj * cout;
  n->o_c2 = o; o += cout;
  n->np = o;
  // initial weights: uniform +-1/sqrt(fan_in) (mode mixing: +-1/(2w)), zero biases
  std::vector<float> P((size_t)n->np, 0.f);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  auto fill = [&](long long off, long long cnt, double a) {
    for (long long i = 0; i < cnt; ++i) P[(size_t)(off + i)] = (float)(a * uni());
  };
  fill(n->o_lw, (long long)cin * w, 1.0 / std::sqrt((double)cin));
  for (int l = 0; l < layers; ++l) {
    fill(n->o_W[l], w * w, 1.0 / std::sqrt((double)w));
    fill(n->o_Wm[l], 2 * m * m * 4 * w * w, 1.0 / (2.0 * w));
  }
  fill(n->o_q1, w * proj, 1.0 / std::sqrt((double)w));
  fill(n->o_q2, (long long)proj * cout, 1.0 / std::sqrt((double)proj));
  // transform tables
  std::vector<float> Fx(2 * m * kN), Fxi(kN * 2 * m), Gy(4 * m * 2 * kN), Giy(2 * kN * 4 * m);
  const double tw = 2.0 * M_PI / kN;
  for (int kx = 0; kx < m; ++kx)
    for (int x = 0; x < kN; ++x) {
      const double ph_ = tw * kx * x, c = std::cos(ph_), sn = std::sin(ph_);
      Fx[(0 * m + kx) * kN + x] = (float)c;
      Fx[(1 * m + kx) * kN + x] = (float)-sn;
      const double al = (kx == 0 ? 1.0 : 2.0) / ((double)kN * kN);
      Fxi[x * 2 * m + 0 * m + kx] = (float)(al * c);
      Fxi[x * 2 * m + 1 * m + kx] = (float)(-al * sn);
    }
  for (int kk = 0; kk < 2 * m; ++kk) {
    const int ky = kk < m ? kk : kk - 2 * (int)m;
    for (int y = 0; y < kN; ++y) {
      const double ph_ = tw * ky * y, c = std::cos(ph_), sn = std::sin(ph_);
      // forward e^{-i ph}: (Yr, Yi) = (c Xr + s Xi, -s Xr + c Xi)
      Gy[(2 * kk + 0) * 2 * kN + 2 * y + 0] = (float)c;
      Gy[(2 * kk + 0) * 2 * kN + 2 * y + 1] = (float)sn;
      Gy[(2 * kk + 1) * 2 * kN + 2 * y + 0] = (float)-sn;
      Gy[(2 * kk + 1) * 2 * kN + 2 * y + 1] = (float)c;
      // inverse e^{+i ph}: (Ur, Ui) = (c Zr - s Zi, s Zr + c Zi)
      Giy[(2 * y + 0) * 4 * m + 2 * kk + 0] = (float)c;
      Giy[(2 * y + 0) * 4 * m + 2 * kk + 1] = (float)-sn;
      Giy[(2 * y + 1) * 4 * m + 2 * kk + 0] = (float)sn;
      Giy[(2 * y + 1) * 4 * m + 2 * kk + 1] = (float)c;
    }
  }
  const long long B = capacity, Pn = kP;
  const long long spec = B * kN * 2 * m * w, modesz = B * 2 * m * m * 2 * w;
  int e;
  if ((e = fno_alloc(n, &n->P, n->np)) || (e = fno_alloc(n, &n->Fx, Fx.size())) ||
      (e = fno_alloc(n, &n->Fxi, Fxi.size())) || (e = fno_alloc(n, &n->Gy, Gy.size())) ||
      (e = fno_alloc(n, &n->Giy, Giy.size())) || (e = fno_alloc(n, &n->mu, cin)) || (e = fno_alloc(n, &n->isd, cin)) ||
      (e = fno_alloc(n, &n->x, B * Pn * cin)) || (e = fno_alloc(n, &n->Xf, spec)) || (e = fno_alloc(n, &n->Zs, modesz)) ||
      (e = fno_alloc(n, &n->Xb, spec)) || (e = fno_alloc(n, &n->zq, B * Pn * proj)) ||
      (e = fno_alloc(n, &n->q, B * Pn * proj)) || (e = fno_alloc(n, &n->out, B * Pn * cout))) {
    felis_fno_destroy(n);
    return e;
  }
  for (int l = 0; l < layers && !e; ++l)
    e = fno_alloc(n, &n->a[l], B * Pn * w) || fno_alloc(n, &n->Y[l], modesz) ? FELIS_ERR_HIP : FELIS_OK;
  for (int l = 0; l < layers && !e; ++l) e = fno_alloc(n, &n->h[l], B * Pn * w);
  if (e) {
    felis_fno_destroy(n);
    return e;
  }
  std::vector<float> mu(cin, 0.f), isd(cin, 1.f);
  if (hipMemcpy(n->P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->Fx, Fx.data(), Fx.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->Fxi, Fxi.data(), Fxi.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->Gy, Gy.data(), Gy.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->Giy, Giy.data(), Giy.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->mu, mu.data(), cin * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->isd, isd.data(), cin * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    felis_fno_destroy(n);
    return FELIS_ERR_HIP;
  }
  *out = n;
  return FELIS_OK;
}

int felis_fno_destroy(felis_fno *n) {
  if (!n) return FELIS_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *p : n->allocs) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return FELIS_OK;
}

const char *felis_fno_last_error(const felis_fno *n) { return n ? n->err.c_str() : "null network"; }

int felis_fno_shape(const felis_fno *n, int *dims) {
  if (!n || !dims) return FELIS_ERR_ARG;
  dims[0] = n->cin;
  dims[1] = n->cout;
  dims[2] = n->w;
  dims[3] = n->m;
  dims[4] = n->L;
  dims[5] = n->ph;
  dims[6] = n->cap;
  return FELIS_OK;
}

int64_t felis_fno_param_count(const felis_fno *n) { return n ? n->np : -1; }

int felis_fno_get_params(felis_fno *n, float *P) {
  if (!n || !P) return FELIS_ERR_ARG;
  if (hipMemcpyAsync(P, n->P, n->np * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
      hipStreamSynchronize(n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_fno_set_params(felis_fno *n, const float *P) {
  if (!n || !P) return FELIS_ERR_ARG;
  if (hipMemcpyAsync(n->P, P, n->np * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess ||
      hipStreamSynchronize(n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_fno_get_grad(felis_fno *n, float *G) {
  if (!n || !G) return FELIS_ERR_ARG;
  if (!n->grads) return FELIS_ERR_STATE;
  if (hipMemcpyAsync(G, n->G, n->np * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
      hipStreamSynchronize(n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_fno_set_input_affine(felis_fno *n, const float *mu, const float *isd) {
  if (!n || !mu || !isd) return FELIS_ERR_ARG;
  if (hipMemcpyAsync(n->mu, mu, n->cin * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess ||
      hipMemcpyAsync(n->isd, isd, n->cin * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess ||
      hipStreamSynchronize(n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

// X [B][P][cin] host, Y [B][P][cout] host; the fused forward where the
// shape allows it (the one the jumps use), else the training path's
int felis_fno_forward(felis_fno *n, int B, const float *X, float *Y) {
  if (!n || !X || !Y || B < 1 || B > n->cap) return FELIS_ERR_ARG;
  const size_t nx = (size_t)B * kP * n->cin, ny = (size_t)B * kP * n->cout;
  float *dX = nullptr;
  if (hipMalloc((void **)&dX, nx * sizeof(float)) != hipSuccess) return FELIS_ERR_HIP;
  int e = hipMemcpyAsync(dX, X, nx * sizeof(float), hipMemcpyHostToDevice, n->stream) == hipSuccess ? FELIS_OK
                                                                                                    : FELIS_ERR_HIP;
  if (!e && !fno_fusable(n)) e = fno_input(n, B, dX);
  if (!e) {
    if (fno_fusable(n)) fno_forward_fused(n, B, dX);
    else fno_forward(n, B);
    if (hipMemcpyAsync(Y, n->out, ny * sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
        hipStreamSynchronize(n->stream) != hipSuccess)
      e = FELIS_ERR_HIP;
  }
  (void)hipStreamSynchronize(n->stream);
  (void)hipFree(dX);
  return e;
}

// One training step on B maps from device or host buffers (dev != 0: device
// pointers): X [B][P][cin], E and R [B][P][cout] (the jump's emission and the
// fine steps' residual, complex per harmonic), wb [B] map weights. Forward,
// the loss, backward and (lr > 0) an Adam step. loss_out: the weighted loss.
int felis_fno_train_step(felis_fno *n, int B, const float *X, const float *E, const float *R, const float *wb,
                         int dev, float lr, double *loss_out) {
  if (!n || !X || !E || !R || !wb || B < 1 || B > n->cap || (n->cout & 1)) return FELIS_ERR_ARG;
  int e = fno_grad_buffers(n);
  if (e) return e;
  const size_t nx = (size_t)B * kP * n->cin, ny = (size_t)B * kP * n->cout;
  const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const float *Xd = X;
  float *tmp = nullptr;
  if (!dev) {
    if (hipMalloc((void **)&tmp, nx * sizeof(float)) != hipSuccess) return FELIS_ERR_HIP;
    if (hipMemcpyAsync(tmp, X, nx * sizeof(float), hipMemcpyHostToDevice, n->stream) != hipSuccess) e = FELIS_ERR_HIP;
    Xd = tmp;
  }
  if (!e && (hipMemcpyAsync(n->E, E, ny * sizeof(float), kind, n->stream) != hipSuccess ||
             hipMemcpyAsync(n->R, R, ny * sizeof(float), kind, n->stream) != hipSuccess ||
             hipMemcpyAsync(n->wb, wb, B * sizeof(float), kind, n->stream) != hipSuccess ||
             hipMemsetAsync(n->loss, 0, sizeof(double), n->stream) != hipSuccess))
    e = FELIS_ERR_HIP;
  if (!e) e = fno_input(n, B, Xd);
  if (!e) {
    fno_forward(n, B);
    const long long cnt = (long long)B * kP * (n->cout / 2);
    hipLaunchKernelGGL(k_fno_loss, dim3(nblk(cnt)), dim3(256), 0, n->stream, B, n->cout / 2, n->out, n->E, n->R,
                       n->wb, n->dout, n->loss);
    fno_backward(n, B);
    if (lr > 0.f) {
      n->step++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
      hipLaunchKernelGGL(k_adam, dim3(nblk(n->np)), dim3(256), 0, n->stream, n->np, n->P, n->G, n->Am, n->Av, lr, b1,
                         b2, 1e-8f, c1, c2);
    }
    double l = 0.0;
    if (hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
        hipStreamSynchronize(n->stream) != hipSuccess)
      e = FELIS_ERR_HIP;
    if (loss_out) *loss_out = l;
  }
  (void)hipStreamSynchronize(n->stream);
  if (tmp) (void)hipFree(tmp);
  if (!e && hipGetLastError() != hipSuccess) e = FELIS_ERR_HIP;
  return e;
}

}  // extern "C"

// ------------------------------------------------------------------------
// The network's maps of a jump's state. The field grid (nx = ny) is cut into
// 64 x 64 cells of b x b field cells (b = nx / 64), centred (first cell at
// offset o = (nx - 64 b) / 2). Per slice, from the x polarisation:
//   E_h  the jump's emission 2 Sw / H at the slot the field was solved at,
//        inside that slot's source window (zero outside);
//   U_h  the jump's field there;
//   R_h  (training) the fine steps' field minus the jump's.
// The input maps are E_h / s_h and U_h / u_h (s_h, u_h their rms over the
// 64 x 64 cells), then 8 of the slice's features (felis_slice_features:
// current, field intensity at the beam, |b_1|, coherent fraction, |b_3|,
// |b_5|, energy spread, energy modulation), log10 s_1, log10 u_1 and the
// cell's coordinates in [-1, 1]; the network's output o_h is the factor by
// which the jump's emission is corrected: U += o_h E_h.
namespace felis {
namespace {

constexpr int kFnoScalars = 12;  // 8 features, 2 scales, 2 coordinates

struct FnoGeom {
  int b, off;  // field cells per map cell, first field cell
};

// one workgroup per (slice, harmonic): the raw maps and their sums.
// s_src: slot of the source (Sw, win2), s_fld: slot of the field (< 1 or >
// slice_n: no field, the maps are zero). scal[sl][3][nh]: sum |E|^2, sum
// |U|^2, and (training) sum over field cells of the fine field's |B|^2.
__global__ void __launch_bounds__(256) k_fno_coarse(GridArgs g, FnoGeom fg, int J, const cpx<float> *__restrict__ A,
                                                    const cpx<float> *__restrict__ Bf,
                                                    const cpx<float> *__restrict__ Sw, const int4 *__restrict__ win2,
                                                    float *__restrict__ Ec, float *__restrict__ Uc,
                                                    float *__restrict__ Rc, float *__restrict__ scal) {
  __shared__ double red[3][4];
  const int sl = blockIdx.x, h = blockIdx.y, nh = g.nh;
  const int s_src = sl + 1, s_fld = sl + 1 + J;
  const bool live = s_fld <= g.slice_n;
  const float k2 = 2.f / (float)g.harm[h];
  const int4 w = win2[s_src];
  const cpx<float> *src = Sw + ((long long)h * g.S + s_src) * g.N;
  const long long fo = ((long long)h * g.S + phys_slot(live ? s_fld : 1, g.ring_base, g.S)) * g.N;
  const float inv = 1.f / (float)(fg.b * fg.b);
  double se = 0.0, su = 0.0, sb = 0.0;
  for (int q = threadIdx.x; q < kP; q += 256) {
    const int I = q % kN, Jc = q / kN;
    cpx<float> e = mk(0.f, 0.f), u = mk(0.f, 0.f), r = mk(0.f, 0.f);
    if (live) {
      for (int dj = 0; dj < fg.b; ++dj)
        for (int di = 0; di < fg.b; ++di) {
          const int i = fg.off + fg.b * I + di, j = fg.off + fg.b * Jc + dj;
          const long long c = (long long)j * g.nx + i;
          const cpx<float> a = A[fo + c];
          if (i >= w.x && i <= w.y && j >= w.z && j <= w.w) e = e + src[c] * k2;
          u = u + a;
          if (Bf) {
            const cpx<float> bv = Bf[fo + c];
            r = r + (bv - a);
            sb += (double)bv.re * bv.re + (double)bv.im * bv.im;
          }
        }
      e = e * inv;
      u = u * inv;
      r = r * inv;
    }
    const long long o = ((long long)sl * kP + q) * 2 * nh + 2 * h;
    Ec[o] = e.re;
    Ec[o + 1] = e.im;
    Uc[o] = u.re;
    Uc[o + 1] = u.im;
    if (Rc) {
      Rc[o] = r.re;
      Rc[o + 1] = r.im;
    }
    se += (double)e.re * e.re + (double)e.im * e.im;
    su += (double)u.re * u.re + (double)u.im * u.im;
  }
  double v[3] = {se, su, sb};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    scal[((long long)sl * 3 + k) * nh + h] = (float)((red[k][0] + red[k][1]) + (red[k][2] + red[k][3]));
  }
}

// The inference path's maps in one pass: one workgroup per (slice,
// harmonic) forms the cells' E_h and U_h exactly as k_fno_coarse does (J = 0:
// the field at the slot the source was solved at), their sums as there,
// then writes the network's input channels of this harmonic (and, for the
// first, the slice's 12 scalar channels) as k_fno_inputs would, with the
// input affine applied and rounded to BF16 as the lift rounds them: the
// BF16 bits the lift would form from k_fno_inputs' FP32 maps. BB: field cells
// per map cell (2 unrolled; 0: fg.b).
template <int BB>
__global__ void __launch_bounds__(256) k_fno_maps_x(GridArgs g, FnoGeom fg, const cpx<float> *__restrict__ A,
                                                    const cpx<float> *__restrict__ Sw, const int4 *__restrict__ win2,
                                                    const float *__restrict__ feat, int nfeat,
                                                    const float *__restrict__ mu, const float *__restrict__ isd,
                                                    uint16_t *__restrict__ X) {
  constexpr int NQ = kP / 256;  // map cells per thread
  __shared__ double red[2][4];
  __shared__ float scs[2];
  const int sl = blockIdx.x, h = blockIdx.y, nh = g.nh, cin = 4 * nh + kFnoScalars;
  const int b = BB > 0 ? BB : fg.b;
  const int s_src = sl + 1;
  const float k2 = 2.f / (float)g.harm[h];
  const int4 w = win2[s_src];
  const cpx<float> *src = Sw + ((long long)h * g.S + s_src) * g.N;
  const long long fo = ((long long)h * g.S + phys_slot(s_src, g.ring_base, g.S)) * g.N;
  const float inv = 1.f / (float)(b * b);
  cpx<float> ev[NQ], uv[NQ];
  double se = 0.0, su = 0.0;
#pragma unroll
  for (int qi = 0; qi < NQ; ++qi) {
    const int q = threadIdx.x + 256 * qi;
    const int I = q % kN, Jc = q / kN;
    cpx<float> e = mk(0.f, 0.f), u = mk(0.f, 0.f);
#pragma unroll
    for (int dj = 0; dj < (BB > 0 ? BB : 64); ++dj) {
      if (BB == 0 && dj >= b) break;
#pragma unroll
      for (int di = 0; di < (BB > 0 ? BB : 64); ++di) {
        if (BB == 0 && di >= b) break;
        const int i = fg.off + b * I + di, j = fg.off + b * Jc + dj;
        const long long c = (long long)j * g.nx + i;
        const cpx<float> a = A[fo + c];
        if (i >= w.x && i <= w.y && j >= w.z && j <= w.w) e = e + src[c] * k2;
        u = u + a;
      }
    }
    e = e * inv;
    u = u * inv;
    ev[qi] = e;
    uv[qi] = u;
    se += (double)e.re * e.re + (double)e.im * e.im;
    su += (double)u.re * u.re + (double)u.im * u.im;
  }
  double v[2] = {se, su};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int k = threadIdx.x;
    scs[k] = (float)((red[k][0] + red[k][1]) + (red[k][2] + red[k][3]));
  }
  __syncthreads();
  // as k_fno_inputs: 1 / rms, the maps scaled (each product rounded to FP32
  // before the affine, as it was stored there)
  const float sE = sqrtf(scs[0] / kP), sU = sqrtf(scs[1] / kP);
  const float is = sE > 0.f ? 1.f / sE : 0.f, iu = sU > 0.f ? 1.f / sU : 0.f;
  auto bf = [&](int ch, float x) {
    const float m0 = mu ? mu[ch] : 0.f, s0 = isd ? isd[ch] : 1.f;
    return (uint16_t)f2bf(__fmul_rn(__fsub_rn(x, m0), s0));
  };
#pragma unroll
  for (int qi = 0; qi < NQ; ++qi) {
    const int q = threadIdx.x + 256 * qi;
    uint16_t *x = X + ((long long)sl * kP + q) * cin;
    x[2 * h] = bf(2 * h, __fmul_rn(ev[qi].re, is));
    x[2 * h + 1] = bf(2 * h + 1, __fmul_rn(ev[qi].im, is));
    x[2 * nh + 2 * h] = bf(2 * nh + 2 * h, __fmul_rn(uv[qi].re, iu));
    x[2 * nh + 2 * h + 1] = bf(2 * nh + 2 * h + 1, __fmul_rn(uv[qi].im, iu));
    if (h == 0) {
      // the slice's scalars (k_fno_inputs' order): 8 features, log10 of the
      // first harmonic's rms emission and field, the cell's coordinates
      const float *f = feat + (long long)sl * nfeat;
      const int c0 = 4 * nh;
      x[c0 + 0] = bf(c0 + 0, f[0]);
      x[c0 + 1] = bf(c0 + 1, f[1]);
      x[c0 + 2] = bf(c0 + 2, f[2]);
      x[c0 + 3] = bf(c0 + 3, f[3]);
      x[c0 + 4] = bf(c0 + 4, nh > 1 ? f[2 + 5] : f[2]);
      x[c0 + 5] = bf(c0 + 5, nh > 2 ? f[2 + 10] : f[2]);
      x[c0 + 6] = bf(c0 + 6, f[1 + 5 * nh]);
      x[c0 + 7] = bf(c0 + 7, f[2 + 5 * nh]);
      x[c0 + 8] = bf(c0 + 8, log10f(sqrtf(scs[0] / kP) + 1e-30f));
      x[c0 + 9] = bf(c0 + 9, log10f(sqrtf(scs[1] / kP) + 1e-30f));
      x[c0 + 10] = bf(c0 + 10, (2.f * (q % kN) + 1.f) / kN - 1.f);
      x[c0 + 11] = bf(c0 + 11, (2.f * (q / kN) + 1.f) / kN - 1.f);
    }
  }
}

// The same for NH harmonics at once (cin = 4 NH + 12 BF16 channels, a 16-byte
// multiple for NH = 1 and 3: 48 bytes a cell at NH = 3), one workgroup of
// 1024 threads per slice: each thread forms its 4 cells' maps of every
// harmonic, the sums go through LDS in a fixed order, and each cell's
// channels leave in 16-byte stores.
template <int NH>
__global__ void __launch_bounds__(1024) k_fno_maps_xs(GridArgs g, FnoGeom fg, const cpx<float> *__restrict__ A,
                                                      const cpx<float> *__restrict__ Sw,
                                                      const int4 *__restrict__ win2, const float *__restrict__ feat,
                                                      int nfeat, const float *__restrict__ mu,
                                                      const float *__restrict__ isd, uint16_t *__restrict__ X) {
  constexpr int NT = 1024, NQ = kP / NT, CIN = 4 * NH + kFnoScalars, NW = NT / 64;
  static_assert(CIN % 8 == 0, "16-byte records");
  __shared__ double red[2 * NH][NW];
  __shared__ float scs[2 * NH];
  const int sl = blockIdx.x, b = fg.b;
  const int s_src = sl + 1;
  const int4 w = win2[s_src];
  const float inv = 1.f / (float)(b * b);
  cpx<float> ev[NQ][NH], uv[NQ][NH];
  double se[NH], su[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    se[h] = su[h] = 0.0;
    const float k2 = 2.f / (float)g.harm[h];
    const cpx<float> *src = Sw + ((long long)h * g.S + s_src) * g.N;
    const cpx<float> *fld = A + ((long long)h * g.S + phys_slot(s_src, g.ring_base, g.S)) * g.N;
#pragma unroll
    for (int qi = 0; qi < NQ; ++qi) {
      const int q = threadIdx.x + NT * qi;
      const int I = q % kN, Jc = q / kN;
      cpx<float> e = mk(0.f, 0.f), u = mk(0.f, 0.f);
      for (int dj = 0; dj < b; ++dj)
        for (int di = 0; di < b; ++di) {
          const int i = fg.off + b * I + di, j = fg.off + b * Jc + dj;
          const long long c = (long long)j * g.nx + i;
          const cpx<float> a = fld[c];
          if (i >= w.x && i <= w.y && j >= w.z && j <= w.w) e = e + src[c] * k2;
          u = u + a;
        }
      e = e * inv;
      u = u * inv;
      ev[qi][h] = e;
      uv[qi][h] = u;
      se[h] += (double)e.re * e.re + (double)e.im * e.im;
      su[h] += (double)u.re * u.re + (double)u.im * u.im;
    }
  }
#pragma unroll
  for (int k = 0; k < 2 * NH; ++k) {
    double v = k < NH ? se[k] : su[k - NH];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2 * NH) {
    double t = 0.0;
    for (int k = 0; k < NW; ++k) t += red[threadIdx.x][k];
    scs[threadIdx.x] = (float)t;
  }
  __syncthreads();
  float is[NH], iu[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const float sE = sqrtf(scs[h] / kP), sU = sqrtf(scs[NH + h] / kP);
    is[h] = sE > 0.f ? 1.f / sE : 0.f;
    iu[h] = sU > 0.f ? 1.f / sU : 0.f;
  }
  const float *f = feat + (long long)sl * nfeat;
  float sc[kFnoScalars];
  sc[0] = f[0];
  sc[1] = f[1];
  sc[2] = f[2];
  sc[3] = f[3];
  sc[4] = NH > 1 ? f[2 + 5] : f[2];
  sc[5] = NH > 2 ? f[2 + 10] : f[2];
  sc[6] = f[1 + 5 * NH];
  sc[7] = f[2 + 5 * NH];
  sc[8] = log10f(sqrtf(scs[0] / kP) + 1e-30f);
  sc[9] = log10f(sqrtf(scs[NH] / kP) + 1e-30f);
  auto bf = [&](int ch, float x) {
    const float m0 = mu ? mu[ch] : 0.f, s0 = isd ? isd[ch] : 1.f;
    return (unsigned)(uint16_t)f2bf(__fmul_rn(__fsub_rn(x, m0), s0));
  };
#pragma unroll
  for (int qi = 0; qi < NQ; ++qi) {
    const int q = threadIdx.x + NT * qi;
    float x[CIN];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      x[2 * h] = __fmul_rn(ev[qi][h].re, is[h]);
      x[2 * h + 1] = __fmul_rn(ev[qi][h].im, is[h]);
      x[2 * NH + 2 * h] = __fmul_rn(uv[qi][h].re, iu[h]);
      x[2 * NH + 2 * h + 1] = __fmul_rn(uv[qi][h].im, iu[h]);
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) x[4 * NH + k] = sc[k];
    x[4 * NH + 10] = (2.f * (q % kN) + 1.f) / kN - 1.f;
    x[4 * NH + 11] = (2.f * (q / kN) + 1.f) / kN - 1.f;
    uint4 *o = reinterpret_cast<uint4 *>(X + ((long long)sl * kP + q) * CIN);
#pragma unroll
    for (int v = 0; v < CIN / 8; ++v) {
      uint4 pk;
      pk.x = bf(8 * v + 0, x[8 * v + 0]) | (bf(8 * v + 1, x[8 * v + 1]) << 16);
      pk.y = bf(8 * v + 2, x[8 * v + 2]) | (bf(8 * v + 3, x[8 * v + 3]) << 16);
      pk.z = bf(8 * v + 4, x[8 * v + 4]) | (bf(8 * v + 5, x[8 * v + 5]) << 16);
      pk.w = bf(8 * v + 6, x[8 * v + 6]) | (bf(8 * v + 7, x[8 * v + 7]) << 16);
      o[v] = pk;
    }
  }
}

// the network's inputs [sl][P][cin] and the normalised E (and R) in place
__global__ void k_fno_inputs(int slice_n, int nh, const float *__restrict__ scal, const float *__restrict__ feat,
                             int nfeat, float *__restrict__ Ec, const float *__restrict__ Uc, float *__restrict__ Rc,
                             float *__restrict__ X) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)slice_n * kP) return;
  const int sl = (int)(i / kP), q = (int)(i % kP);
  const int cin = 4 * nh + kFnoScalars;
  float *x = X + i * cin;
  const float *sc = scal + (long long)sl * 3 * nh;
  for (int h = 0; h < nh; ++h) {
    const float s = sqrtf(sc[h] / kP), u = sqrtf(sc[nh + h] / kP);
    const float is = s > 0.f ? 1.f / s : 0.f, iu = u > 0.f ? 1.f / u : 0.f;
    float *e = Ec + i * 2 * nh + 2 * h;
    e[0] *= is;
    e[1] *= is;
    if (Rc) {
      float *r = Rc + i * 2 * nh + 2 * h;
      r[0] *= is;
      r[1] *= is;
    }
    x[2 * h] = e[0];
    x[2 * h + 1] = e[1];
    x[2 * nh + 2 * h] = Uc[i * 2 * nh + 2 * h] * iu;
    x[2 * nh + 2 * h + 1] = Uc[i * 2 * nh + 2 * h + 1] * iu;
  }
  // features: current, intensity h1, |b1|, coherent fraction h1, |b3|, |b5|
  // (|b1| where there are fewer harmonics), energy spread, modulation
  const float *f = feat + (long long)sl * nfeat;
  float *xs = x + 4 * nh;
  xs[0] = f[0];
  xs[1] = f[1];
  xs[2] = f[2];
  xs[3] = f[3];
  xs[4] = nh > 1 ? f[2 + 5] : f[2];
  xs[5] = nh > 2 ? f[2 + 10] : f[2];
  xs[6] = f[1 + 5 * nh];
  xs[7] = f[2 + 5 * nh];
  xs[8] = log10f(sqrtf(sc[0] / kP) + 1e-30f);
  xs[9] = log10f(sqrtf(sc[nh] / kP) + 1e-30f);
  xs[10] = (2.f * (q % kN) + 1.f) / kN - 1.f;
  xs[11] = (2.f * (q / kN) + 1.f) / kN - 1.f;
}

// the correction of a jump's emission from the network's factor maps, one
// workgroup per (pol x harmonic, slot): on the slot's source window,
// U += o_h(x, y) 2 Sw / H with o_h bilinear between the map cells' centres
// (held at the edge cells beyond them), and the field energy's change added
// to the slot's slice energy in a fixed order (as k_emit_correct)
__global__ void __launch_bounds__(256) k_fno_correct(GridArgs g, FnoGeom fg, const int4 *__restrict__ win2,
                                                     const float *__restrict__ o, const cpx<float> *__restrict__ Sw,
                                                     cpx<float> *__restrict__ U, double *__restrict__ slice_e) {
  __shared__ double red[4];
  const int s = 1 + blockIdx.x, ph = blockIdx.y, h = ph % g.nh, nh = g.nh;
  const int4 b = win2[s];
  const int bw = b.y - b.x + 1, bh = b.w - b.z + 1;
  const int ncell = bw > 0 && bh > 0 ? bw * bh : 0;
  const float k2 = 2.f / (float)g.harm[h];
  const float *om = o + (long long)(s - 1) * kP * 2 * nh + 2 * h;
  const cpx<float> *src = Sw + ((long long)ph * g.S + s) * g.N;
  cpx<float> *u = U + ((long long)ph * g.S + phys_slot(s, g.ring_base, g.S)) * g.N;
  const float cen = 0.5f * (fg.b - 1), ib = 1.f / fg.b;
  double de = 0.0;
  for (int q = threadIdx.x; q < ncell; q += blockDim.x) {
    const int j = b.z + q / bw, i = b.x + q % bw;
    const long long c = (long long)j * g.nx + i;
    float tx = ((float)(i - fg.off) - cen) * ib, ty = ((float)(j - fg.off) - cen) * ib;
    tx = fminf(fmaxf(tx, 0.f), (float)(kN - 1));
    ty = fminf(fmaxf(ty, 0.f), (float)(kN - 1));
    const int x0 = min((int)tx, kN - 2), y0 = min((int)ty, kN - 2);
    const float wx = tx - x0, wy = ty - y0;
    auto at = [&](int xx, int yy) {
      const float *p = om + (long long)(yy * kN + xx) * 2 * nh;
      return mk(p[0], p[1]);
    };
    const cpx<float> f = (at(x0, y0) * (1.f - wx) + at(x0 + 1, y0) * wx) * (1.f - wy) +
                         (at(x0, y0 + 1) * (1.f - wx) + at(x0 + 1, y0 + 1) * wx) * wy;
    const cpx<float> old = u[c];
    const cpx<float> nw = old + f * (src[c] * k2);
    u[c] = nw;
    de += ((double)nw.re * nw.re + (double)nw.im * nw.im) - ((double)old.re * old.re + (double)old.im * old.im);
  }
  for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = de;
  __syncthreads();
  if (threadIdx.x == 0) slice_e[(long long)ph * g.slice_n + s - 1] += (red[0] + red[1]) + (red[2] + red[3]);
}

FnoGeom fno_geom(const felis_desc &d) {
  FnoGeom fg;
  fg.b = d.nx / kN;
  fg.off = (d.nx - kN * fg.b) / 2;
  return fg;
}

}  // namespace
}  // namespace felis

using felis::FnoGeom;

int felis_fno_maps_ok(const felis_ctx *ctx) {
  const felis_desc &d = ctx->d;
  return d.nx == d.ny && d.nx >= kN && d.dimension == 3 && ctx->mixed();
}

size_t felis_fno_work_floats(const felis_ctx *ctx) {
  const size_t sn = ctx->d.slice_n, nh = ctx->d.n_harm;
  return sn * kP * 2 * nh * 2 + ((sn * 3 * nh + 3) & ~(size_t)3) + sn * kP * (4 * nh + felis::kFnoScalars);
}

// The learned correction of a jump in the run (after its push and deposit,
// before the slippage): maps of every slot, the network, the correction.
// work: felis_fno_work_floats(ctx) floats; feat: the slices' features
// (device, as felis_slice_features) before the jump.
int felis_fno_correct_dev(felis_ctx *ctx, felis_fno *net, const float *feat, float *work) {
  const felis_desc &d = ctx->d;
  const int sn = d.slice_n, nh = d.n_harm;
  int cin = 0, cout = 0;
  felis_fno_channels(net, &cin, &cout);
  if (cin != 4 * nh + felis::kFnoScalars || cout != 2 * nh) {
    ctx->err = "FNO channels do not match the harmonics";
    return FELIS_ERR_ARG;
  }
  // (X 16-byte aligned for the records of k_fno_maps_xs)
  float *Ec = work, *Uc = Ec + (size_t)sn * kP * 2 * nh, *scal = Uc + (size_t)sn * kP * 2 * nh,
        *X = scal + (((size_t)sn * 3 * nh + 3) & ~(size_t)3);
  felis::GridArgs g = ctx->grid();
  const FnoGeom fg = felis::fno_geom(d);
  const float *o = nullptr;
  const float *mu = nullptr, *isd = nullptr;
  int e;
  if (felis_fno_input_affine_dev(net, &mu, &isd) == FELIS_OK) {
    // the maps straight into the lift's BF16 inputs (in X's space)
    uint16_t *XB = reinterpret_cast<uint16_t *>(X);
    if (nh == 3)
      hipLaunchKernelGGL((felis::k_fno_maps_xs<3>), dim3(sn), dim3(1024), 0, ctx->stream, g, fg,
                         (const felis::cpx<float> *)ctx->U, (const felis::cpx<float> *)ctx->Sw, ctx->win2, feat,
                         felis::features_per_slice(nh), mu, isd, XB);
    else if (nh == 1)
      hipLaunchKernelGGL((felis::k_fno_maps_xs<1>), dim3(sn), dim3(1024), 0, ctx->stream, g, fg,
                         (const felis::cpx<float> *)ctx->U, (const felis::cpx<float> *)ctx->Sw, ctx->win2, feat,
                         felis::features_per_slice(nh), mu, isd, XB);
    else if (fg.b == 2)
      hipLaunchKernelGGL((felis::k_fno_maps_x<2>), dim3(sn, nh), dim3(256), 0, ctx->stream, g, fg,
                         (const felis::cpx<float> *)ctx->U, (const felis::cpx<float> *)ctx->Sw, ctx->win2, feat,
                         felis::features_per_slice(nh), mu, isd, XB);
    else
      hipLaunchKernelGGL((felis::k_fno_maps_x<0>), dim3(sn, nh), dim3(256), 0, ctx->stream, g, fg,
                         (const felis::cpx<float> *)ctx->U, (const felis::cpx<float> *)ctx->Sw, ctx->win2, feat,
                         felis::features_per_slice(nh), mu, isd, XB);
    e = felis_fno_forward_bf16_dev(net, sn, XB, &o, ctx->stream);
  } else {
    hipLaunchKernelGGL(felis::k_fno_coarse, dim3(sn, nh), dim3(256), 0, ctx->stream, g, fg, 0,
                       (const felis::cpx<float> *)ctx->U, nullptr, (const felis::cpx<float> *)ctx->Sw, ctx->win2, Ec,
                       Uc, nullptr, scal);
    hipLaunchKernelGGL(felis::k_fno_inputs, dim3(nblk((long long)sn * kP)), dim3(256), 0, ctx->stream, sn, nh, scal,
                       feat, felis::features_per_slice(nh), Ec, Uc, nullptr, X);
    e = felis_fno_forward_dev(net, sn, X, &o, ctx->stream);
  }
  if (e) {
    ctx->err = "FNO forward";
    return e;
  }
  hipLaunchKernelGGL(felis::k_fno_correct, dim3(sn, ctx->npol() * nh), dim3(256), 0, ctx->stream, g, fg, ctx->win2, o,
                     (const felis::cpx<float> *)ctx->Sw, (felis::cpx<float> *)ctx->U, ctx->slice_e);
  return hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP;
}

extern "C" int felis_fno_jump_maps(felis_ctx *ctx, int slot, const float *feat, int nrows, const int *rows, float *X,
                                   float *E, float *R, float *info) {
  if (!ctx || slot < 0 || slot > 1 || !feat || nrows < 0 || (nrows && (!rows || !X || !E || !R || !info)))
    return FELIS_ERR_ARG;
  if (!felis_fno_maps_ok(ctx)) {
    ctx->err = "FNO maps: batched 3D path on a square grid of 64 cells or more";
    return FELIS_ERR_UNSUPPORTED;
  }
  felis_ctx::Snap &sn_ = ctx->snap[slot];
  if (!sn_.valid || sn_.np != ctx->np) {
    ctx->err = "no snapshot of this state in the slot";
    return FELIS_ERR_STATE;
  }
  if (sn_.ring_base != ctx->ring_base) {
    ctx->err = "snapshot and state differ in slippage";
    return FELIS_ERR_STATE;
  }
  if (!ctx->last_jump || !ctx->Sw) {
    ctx->err = "no jump has run on this context";
    return FELIS_ERR_STATE;
  }
  const felis_desc &d = ctx->d;
  const int sn = d.slice_n, nh = d.n_harm, cin = 4 * nh + felis::kFnoScalars, F = felis::features_per_slice(nh);
  for (int k = 0; k < nrows; ++k)
    if (rows[k] < 0 || rows[k] >= sn) return FELIS_ERR_ARG;
  float *dev = nullptr;
  const size_t nmap = (size_t)sn * kP * 2 * nh;
  const size_t total = 3 * nmap + (size_t)sn * 3 * nh + (size_t)sn * kP * cin + (size_t)sn * F;
  if (hipMalloc((void **)&dev, total * sizeof(float)) != hipSuccess) {
    ctx->err = "FNO maps: hipMalloc";
    return FELIS_ERR_HIP;
  }
  float *Ec = dev, *Uc = Ec + nmap, *Rc = Uc + nmap, *scal = Rc + nmap, *Xd = scal + (size_t)sn * 3 * nh,
        *fd = Xd + (size_t)sn * kP * cin;
  int e = FELIS_OK;
  felis::GridArgs g = ctx->grid();
  const FnoGeom fg = felis::fno_geom(d);
  if (hipMemcpyAsync(fd, feat, (size_t)sn * F * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    e = FELIS_ERR_HIP;
  if (!e) {
    hipLaunchKernelGGL(felis::k_fno_coarse, dim3(sn, nh), dim3(256), 0, ctx->stream, g, fg, ctx->last_jump,
                       (const felis::cpx<float> *)ctx->U, (const felis::cpx<float> *)sn_.U,
                       (const felis::cpx<float> *)ctx->Sw, ctx->win2, Ec, Uc, Rc, scal);
    hipLaunchKernelGGL(felis::k_fno_inputs, dim3(nblk((long long)sn * kP)), dim3(256), 0, ctx->stream, sn, nh, scal,
                       fd, F, Ec, Uc, Rc, Xd);
    if (hipGetLastError() != hipSuccess) e = FELIS_ERR_HIP;
  }
  for (int k = 0; k < nrows && !e; ++k) {
    const size_t r = rows[k];
    if (hipMemcpyAsync(X + (size_t)k * kP * cin, Xd + r * kP * cin, (size_t)kP * cin * sizeof(float),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(E + (size_t)k * kP * 2 * nh, Ec + r * kP * 2 * nh, (size_t)kP * 2 * nh * sizeof(float),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(R + (size_t)k * kP * 2 * nh, Rc + r * kP * 2 * nh, (size_t)kP * 2 * nh * sizeof(float),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(info + (size_t)k * 3 * nh, scal + r * 3 * nh, (size_t)3 * nh * sizeof(float),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
      e = FELIS_ERR_HIP;
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) e = FELIS_ERR_HIP;
  (void)hipFree(dev);
  if (e) ctx->err = "FNO maps";
  return e;
}
