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

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kN = 64;        // map side
constexpr int kP = kN * kN;   // positions per map

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet)
__device__ inline short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

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
__global__ void __launch_bounds__(256) k_bgemm(int batch, int M, int N, int K, Op A, Op B, OpW C, int acc,
                                               const float *__restrict__ bias) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tm_n = (M + 31) >> 5, tn_n = (N + 31) >> 5;
  const long long tile = (long long)blockIdx.x * 4 + wave, per = (long long)tm_n * tn_n;
  if (tile >= per * batch) return;
  const long long bi = tile / per;
  const int t = (int)(tile % per), m0 = (t / tn_n) * 32, n0 = (t % tn_n) * 32;
  const int r = l & 15, kq = (l >> 4) * 4;
  const float *pa = A.p + ix(A.b, bi), *pb = B.p + ix(B.b, bi);
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
        a[u][q] = (vk && va[u]) ? f2bf(pa[ra[u] + ka]) : (short)0;
        b[u][q] = (vk && vb[u]) ? f2bf(pb[kb + cb[u]]) : (short)0;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) d[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], d[i][j], 0, 0, 0);
  }
  float *pc = C.p + ix(C.b, bi);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + 16 * i + kq + q, n = n0 + 16 * j + r;
        if (m < M && n < N) {
          float *o = pc + ix(C.r, m) + ix(C.c, n);
          float v = d[i][j][q];
          if (bias) v += bias[n];
          if (acc) v += *o;
          *o = v;
        }
      }
}

__device__ inline float gelu(float z) {  // tanh form
  const float c = 0.7978845608f;
  return 0.5f * z * (1.f + tanhf(c * (z + 0.044715f * z * z * z)));
}
__device__ inline float gelu_d(float z) {
  const float c = 0.7978845608f;
  const float u = c * (z + 0.044715f * z * z * z), t = tanhf(u);
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

Ix U1(long long s) { return Ix{31, 0, s}; }
Ix U2(int sh, long long hi, long long lo) { return Ix{sh, hi, lo}; }
int lg2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}

void bgemm(hipStream_t st, int batch, int M, int N, int K, Op A, Op B, OpW C, int acc, const float *bias) {
  const long long tiles = (long long)batch * ((M + 31) / 32) * ((N + 31) / 32);
  if (tiles <= 0) return;
  hipLaunchKernelGGL(k_bgemm, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, batch, M, N, K, A, B, C, acc, bias);
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
  int e = fno_input(n, B, X);
  if (!e) fno_forward(n, B);
  (void)hipEventRecord(ev, n->stream);
  (void)hipStreamWaitEvent(st, ev, 0);
  (void)hipEventDestroy(ev);
  if (out) *out = n->out;
  return e ? e : (hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP);
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

// X [B][P][cin] host, Y [B][P][cout] host
int felis_fno_forward(felis_fno *n, int B, const float *X, float *Y) {
  if (!n || !X || !Y || B < 1 || B > n->cap) return FELIS_ERR_ARG;
  const size_t nx = (size_t)B * kP * n->cin, ny = (size_t)B * kP * n->cout;
  float *dX = nullptr;
  if (hipMalloc((void **)&dX, nx * sizeof(float)) != hipSuccess) return FELIS_ERR_HIP;
  int e = hipMemcpyAsync(dX, X, nx * sizeof(float), hipMemcpyHostToDevice, n->stream) == hipSuccess ? FELIS_OK
                                                                                                    : FELIS_ERR_HIP;
  if (!e) e = fno_input(n, B, dX);
  if (!e) {
    fno_forward(n, B);
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
