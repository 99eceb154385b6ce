// A small multilayer perceptron on gfx950 matrix cores: BF16 operands, FP32
// accumulation (v_mfma_f32_16x16x16_bf16), FP32 master weights and Adam.
// It is the learned part of the batched step (felis_sur_*): one sample per
// slice, so every layer is a (slices x width) x (width x width) product.
//
// Products: one wavefront per 16 x 16 output tile, K in steps of 16. Lane l
// holds A[l%16][4(l/16)..4(l/16)+3], B[4(l/16)..+3][l%16] and returns
// D[4(l/16)..+3][l%16]. Operands are read with arbitrary strides, so the
// same kernel forms X W, dZ W^T and X^T dZ.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "felis.h"
#include "felis_mlp.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// C[m*ldc + n] = sum_k A(m,k) B(k,n) (+ C if acc), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; BF16 operands, FP32 result.
__global__ void __launch_bounds__(64) k_gemm_bf16(int M, int N, int K, const uint16_t *__restrict__ A, long long sam,
                                                  long long sak, const uint16_t *__restrict__ B, long long sbk,
                                                  long long sbn, float *__restrict__ C, long long ldc, int acc) {
  const int tm = blockIdx.y * 16, tn = blockIdx.x * 16, l = threadIdx.x;
  const int r = l & 15, kq = (l >> 4) * 4;
  v4f d = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a, b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kq + q, m = tm + r, n = tn + r;
      a[q] = (short)((m < M && k < K) ? A[m * sam + k * sak] : 0);
      b[q] = (short)((n < N && k < K) ? B[k * sbk + n * sbn] : 0);
    }
    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = tm + kq + q, n = tn + r;
    if (m < M && n < N) {
      float v = d[q];
      if (acc) v += C[m * ldc + n];
      C[m * ldc + n] = v;
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

// Z += bias; H = gelu(Z) (hidden) or Z (output), stored as BF16 and FP32
__global__ void k_bias_act(int M, int N, float *__restrict__ Z, const float *__restrict__ bias, int act,
                           uint16_t *__restrict__ Hb, float *__restrict__ Hf) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)M * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  const float h = act ? gelu(z) : z;
  if (Hb) Hb[i] = f2bf(h);
  if (Hf) Hf[i] = h;
}

// mean squared error over M x N and its gradient 2 (Y - T) / (M N) in BF16
__global__ void k_mse(int M, int N, const float *__restrict__ Y, const float *__restrict__ T,
                      const float *__restrict__ wrow, uint16_t *__restrict__ dY, float *__restrict__ dYf,
                      double *__restrict__ loss) {
  __shared__ double sh[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < (long long)M * N) {
    const float w = wrow ? wrow[i / N] : 1.f;
    const float r = Y[i] - T[i];
    e = (double)w * r * r;
    const float g = 2.f * w * r / (float)((long long)M * N);
    dY[i] = f2bf(g);
    dYf[i] = g;
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, sh[0] / ((double)M * N));
}

// dZ = dH * gelu'(Z), in BF16 (and FP32 for the bias gradient)
__global__ void k_act_back(int M, int N, const float *__restrict__ dH, const float *__restrict__ Z,
                           uint16_t *__restrict__ dZb, float *__restrict__ dZf) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)M * N) return;
  const float g = dH[i] * gelu_d(Z[i]);
  dZb[i] = f2bf(g);
  dZf[i] = g;
}

// db[n] = sum_m dZ[m, n], one thread per column in a fixed order
__global__ void k_col_sum(int M, int N, const float *__restrict__ dZ, float *__restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(long long)m * N + n];
  db[n] = s;
}

// Adam on FP32 master weights, then the BF16 copy
__global__ void k_adam(long long n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// network input: (x - mu) * isd per column in FP32, then BF16
__global__ void k_input_bf16(long long n, int cols, const float *__restrict__ x, const float *__restrict__ mu,
                             const float *__restrict__ isd, uint16_t *__restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int c = (int)(i % cols);
    y[i] = f2bf((x[i] - mu[c]) * isd[c]);
  }
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------------ network

struct felis_mlp {
  int L = 0;                    // layers
  std::vector<int> dims;        // L+1 widths
  int cap = 0;                  // rows the work buffers hold
  std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;  // FP32, device
  std::vector<uint16_t *> Wb;                         // BF16 weights, device
  std::vector<float *> Z, Hf;                         // per layer pre-/post-activation FP32
  std::vector<uint16_t *> Hb;                         // per layer activation BF16 (index 0 = input)
  float *dHf = nullptr, *dZf = nullptr, *Tf = nullptr, *wrow = nullptr;
  float *in_mu = nullptr, *in_isd = nullptr;  // input affine, applied in FP32 before BF16
  uint16_t *dZb = nullptr;
  double *loss = nullptr;
  long long step = 0;
  std::vector<void *> allocs;
  std::string err;
  hipStream_t stream = nullptr;
};

namespace {

template <typename P>
int mlp_alloc(felis_mlp *n, P **p, size_t bytes) {
  if (hipMalloc((void **)p, bytes ? bytes : 16) != hipSuccess) {
    n->err = "hipMalloc";
    return FELIS_ERR_HIP;
  }
  n->allocs.push_back((void *)*p);
  return FELIS_OK;
}

void gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, long long sam, long long sak, const uint16_t *B,
          long long sbk, long long sbn, float *C, long long ldc, int acc) {
  dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(64), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, acc);
}

void forward(felis_mlp *n, int M) {
  for (int l = 0; l < n->L; ++l) {
    const int K = n->dims[l], N = n->dims[l + 1];
    // Z = H W: H (M x K) row-major, W (K x N) row-major
    gemm(n->stream, M, N, K, n->Hb[l], K, 1, n->Wb[l], N, 1, n->Z[l], N, 0);
    const bool last = l == n->L - 1;
    hipLaunchKernelGGL(k_bias_act, dim3(nblk((long long)M * N)), dim3(256), 0, n->stream, M, N, n->Z[l], n->b[l],
                       last ? 0 : 1, last ? nullptr : n->Hb[l + 1], n->Hf[l + 1]);
  }
}

}  // namespace

extern "C" {

// dims[0..layers]: input, hidden..., output widths; capacity: rows per call.
int felis_mlp_create(int layers, const int *dims, int capacity, uint64_t seed, felis_mlp **out) {
  if (!dims || !out || layers < 1 || layers > 8 || capacity < 1) return FELIS_ERR_ARG;
  for (int l = 0; l <= layers; ++l)
    if (dims[l] < 1 || dims[l] > 4096) return FELIS_ERR_ARG;
  felis_mlp *n = new felis_mlp();
  n->L = layers;
  n->dims.assign(dims, dims + layers + 1);
  n->cap = capacity;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
    delete n;
    return FELIS_ERR_HIP;
  }
  n->W.resize(layers); n->b.resize(layers); n->gW.resize(layers); n->gb.resize(layers);
  n->mW.resize(layers); n->vW.resize(layers); n->mb.resize(layers); n->vb.resize(layers);
  n->Wb.resize(layers); n->Z.resize(layers); n->Hf.resize(layers + 1); n->Hb.resize(layers + 1);
  int e = FELIS_OK;
  int maxw = 0;
  for (int l = 0; l <= layers; ++l) maxw = std::max(maxw, dims[l]);
  for (int l = 0; l < layers && !e; ++l) {
    const size_t nw = (size_t)dims[l] * dims[l + 1], nb = dims[l + 1];
    if ((e = mlp_alloc(n, &n->W[l], nw * 4)) || (e = mlp_alloc(n, &n->gW[l], nw * 4)) ||
        (e = mlp_alloc(n, &n->mW[l], nw * 4)) || (e = mlp_alloc(n, &n->vW[l], nw * 4)) ||
        (e = mlp_alloc(n, &n->Wb[l], nw * 2)) || (e = mlp_alloc(n, &n->b[l], nb * 4)) ||
        (e = mlp_alloc(n, &n->gb[l], nb * 4)) || (e = mlp_alloc(n, &n->mb[l], nb * 4)) ||
        (e = mlp_alloc(n, &n->vb[l], nb * 4)) || (e = mlp_alloc(n, &n->Z[l], (size_t)capacity * nb * 4)))
      break;
  }
  for (int l = 0; l <= layers && !e; ++l) {
    if ((e = mlp_alloc(n, &n->Hb[l], (size_t)capacity * dims[l] * 2)) ||
        (e = mlp_alloc(n, &n->Hf[l], (size_t)capacity * dims[l] * 4)))
      break;
  }
  if (!e && ((e = mlp_alloc(n, &n->dHf, (size_t)capacity * maxw * 4)) ||
             (e = mlp_alloc(n, &n->dZf, (size_t)capacity * maxw * 4)) ||
             (e = mlp_alloc(n, &n->dZb, (size_t)capacity * maxw * 2)) ||
             (e = mlp_alloc(n, &n->Tf, (size_t)capacity * dims[layers] * 4)) ||
             (e = mlp_alloc(n, &n->wrow, (size_t)capacity * 4)) || (e = mlp_alloc(n, &n->loss, 8)) ||
             (e = mlp_alloc(n, &n->in_mu, (size_t)dims[0] * 4)) || (e = mlp_alloc(n, &n->in_isd, (size_t)dims[0] * 4)))) {
  }
  if (e) {
    for (void *p : n->allocs) (void)hipFree(p);
    (void)hipStreamDestroy(n->stream);
    delete n;
    return e;
  }
  {
    std::vector<float> z(dims[0], 0.f), o(dims[0], 1.f);
    (void)hipMemcpy(n->in_mu, z.data(), dims[0] * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(n->in_isd, o.data(), dims[0] * 4, hipMemcpyHostToDevice);
  }
  // Glorot-uniform weights from a 64-bit LCG, zero biases and moments
  uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
  for (int l = 0; l < layers; ++l) {
    const size_t nw = (size_t)dims[l] * dims[l + 1];
    std::vector<float> w(nw);
    const double lim = std::sqrt(6.0 / (dims[l] + dims[l + 1]));
    for (size_t i = 0; i < nw; ++i) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      w[i] = (float)(lim * (2.0 * ((s >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
    }
    std::vector<uint16_t> wb(nw);
    for (size_t i = 0; i < nw; ++i) wb[i] = f2bf(w[i]);
    (void)hipMemcpy(n->W[l], w.data(), nw * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(n->Wb[l], wb.data(), nw * 2, hipMemcpyHostToDevice);
    (void)hipMemset(n->mW[l], 0, nw * 4);
    (void)hipMemset(n->vW[l], 0, nw * 4);
    (void)hipMemset(n->b[l], 0, dims[l + 1] * 4);
    (void)hipMemset(n->mb[l], 0, dims[l + 1] * 4);
    (void)hipMemset(n->vb[l], 0, dims[l + 1] * 4);
  }
  *out = n;
  return FELIS_OK;
}

int felis_mlp_destroy(felis_mlp *n) {
  if (!n) return FELIS_OK;
  (void)hipStreamSynchronize(n->stream);
  for (void *p : n->allocs) (void)hipFree(p);
  (void)hipStreamDestroy(n->stream);
  delete n;
  return FELIS_OK;
}

const char *felis_mlp_last_error(const felis_mlp *n) { return n ? n->err.c_str() : "null network"; }

int felis_mlp_shape(const felis_mlp *n, int *in, int *out, int *capacity) {
  if (!n) return FELIS_ERR_ARG;
  if (in) *in = n->dims[0];
  if (out) *out = n->dims[n->L];
  if (capacity) *capacity = n->cap;
  return FELIS_OK;
}

// Weights and biases of layer l as FP32 (host arrays, W row-major in x out).
int felis_mlp_get_layer(felis_mlp *n, int l, float *W, float *b) {
  if (!n || l < 0 || l >= n->L) return FELIS_ERR_ARG;
  const size_t nw = (size_t)n->dims[l] * n->dims[l + 1];
  if (W && hipMemcpy(W, n->W[l], nw * 4, hipMemcpyDeviceToHost) != hipSuccess) return FELIS_ERR_HIP;
  if (b && hipMemcpy(b, n->b[l], n->dims[l + 1] * 4, hipMemcpyDeviceToHost) != hipSuccess) return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_mlp_set_layer(felis_mlp *n, int l, const float *W, const float *b) {
  if (!n || l < 0 || l >= n->L) return FELIS_ERR_ARG;
  const size_t nw = (size_t)n->dims[l] * n->dims[l + 1];
  if (W) {
    std::vector<uint16_t> wb(nw);
    for (size_t i = 0; i < nw; ++i) wb[i] = f2bf(W[i]);
    if (hipMemcpy(n->W[l], W, nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->Wb[l], wb.data(), nw * 2, hipMemcpyHostToDevice) != hipSuccess)
      return FELIS_ERR_HIP;
  }
  if (b && hipMemcpy(n->b[l], b, n->dims[l + 1] * 4, hipMemcpyHostToDevice) != hipSuccess) return FELIS_ERR_HIP;
  return FELIS_OK;
}

// Y[rows][out] = net(X[rows][in]); X, Y device pointers (FP32) on the
// network's stream ordered after `after` (may be null).
int felis_mlp_forward_dev(felis_mlp *n, int rows, const float *X, float *Y, hipStream_t after) {
  if (!n || !X || rows < 1 || rows > n->cap) return FELIS_ERR_ARG;
  hipStream_t s = after ? after : n->stream;
  hipStream_t keep = n->stream;
  n->stream = s;
  const long long ni = (long long)rows * n->dims[0];
  hipLaunchKernelGGL(k_input_bf16, dim3(nblk(ni)), dim3(256), 0, s, ni, n->dims[0], X, n->in_mu, n->in_isd,
                     n->Hb[0]);
  forward(n, rows);
  if (Y)
    (void)hipMemcpyAsync(Y, n->Hf[n->L], (size_t)rows * n->dims[n->L] * 4, hipMemcpyDeviceToDevice, s);
  n->stream = keep;
  return hipGetLastError() == hipSuccess ? FELIS_OK : FELIS_ERR_HIP;
}

// host arrays
int felis_mlp_forward(felis_mlp *n, int rows, const float *X, float *Y) {
  if (!n || !X || !Y || rows < 1 || rows > n->cap) return FELIS_ERR_ARG;
  if (hipMemcpyAsync(n->Hf[0], X, (size_t)rows * n->dims[0] * 4, hipMemcpyHostToDevice, n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  const long long ni = (long long)rows * n->dims[0];
  hipLaunchKernelGGL(k_input_bf16, dim3(nblk(ni)), dim3(256), 0, n->stream, ni, n->dims[0], n->Hf[0], n->in_mu,
                     n->in_isd, n->Hb[0]);
  forward(n, rows);
  if (hipMemcpyAsync(Y, n->Hf[n->L], (size_t)rows * n->dims[n->L] * 4, hipMemcpyDeviceToHost, n->stream) !=
          hipSuccess ||
      hipStreamSynchronize(n->stream) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

// One gradient step on the mean squared error (row weights optional).
// Gradients are left in place (felis_mlp_get_grad) when lr == 0.
int felis_mlp_train_step(felis_mlp *n, int rows, const float *X, const float *T, const float *wrow, float lr,
                         double *loss_out) {
  if (!n || !X || !T || rows < 1 || rows > n->cap) return FELIS_ERR_ARG;
  hipStream_t s = n->stream;
  const int L = n->L, dout = n->dims[L];
  if (hipMemcpyAsync(n->Hf[0], X, (size_t)rows * n->dims[0] * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(n->Tf, T, (size_t)rows * dout * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      (wrow && hipMemcpyAsync(n->wrow, wrow, (size_t)rows * 4, hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipMemsetAsync(n->loss, 0, 8, s) != hipSuccess)
    return FELIS_ERR_HIP;
  const long long ni = (long long)rows * n->dims[0];
  hipLaunchKernelGGL(k_input_bf16, dim3(nblk(ni)), dim3(256), 0, s, ni, n->dims[0], n->Hf[0], n->in_mu, n->in_isd,
                     n->Hb[0]);
  forward(n, rows);
  // output layer: dZ = dY
  const long long no = (long long)rows * dout;
  hipLaunchKernelGGL(k_mse, dim3(nblk(no)), dim3(256), 0, s, rows, dout, n->Hf[L], n->Tf,
                     wrow ? n->wrow : nullptr, n->dZb, n->dZf, n->loss);
  for (int l = L - 1; l >= 0; --l) {
    const int K = n->dims[l], N = n->dims[l + 1];
    if (l < L - 1) {
      // dZ = dH * gelu'(Z)
      hipLaunchKernelGGL(k_act_back, dim3(nblk((long long)rows * N)), dim3(256), 0, s, rows, N, n->dHf, n->Z[l],
                         n->dZb, n->dZf);
    }
    // gW = H^T dZ : (K x rows) (rows x N)
    gemm(s, K, N, rows, n->Hb[l], 1, K, n->dZb, N, 1, n->gW[l], N, 0);
    hipLaunchKernelGGL(k_col_sum, dim3(nblk(N)), dim3(256), 0, s, rows, N, n->dZf, n->gb[l]);
    // dH_{l} = dZ W^T : (rows x N) (N x K), into dHf for the layer below
    if (l > 0) gemm(s, rows, K, N, n->dZb, N, 1, n->Wb[l], 1, N, n->dHf, K, 0);
  }
  if (lr > 0.f) {
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    for (int l = 0; l < L; ++l) {
      const long long nw = (long long)n->dims[l] * n->dims[l + 1];
      hipLaunchKernelGGL(k_adam, dim3(nblk(nw)), dim3(256), 0, s, nw, n->W[l], n->gW[l], n->mW[l], n->vW[l],
                         n->Wb[l], lr, b1, b2, eps, c1, c2);
      hipLaunchKernelGGL(k_adam, dim3(nblk(n->dims[l + 1])), dim3(256), 0, s, (long long)n->dims[l + 1], n->b[l],
                         n->gb[l], n->mb[l], n->vb[l], (uint16_t *)nullptr, lr, b1, b2, eps, c1, c2);
    }
  }
  double lh = 0.0;
  if (hipMemcpyAsync(&lh, n->loss, 8, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return FELIS_ERR_HIP;
  if (loss_out) *loss_out = lh;
  return FELIS_OK;
}

// Input affine: the network reads (x - mu) * inv_sd (FP32) per input column.
int felis_mlp_set_input_affine(felis_mlp *n, const float *mu, const float *inv_sd) {
  if (!n || !mu || !inv_sd) return FELIS_ERR_ARG;
  if (hipMemcpy(n->in_mu, mu, n->dims[0] * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->in_isd, inv_sd, n->dims[0] * 4, hipMemcpyHostToDevice) != hipSuccess)
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_mlp_get_input_affine(felis_mlp *n, float *mu, float *inv_sd) {
  if (!n) return FELIS_ERR_ARG;
  if ((mu && hipMemcpy(mu, n->in_mu, n->dims[0] * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
      (inv_sd && hipMemcpy(inv_sd, n->in_isd, n->dims[0] * 4, hipMemcpyDeviceToHost) != hipSuccess))
    return FELIS_ERR_HIP;
  return FELIS_OK;
}

int felis_mlp_get_grad(felis_mlp *n, int l, float *gW, float *gb) {
  if (!n || l < 0 || l >= n->L) return FELIS_ERR_ARG;
  const size_t nw = (size_t)n->dims[l] * n->dims[l + 1];
  if (gW && hipMemcpy(gW, n->gW[l], nw * 4, hipMemcpyDeviceToHost) != hipSuccess) return FELIS_ERR_HIP;
  if (gb && hipMemcpy(gb, n->gb[l], n->dims[l + 1] * 4, hipMemcpyDeviceToHost) != hipSuccess) return FELIS_ERR_HIP;
  return FELIS_OK;
}

// C (M x N) = A (M x K) B (K x N), all row-major host arrays, operands
// rounded to BF16: the matrix-core product on its own (layout check).
int felis_gemm_bf16(int M, int N, int K, const float *A, const float *B, float *C) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1) return FELIS_ERR_ARG;
  std::vector<uint16_t> a((size_t)M * K), b((size_t)K * N);
  for (size_t i = 0; i < a.size(); ++i) a[i] = f2bf(A[i]);
  for (size_t i = 0; i < b.size(); ++i) b[i] = f2bf(B[i]);
  uint16_t *da = nullptr, *db = nullptr;
  float *dc = nullptr;
  int e = FELIS_OK;
  if (hipMalloc((void **)&da, a.size() * 2) != hipSuccess || hipMalloc((void **)&db, b.size() * 2) != hipSuccess ||
      hipMalloc((void **)&dc, (size_t)M * N * 4) != hipSuccess)
    e = FELIS_ERR_HIP;
  if (!e && (hipMemcpy(da, a.data(), a.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(db, b.data(), b.size() * 2, hipMemcpyHostToDevice) != hipSuccess))
    e = FELIS_ERR_HIP;
  if (!e) {
    gemm(nullptr, M, N, K, da, K, 1, db, N, 1, dc, N, 0);
    if (hipGetLastError() != hipSuccess || hipMemcpy(C, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost) != hipSuccess)
      e = FELIS_ERR_HIP;
  }
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dc);
  return e;
}

}  // extern "C"
