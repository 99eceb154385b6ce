// The wake-function table of wakeField.cpp:53-159 on the device: the
// resistive-wall impedance impZ(k) on 1000 wavenumbers (for a flat pipe an
// integral over 50001 points each, which the reference evaluates serially at
// start-up) and its cosine transform to the slice wake wakeF.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "felis.h"
#include "felis_common.h"
#include "felis_deck.h"

using namespace felis;

namespace {

constexpr int kNumK = 1000, kNumX = 50000;    // wakeField.cpp:64-66
constexpr double kMaxK = 20.0, kMaxX = 2000.0;
constexpr int kThreads = 256;

__device__ inline cpx<double> cdiv(cpx<double> a, cpx<double> b) {
  const double d = b.re * b.re + b.im * b.im;
  return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

// cosh(x_j) and sinh(x_j) / x_j for x_j = j dX, j = 1..numX (the same for
// every wavenumber)
__global__ void k_wake_hyp(double dX, double *__restrict__ ch, double *__restrict__ shx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < 1 || j > kNumX) return;
  const double x = j * dX;
  ch[j] = cosh(x);
  shx[j] = sinh(x) / x;
}

// impZ[i], i = 1..numK, one workgroup each. Flat pipe (shape 1,
// wakeField.cpp:82-105): the real part of sum_j ccal(x_j) dX with
// ccal = 1 / cosh(x) / (a cosh(x) - i k sinh(x) / x), a = 2 / (1 - i) / root,
// root = sqrt(k - i k^2 ctau / s0) (computed on the host), ccal(0) =
// sqrt(k) / (2 + k^3 - 2 sqrt(k) k) and ccal = 0 beyond x = 320; the
// per-thread partial sums are added in a fixed tree. Circular pipe (shape 2,
// 106-118): the closed forms of the AC and DC conductivities.
__global__ void __launch_bounds__(kThreads)
    k_wake_imp(int shape, int acdc, double s0, double ctau, double dK, double dX, const cpx<double> *__restrict__ a_k,
               const double *__restrict__ ch, const double *__restrict__ shx, double *__restrict__ impZ,
               unsigned *__restrict__ nan_flag) {
  __shared__ double red[kThreads];
  const int i = blockIdx.x + 1, t = threadIdx.x;
  const double k = i * dK;
  if (shape == 2) {
    if (t == 0) {
      if (acdc == 1) {
        // sqrt(2 s0 / ctau) / k (i + s0 / (2 k ctau)) - i k / 2
        const double sq = sqrt(2.0 * s0 / ctau) / k;
        const cpx<double> cd = mk(sq * (s0 * 0.5 / k / ctau), sq - 0.5 * k);
        impZ[i] = cdiv(mk(1.0, 0.0), cd).re;
      } else {
        const double dino = 2.0 / k + 0.25 * k * k - sqrt(k);
        impZ[i] = 1.0 / sqrt(k) / dino;
      }
    }
    return;
  }
  const cpx<double> a = a_k[i];
  double s = 0.0;
  bool bad = false;
  for (int j = t; j <= kNumX; j += kThreads) {
    double re;
    if (j == 0) {
      re = sqrt(k) / (2.0 + k * k * k - 2.0 * sqrt(k) * k);
    } else if (j * dX > 320.0) {
      re = 0.0;
    } else {
      const double c = ch[j];
      const cpx<double> cd = mk(a.re * c, a.im * c - k * shx[j]);
      re = cdiv(mk(1.0 / c, 0.0), cd).re;
      bad |= re != re;
    }
    s += re * dX;
  }
  if (bad) atomicOr(nan_flag, 1u);
  red[t] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) impZ[i] = red[0];
}

// wakeF[i] = coef sum_{j=0..numK} impZ[j] cos(j dK z_i) dK, z_i = i dZ / s0
// (wakeField.cpp:150-158), one thread per slice, in the reference's order
__global__ void k_wake_cos(const double *__restrict__ impZ, int slice_n, double dZ, double s0, double dK, double coef,
                           double *__restrict__ wakeF) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slice_n) return;
  const double z = (i * dZ) / s0;
  double sum = 0.0;
  for (int j = 0; j <= kNumK; ++j) sum += impZ[j] * cos(j * dK * z) * dK;
  wakeF[i] = 0.0 + sum * coef;
}

}  // namespace

extern "C" int felis_wake_function(int shape, int acdc, double radius, double cond, double ctau, int slice_n,
                                   double dZ, double *wakeF, double *impZ_out) {
  if (!wakeF || slice_n <= 0 || radius <= 0 || cond <= 0 || (shape != 1 && shape != 2)) return FELIS_ERR_ARG;
  const double s0 = std::pow(2.0 * radius * radius / kZ0 / cond, 1.0 / 3.0);
  const double dK = kMaxK / kNumK, dX = kMaxX / kNumX;
  // a = 2 / (1 - i) / sqrt(k - i k^2 ctau / s0) per wavenumber (host complex
  // arithmetic, as the reference evaluates root once per k)
  std::vector<cpx<double>> a(kNumK + 1, mk(0.0, 0.0));
  if (shape == 1)
    for (int i = 1; i <= kNumK; ++i) {
      const double k = i * dK;
      const std::complex<double> root = std::sqrt(std::complex<double>(k, -k * k * ctau / s0));
      const std::complex<double> v = 2.0 / std::complex<double>(1.0, -1.0) / root;
      a[i] = mk(v.real(), v.imag());
    }
  char *buf = nullptr;
  const size_t n_hyp = kNumX + 1, n_imp = kNumK + 1;
  const size_t bytes = (2 * n_hyp + n_imp + (size_t)slice_n) * sizeof(double) + n_imp * sizeof(cpx<double>) + 256;
  if (hipMalloc((void **)&buf, bytes) != hipSuccess) return FELIS_ERR_HIP;
  double *ch = (double *)buf, *shx = ch + n_hyp, *imp = shx + n_hyp, *wf = imp + n_imp;
  cpx<double> *ak = (cpx<double> *)(wf + slice_n);
  unsigned *flag = (unsigned *)(ak + n_imp);
  hipError_t e = hipMemcpy(ak, a.data(), n_imp * sizeof(cpx<double>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(imp, 0, n_imp * sizeof(double));
  if (e == hipSuccess) e = hipMemset(flag, 0, sizeof(unsigned));
  if (e == hipSuccess) {
    if (shape == 1) hipLaunchKernelGGL(k_wake_hyp, dim3((kNumX + 256) / 256), dim3(256), 0, 0, dX, ch, shx);
    hipLaunchKernelGGL(k_wake_imp, dim3(kNumK), dim3(kThreads), 0, 0, shape, acdc, s0, ctau, dK, dX, ak, ch, shx,
                       imp, flag);
    const double coef = kZ0 * 3e8 / M_PI / radius / radius / M_PI;
    hipLaunchKernelGGL(k_wake_cos, dim3((slice_n + 255) / 256), dim3(256), 0, 0, imp, slice_n, dZ, s0, dK, coef, wf);
    e = hipGetLastError();
  }
  unsigned nan = 0;
  if (e == hipSuccess) e = hipMemcpy(wakeF, wf, (size_t)slice_n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && impZ_out) e = hipMemcpy(impZ_out, imp, n_imp * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&nan, flag, sizeof(unsigned), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) return FELIS_ERR_HIP;
  return nan ? FELIS_ERR_NAN : FELIS_OK;
}

extern "C" int felis_deck_wake_function_device(const felis_deck *D, double *wakeF, double *impZ) {
  if (!D) return FELIS_ERR_ARG;
  return felis_wake_function(D->wake_shape, D->wake_acdc, D->radius, D->cond, D->ctau, D->slice_n,
                             D->num_slice * D->lambda0, wakeF, impZ);
}
