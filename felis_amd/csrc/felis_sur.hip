// Support of the learned jump correction: device-side state snapshots, the
// per-slice moments the correction network reads, and the residual of a
// jump against the fine steps from the same state (its training targets).
// None of this has a counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "felis_ctx.h"
#include "felis_sur.h"

using namespace felis;

namespace {

int fail(felis_ctx *ctx, int code, const char *msg) {
  ctx->err = msg;
  return code;
}

#define SCHK(call)                                                   \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);  \
      return FELIS_ERR_HIP;                                          \
    }                                                                \
  } while (0)

size_t field_bytes(const felis_ctx *ctx) { return (size_t)2 * ctx->d.n_harm * ctx->S * ctx->N * ctx->esize(); }
size_t part_bytes(const felis_ctx *ctx) { return (size_t)ctx->np * (sizeof(double) + 5 * ctx->rsize()); }

// particle arrays in snapshot order: theta, then x, y, px, py, gamma
void part_arrays(felis_ctx *ctx, void **p, size_t *bytes) {
  p[0] = ctx->theta; bytes[0] = (size_t)ctx->np * sizeof(double);
  p[1] = ctx->x; p[2] = ctx->y; p[3] = ctx->px; p[4] = ctx->py; p[5] = ctx->gamma;
  for (int k = 1; k < 6; ++k) bytes[k] = (size_t)ctx->np * ctx->rsize();
}

template <int NH>
void launch_moments(felis_ctx *ctx, double *out) {
  GridArgs g = ctx->grid();
  const cpx<float> *Ux = (const cpx<float> *)ctx->U, *Uy = (const cpx<float> *)((char *)ctx->U + ctx->pol_bytes());
  const double inv_dx = 1.0 / ctx->d.dx, inv_dy = 1.0 / ctx->d.dy;
  if (ctx->uy_live)
    hipLaunchKernelGGL((k_slice_moments<NH, true>), dim3(ctx->d.slice_n), dim3(256), 0, ctx->stream, g,
                       ctx->d.n_species, ctx->seg_off, ctx->seg_w, ctx->theta, (const float *)ctx->gamma,
                       (const float *)ctx->x, (const float *)ctx->y, Ux, Uy, inv_dx, inv_dy, out);
  else
    hipLaunchKernelGGL((k_slice_moments<NH, false>), dim3(ctx->d.slice_n), dim3(256), 0, ctx->stream, g,
                       ctx->d.n_species, ctx->seg_off, ctx->seg_w, ctx->theta, (const float *)ctx->gamma,
                       (const float *)ctx->x, (const float *)ctx->y, Ux, Uy, inv_dx, inv_dy, out);
}

}  // namespace

// enqueues the per-slice moments of the current state into dev_out
// [slice_n][moments_per_slice(n_harm)] on the context's stream
int felis_slice_moments_dev(felis_ctx *ctx, double *dev_out) {
  switch (ctx->d.n_harm) {
    case 1: launch_moments<1>(ctx, dev_out); break;
    case 2: launch_moments<2>(ctx, dev_out); break;
    case 3: launch_moments<3>(ctx, dev_out); break;
    case 4: launch_moments<4>(ctx, dev_out); break;
    case 5: launch_moments<5>(ctx, dev_out); break;
    default: return fail(ctx, FELIS_ERR_UNSUPPORTED, "slice moments support up to 5 harmonics");
  }
  SCHK(hipGetLastError());
  return FELIS_OK;
}

// moments into M, then the network inputs into X (device buffers)
int felis_slice_features_dev(felis_ctx *ctx, double *M, float *X) {
  int e = felis_slice_moments_dev(ctx, M);
  if (e) return e;
  const felis_desc &d = ctx->d;
  const double amp_kA = kECharge * kC / (d.lambda0 * d.num_slice) * 1e-3;
  hipLaunchKernelGGL(k_slice_features, dim3((unsigned)((d.slice_n + 127) / 128)), dim3(128), 0, ctx->stream,
                     d.slice_n, d.n_harm, amp_kA, d.gamma_ref, M, X);
  SCHK(hipGetLastError());
  return FELIS_OK;
}

// the network's outputs [slice_n][outputs_per_slice] into the per-system
// emission factor
int felis_decode_corr_dev(felis_ctx *ctx, const float *Y, void *gain) {
  hipLaunchKernelGGL(k_decode_corr, dim3((unsigned)((ctx->d.slice_n + 127) / 128)), dim3(128), 0, ctx->stream,
                     ctx->d.slice_n, ctx->d.n_harm, Y, (cpx<float> *)gain);
  SCHK(hipGetLastError());
  return FELIS_OK;
}

// sums over slices of the first three moments (sum w, sum w dg, sum w dg^2)
__global__ void k_moment_totals(int slice_n, int nq, const double *__restrict__ M, double *__restrict__ out) {
  __shared__ double sh[3][256];
  double a[3] = {0.0, 0.0, 0.0};
  for (int s = threadIdx.x; s < slice_n; s += blockDim.x)
    for (int q = 0; q < 3; ++q) a[q] += M[(long long)s * nq + q];
  for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] = sh[threadIdx.x][0];
}

int felis_moment_totals_dev(felis_ctx *ctx, const double *M, double *out3) {
  hipLaunchKernelGGL(k_moment_totals, dim3(1), dim3(256), 0, ctx->stream, ctx->d.slice_n,
                     moments_per_slice(ctx->d.n_harm), M, out3);
  SCHK(hipGetLastError());
  return FELIS_OK;
}

extern "C" {

int felis_state_save(felis_ctx *ctx, int slot) {
  if (!ctx || slot < 0 || slot > 1) return FELIS_ERR_ARG;
  if (!ctx->mixed() || ctx->d.dimension != 3) return fail(ctx, FELIS_ERR_UNSUPPORTED, "snapshots: batched 3D path");
  if (!ctx->np) return fail(ctx, FELIS_ERR_STATE, "no particles uploaded");
  felis_ctx::Snap &sn = ctx->snap[slot];
  if (sn.np >= 0 && sn.np != ctx->np) return fail(ctx, FELIS_ERR_STATE, "snapshot holds another particle count");
  if (!sn.U) {
    SCHK(hipMalloc(&sn.U, field_bytes(ctx)));
    ctx->allocs.push_back(sn.U);
    SCHK(hipMalloc(&sn.part, part_bytes(ctx)));
    ctx->allocs.push_back(sn.part);
    SCHK(hipMalloc((void **)&sn.wake, (size_t)ctx->d.slice_n * sizeof(double)));
    ctx->allocs.push_back(sn.wake);
    sn.np = ctx->np;
  }
  SCHK(hipMemcpyAsync(sn.U, ctx->U, field_bytes(ctx), hipMemcpyDeviceToDevice, ctx->stream));
  void *p[6];
  size_t b[6];
  part_arrays(ctx, p, b);
  char *dst = (char *)sn.part;
  for (int k = 0; k < 6; ++k) {
    SCHK(hipMemcpyAsync(dst, p[k], b[k], hipMemcpyDeviceToDevice, ctx->stream));
    dst += b[k];
  }
  SCHK(hipMemcpyAsync(sn.wake, ctx->wakeE, (size_t)ctx->d.slice_n * sizeof(double), hipMemcpyDeviceToDevice,
                      ctx->stream));
  sn.ring_base = ctx->ring_base;
  sn.valid = true;
  return FELIS_OK;
}

int felis_state_restore(felis_ctx *ctx, int slot) {
  if (!ctx || slot < 0 || slot > 1) return FELIS_ERR_ARG;
  felis_ctx::Snap &sn = ctx->snap[slot];
  if (!sn.valid || sn.np != ctx->np) return fail(ctx, FELIS_ERR_STATE, "no snapshot of this state in the slot");
  SCHK(hipMemcpyAsync(ctx->U, sn.U, field_bytes(ctx), hipMemcpyDeviceToDevice, ctx->stream));
  void *p[6];
  size_t b[6];
  part_arrays(ctx, p, b);
  const char *src = (const char *)sn.part;
  for (int k = 0; k < 6; ++k) {
    SCHK(hipMemcpyAsync(p[k], src, b[k], hipMemcpyDeviceToDevice, ctx->stream));
    src += b[k];
  }
  SCHK(hipMemcpyAsync(ctx->wakeE, sn.wake, (size_t)ctx->d.slice_n * sizeof(double), hipMemcpyDeviceToDevice,
                      ctx->stream));
  ctx->ring_base = sn.ring_base;
  return FELIS_OK;
}

int felis_slice_moments(felis_ctx *ctx, double *out) {
  if (!ctx || !out) return FELIS_ERR_ARG;
  if (!ctx->mixed() || ctx->d.dimension != 3) return fail(ctx, FELIS_ERR_UNSUPPORTED, "moments: batched 3D path");
  if (!ctx->np) return fail(ctx, FELIS_ERR_STATE, "no particles uploaded");
  const size_t n = (size_t)ctx->d.slice_n * moments_per_slice(ctx->d.n_harm);
  double *dev = nullptr;
  SCHK(hipMalloc((void **)&dev, n * sizeof(double)));
  int e = felis_slice_moments_dev(ctx, dev);
  if (!e && (hipMemcpyAsync(out, dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
    e = fail(ctx, FELIS_ERR_HIP, "slice moments read");
  (void)hipFree(dev);
  return e;
}

int felis_slice_features(felis_ctx *ctx, float *out) {
  if (!ctx || !out) return FELIS_ERR_ARG;
  if (!ctx->mixed() || ctx->d.dimension != 3) return fail(ctx, FELIS_ERR_UNSUPPORTED, "features: batched 3D path");
  if (!ctx->np) return fail(ctx, FELIS_ERR_STATE, "no particles uploaded");
  const int nh = ctx->d.n_harm, sn = ctx->d.slice_n;
  const size_t nm = (size_t)sn * moments_per_slice(nh), nf = (size_t)sn * features_per_slice(nh);
  double *dev = nullptr;
  SCHK(hipMalloc((void **)&dev, nm * sizeof(double) + nf * sizeof(float)));
  float *X = (float *)(dev + nm);
  int e = felis_slice_features_dev(ctx, dev, X);
  if (!e && (hipMemcpyAsync(out, X, nf * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
    e = fail(ctx, FELIS_ERR_HIP, "slice features read");
  (void)hipFree(dev);
  return e;
}

int felis_jump_residual(felis_ctx *ctx, int slot, double *field_out, double *part_out) {
  if (!ctx || slot < 0 || slot > 1) return FELIS_ERR_ARG;
  felis_ctx::Snap &sn = ctx->snap[slot];
  if (!sn.valid || sn.np != ctx->np) return fail(ctx, FELIS_ERR_STATE, "no snapshot of this state in the slot");
  if (sn.ring_base != ctx->ring_base) return fail(ctx, FELIS_ERR_STATE, "snapshot and state differ in slippage");
  const felis_desc &d = ctx->d;
  const int nsys = 2 * d.n_harm * d.slice_n;
  const size_t nf = (size_t)nsys * 4, np_ = (size_t)d.slice_n * kPartCmp;
  double *dev = nullptr;
  SCHK(hipMalloc((void **)&dev, (nf + np_) * sizeof(double)));
  GridArgs g = ctx->grid();
  hipLaunchKernelGGL(k_field_cmp, dim3(nsys), dim3(256), 0, ctx->stream, g, (const cpx<float> *)ctx->U,
                     (const cpx<float> *)sn.U, dev);
  const double *thb = (const double *)sn.part;
  const float *gb = (const float *)((const char *)sn.part + (size_t)ctx->np * (sizeof(double) + 4 * ctx->rsize()));
  hipLaunchKernelGGL(k_part_cmp, dim3(d.slice_n), dim3(256), 0, ctx->stream, g, d.n_species, ctx->seg_off,
                     ctx->seg_w, ctx->theta, (const float *)ctx->gamma, thb, gb, dev + nf);
  int e = FELIS_OK;
  if (hipGetLastError() != hipSuccess ||
      (field_out && hipMemcpyAsync(field_out, dev, nf * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) !=
                        hipSuccess) ||
      (part_out && hipMemcpyAsync(part_out, dev + nf, np_ * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) !=
                       hipSuccess) ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    e = fail(ctx, FELIS_ERR_HIP, "jump residual");
  (void)hipFree(dev);
  return e;
}

int felis_jump_emission_residual(felis_ctx *ctx, int slot, double *out) {
  if (!ctx || slot < 0 || slot > 1 || !out) return FELIS_ERR_ARG;
  felis_ctx::Snap &sn = ctx->snap[slot];
  if (!sn.valid || sn.np != ctx->np) return fail(ctx, FELIS_ERR_STATE, "no snapshot of this state in the slot");
  if (sn.ring_base != ctx->ring_base) return fail(ctx, FELIS_ERR_STATE, "snapshot and state differ in slippage");
  if (!ctx->last_jump || !ctx->Sw) return fail(ctx, FELIS_ERR_STATE, "no jump has run on this context");
  const felis_desc &d = ctx->d;
  const int nsys = 2 * d.n_harm * d.slice_n;
  double *dev = nullptr;
  SCHK(hipMalloc((void **)&dev, (size_t)nsys * 3 * sizeof(double)));
  GridArgs g = ctx->grid();
  hipLaunchKernelGGL(k_emit_cmp, dim3(nsys), dim3(256), 0, ctx->stream, g, ctx->last_jump,
                     (const cpx<float> *)ctx->U, (const cpx<float> *)sn.U, (const cpx<float> *)ctx->Sw, ctx->win2, dev);
  int e = FELIS_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, dev, (size_t)nsys * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) !=
          hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    e = fail(ctx, FELIS_ERR_HIP, "jump emission residual");
  (void)hipFree(dev);
  return e;
}

}  // extern "C"
