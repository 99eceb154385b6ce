// Field diagnostics of diag/felField.cpp computed on the device from the live
// field instead of a field dump: the slice power and the transverse fluence
// of every polarisation and harmonic, and the normalisations the dump records.
#include <cmath>

#include "felis_ctx.h"

using namespace felis;

namespace felis {

// fluence[ph][c] = coef * sum over slices 1..slice_n of |U[ph][c]|^2, summed
// in slice order in FP64 (felField.cpp:142-158 sums the same way per cell).
// grid: x = cells / 256, y = pol x harmonic
template <typename T>
static __global__ void __launch_bounds__(256)
    k_fluence(GridArgs g, const cpx<T> *__restrict__ U, double coef, double *__restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const int ph = blockIdx.y;
  const cpx<T> *u = U + (long long)ph * g.S * g.N + c;
  double s = 0.0;
  for (int sl = 1; sl <= g.slice_n; ++sl) {
    const cpx<T> v = u[(long long)phys_slot(sl, g.ring_base, g.S) * g.N];
    s += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
  }
  out[(long long)ph * g.N + c] = s * coef;
}

}  // namespace felis

// saveFieldHDF.cpp:66-73: powerNorm (W per sum of |U|^2 over a slice's cells)
// and intensityNorm (J/m^2 per sum of |U|^2 over the slices at one cell).
// Constants of the reference's constants.h.
extern "C" int felis_field_norms(const felis_ctx *ctx, double spot_sig_r, double *power_norm,
                                 double *intensity_norm) {
  if (!ctx) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  const double eMass = 9.10938291e-31, velocityC = 299792458.0, eCharge = 1.60217662e-19, Z0 = 376.730313668;
  const double area = d.dimension == 1 ? 2.0 * M_PI * spot_sig_r * spot_sig_r : d.dx * d.dy;
  const double coef = eMass * velocityC * velocityC * d.ks / eCharge;
  const double bucket_z = d.num_slice * d.lambda0;
  if (power_norm) *power_norm = coef * coef / (2.0 * Z0) * area;
  if (intensity_norm) *intensity_norm = coef * coef / (2.0 * Z0) * bucket_z / velocityC;
  return FELIS_OK;
}

extern "C" int felis_field_diag(felis_ctx *ctx, double spot_sig_r, double *power, double *fluence) {
  if (!ctx) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  double pn = 0.0, in = 0.0;
  felis_field_norms(ctx, spot_sig_r, &pn, &in);
  const int nph = 2 * d.n_harm;
  if (power) {
    // per-slice sums of |U|^2 over every cell (felField.cpp:142-159)
    int e = felis_slice_energy(ctx, power);
    if (e) return e;
    for (long long k = 0; k < (long long)nph * d.slice_n; ++k) power[k] *= pn;
  }
  if (fluence) {
    double *dev = nullptr;
    if (hipMalloc((void **)&dev, (size_t)nph * ctx->N * sizeof(double)) != hipSuccess) {
      ctx->err = "felis_field_diag: device buffer";
      return FELIS_ERR_HIP;
    }
    GridArgs g = ctx->grid();
    const dim3 grid((unsigned)((ctx->N + 255) / 256), (unsigned)nph);
    if (ctx->mixed())
      hipLaunchKernelGGL(k_fluence<float>, grid, dim3(256), 0, ctx->stream, g, (const cpx<float> *)ctx->U, in, dev);
    else
      hipLaunchKernelGGL(k_fluence<double>, grid, dim3(256), 0, ctx->stream, g, (const cpx<double> *)ctx->U, in, dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(fluence, dev, (size_t)nph * ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) {
      ctx->err = std::string("felis_field_diag: ") + hipGetErrorString(e);
      return FELIS_ERR_HIP;
    }
  }
  return FELIS_OK;
}
