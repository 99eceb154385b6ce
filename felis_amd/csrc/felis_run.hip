// felis_run: the step loop of main.cpp:92-156 (I/O off) on the device.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "felis_ctx.h"

using namespace felis;

#define RCHK(call)                                                              \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
      return FELIS_ERR_HIP;                                                     \
    }                                                                           \
  } while (0)

namespace {

struct RunBuffers {
  double *energy = nullptr, *bfac = nullptr, *twiss = nullptr;
};

// One loop iteration, enqueued on ctx->stream. lat points at the three
// lattice entries of this step: solve (previous half step), (it,0), (it,0.5).
int enqueue_step(felis_ctx *ctx, const felis_lattice *lat, double *e_out, double *b_out, double *t_out) {
  int e;
  const felis_desc &d = ctx->d;
  if (e_out && (e = ctx->energy(e_out))) return e;
  if (t_out && (e = ctx->twiss(t_out))) return e;
  if (b_out && (e = ctx->bfactor(b_out))) return e;
  if ((e = ctx->solve_field(&lat[0]))) return e;
  if ((e = ctx->transverse(&lat[1]))) return e;
  if ((e = ctx->transverse(&lat[2]))) return e;
  if (!lat[2].drift_flag) e = ctx->push_theta_gamma(&lat[2]);
  else e = ctx->drift(&lat[2]);
  if (e) return e;
  if ((e = ctx->periodic())) return e;
  if (!lat[2].drift_flag && d.mode == 2 && (e = ctx->shift())) return e;
  return FELIS_OK;
}

}  // namespace

extern "C" int felis_run(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
                         const double *wakeF, double *energy, double *bfac, double *twiss) {
  if (!ctx || !lat_tab || n_steps < 0) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  const int nh = d.n_harm, ns = d.n_species;
  const size_t ne = 2 * (size_t)nh, nb = 2 * (size_t)nh + 1, nt = 7 * (size_t)ns;
  double *dev = nullptr;
  size_t per = ne + nb + nt;
  RCHK(hipMalloc((void **)&dev, std::max<size_t>(1, (size_t)n_steps * per) * sizeof(double)));
  int err = FELIS_OK;
  for (int k = 0; k < n_steps && !err; ++k) {
    int it = it0 + k;
    if (d.mode == 2 && wakeF && wake_step > 0 && it % wake_step == 0) {
      err = ctx->update_wake(wakeF, nullptr);
      if (err) break;
    }
    double *row = dev + (size_t)k * per;
    err = enqueue_step(ctx, &lat_tab[2 * k], energy ? row : nullptr, bfac ? row + ne : nullptr,
                       twiss ? row + ne + nb : nullptr);
    if (!err && ((k & 31) == 31 || k == n_steps - 1)) err = ctx->check_err();
  }
  if (!err) {
    std::vector<double> h((size_t)n_steps * per);
    hipError_t he = hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
      ctx->err = std::string("run outputs: ") + hipGetErrorString(he);
      err = FELIS_ERR_HIP;
    } else {
      for (int k = 0; k < n_steps; ++k) {
        const double *row = h.data() + (size_t)k * per;
        if (energy) memcpy(energy + (size_t)k * ne, row, ne * sizeof(double));
        if (bfac) memcpy(bfac + (size_t)k * nb, row + ne, nb * sizeof(double));
        if (twiss) memcpy(twiss + (size_t)k * nt, row + ne + nb, nt * sizeof(double));
      }
    }
  } else {
    (void)hipStreamSynchronize(ctx->stream);
  }
  (void)hipFree(dev);
  return err;
}
