// C-ABI of libfelis_hip.so: context, state transfer and the per-step operator
// surface of the reference (mesh.h:150-178) on one MI355X.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "felis_ctx.h"
#include "felis_kernels.h"

using namespace felis;

// ------------------------------------------------------------------ helpers

#define HIPCHK(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
      return FELIS_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

#define LAUNCHCHK()                                                           \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);      \
      return FELIS_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

static int fail(felis_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// Clears device memory with a kernel (hipMemsetAsync nodes are kept out of
// code that may be captured into a graph).
static int zero_async(felis_ctx *ctx, void *p, size_t bytes) {
  size_t n = bytes / sizeof(unsigned);
  if (n == 0) return FELIS_OK;
  KTimer kt(ctx, FELIS_K_ZERO);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65535 * 8)), dim3(256), 0,
                     ctx->stream, (unsigned *)p, (long long)n, 0u);
  LAUNCHCHK();
  return FELIS_OK;
}

template <typename P>
static int dalloc(felis_ctx *ctx, P **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  HIPCHK(hipMalloc((void **)p, bytes));
  ctx->allocs.push_back((void *)*p);
  return FELIS_OK;
}

GridArgs felis_ctx::grid() const {
  GridArgs g{};
  g.nx = d.nx;
  g.ny = d.ny;
  g.S = S;
  g.slice_n = d.slice_n;
  g.nh = d.n_harm;
  g.nr = d.nr;
  g.L = d.sc_l;
  g.F = d.sc_f;
  g.bo = bo;
  g.bn = d.bessel_n;
  g.ring_base = ring_base;
  g.wake_on = d.wake_on;
  for (int h = 0; h < FELIS_MAX_HARM; ++h) g.harm[h] = h < d.n_harm ? d.harmony[h] : 1;
  g.N = (long long)N;
  g.dx = d.dx;
  g.dy = d.dy;
  g.dr = d.dr;
  g.minX = d.min_x;
  g.minY = d.min_y;
  g.dz = d.dz * dz_mul;
  g.ks = d.ks;
  g.dB = dB;
  g.dPhi = 2 * M_PI * d.num_slice;
  g.gref = mixed() ? d.gamma_ref : 0.0;
  g.ring_dev = ring_dev_active;
  return g;
}

// particlePush.cpp:222-238 (and solveField.cpp:216-231)
int lat_args(felis_ctx *ctx, const felis_lattice *lat, LatArgs *out) {
  LatArgs l{};
  l.K0 = lat->K0;
  l.ue = lat->ue;
  l.K0_alpha = lat->K0_alpha;
  l.ku = lat->ku;
  l.g = lat->g;
  l.und_type = lat->und_type;
  l.current_flag = lat->current_flag;
  double ue = lat->ue;
  std::complex<double> U((1.0 - ue * ue) / (1.0 + ue * ue), 2.0 * ue / (1.0 + ue * ue));
  l.Phi = std::arg(U);
  if (lat->K0_alpha == 1) {
    l.etaXr = std::cos(l.Phi * 0.5);
    l.etaYi = std::sin(l.Phi * 0.5);
  } else if (lat->K0_alpha == -1) {
    l.etaXi = std::sin(l.Phi * 0.5);
    l.etaYr = std::cos(l.Phi * 0.5);
  } else {
    return fail(ctx, FELIS_ERR_ARG, "K0_alpha must be 1 or -1 (particlePush.cpp:228-238)");
  }
  *out = l;
  return FELIS_OK;
}

int dep_args(felis_ctx *ctx, const felis_lattice *lat, DepArgs *out) {
  const felis_desc &d = ctx->d;
  LatArgs l;
  int e = lat_args(ctx, lat, &l);
  if (e) return e;
  double K0 = lat->K0, ue = lat->ue, ks = d.ks;
  DepArgs a{};
  a.coef_U = d.dz * kECharge * kECharge * kMu0 * K0 * std::sqrt(1.0 + ue * ue) /
             (4.0 * ks * kEMass * (d.lambda0 * d.num_slice) * d.dx * d.dy);
  a.xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  a.coef_Ez = kECharge * kC * kC * kMu0 * ks * (1.0 + 0.5 * (1.0 + ue * ue) * K0 * K0) /
              (2.0 * M_PI * d.lambda0 * d.num_slice * d.gamR * d.gamR);
  a.Phi = l.Phi;
  a.etaXr = l.etaXr;
  a.etaXi = l.etaXi;
  a.etaYr = l.etaYr;
  a.etaYi = l.etaYi;
  a.xi2c = std::sqrt(2.0) * K0 / (1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue));
  a.ue = ue;
  a.ka = lat->K0_alpha;
  a.inv_dx = 1.0 / d.dx;
  a.inv_dy = 1.0 / d.dy;
  a.inv_dB = 1.0 / ctx->dB;
  a.run = ctx->tune.deposit_run > 0 ? ctx->tune.deposit_run : 8;  // the default run of the batched deposit
  *out = a;
  return FELIS_OK;
}

int felis_ctx::check_err() {
  unsigned h = 0;
  hipError_t e = hipMemcpyAsync(&h, errw, sizeof(unsigned), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    err = std::string("error word read: ") + hipGetErrorString(e);
    return FELIS_ERR_HIP;
  }
  if (h == 0) return FELIS_OK;
  (void)hipMemsetAsync(errw, 0, sizeof(unsigned), stream);
  if (h & EB_GUARD) {
    err = "dTheta exceeded 2*pi*numSlice (particlePush.cpp:370-374)";
    return FELIS_ERR_GUARD;
  }
  if (h & EB_NAN) {
    err = "NaN theta at the wrap";
    return FELIS_ERR_NAN;
  }
  if (h & EB_RANGE) {
    err = "theta outside [0, 2*pi*numSlice] after the wrap (rearrangeParticles.cpp:105-108)";
    return FELIS_ERR_RANGE;
  }
  if (h & EB_BESSEL) {
    err = "Bessel table index beyond num_grids";
    return FELIS_ERR_UNSUPPORTED;
  }
  if (h & EB_EMCARRY) {
    err = "in-grid particle with idxR >= nr: the reference carries Em from the previous particle "
          "(particlePush.cpp:248), which the parallel push does not model";
    return FELIS_ERR_UNSUPPORTED;
  }
  err = "unknown device error";
  return FELIS_ERR_STATE;
}

// ---------------------------------------------------------------- tables

// boundary.cpp:73-106 and the hoisted Thomas coefficients of
// solveField.cpp:137-152 / 165-180 (field independent: H and the ABC only).
static void build_bessel(felis_ctx *ctx);

// The ABC coefficients alpha, alphaP, beta, betaP (abc) and the hoisted
// Thomas factors per harmonic (adix, adiy) for a field step of length dz.
void abc_tables(const felis_desc &d, double dz, std::vector<cpx<double>> &abc, std::vector<cpx<double>> &adix,
                std::vector<cpx<double>> &adiy) {
  using C = std::complex<double>;
  const C I(0.0, 1.0);
  int nx = d.nx, ny = d.ny, L = d.abc_n;
  std::vector<double> sx(nx, 0.0), sy(ny, 0.0);
  for (int i = 0; i < L && i < nx; ++i) sx[i] = d.abc_sig * (L - i) * (L - i) / (1.0 * L * L);
  for (int i = std::max(0, nx - L); i < nx; ++i) sx[i] = d.abc_sig * (i - nx + L) * (i - nx + L) / (1.0 * L * L);
  for (int j = 0; j < L && j < ny; ++j) sy[j] = d.abc_sig * (L - j) * (L - j) / (1.0 * L * L);
  for (int j = std::max(0, ny - L); j < ny; ++j) sy[j] = d.abc_sig * (j - ny + L) * (j - ny + L) / (1.0 * L * L);
  std::vector<C> cx(nx), cy(ny);
  for (int i = 0; i < nx; ++i) cx[i] = 1.0 + I * sx[i];
  for (int j = 0; j < ny; ++j) cy[j] = 1.0 + I * sy[j];
  std::vector<C> al(nx, 0.0), alP(nx, 0.0), be(ny, 0.0), beP(ny, 0.0);
  for (int i = 1; i < nx - 1; ++i) {
    al[i] = -I * dz / (2.0 * d.ks * d.dx * d.dx * cx[i] * (cx[i] + cx[i - 1]));
    alP[i] = -I * dz / (2.0 * d.ks * d.dx * d.dx * cx[i] * (cx[i] + cx[i + 1]));
  }
  for (int j = 1; j < ny - 1; ++j) {
    be[j] = -I * dz / (2.0 * d.ks * d.dy * d.dy * cy[j] * (cy[j] + cy[j - 1]));
    beP[j] = -I * dz / (2.0 * d.ks * d.dy * d.dy * cy[j] * (cy[j] + cy[j + 1]));
  }
  abc.assign(2 * nx + 2 * ny, mk(0.0, 0.0));
  for (int i = 0; i < nx; ++i) {
    abc[i] = mk(al[i].real(), al[i].imag());
    abc[nx + i] = mk(alP[i].real(), alP[i].imag());
  }
  for (int j = 0; j < ny; ++j) {
    abc[2 * nx + j] = mk(be[j].real(), be[j].imag());
    abc[2 * nx + ny + j] = mk(beP[j].real(), beP[j].imag());
  }
  auto coef = [&](int n, const std::vector<C> &a, const std::vector<C> &aP, double H,
                   cpx<double> *out) {
    // out[0..n) r, [n..2n) a, [2n..3n) q, [3n..4n) p
    std::vector<C> CC(n, 0.0);
    for (int k = 0; k < 4 * n; ++k) out[k] = mk(0.0, 0.0);
    CC[1] = (H - aP[1] - a[1]) / aP[1];
    out[1] = mk((1.0 / aP[1]).real(), (1.0 / aP[1]).imag());
    for (int i = 2; i < n - 2; ++i) {
      CC[i] = (H - aP[i] - a[i]) / aP[i] - a[i] / (aP[i] * CC[i - 1]);
      C r = 1.0 / aP[i], m = -a[i] / (aP[i] * CC[i - 1]);
      out[i] = mk(r.real(), r.imag());
      out[n + i] = mk(m.real(), m.imag());
    }
    CC[n - 2] = H - aP[n - 2] - a[n - 2] - 1.0 / CC[n - 3];
    out[n - 2] = mk(1.0, 0.0);
    C m = -1.0 / CC[n - 3];
    out[n + n - 2] = mk(m.real(), m.imag());
    for (int i = 1; i <= n - 2; ++i) {
      C q = 1.0 / CC[i];
      out[2 * n + i] = mk(q.real(), q.imag());
      C p = (i == n - 2) ? C(0.0) : -q;
      out[3 * n + i] = mk(p.real(), p.imag());
    }
  };
  adix.assign((size_t)d.n_harm * 4 * nx, mk(0.0, 0.0));
  adiy.assign((size_t)d.n_harm * 4 * ny, mk(0.0, 0.0));
  for (int h = 0; h < d.n_harm; ++h) {
    coef(nx, al, alP, (double)d.harmony[h], &adix[(size_t)h * 4 * nx]);
    coef(ny, be, beP, (double)d.harmony[h], &adiy[(size_t)h * 4 * ny]);
  }
}

static void build_tables(felis_ctx *ctx) {
  if (ctx->d.dimension == 3) abc_tables(ctx->d, ctx->d.dz, ctx->h_abc, ctx->h_adix, ctx->h_adiy);
  build_bessel(ctx);
}

// Bessel table, boundary.cpp:109-123 (J_n via the C++17 special functions)
static void build_bessel(felis_ctx *ctx) {
  const felis_desc &d = ctx->d;
  int maxH = d.harmony[d.n_harm - 1];
  double bmax = maxH * 2.0 * M_PI;
  ctx->dB = bmax / (double)d.bessel_n;
  ctx->bo = maxH / 2 + 2;  // one more than the reference's rows for even maxH
  ctx->h_bessel.assign((size_t)d.bessel_n * ctx->bo, 0.0);
  for (int i = 0; i < d.bessel_n; ++i) {
    double xi = ctx->dB * i;
    for (int j = 0; j < ctx->bo; ++j) ctx->h_bessel[(size_t)i * ctx->bo + j] = std::cyl_bessel_j((double)j, xi);
  }
}

// ---------------------------------------------------------------- lifetime

extern "C" int felis_create(const felis_desc *desc, felis_ctx **out) {
  if (!desc || !out) return FELIS_ERR_ARG;
  *out = nullptr;
  const felis_desc &d = *desc;
  if (d.dimension == 1) {
    // solveField.cpp:409-549 / particlePush.cpp:388-524: one cell per slice,
    // FP64 only (parameterSetting.cpp:112, 196-200 fix nx = ny = nr = 1)
    if (d.nx != 1 || d.ny != 1 || d.nr != 1 || d.sc_f != 1 || d.sc_l < 1 || d.sc_l > 16 || d.slice_n < 1 ||
        d.n_harm < 1 || d.n_harm > FELIS_MAX_HARM || d.num_slice < 1 || d.n_species < 1 || d.bessel_n < 2)
      return FELIS_ERR_ARG;
    if (d.precision != FELIS_PLAIN_FP64) return FELIS_ERR_UNSUPPORTED;
  } else if (d.dimension != 3) {
    return FELIS_ERR_UNSUPPORTED;
  } else if (d.nx < 6 || d.ny < 6 || d.slice_n < 1 || d.n_harm < 1 || d.n_harm > FELIS_MAX_HARM ||
             d.num_slice < 1 || d.n_species < 1 || d.nr < 3 || d.nr > 256 || d.sc_l < 1 || d.sc_l > 4 ||
             d.sc_f < 1 || d.bessel_n < 2 || d.abc_n < 1 || d.nx > 64 * 4 + 2 || d.ny > 64 * 4 + 2) {
    return FELIS_ERR_ARG;
  }
  if (d.precision != FELIS_PLAIN_FP64 && d.precision != FELIS_MIXED_FP32) return FELIS_ERR_ARG;
  for (int h = 0; h < d.n_harm; ++h)
    if (d.harmony[h] < 1) return FELIS_ERR_ARG;
  felis_ctx *ctx = new felis_ctx();
  ctx->d = d;
  ctx->S = d.slice_n + 2;
  ctx->N = (size_t)d.nx * d.ny;
  int e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess ? FELIS_OK : FELIS_ERR_HIP;
  if (e) {
    ctx->err = "hipStreamCreate failed";
    delete ctx;
    return e;
  }
  build_tables(ctx);
  {
    int dev = 0, lds = 65536;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      ctx->lds_max = lds - 1024;  // room for the kernel's static LDS
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      ctx->n_cu = ncu;
    if (ctx->lds_max > 65536 - 1024) {
      // opt in to more than 64 KiB of dynamic LDS where the device allows it
      if (hipFuncSetAttribute((const void *)k_deposit_win<double, double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, ctx->lds_max) != hipSuccess ||
          hipFuncSetAttribute((const void *)k_deposit_win<float, float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, ctx->lds_max) != hipSuccess)
        ctx->lds_max = 65536 - 1024;
    }
    (void)hipGetLastError();
  }
  const size_t es = ctx->esize();
  const size_t fsz = (size_t)2 * d.n_harm * ctx->S * ctx->N * es;
  const int nsys = 2 * d.n_harm * d.slice_n;
  const size_t nrow = (size_t)d.sc_f * d.nr + 1;
  if ((e = dalloc(ctx, &ctx->U, fsz)) || (e = dalloc(ctx, &ctx->Sc, fsz)) ||
      (e = dalloc(ctx, &ctx->Uc, (size_t)nsys * ctx->N * es)) ||
      (e = dalloc(ctx, &ctx->Ez, (size_t)d.sc_l * ctx->S * d.sc_f * d.nr * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->ScR, (size_t)ctx->S * d.sc_l * nrow * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->ez_b, (size_t)2 * d.sc_l * d.sc_f * d.nr * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->ez_d, (size_t)d.slice_n * d.sc_l * d.sc_f * d.nr * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->abc, ctx->h_abc.size() * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->adix, ctx->h_adix.size() * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->adiy, ctx->h_adiy.size() * sizeof(cpx<double>))) ||
      (e = dalloc(ctx, &ctx->bessel, ctx->h_bessel.size() * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->bessel_f, ctx->h_bessel.size() * sizeof(float))) ||
      (e = dalloc(ctx, &ctx->wakeE, (size_t)d.slice_n * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->wakeF, (size_t)d.slice_n * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->den, (size_t)d.slice_n * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->slice_e, (size_t)nsys * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->win, (size_t)ctx->S * sizeof(int4))) ||
      (e = dalloc(ctx, &ctx->beam_sig, (size_t)2 * d.n_species * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->red_out, (size_t)1024 * sizeof(double))) ||
      (e = dalloc(ctx, &ctx->errw, sizeof(unsigned)))) {
    felis_destroy(ctx);
    return e;
  }
  (void)hipMemsetAsync(ctx->U, 0, fsz, ctx->stream);
  (void)hipMemsetAsync(ctx->Sc, 0, fsz, ctx->stream);
  (void)hipMemsetAsync(ctx->Uc, 0, (size_t)nsys * ctx->N * es, ctx->stream);
  (void)hipMemsetAsync(ctx->Ez, 0, (size_t)d.sc_l * ctx->S * d.sc_f * d.nr * sizeof(cpx<double>), ctx->stream);
  (void)hipMemsetAsync(ctx->wakeE, 0, (size_t)d.slice_n * sizeof(double), ctx->stream);
  (void)hipMemsetAsync(ctx->errw, 0, sizeof(unsigned), ctx->stream);
  (void)hipMemcpyAsync(ctx->abc, ctx->h_abc.data(), ctx->h_abc.size() * sizeof(cpx<double>), hipMemcpyHostToDevice, ctx->stream);
  (void)hipMemcpyAsync(ctx->adix, ctx->h_adix.data(), ctx->h_adix.size() * sizeof(cpx<double>), hipMemcpyHostToDevice, ctx->stream);
  (void)hipMemcpyAsync(ctx->adiy, ctx->h_adiy.data(), ctx->h_adiy.size() * sizeof(cpx<double>), hipMemcpyHostToDevice, ctx->stream);
  (void)hipMemcpyAsync(ctx->bessel, ctx->h_bessel.data(), ctx->h_bessel.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  const std::vector<float> bes_f(ctx->h_bessel.begin(), ctx->h_bessel.end());
  (void)hipMemcpyAsync(ctx->bessel_f, bes_f.data(), bes_f.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    felis_destroy(ctx);
    return FELIS_ERR_HIP;
  }
  *out = ctx;
  return FELIS_OK;
}

void felis_ctx::free_particles() {
  for (void *p : part_allocs) (void)hipFree(p);
  part_allocs.clear();
  np = 0;
  ntiles = 0;
}

hipEvent_t felis_ctx::ev_get() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

// Folds recorded event pairs into the per-kernel totals.
void felis_ctx::ev_flush() {
  if (ev_marks.empty()) {
    ev_used = 0;
    return;
  }
  (void)hipStreamSynchronize(stream);
  for (auto &m : ev_marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.second.first, m.second.second) == hipSuccess) {
      prof_ms[m.first] += ms;
      prof_n[m.first] += 1;
    }
  }
  ev_marks.clear();
  ev_used = 0;
}

extern "C" int felis_profile(felis_ctx *ctx, int enable) {
  if (!ctx) return FELIS_ERR_ARG;
  ctx->ev_flush();
  if (enable) {
    for (int k = 0; k < FELIS_K_COUNT; ++k) {
      ctx->prof_ms[k] = 0;
      ctx->prof_n[k] = 0;
    }
  }
  ctx->profiling = enable != 0;
  return FELIS_OK;
}

extern "C" int felis_profile_read(felis_ctx *ctx, double *ms, int64_t *launches) {
  if (!ctx) return FELIS_ERR_ARG;
  ctx->ev_flush();
  for (int k = 0; k < FELIS_K_COUNT; ++k) {
    if (ms) ms[k] = ctx->prof_ms[k];
    if (launches) launches[k] = ctx->prof_n[k];
  }
  return FELIS_OK;
}

extern "C" int felis_destroy(felis_ctx *ctx) {
  if (!ctx) return FELIS_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->ev_marks.clear();
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  ctx->release_graph();
  ctx->free_particles();
  for (void *p : ctx->allocs) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FELIS_OK;
}

extern "C" const char *felis_last_error(const felis_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int felis_get_desc(const felis_ctx *ctx, felis_desc *out) {
  if (!ctx || !out) return FELIS_ERR_ARG;
  *out = ctx->d;
  return FELIS_OK;
}

extern "C" int felis_particle_layout(felis_ctx *ctx, int64_t *seg_off, double *seg_weight, int64_t *n_particles) {
  if (!ctx) return FELIS_ERR_ARG;
  if (n_particles) *n_particles = ctx->np;
  if (ctx->h_seg_off.empty()) {
    if (!seg_off && !seg_weight) return FELIS_OK;
    ctx->err = "felis_particle_layout: no particles uploaded";
    return FELIS_ERR_STATE;
  }
  if (seg_off)
    for (size_t k = 0; k < ctx->h_seg_off.size(); ++k) seg_off[k] = ctx->h_seg_off[k];
  if (seg_weight)
    for (size_t k = 0; k < ctx->h_seg_w.size(); ++k) seg_weight[k] = ctx->h_seg_w[k];
  return FELIS_OK;
}

extern "C" int felis_device_info(int *info) {
  if (!info) return FELIS_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FELIS_ERR_HIP;
  const hipDeviceAttribute_t at[8] = {hipDeviceAttributeMultiprocessorCount,
                                      hipDeviceAttributeMaxSharedMemoryPerBlock,
                                      hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                      hipDeviceAttributeWarpSize,
                                      hipDeviceAttributeL2CacheSize,
                                      hipDeviceAttributeClockRate,
                                      hipDeviceAttributeMemoryClockRate,
                                      hipDeviceAttributeMemoryBusWidth};
  for (int k = 0; k < 8; ++k) {
    info[k] = 0;
    if (hipDeviceGetAttribute(&info[k], at[k], dev) != hipSuccess) return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}

extern "C" int felis_sync(felis_ctx *ctx) {
  if (!ctx) return FELIS_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ctx->check_err();
}

extern "C" int felis_get_tables(felis_ctx *ctx, double *alpha, double *alphaP, double *beta,
                                double *betaP, double *bessel) {
  if (!ctx) return FELIS_ERR_ARG;
  int nx = ctx->d.nx, ny = ctx->d.ny;
  if (alpha) memcpy(alpha, &ctx->h_abc[0], nx * sizeof(cpx<double>));
  if (alphaP) memcpy(alphaP, &ctx->h_abc[nx], nx * sizeof(cpx<double>));
  if (beta) memcpy(beta, &ctx->h_abc[2 * nx], ny * sizeof(cpx<double>));
  if (betaP) memcpy(betaP, &ctx->h_abc[2 * nx + ny], ny * sizeof(cpx<double>));
  if (bessel) memcpy(bessel, ctx->h_bessel.data(), ctx->h_bessel.size() * sizeof(double));
  return FELIS_OK;
}

// ---------------------------------------------------------------- transfers

template <typename R>
static int upload_coord(felis_ctx *ctx, void *dst, const double *src, double off) {
  size_t n = (size_t)ctx->np;
  if (!src) return fail(ctx, FELIS_ERR_ARG, "null particle array");
  if (sizeof(R) == sizeof(double) && off == 0.0) {
    HIPCHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    return FELIS_OK;
  }
  std::vector<R> tmp(n);
  for (size_t k = 0; k < n; ++k) tmp[k] = (R)(src[k] - off);
  HIPCHK(hipMemcpy(dst, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice));
  return FELIS_OK;
}

template <typename R>
static int download_coord(felis_ctx *ctx, const void *src, double *dst, double off) {
  size_t n = (size_t)ctx->np;
  if (!dst) return FELIS_OK;
  if (sizeof(R) == sizeof(double) && off == 0.0) {
    HIPCHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    return FELIS_OK;
  }
  std::vector<R> tmp(n);
  HIPCHK(hipMemcpy(tmp.data(), src, n * sizeof(R), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < n; ++k) dst[k] = (double)tmp[k] + off;
  return FELIS_OK;
}

// Whether an in-grid particle (idxI in [0, nx-2], idxJ in [0, ny-2] after
// truncation toward zero, so x from minX - dx on) can have idxR >= nr, where
// the reference carries Em (particlePush.cpp:283-304). With the default
// radial grid (parameterSetting.cpp:193-195) it cannot; a small
// radial_grids can. Decided with a margin, so that no such particle goes
// unprepared.
static bool em_carry_possible(const felis_desc &d) {
  if (d.dr <= 0) return false;
  const double xm = std::max(std::fabs(d.min_x - d.dx), std::fabs(d.min_x + (d.nx - 1) * d.dx));
  const double ym = std::max(std::fabs(d.min_y - d.dy), std::fabs(d.min_y + (d.ny - 1) * d.dy));
  return std::sqrt(xm * xm + ym * ym) / d.dr + 0.5 >= d.nr * (1.0 - 1e-9);
}

// Allocates the particle arrays and the tile table for a segment layout.
int felis_ctx::alloc_particles(const int64_t *seg_off, const double *seg_weight) {
  felis_ctx *ctx = this;
  int nseg = d.n_species * d.slice_n;
  if (seg_off[0] != 0) return fail(ctx, FELIS_ERR_ARG, "seg_off[0] must be 0");
  for (int s = 0; s < nseg; ++s)
    if (seg_off[s + 1] < seg_off[s]) return fail(ctx, FELIS_ERR_ARG, "seg_off must be non-decreasing");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->release_graph();
  ctx->free_particles();
  ctx->np = seg_off[nseg];
  ctx->h_seg_off.assign(seg_off, seg_off + nseg + 1);
  ctx->h_seg_w.assign(seg_weight, seg_weight + nseg);
  // tiles of <= kTile particles, never straddling a segment
  std::vector<int> tseg, tcnt, segt(nseg + 1, 0);
  std::vector<long long> tstart;
  ctx->species_tile.assign(d.n_species + 1, 0);
  for (int s = 0; s < nseg; ++s) {
    if (s % d.slice_n == 0) ctx->species_tile[s / d.slice_n] = (int)tseg.size();
    segt[s] = (int)tseg.size();
    for (long long b = seg_off[s]; b < seg_off[s + 1]; b += kTile) {
      tseg.push_back(s);
      tstart.push_back(b);
      tcnt.push_back((int)std::min<long long>(kTile, seg_off[s + 1] - b));
    }
  }
  ctx->species_tile[d.n_species] = (int)tseg.size();
  segt[nseg] = (int)tseg.size();
  ctx->ntiles = (int)tseg.size();
  std::vector<int> cseg, ccnt;
  std::vector<long long> cstart;
  for (int s = 0; s < nseg; ++s)
    for (long long b = seg_off[s]; b < seg_off[s + 1]; b += kChunk) {
      cseg.push_back(s);
      cstart.push_back(b);
      ccnt.push_back((int)std::min<long long>(kChunk, seg_off[s + 1] - b));
    }
  ctx->nchunks = (int)cseg.size();
  const size_t np = (size_t)ctx->np, rs = ctx->rsize();
  auto pal = [&](void **p, size_t bytes) -> int {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      ctx->err = std::string("hipMalloc particles: ") + hipGetErrorString(e);
      return FELIS_ERR_HIP;
    }
    ctx->part_allocs.push_back(*p);
    return FELIS_OK;
  };
  int e;
  if ((e = pal((void **)&ctx->theta, np * sizeof(double))) || (e = pal(&ctx->gamma, np * rs)) ||
      (e = pal(&ctx->x, np * rs)) || (e = pal(&ctx->y, np * rs)) || (e = pal(&ctx->px, np * rs)) ||
      (e = pal(&ctx->py, np * rs)) || (e = pal((void **)&ctx->seg_off, (nseg + 1) * sizeof(long long))) ||
      (e = pal((void **)&ctx->seg_w, nseg * sizeof(double))) ||
      (e = pal((void **)&ctx->tile_seg, tseg.size() * sizeof(int))) ||
      (e = pal((void **)&ctx->tile_start, tstart.size() * sizeof(long long))) ||
      (e = pal((void **)&ctx->tile_cnt, tcnt.size() * sizeof(int))) ||
      (e = pal((void **)&ctx->seg_tile, segt.size() * sizeof(int))) ||
      (e = pal((void **)&ctx->tile_bb, tseg.size() * sizeof(int4))) ||
      (e = pal((void **)&ctx->tile_red, std::max<size_t>(tseg.size(), 1) * (2 * FELIS_MAX_HARM + 1) * sizeof(double))) ||
      (e = pal((void **)&ctx->chunk_seg, cseg.size() * sizeof(int))) ||
      (e = pal((void **)&ctx->chunk_start, cstart.size() * sizeof(long long))) ||
      (e = pal((void **)&ctx->chunk_cnt, ccnt.size() * sizeof(int))))
    return e;
  ctx->em_carry = d.dimension == 3 && em_carry_possible(d);
  ctx->emc = nullptr;
  ctx->em_last = ctx->em_src = nullptr;
  if (ctx->em_carry &&
      ((e = pal(&ctx->emc, np * d.sc_l * 2 * sizeof(double))) ||
       (e = pal((void **)&ctx->em_last, tseg.size() * sizeof(int))) ||
       (e = pal((void **)&ctx->em_src, tseg.size() * sizeof(int)))))
    return e;
  if (ctx->em_carry) HIPCHK(hipMemset(ctx->emc, 0, np * d.sc_l * 2 * sizeof(double)));
  HIPCHK(hipMemcpy(ctx->seg_off, seg_off, (nseg + 1) * sizeof(long long), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->seg_w, seg_weight, nseg * sizeof(double), hipMemcpyHostToDevice));
  if (!tseg.empty()) {
    HIPCHK(hipMemcpy(ctx->tile_seg, tseg.data(), tseg.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tile_start, tstart.data(), tstart.size() * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tile_cnt, tcnt.data(), tcnt.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->seg_tile, segt.data(), segt.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->chunk_seg, cseg.data(), cseg.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->chunk_start, cstart.data(), cstart.size() * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->chunk_cnt, ccnt.data(), ccnt.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return FELIS_OK;
}

extern "C" int felis_upload_particles(felis_ctx *ctx, const int64_t *seg_off, const double *seg_weight,
                                      const double *x, const double *y, const double *px,
                                      const double *py, const double *theta, const double *gamma) {
  if (!ctx || !seg_off || !seg_weight) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  int e = ctx->alloc_particles(seg_off, seg_weight);
  if (e) return e;
  const size_t np = (size_t)ctx->np;
  if (np) {
    if (!theta) return fail(ctx, FELIS_ERR_ARG, "null theta");
    HIPCHK(hipMemcpy(ctx->theta, theta, np * sizeof(double), hipMemcpyHostToDevice));
    double gref = ctx->mixed() ? d.gamma_ref : 0.0;
    if (ctx->mixed()) {
      if ((e = upload_coord<float>(ctx, ctx->gamma, gamma, gref)) || (e = upload_coord<float>(ctx, ctx->x, x, 0)) ||
          (e = upload_coord<float>(ctx, ctx->y, y, 0)) || (e = upload_coord<float>(ctx, ctx->px, px, 0)) ||
          (e = upload_coord<float>(ctx, ctx->py, py, 0)))
        return e;
    } else {
      if ((e = upload_coord<double>(ctx, ctx->gamma, gamma, 0)) || (e = upload_coord<double>(ctx, ctx->x, x, 0)) ||
          (e = upload_coord<double>(ctx, ctx->y, y, 0)) || (e = upload_coord<double>(ctx, ctx->px, px, 0)) ||
          (e = upload_coord<double>(ctx, ctx->py, py, 0)))
        return e;
    }
  }
  return FELIS_OK;
}

extern "C" int felis_download_particles(felis_ctx *ctx, double *x, double *y, double *px, double *py,
                                        double *theta, double *gamma) {
  if (!ctx) return FELIS_ERR_ARG;
  if (!ctx->np) return FELIS_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (theta) HIPCHK(hipMemcpy(theta, ctx->theta, ctx->np * sizeof(double), hipMemcpyDeviceToHost));
  int e;
  if (ctx->mixed()) {
    double gref = ctx->d.gamma_ref;
    if ((e = download_coord<float>(ctx, ctx->gamma, gamma, gref)) || (e = download_coord<float>(ctx, ctx->x, x, 0)) ||
        (e = download_coord<float>(ctx, ctx->y, y, 0)) || (e = download_coord<float>(ctx, ctx->px, px, 0)) ||
        (e = download_coord<float>(ctx, ctx->py, py, 0)))
      return e;
  } else {
    if ((e = download_coord<double>(ctx, ctx->gamma, gamma, 0)) || (e = download_coord<double>(ctx, ctx->x, x, 0)) ||
        (e = download_coord<double>(ctx, ctx->y, y, 0)) || (e = download_coord<double>(ctx, ctx->px, px, 0)) ||
        (e = download_coord<double>(ctx, ctx->py, py, 0)))
      return e;
  }
  return FELIS_OK;
}

// Field transfer between the reference layout [h][slot*N + cell] (complex
// double) and the device layout [pol][h][phys slot][cell] of T.
extern "C" int felis_upload_field(felis_ctx *ctx, int which, const double *data) {
  if (!ctx || !data) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (which == FELIS_EZ) {
    HIPCHK(hipMemcpy(ctx->Ez, data, (size_t)d.sc_l * ctx->S * d.sc_f * d.nr * sizeof(cpx<double>), hipMemcpyHostToDevice));
    return FELIS_OK;
  }
  if (which < 0 || which > 3) return FELIS_ERR_ARG;
  const bool isU = which < 2;
  const int pol = which & 1;
  if (pol == 1) {
    const size_t n = 2 * (size_t)d.n_harm * ctx->S * ctx->N;
    for (size_t k = 0; k < n && !ctx->uy_live; ++k)
      if (data[k] != 0.0) ctx->uy_live = true;
  }
  const size_t N = ctx->N, S = ctx->S, es = ctx->esize();
  char *base = (char *)(isU ? ctx->U : ctx->Sc);
  std::vector<char> tmp(N * es);
  for (int h = 0; h < d.n_harm; ++h)
    for (int slot = 0; slot < (int)S; ++slot) {
      int ph = isU ? (slot + ctx->ring_base) % (int)S : slot;
      const double *src = data + 2 * ((size_t)h * S * N + (size_t)slot * N);
      if (ctx->mixed()) {
        float *t = (float *)tmp.data();
        for (size_t k = 0; k < 2 * N; ++k) t[k] = (float)src[k];
      } else {
        memcpy(tmp.data(), src, N * es);
      }
      size_t off = (((size_t)pol * d.n_harm + h) * S + ph) * N * es;
      HIPCHK(hipMemcpy(base + off, tmp.data(), N * es, hipMemcpyHostToDevice));
    }
  return FELIS_OK;
}

extern "C" int felis_download_field(felis_ctx *ctx, int which, double *data) {
  if (!ctx || !data) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (which == FELIS_EZ) {
    HIPCHK(hipMemcpy(data, ctx->Ez, (size_t)d.sc_l * ctx->S * d.sc_f * d.nr * sizeof(cpx<double>), hipMemcpyDeviceToHost));
    return FELIS_OK;
  }
  if (which < 0 || which > 3) return FELIS_ERR_ARG;
  const bool isU = which < 2;
  const int pol = which & 1;
  if (pol == 1) {
    const size_t n = 2 * (size_t)d.n_harm * ctx->S * ctx->N;
    for (size_t k = 0; k < n && !ctx->uy_live; ++k)
      if (data[k] != 0.0) ctx->uy_live = true;
  }
  const size_t N = ctx->N, S = ctx->S, es = ctx->esize();
  const char *base = (const char *)(isU ? ctx->U : ctx->Sc);
  std::vector<char> tmp(N * es);
  for (int h = 0; h < d.n_harm; ++h)
    for (int slot = 0; slot < (int)S; ++slot) {
      int ph = isU ? (slot + ctx->ring_base) % (int)S : slot;
      size_t off = (((size_t)pol * d.n_harm + h) * S + ph) * N * es;
      HIPCHK(hipMemcpy(tmp.data(), base + off, N * es, hipMemcpyDeviceToHost));
      double *dst = data + 2 * ((size_t)h * S * N + (size_t)slot * N);
      if (ctx->mixed()) {
        const float *t = (const float *)tmp.data();
        for (size_t k = 0; k < 2 * N; ++k) dst[k] = t[k];
      } else {
        memcpy(dst, tmp.data(), N * es);
      }
    }
  return FELIS_OK;
}

extern "C" int felis_set_wake(felis_ctx *ctx, const double *wakeE) {
  if (!ctx || !wakeE) return FELIS_ERR_ARG;
  HIPCHK(hipMemcpyAsync(ctx->wakeE, wakeE, ctx->d.slice_n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return FELIS_OK;
}

// solveField.cpp:463-468: the rms beam sizes whose Gaussian area normalises
// the 1D source (sigma = sqrt(emittance / gamma0 / gamma_twiss) per species).
extern "C" int felis_set_beam_sigma(felis_ctx *ctx, const double *sig) {
  if (!ctx || !sig) return FELIS_ERR_ARG;
  for (int s = 0; s < 2 * ctx->d.n_species; ++s)
    if (!(sig[s] > 0.0)) {
      ctx->err = "beam sigma must be positive";
      return FELIS_ERR_ARG;
    }
  HIPCHK(hipMemcpyAsync(ctx->beam_sig, sig, 2 * ctx->d.n_species * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->sig_set = true;
  return FELIS_OK;
}

// ---------------------------------------------------------------- operators

template <typename R, typename T>
static int launch_transverse(felis_ctx *ctx, const LatArgs &l) {
  if (!ctx->np) return FELIS_OK;
  GridArgs g = ctx->grid();
  unsigned nb = (unsigned)((ctx->np + 255) / 256);
  KTimer kt(ctx, FELIS_K_TRANSVERSE);
  hipLaunchKernelGGL((k_transverse<R>), dim3(nb), dim3(256), 0, ctx->stream, (long long)ctx->np, g, l,
                     (R *)ctx->x, (R *)ctx->y, (R *)ctx->px, (R *)ctx->py, (const R *)ctx->gamma);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::transverse(const felis_lattice *lat) {
  if (lat->und_type != 1 && lat->und_type != 2)
    return fail(this, FELIS_ERR_ARG, "undulator_type must be BiPolar or QuadPolar (particlePush.cpp:53-68)");
  // transversePush reads only K0, ue, K0_alpha, ku, g and the type
  LatArgs l{};
  l.K0 = lat->K0;
  l.ue = lat->ue;
  l.K0_alpha = lat->K0_alpha;
  l.ku = lat->ku;
  l.g = lat->g;
  l.und_type = lat->und_type;
  return mixed() ? launch_transverse<float, float>(this, l) : launch_transverse<double, double>(this, l);
}

template <typename R, typename T>
static int launch_push(felis_ctx *ctx, const LatArgs &l) {
  if (!ctx->ntiles) return FELIS_OK;
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_PUSH);
  if (ctx->em_carry) {
    const EmCarryPlain<R> pol{g, (const R *)ctx->x, (const R *)ctx->y, (const cpx<double> *)ctx->Ez};
    hipLaunchKernelGGL((k_em_last<EmCarryPlain<R>>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, pol,
                       ctx->tiles(), ctx->em_last);
    hipLaunchKernelGGL(k_em_scan, dim3(1), dim3(1024), 0, ctx->stream, ctx->ntiles, ctx->em_last, ctx->em_src);
    hipLaunchKernelGGL((k_em_fill<EmCarryPlain<R>>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, pol,
                       ctx->tiles(), ctx->em_last, ctx->em_src, (cpx<double> *)ctx->emc);
  }
  hipLaunchKernelGGL((k_push_tg<R, T>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, g, l, ctx->tiles(),
                     ctx->theta, (R *)ctx->gamma, (const R *)ctx->x, (const R *)ctx->y, (const R *)ctx->px,
                     (const R *)ctx->py, (const cpx<T> *)ctx->U,
                     (const cpx<T> *)((char *)ctx->U + ctx->pol_bytes()), (const cpx<double> *)ctx->Ez,
                     ctx->bessel, ctx->wakeE, ctx->em_carry ? (const cpx<double> *)ctx->emc : nullptr, ctx->errw);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::push_theta_gamma(const felis_lattice *lat) {
  LatArgs l;
  int e = lat_args(this, lat, &l);
  if (e) return e;
  if (d.dimension == 1) {
    if (!ntiles) return FELIS_OK;
    GridArgs g = grid();
    felis_ctx *ctx = this;
    KTimer kt(ctx, FELIS_K_PUSH);
    hipLaunchKernelGGL(k_push_tg_1d, dim3(ntiles), dim3(kTile), 0, stream, g, l, tiles(), theta, (double *)gamma,
                       (const cpx<double> *)U, (const cpx<double> *)((char *)U + pol_bytes()),
                       (const cpx<double> *)Ez, bessel, wakeE, errw);
    LAUNCHCHK();
    return FELIS_OK;
  }
  return mixed() ? launch_push<float, float>(this, l) : launch_push<double, double>(this, l);
}

template <typename R, typename T>
static int launch_drift(felis_ctx *ctx, const LatArgs &l) {
  if (!ctx->ntiles) return FELIS_OK;
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_DRIFT);
  hipLaunchKernelGGL((k_drift<R, T>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, g, l, ctx->tiles(),
                     ctx->theta, (const R *)ctx->x, (const R *)ctx->y, (const R *)ctx->px, (const R *)ctx->py,
                     (const cpx<T> *)ctx->U, (const cpx<T> *)((char *)ctx->U + ctx->pol_bytes()), ctx->wakeE,
                     ctx->d.gamR);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::drift(const felis_lattice *lat) {
  LatArgs l{};
  l.ku = lat->ku;
  return mixed() ? launch_drift<float, float>(this, l) : launch_drift<double, double>(this, l);
}

int felis_ctx::periodic() {
  felis_ctx *ctx = this;
  if (!np) return FELIS_OK;
  unsigned nb = (unsigned)((np + 255) / 256);
  KTimer kt(ctx, FELIS_K_PERIODIC);
  hipLaunchKernelGGL(k_periodic, dim3(nb), dim3(256), 0, stream, (long long)np, d.num_slice * 2 * M_PI, theta, errw);
  LAUNCHCHK();
  return FELIS_OK;
}

template <typename T>
static int launch_zero_slot(felis_ctx *ctx, int phys) {
  GridArgs g = ctx->grid();
  dim3 grid((unsigned)((ctx->N + 255) / 256), 2 * ctx->d.n_harm);
  KTimer kt(ctx, FELIS_K_SHIFT);
  hipLaunchKernelGGL((k_zero_slot<T>), grid, dim3(256), 0, ctx->stream, (cpx<T> *)ctx->U, phys, g.S,
                     (long long)ctx->N, 2 * ctx->d.n_harm);
  LAUNCHCHK();
  return FELIS_OK;
}

// solveField.cpp:19-47: slot s <- s-1 by rotating the ring; the slot that
// becomes logical slot 0 is cleared (it held the dropped ghost slot).
int felis_ctx::shift() {
  ring_base = (ring_base - 1 + S) % S;
  return mixed() ? launch_zero_slot<float>(this, ring_base) : launch_zero_slot<double>(this, ring_base);
}

template <typename R, typename T>
static int launch_deposit(felis_ctx *ctx, const DepArgs &da, double A) {
  GridArgs g = ctx->grid();
  const felis_desc &d = ctx->d;
  size_t fsz = (size_t)2 * d.n_harm * ctx->S * ctx->N * ctx->esize();
  int ez = zero_async(ctx, ctx->Sc, fsz);
  if (ez) return ez;
  if (ctx->nchunks) {
    KTimer kt(ctx, FELIS_K_DEPOSIT);
    hipLaunchKernelGGL((k_deposit_win<R, T>), dim3(ctx->nchunks), dim3(512), ctx->lds_max, ctx->stream, g, da,
                       ctx->chunks(), ctx->theta, (const R *)ctx->gamma, (const R *)ctx->x, (const R *)ctx->y,
                       (const R *)ctx->px, (const R *)ctx->py, ctx->seg_w, ctx->bessel, (cpx<T> *)ctx->Sc,
                       (cpx<T> *)((char *)ctx->Sc + ctx->pol_bytes()), ctx->lds_max, ctx->errw);
    LAUNCHCHK();
  }
  // radial space charge: only the last species survives (solveField.cpp:246-400)
  size_t nrow = (size_t)d.sc_f * d.nr + 1;
  ez = zero_async(ctx, ctx->ScR, (size_t)ctx->S * d.sc_l * nrow * sizeof(cpx<double>));
  if (ez) return ez;
  int t0 = ctx->species_tile[d.n_species - 1], t1 = ctx->species_tile[d.n_species];
  if (t1 > t0) {
    size_t lds = (size_t)d.sc_l * nrow * sizeof(cpx<double>);
    KTimer kt(ctx, FELIS_K_EZ_DEPOSIT);
    hipLaunchKernelGGL((k_ez_deposit<R>), dim3(t1 - t0), dim3(kTile), lds, ctx->stream, g, da, ctx->tiles(), t0,
                       ctx->theta, (const R *)ctx->x, (const R *)ctx->y, ctx->seg_w, (cpx<double> *)ctx->ScR);
    LAUNCHCHK();
  }
  (void)A;
  return FELIS_OK;
}

// solveField.cpp:361-397 for every slice from the radial source in ScR.
int felis_ctx::ez_solve(const felis_lattice *lat) {
  felis_ctx *ctx = this;
  GridArgs g = grid();
  double K0 = lat->K0, ue = lat->ue;
  double A = d.ks * d.ks * (1.0 + 0.5 * (1.0 + ue * ue) * K0 * K0) / (d.gamR * d.gamR) * d.dr * d.dr;
  int nsys = d.slice_n * d.sc_l * d.sc_f;
  KTimer kt(ctx, FELIS_K_EZ_SOLVE);
  // the elimination coefficients depend on the lattice only through A
  if (A != ez_A) {
    hipLaunchKernelGGL(k_ez_coef, dim3((unsigned)((d.sc_l * d.sc_f + 63) / 64)), dim3(64), 0, stream, g, A, ez_b);
    LAUNCHCHK();
    ez_A = A;
  }
  const int E = (d.nr + 63) / 64;
  const cpx<double> *src = ScR;
  cpx<double> *ez = Ez;
  switch (E) {
    case 1: hipLaunchKernelGGL((k_ez_solve<1>), dim3(nsys), dim3(64), 0, stream, g, ez_b, src, ez); break;
    case 2: hipLaunchKernelGGL((k_ez_solve<2>), dim3(nsys), dim3(64), 0, stream, g, ez_b, src, ez); break;
    case 3: hipLaunchKernelGGL((k_ez_solve<3>), dim3(nsys), dim3(64), 0, stream, g, ez_b, src, ez); break;
    case 4: hipLaunchKernelGGL((k_ez_solve<4>), dim3(nsys), dim3(64), 0, stream, g, ez_b, src, ez); break;
    default:
      err = "number of radial grid points above 256";
      return FELIS_ERR_UNSUPPORTED;
  }
  LAUNCHCHK();
  return FELIS_OK;
}

// The deposit half of solve_Sc_3D (solveField.cpp:187-359) without the solve.
int felis_ctx::deposit_only(const felis_lattice *lat) {
  DepArgs da;
  int e = dep_args(this, lat, &da);
  if (e) return e;
  return mixed() ? launch_deposit<float, float>(this, da, 0.0) : launch_deposit<double, double>(this, da, 0.0);
}

template <typename T>
static int launch_adi(felis_ctx *ctx, int cf) {
  GridArgs g = ctx->grid();
  const felis_desc &d = ctx->d;
  AdiCoef c{ctx->adix, ctx->adiy, ctx->abc};
  int nsys = 2 * d.n_harm * d.slice_n;
  int n = d.nx - 2;
  dim3 gx((unsigned)((d.ny - 2 + 3) / 4), (unsigned)nsys);
  const cpx<T> *U = (const cpx<T> *)ctx->U;
  const cpx<T> *Sc = (const cpx<T> *)ctx->Sc;
  cpx<T> *Uc = (cpx<T> *)ctx->Uc;
  {
  KTimer kt(ctx, FELIS_K_ADI_X);
  if (n <= 64) hipLaunchKernelGGL((k_adi_x<T, 1>), gx, dim3(256), 0, ctx->stream, g, c, cf, U, Sc, Uc);
  else if (n <= 128) hipLaunchKernelGGL((k_adi_x<T, 2>), gx, dim3(256), 0, ctx->stream, g, c, cf, U, Sc, Uc);
  else if (n <= 192) hipLaunchKernelGGL((k_adi_x<T, 3>), gx, dim3(256), 0, ctx->stream, g, c, cf, U, Sc, Uc);
  else hipLaunchKernelGGL((k_adi_x<T, 4>), gx, dim3(256), 0, ctx->stream, g, c, cf, U, Sc, Uc);
  LAUNCHCHK();
  }
  dim3 gy((unsigned)nsys, (unsigned)((n + 127) / 128));
  KTimer kt(ctx, FELIS_K_ADI_Y);
  hipLaunchKernelGGL((k_adi_y<T>), gy, dim3(128), 0, ctx->stream, g, c, cf, (cpx<T> *)ctx->U, Sc,
                     (const cpx<T> *)Uc);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::adi_only(int current_flag) {
  return mixed() ? launch_adi<float>(this, current_flag) : launch_adi<double>(this, current_flag);
}

int felis_ctx::solve_field(const felis_lattice *lat) {
  felis_ctx *ctx = this;
  if (d.dimension == 1) {
    // solveField.cpp:49-66: Sc cleared, the 1D source if current, U += Sc
    int e = zero_async(ctx, Sc, (size_t)2 * d.n_harm * S * esize());
    if (e) return e;
    GridArgs g = grid();
    if (lat->current_flag) {
      if (!sig_set) {
        err = "1D source needs the beam sizes (felis_set_beam_sigma)";
        return FELIS_ERR_STATE;
      }
      LatArgs l;
      if ((e = lat_args(this, lat, &l))) return e;
      KTimer kt(ctx, FELIS_K_DEPOSIT);
      hipLaunchKernelGGL(k_sc_1d, dim3(d.slice_n), dim3(256), 0, stream, g, l, d.n_species, seg_off, seg_w,
                         beam_sig, theta, (const double *)gamma, bessel, d.lambda0 * d.num_slice,
                         (cpx<double> *)Sc, (cpx<double> *)((char *)Sc + pol_bytes()), Ez, errw);
      LAUNCHCHK();
    }
    KTimer kt(ctx, FELIS_K_ADI_X);
    const int n = 2 * d.n_harm * d.slice_n;
    hipLaunchKernelGGL(k_field_1d, dim3((n + 255) / 256), dim3(256), 0, stream, g, (cpx<double> *)U,
                       (const cpx<double> *)Sc);
    LAUNCHCHK();
    return FELIS_OK;
  }
  // solveField.cpp:54-58 zeroes only the first slice_n+2 entries of each Sc;
  // the deposit below clears everything when it runs.
  if (!lat->current_flag) {
    for (int pol = 0; pol < 2; ++pol)
      for (int h = 0; h < d.n_harm; ++h) {
        size_t off = (((size_t)pol * d.n_harm + h) * S) * N * esize();
        int e = zero_async(ctx, (char *)Sc + off, std::min<size_t>(S, N * S) * esize());
        if (e) return e;
      }
  } else {
    int e = deposit_only(lat);
    if (e) return e;
    if ((e = ez_solve(lat))) return e;
  }
  return mixed() ? launch_adi<float>(ctx, lat->current_flag) : launch_adi<double>(ctx, lat->current_flag);
}

template <typename T>
static int launch_energy(felis_ctx *ctx, double *dev_out) {
  GridArgs g = ctx->grid();
  // the batched path keeps a dead y polarisation exactly zero (felis_ctx::lat_uy):
  // its slice energies are zero without a pass over it
  const int npol = ctx->mixed() ? ctx->npol() : 2, nsl = ctx->d.n_harm * ctx->d.slice_n;
  if (npol < 2) {
    int e = zero_async(ctx, ctx->slice_e + nsl, (size_t)nsl * sizeof(double));
    if (e) return e;
  }
  KTimer kt(ctx, FELIS_K_ENERGY);
  hipLaunchKernelGGL((k_energy_slices<T>), dim3(npol * nsl), dim3(256), 0, ctx->stream, g, (const cpx<T> *)ctx->U,
                     ctx->slice_e);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_sum_rows, dim3(2 * ctx->d.n_harm), dim3(kSumRowThreads), 0, ctx->stream, ctx->slice_e,
                     ctx->d.slice_n, ctx->d.slice_n, 1, dev_out);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::energy(double *dev_out) {
  return mixed() ? launch_energy<float>(this, dev_out) : launch_energy<double>(this, dev_out);
}

template <typename R>
static int launch_bfactor(felis_ctx *ctx, double *dev_out) {
  GridArgs g = ctx->grid();
  int nv = 2 * ctx->d.n_harm + 1;
  if (!ctx->ntiles) return zero_async(ctx, dev_out, nv * sizeof(double));
  KTimer kt(ctx, FELIS_K_BFACTOR);
  hipLaunchKernelGGL((k_bfactor_tiles<R>), dim3((ctx->ntiles + kTilesPerBlock - 1) / kTilesPerBlock),
                     dim3(64 * kTilesPerBlock), 0, ctx->stream, g, ctx->tiles(), ctx->ntiles, ctx->theta, ctx->seg_w,
                     ctx->tile_red);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_sum_rows, dim3(nv), dim3(kSumRowThreads), 0, ctx->stream, ctx->tile_red, ctx->ntiles, 1, nv,
                     dev_out);
  LAUNCHCHK();
  return FELIS_OK;
}

int felis_ctx::bfactor(double *dev_out) {
  return mixed() ? launch_bfactor<float>(this, dev_out) : launch_bfactor<double>(this, dev_out);
}

template <typename R>
static int launch_twiss(felis_ctx *ctx, double *dev_out) {
  GridArgs g = ctx->grid();
  const felis_desc &d = ctx->d;
  int e0 = zero_async(ctx, dev_out, 7 * d.n_species * sizeof(double));
  if (e0 || !ctx->ntiles) return e0;
  KTimer kt(ctx, FELIS_K_TWISS);
  hipLaunchKernelGGL((k_twiss_tiles<R>), dim3((ctx->ntiles + kTilesPerBlock - 1) / kTilesPerBlock),
                     dim3(64 * kTilesPerBlock), 0, ctx->stream, g, ctx->tiles(), ctx->ntiles, (const R *)ctx->x, (const R *)ctx->y, (const R *)ctx->px, (const R *)ctx->py,
                     (const R *)ctx->gamma, ctx->tile_red);
  LAUNCHCHK();
  for (int s = 0; s < d.n_species; ++s) {
    int t0 = ctx->species_tile[s], t1 = ctx->species_tile[s + 1];
    if (t1 <= t0) continue;
    hipLaunchKernelGGL(k_sum_rows, dim3(7), dim3(kSumRowThreads), 0, ctx->stream, ctx->tile_red + (size_t)t0 * 7,
                       t1 - t0, 1, 7,
                       dev_out + 7 * s);
    LAUNCHCHK();
  }
  return FELIS_OK;
}

int felis_ctx::twiss(double *dev_out) {
  return mixed() ? launch_twiss<float>(this, dev_out) : launch_twiss<double>(this, dev_out);
}

// makeDensity (wakeField.cpp:161-223) needs only counts and weights, which
// never change after the load: it is evaluated on the host, the O(n^2)
// convolution (wakeField.cpp:28-35) on the device.
int felis_ctx::update_wake(const double *wakeF_host, double *total_cnt) {
  felis_ctx *ctx = this;
  int sn = d.slice_n;
  std::vector<double> den(sn, 0.0);
  for (int s = 0; s < d.n_species; ++s)
    for (int sl = 0; sl < sn; ++sl) {
      size_t seg = (size_t)s * sn + sl;
      double cnt = h_seg_off.empty() ? 0.0 : (double)(h_seg_off[seg + 1] - h_seg_off[seg]);
      double w = h_seg_w.empty() ? 0.0 : h_seg_w[seg];
      den[sl] += w * cnt;
    }
  double tot = 0.0;
  for (double v : den) tot += v;
  double dZ = d.num_slice * d.lambda0;
  if (tot > 0.0) {
    double norm = 1.0 / (tot * dZ);
    for (double &v : den) v *= norm;
  }
  if (total_cnt) *total_cnt = tot;
  h_den = den;
  HIPCHK(hipMemcpyAsync(ctx->den, den.data(), sn * sizeof(double), hipMemcpyHostToDevice, stream));
  if (wakeF_host) HIPCHK(hipMemcpyAsync(wakeF, wakeF_host, sn * sizeof(double), hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  KTimer kt(ctx, FELIS_K_WAKE);
  hipLaunchKernelGGL(k_wake, dim3((sn + 127) / 128), dim3(128), 0, stream, sn, dZ, tot * 1.602e-19, ctx->den, wakeF,
                     wakeE);
  LAUNCHCHK();
  return FELIS_OK;
}

// ---------------------------------------------------------------- C-ABI

#define NEED_CTX() \
  if (!ctx) return FELIS_ERR_ARG
#define NEED_LAT() \
  if (!lat) return fail(ctx, FELIS_ERR_ARG, "null lattice")

extern "C" int felis_solve_field(felis_ctx *ctx, const felis_lattice *lat) {
  if (ctx && lat) ctx->lat_uy(lat);
  NEED_CTX();
  NEED_LAT();
  int e = ctx->solve_field(lat);
  return e ? e : ctx->check_err();
}

extern "C" int felis_transverse_push(felis_ctx *ctx, const felis_lattice *lat) {
  NEED_CTX();
  NEED_LAT();
  int e = ctx->transverse(lat);
  return e ? e : ctx->check_err();
}

extern "C" int felis_push_theta_gamma(felis_ctx *ctx, const felis_lattice *lat, int iteration) {
  NEED_CTX();
  NEED_LAT();
  (void)iteration;
  int e = ctx->push_theta_gamma(lat);
  return e ? e : ctx->check_err();
}

extern "C" int felis_drift_theta_gamma(felis_ctx *ctx, const felis_lattice *lat, int iteration) {
  NEED_CTX();
  NEED_LAT();
  (void)iteration;
  int e = ctx->drift(lat);
  return e ? e : ctx->check_err();
}

extern "C" int felis_periodic_particles(felis_ctx *ctx, int iteration) {
  NEED_CTX();
  (void)iteration;
  int e = ctx->periodic();
  return e ? e : ctx->check_err();
}

extern "C" int felis_shift_field(felis_ctx *ctx, int iteration) {
  NEED_CTX();
  (void)iteration;
  int e = ctx->shift();
  return e ? e : ctx->check_err();
}

extern "C" int felis_update_wake(felis_ctx *ctx, const double *wakeF, double *wakeE, double *den,
                                 double *total_cnt) {
  NEED_CTX();
  if (!wakeF) return fail(ctx, FELIS_ERR_ARG, "null wakeF");
  int e = ctx->update_wake(wakeF, total_cnt);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (wakeE) HIPCHK(hipMemcpy(wakeE, ctx->wakeE, ctx->d.slice_n * sizeof(double), hipMemcpyDeviceToHost));
  if (den) memcpy(den, ctx->h_den.data(), ctx->d.slice_n * sizeof(double));
  return FELIS_OK;
}

extern "C" int felis_update_total_energy(felis_ctx *ctx, double *ex, double *ey) {
  NEED_CTX();
  int e = ctx->energy(ctx->red_out);
  if (e) return e;
  int nh = ctx->d.n_harm;
  std::vector<double> h(2 * nh);
  HIPCHK(hipMemcpyAsync(h.data(), ctx->red_out, 2 * nh * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nh; ++k) {
    if (ex) ex[k] = h[k];
    if (ey) ey[k] = h[nh + k];
  }
  return FELIS_OK;
}

extern "C" int felis_slice_energy(felis_ctx *ctx, double *out) {
  NEED_CTX();
  if (!out) return FELIS_ERR_ARG;
  int e = ctx->energy(ctx->red_out);
  if (e) return e;
  HIPCHK(hipMemcpyAsync(out, ctx->slice_e, 2 * ctx->d.n_harm * ctx->d.slice_n * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return FELIS_OK;
}

extern "C" int felis_update_bfactor(felis_ctx *ctx, double *b, double *total_n) {
  NEED_CTX();
  int e = ctx->bfactor(ctx->red_out);
  if (e) return e;
  int nv = 2 * ctx->d.n_harm + 1;
  std::vector<double> h(nv);
  HIPCHK(hipMemcpyAsync(h.data(), ctx->red_out, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (b) memcpy(b, h.data(), 2 * ctx->d.n_harm * sizeof(double));
  if (total_n) *total_n = h[nv - 1];
  return FELIS_OK;
}

extern "C" int felis_calculate_twiss(felis_ctx *ctx, double *sums) {
  NEED_CTX();
  int e = ctx->twiss(ctx->red_out);
  if (e) return e;
  int nv = 7 * ctx->d.n_species;
  HIPCHK(hipMemcpyAsync(sums, ctx->red_out, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return FELIS_OK;
}
