// felis_run: the step loop of main.cpp:92-156 (I/O off) on the device.
//
// FELIS_PLAIN_FP64 replays the per-phase operators in the reference's order.
// FELIS_MIXED_FP32 runs the batched step: per undulator period one fused ADI
// launch over every (slice, harmonic, pol) system, one particle launch over
// every tile of every slice (uneven slices are split into equal tiles), one
// deposit launch per slice for the next period, and one reduction launch
// for the diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "felis_ctx.h"
#include "felis_adi.h"
#include "felis_pstep.h"

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

__global__ void k_win_empty(int4 *win, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) win[i] = make_int4(1, 0, 1, 0);
}

// One loop iteration of the plain path, enqueued on ctx->stream. lat points
// at the three lattice entries of this step: solve (previous half step),
// (it,0), (it,0.5).
int enqueue_step(felis_ctx *ctx, const felis_lattice *lat, double *e_out, double *b_out, double *t_out) {
  int e;
  const felis_desc &d = ctx->d;
  if (e_out && (e = ctx->energy(e_out))) return e;
  if (t_out && (e = ctx->twiss(t_out))) return e;
  if (b_out && (e = ctx->bfactor(b_out))) return e;
  if ((e = ctx->solve_field(&lat[0]))) return e;
  // main.cpp:133-140: no transverse motion in 1D
  if (d.dimension == 3 && (e = ctx->transverse(&lat[1]))) return e;
  if (d.dimension == 3 && (e = ctx->transverse(&lat[2]))) return e;
  if (!lat[2].drift_flag) e = ctx->push_theta_gamma(&lat[2]);
  else e = ctx->drift(&lat[2]);
  if (e) return e;
  if ((e = ctx->periodic())) return e;
  if (!lat[2].drift_flag && d.mode == 2 && (e = ctx->shift())) return e;
  return FELIS_OK;
}

// The batched ADI over every (slice, harmonic, pol) system with the scan
// tables of adi_scan_tables (device copy in tab).
int launch_adi_scan(felis_ctx *ctx, const cpx<float> *tab, int cf) {
  GridArgs g = ctx->grid();
  const felis_desc &d = ctx->d;
  const int nsys = 2 * d.n_harm * d.slice_n;
  const size_t lds = adi_scan_lds(d.nx, d.ny);
  const int E = adi_scan_E(d.nx, d.ny);
  cpx<float> *U = (cpx<float> *)ctx->U, *Sc = (cpx<float> *)ctx->Sc;
  KTimer kt(ctx, FELIS_K_ADI_FUSED);
#define ADI_CASE(e)                                                                                       \
  case e:                                                                                                 \
    hipLaunchKernelGGL((k_adi_scan<e>), dim3(nsys), dim3(512), lds, ctx->stream, g, tab, cf, U, Sc, ctx->win, \
                       ctx->slice_e);                                                                      \
    break;
  switch (E) {
    ADI_CASE(1) ADI_CASE(2) ADI_CASE(3) ADI_CASE(4) ADI_CASE(5) ADI_CASE(6) ADI_CASE(7) ADI_CASE(8) ADI_CASE(9)
    default:
      ctx->err = "grid too wide for the LDS-resident ADI";
      return FELIS_ERR_UNSUPPORTED;
  }
#undef ADI_CASE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("adi launch: ") + hipGetErrorString(e);
    return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}

// true when the system's right-hand side and tables fit the LDS of one
// workgroup; sets the kernel's dynamic LDS limit
bool adi_scan_fits(felis_ctx *ctx) {
  const felis_desc &d = ctx->d;
  const size_t lds = adi_scan_lds(d.nx, d.ny);
  const int E = adi_scan_E(d.nx, d.ny);
  if (E > 9 || lds + 256 > (size_t)ctx->lds_max + 1024) return false;
  const void *fn = nullptr;
  switch (E) {
    case 1: fn = (const void *)k_adi_scan<1>; break;
    case 2: fn = (const void *)k_adi_scan<2>; break;
    case 3: fn = (const void *)k_adi_scan<3>; break;
    case 4: fn = (const void *)k_adi_scan<4>; break;
    case 5: fn = (const void *)k_adi_scan<5>; break;
    case 6: fn = (const void *)k_adi_scan<6>; break;
    case 7: fn = (const void *)k_adi_scan<7>; break;
    case 8: fn = (const void *)k_adi_scan<8>; break;
    default: fn = (const void *)k_adi_scan<9>; break;
  }
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
}

template <int NH>
int launch_particles(felis_ctx *ctx, const PStepArgs &a, double *tile_red) {
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_PARTICLE_STEP);
  hipLaunchKernelGGL((k_pstep<NH>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, g, a, ctx->tiles(), ctx->theta,
                     (float *)ctx->gamma, (float *)ctx->x, (float *)ctx->y, (float *)ctx->px, (float *)ctx->py,
                     ctx->seg_w, (const cpx<float> *)ctx->U, (const cpx<float> *)((char *)ctx->U + ctx->pol_bytes()),
                     (const cpx<double> *)ctx->Ez, ctx->bessel, ctx->wakeE, tile_red, ctx->errw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("particle step launch: ") + hipGetErrorString(e);
    return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}

template <int NH>
int launch_deposit_slot(felis_ctx *ctx, const DepArgs &dep, int lds) {
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_DEPOSIT);
  hipLaunchKernelGGL((k_deposit_slot<NH>), dim3(ctx->d.slice_n), dim3(1024), lds, ctx->stream, g, dep,
                     ctx->d.n_species, ctx->seg_off, ctx->seg_w, ctx->theta, (const float *)ctx->gamma,
                     (const float *)ctx->x, (const float *)ctx->y, (const float *)ctx->px, (const float *)ctx->py,
                     ctx->bessel, (cpx<float> *)ctx->Sc, (cpx<float> *)((char *)ctx->Sc + ctx->pol_bytes()),
                     ctx->ScR, ctx->win, lds, ctx->errw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("deposit launch: ") + hipGetErrorString(e);
    return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}

#define NH_DISPATCH(fn, nh, ...)                 \
  switch (nh) {                                  \
    case 1: return fn<1>(__VA_ARGS__);           \
    case 2: return fn<2>(__VA_ARGS__);           \
    case 3: return fn<3>(__VA_ARGS__);           \
    case 4: return fn<4>(__VA_ARGS__);           \
    case 5: return fn<5>(__VA_ARGS__);           \
    case 6: return fn<6>(__VA_ARGS__);           \
    case 7: return fn<7>(__VA_ARGS__);           \
    case 8: return fn<8>(__VA_ARGS__);           \
    default: return FELIS_ERR_UNSUPPORTED;       \
  }

int particles(felis_ctx *ctx, const PStepArgs &a, double *tile_red) {
  NH_DISPATCH(launch_particles, ctx->d.n_harm, ctx, a, tile_red)
}
int deposit_slot(felis_ctx *ctx, const DepArgs &dep, int lds) {
  NH_DISPATCH(launch_deposit_slot, ctx->d.n_harm, ctx, dep, lds)
}

template <int NH>
hipError_t set_deposit_lds(int lds) {
  return hipFuncSetAttribute((const void *)k_deposit_slot<NH>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}
hipError_t deposit_lds_attr(int nh, int lds) {
  switch (nh) {
    case 1: return set_deposit_lds<1>(lds);
    case 2: return set_deposit_lds<2>(lds);
    case 3: return set_deposit_lds<3>(lds);
    case 4: return set_deposit_lds<4>(lds);
    case 5: return set_deposit_lds<5>(lds);
    case 6: return set_deposit_lds<6>(lds);
    case 7: return set_deposit_lds<7>(lds);
    case 8: return set_deposit_lds<8>(lds);
    default: return hipErrorInvalidValue;
  }
}

// The batched FP32/FP64 loop.
int run_mixed(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
              const double *wakeF, double *dev_out, size_t per) {
  const felis_desc &d = ctx->d;
  if (d.sc_l != 1) {
    ctx->err = "the batched path models one longitudinal space-charge mode (number_longitudinal_mode = 1)";
    return FELIS_ERR_UNSUPPORTED;
  }
  // LDS-resident scan ADI when the system fits, else the two-pass ADI with a
  // full clear of Sc after each use
  const bool fused = adi_scan_fits(ctx);
  // the batched deposit adapts to any LDS budget (harmonics in chunks, else
  // global atomics); FELIS_DEPOSIT_LDS overrides the default for tuning
  int dep_lds = ctx->lds_max;
  if (const char *v = std::getenv("FELIS_DEPOSIT_LDS")) dep_lds = std::max(4096, std::min(ctx->lds_max, std::atoi(v)));
  if (deposit_lds_attr(d.n_harm, dep_lds) != hipSuccess) {
    ctx->err = "hipFuncSetAttribute(k_deposit_slot)";
    return FELIS_ERR_HIP;
  }
  const int nh = d.n_harm, ns = d.n_species;
  const int nred = 2 * nh + 1 + 7;
  double *tile_red = nullptr, *scratch_e = nullptr;
  int *species_tile = nullptr;
  cpx<float> *adi_tab = nullptr;
  std::vector<cpx<float>> h_tab;
  adi_scan_tables(d.nx, d.ny, nh, ctx->h_adix, ctx->h_adiy, ctx->h_abc, h_tab);
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tile_red);
    (void)hipFree(species_tile);
    (void)hipFree(adi_tab);
    (void)hipFree(scratch_e);
  };
  if (hipMalloc((void **)&tile_red, (size_t)std::max(1, ctx->ntiles) * nred * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&species_tile, (ns + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&adi_tab, h_tab.size() * sizeof(cpx<float>)) != hipSuccess ||
      hipMalloc((void **)&scratch_e, 2 * nh * sizeof(double)) != hipSuccess ||
      hipMemcpy(adi_tab, h_tab.data(), h_tab.size() * sizeof(cpx<float>), hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    ctx->err = "batched step buffers";
    return FELIS_ERR_HIP;
  }
  if (hipMemcpy(species_tile, ctx->species_tile.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice) !=
      hipSuccess) {
    cleanup();
    ctx->err = "species_tile upload";
    return FELIS_ERR_HIP;
  }
  const size_t ne = 2 * (size_t)nh, nbf = 2 * (size_t)nh + 1;
  // Sc starts clear; the fused ADI clears every interior cell it reads
  const size_t fsz = (size_t)2 * nh * ctx->S * ctx->N * ctx->esize();
  hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)ctx->Sc,
                     (long long)(fsz / sizeof(unsigned)), 0u);
  hipLaunchKernelGGL(k_win_empty, dim3((unsigned)((ctx->S + 255) / 256)), dim3(256), 0, ctx->stream, ctx->win,
                     ctx->S);
  bool prev_shift = false;
  int err = FELIS_OK;
  for (int k = 0; k < n_steps && !err; ++k) {
    const int it = it0 + k;
    const felis_lattice *ls = &lat_tab[2 * k], *l0 = &lat_tab[2 * k + 1], *l1 = &lat_tab[2 * k + 2];
    if (d.mode == 2 && wakeF && wake_step > 0 && it % wake_step == 0) {
      if ((err = ctx->update_wake(wakeF, nullptr))) break;
    }
    double *row = dev_out + (size_t)k * per;
    if (k == 0) {
      // diagnostics of the state handed in and, when the first solve needs one, its deposit
      if ((err = ctx->energy(row)) || (err = ctx->bfactor(row + ne)) || (err = ctx->twiss(row + ne + nbf))) break;
      if (ls->current_flag) {
        DepArgs dep;
        if ((err = dep_args(ctx, ls, &dep)) || (err = deposit_slot(ctx, dep, dep_lds))) break;
      }
    } else {
      GridArgs g = ctx->grid();
      KTimer kt(ctx, FELIS_K_ENERGY);
      hipLaunchKernelGGL(k_reduce_diag, dim3((unsigned)(ne + nbf + 7 * ns)), dim3(256), 0, ctx->stream, g,
                         ctx->slice_e, prev_shift ? 1 : 0, tile_red, nred, species_tile, ns, row);
    }
    // solveField: the radial Ez from the last deposit, then the ADI
    if (ls->current_flag && (err = ctx->ez_solve(ls))) break;
    if (fused) {
      if ((err = launch_adi_scan(ctx, adi_tab, ls->current_flag))) break;
    } else {
      if ((err = ctx->adi_only(ls->current_flag)) || (err = ctx->energy(scratch_e))) break;
      if (ls->current_flag) {
        hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)ctx->Sc,
                           (long long)(fsz / sizeof(unsigned)), 0u);
      }
    }
    // particles: transverse x2, theta/gamma or drift, wrap, diagnostic moments
    PStepArgs a{};
    if (l0->und_type != 1 && l0->und_type != 2) {
      ctx->err = "undulator_type must be BiPolar or QuadPolar (particlePush.cpp:53-68)";
      err = FELIS_ERR_ARG;
      break;
    }
    if (!l1->drift_flag) {
      if ((err = lat_args(ctx, l1, &a.l1))) break;
    } else {
      a.l1.ku = l1->ku;
    }
    a.l0.K0 = l0->K0; a.l0.ue = l0->ue; a.l0.K0_alpha = l0->K0_alpha; a.l0.ku = l0->ku; a.l0.g = l0->g;
    a.l0.und_type = l0->und_type;
    a.l1.K0 = l1->K0; a.l1.ue = l1->ue; a.l1.K0_alpha = l1->K0_alpha; a.l1.ku = l1->ku; a.l1.g = l1->g;
    a.l1.und_type = l1->und_type;
    a.drift = l1->drift_flag;
    a.nred = nred;
    a.gamR = d.gamR;
    a.inv_dx = 1.0 / d.dx;
    a.inv_dy = 1.0 / d.dy;
    a.inv_dr = 1.0 / d.dr;
    for (int h = 0; h < nh; ++h) {
      const int H = d.harmony[h];
      a.hinv[h] = (float)(1.0 / H);
      a.hdb[h] = (float)(H / ctx->dB);
      a.hu2[h] = (float)(1.0 / (2.0 * H * H));
      a.hstep[h] = H - (h ? d.harmony[h - 1] : 0);
    }
    if ((err = particles(ctx, a, tile_red))) break;
    // the next solveField's deposit, with the lattice it will run under
    if (l1->current_flag) {
      DepArgs dep;
      if ((err = dep_args(ctx, l1, &dep))) break;
      if ((err = deposit_slot(ctx, dep, dep_lds))) break;
    }
    prev_shift = !l1->drift_flag && d.mode == 2;
    if (prev_shift && (err = ctx->shift())) break;
    if ((k & 31) == 31 || k == n_steps - 1) err = ctx->check_err();
  }
  cleanup();
  return err;
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
  if (ctx->mixed()) {
    if (!ctx->np) {
      ctx->err = "no particles uploaded";
      err = FELIS_ERR_STATE;
    } else {
      err = run_mixed(ctx, lat_tab, it0, n_steps, wake_step, wakeF, dev, per);
    }
  } else {
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
