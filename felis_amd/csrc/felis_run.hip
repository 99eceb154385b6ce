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
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "felis_ctx.h"
#include "felis_adi.h"
#include "felis_fno.h"
#include "felis_pstep.h"
#include "felis_jump.h"
#include "felis_mlp.h"
#include "felis_sur.h"

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
int launch_adi_scan(felis_ctx *ctx, const cpx<float> *tab, int cf, const cpx<float> *Sc, const int4 *win) {
  GridArgs g = ctx->grid();
  const felis_desc &d = ctx->d;
  const int nsys = ctx->npol() * d.n_harm * d.slice_n;  // the x systems first: a dead Uy is skipped
  const int cw = ctx->adi_cols;
  const size_t lds = adi_scan_lds(d.nx, d.ny, cw);
  const int E = adi_scan_E(d.nx, d.ny);
  cpx<float> *U = (cpx<float> *)ctx->U;
  cpx<float> *split = (cpx<float> *)ctx->Uc;  // columns solved second when a system is split
  KTimer kt(ctx, FELIS_K_ADI_FUSED);
  // 1024 threads need the kernel in 128 registers: whole systems of E <= 9
  const bool wide = ctx->tune.adi_threads == 1024 && E <= 9 && cw == d.nx - 2;
#define ADI_CASE(e)                                                                                         \
  case e:                                                                                                   \
    if (wide)                                                                                               \
      hipLaunchKernelGGL((k_adi_scan<e, 1024>), dim3(nsys), dim3(1024), lds, ctx->stream, g, tab, cf, U, Sc, \
                         win, ctx->slice_e, cw, split);                                               \
    else                                                                                                    \
      hipLaunchKernelGGL((k_adi_scan<e, 512>), dim3(nsys), dim3(512), lds, ctx->stream, g, tab, cf, U, Sc,   \
                         win, ctx->slice_e, cw, split);                                               \
    break;
  switch (E) {
    ADI_CASE(1) ADI_CASE(2) ADI_CASE(3) ADI_CASE(4) ADI_CASE(5) ADI_CASE(6) ADI_CASE(7) ADI_CASE(8) ADI_CASE(9)
    case 10:
      hipLaunchKernelGGL((k_adi_scan<10, 512>), dim3(nsys), dim3(512), lds, ctx->stream, g, tab, cf, U, Sc, win,
                         ctx->slice_e, cw, split);
      break;
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

// true when the system's tables and its y right-hand side (all of it, or
// the first 8E columns with the rest through scratch) fit the LDS of one
// workgroup; records the columns kept in LDS and sets the kernel's dynamic
// LDS limit
bool adi_scan_fits(felis_ctx *ctx) {
  const felis_desc &d = ctx->d;
  int cols = adi_scan_cols(d.nx, d.ny, (size_t)ctx->lds_max);
  if (cols == 0) return false;
  // tuned: fewer lanes' columns in LDS (the rest through scratch)
  const int Et = adi_scan_E(d.nx, d.ny);
  if (ctx->tune.adi_lanes > 0 && ctx->tune.adi_lanes * Et < cols) cols = ctx->tune.adi_lanes * Et;
  ctx->adi_cols = cols;
  const size_t lds = adi_scan_lds(d.nx, d.ny, cols);
  const int E = adi_scan_E(d.nx, d.ny);
  const void *fn[2] = {nullptr, nullptr};
  switch (E) {
#define ADI_FN(e)                                   \
  case e:                                           \
    fn[0] = (const void *)k_adi_scan<e, 512>;       \
    fn[1] = (const void *)k_adi_scan<e, 1024>;      \
    break;
    ADI_FN(1) ADI_FN(2) ADI_FN(3) ADI_FN(4) ADI_FN(5) ADI_FN(6) ADI_FN(7) ADI_FN(8) ADI_FN(9)
    default: fn[0] = (const void *)k_adi_scan<10, 512>; break;
#undef ADI_FN
  }
  return hipFuncSetAttribute(fn[0], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess &&
         (!fn[1] || hipFuncSetAttribute(fn[1], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess);
}

// the Em carry's three launches ahead of a push (felis_emcarry.h)
template <bool LM>
void em_carry_mixed(felis_ctx *ctx, const GridArgs &g, const PStepArgs &a) {
  const EmCarryMixed<LM> pol{g,
                             a,
                             (const float *)ctx->x,
                             (const float *)ctx->y,
                             (const float *)ctx->px,
                             (const float *)ctx->py,
                             (const float *)ctx->gamma,
                             (const cpx<double> *)ctx->Ez};
  hipLaunchKernelGGL((k_em_last<EmCarryMixed<LM>>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, pol,
                     ctx->tiles(), ctx->em_last);
  hipLaunchKernelGGL(k_em_scan, dim3(1), dim3(1024), 0, ctx->stream, ctx->ntiles, ctx->em_last, ctx->em_src);
  hipLaunchKernelGGL((k_em_fill<EmCarryMixed<LM>>), dim3(ctx->ntiles), dim3(kTile), 0, ctx->stream, pol,
                     ctx->tiles(), ctx->em_last, ctx->em_src, (cpx<float> *)ctx->emc);
}

// k_pstep over every tile; uy: the Uy-live form
template <int NH>
int launch_particles(felis_ctx *ctx, const PStepArgs &a, double *tile_red, void *Ufield, int4 *tile_bb, bool uy,
                     hipStream_t st) {
  if (ctx->ntiles <= 0) return FELIS_OK;
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_PARTICLE_STEP);
  const cpx<float> *Ux = (const cpx<float> *)Ufield, *Uy = (const cpx<float> *)((char *)Ufield + ctx->pol_bytes());
  const bool lm = ctx->d.sc_l > 1;
  const bool carry = ctx->em_carry && !a.drift;
  const cpx<float> *emc = carry ? (const cpx<float> *)ctx->emc : nullptr;
#define PSTEP_LAUNCH(UY, LM)                                                                                    \
  hipLaunchKernelGGL((k_pstep<NH, UY, LM>), dim3(ctx->ntiles), dim3(kTile), 0, st, g, a, ctx->tiles(), ctx->theta, \
                     (float *)ctx->gamma, (float *)ctx->x, (float *)ctx->y, (float *)ctx->px, (float *)ctx->py,    \
                     ctx->seg_w, Ux, Uy, (const cpx<double> *)ctx->Ez, ctx->bessel_f, ctx->wakeE, emc, tile_red,   \
                     tile_bb, ctx->errw)
  if (uy) {
    if (lm) PSTEP_LAUNCH(true, true);
    else PSTEP_LAUNCH(true, false);
  } else {
    if (lm) PSTEP_LAUNCH(false, true);
    else PSTEP_LAUNCH(false, false);
  }
#undef PSTEP_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("particle step launch: ") + hipGetErrorString(e);
    return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}

// k_deposit_slot over slices [s0, s1) on stream st; uy: both polarisations;
// tile_boxes: the window from the boxes the push left per tile
template <int NH>
int launch_deposit_slot(felis_ctx *ctx, const DepArgs &dep, int lds, bool tile_boxes, int s0, int s1, bool uy,
                        hipStream_t st) {
  if (s1 <= s0) return FELIS_OK;
  GridArgs g = ctx->grid();
  KTimer kt(ctx, FELIS_K_DEPOSIT);
  cpx<float> *Scx = (cpx<float> *)ctx->Sc, *Scy = (cpx<float> *)((char *)ctx->Sc + ctx->pol_bytes());
  const int nt = ctx->tune.deposit_threads ? ctx->tune.deposit_threads : 1024;
#define DEPOSIT_LAUNCH(NPOL, LM, DET)                                                                        \
  hipLaunchKernelGGL((k_deposit_slot<NH, NPOL, LM, DET>), dim3(s1 - s0), dim3(nt), lds, st, g, dep,              \
                     ctx->d.n_species, ctx->seg_off, ctx->seg_w, ctx->theta, (const float *)ctx->gamma,           \
                     (const float *)ctx->x, (const float *)ctx->y, (const float *)ctx->px, (const float *)ctx->py, \
                     ctx->bessel, Scx, Scy, ctx->ScR, ctx->win, lds, tile_boxes ? ctx->tile_bb : nullptr,          \
                     ctx->seg_tile, s0, ctx->errw)
#define DEPOSIT_LAUNCH_DET(NPOL, LM)                 \
  if (ctx->dep_mode == FELIS_DEPOSIT_FIXED) {        \
    DEPOSIT_LAUNCH(NPOL, LM, true);                  \
  } else {                                           \
    DEPOSIT_LAUNCH(NPOL, LM, false);                 \
  }
  const bool lm = ctx->d.sc_l > 1;
  if (uy) {
    if (lm) {
      DEPOSIT_LAUNCH_DET(2, true);
    } else {
      DEPOSIT_LAUNCH_DET(2, false);
    }
  } else {
    if (lm) {
      DEPOSIT_LAUNCH_DET(1, true);
    } else {
      DEPOSIT_LAUNCH_DET(1, false);
    }
  }
#undef DEPOSIT_LAUNCH_DET
#undef DEPOSIT_LAUNCH
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

int particles_all(felis_ctx *ctx, const PStepArgs &a, double *tile_red, void *Ufield, int4 *tile_bb, bool uy,
                  hipStream_t st) {
  NH_DISPATCH(launch_particles, ctx->d.n_harm, ctx, a, tile_red, Ufield, tile_bb, uy, st)
}
int deposit_range(felis_ctx *ctx, const DepArgs &dep, int lds, bool tile_boxes, int s0, int s1, bool uy,
                  hipStream_t st) {
  NH_DISPATCH(launch_deposit_slot, ctx->d.n_harm, ctx, dep, lds, tile_boxes, s0, s1, uy, st)
}
// the deposit of every slice on the context's stream (the window from the
// particles' positions)
int deposit_slot(felis_ctx *ctx, const DepArgs &dep, int lds) {
  return deposit_range(ctx, dep, lds, false, 0, ctx->d.slice_n, ctx->uy_live, ctx->stream);
}

// The particle step of every tile and, when dep is given, the next deposit,
// which takes its window from the boxes the push leaves per tile. The push's
// variant is chosen before lat_uy(ln) (the Uy state it runs under), the
// deposit's after it. dzm: the push's step in periods (a jump's J; the
// deposit's is one).
int particles_and_deposit(felis_ctx *ctx, const PStepArgs &a, double *tile_red, void *Ufield, double dzm,
                          const felis_lattice *ln, const DepArgs *dep, int lds) {
  const felis_desc &d = ctx->d;
  const bool uy_push = ctx->uy_live;
  if (dep) ctx->lat_uy(ln);
  const bool uy_dep = ctx->uy_live;
  ctx->dz_mul = dzm;
  if (ctx->em_carry && !a.drift) {
    GridArgs g = ctx->grid();
    if (d.sc_l > 1) em_carry_mixed<true>(ctx, g, a);
    else em_carry_mixed<false>(ctx, g, a);
  }
  int e = particles_all(ctx, a, tile_red, Ufield, ctx->tile_bb, uy_push, ctx->stream);
  ctx->dz_mul = 1.0;
  if (e) return e;
  return dep ? deposit_range(ctx, *dep, lds, true, 0, d.slice_n, uy_dep, ctx->stream) : FELIS_OK;
}

// the step's diagnostics row (energy, bFactor, twiss) from the slice
// energies and the tile moments: parts, then their fixed-order sums
void reduce_diag(felis_ctx *ctx, int skip, const double *tile_red, const int *species_tile, double *parts,
                 double *out, const int *step_dev, int per) {
  const felis_desc &d = ctx->d;
  const int nout = 4 * d.n_harm + 1 + 7 * d.n_species;
  // parts per output: one workgroup per output is latency-bound on long
  // rows (measured: 16 parts save 24 us a step at 45,600 tiles, and cost
  // 1.4 us at 4,100)
  const int np = ctx->ntiles >= 16384 ? kRedParts : 1;
  GridArgs g = ctx->grid();
  hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)nout, (unsigned)np), dim3(256), 0, ctx->stream, g, ctx->slice_e,
                     skip, tile_red, species_tile, d.n_species, parts, out, step_dev, per);
  if (np > 1)
    hipLaunchKernelGGL(k_reduce_final, dim3((unsigned)((nout + 63) / 64)), dim3(64), 0, ctx->stream, nout, np, parts,
                       out, step_dev, per);
}

template <int NH>
hipError_t set_deposit_lds(int lds) {
  const void *fn[8] = {(const void *)k_deposit_slot<NH, 1, false, false>,
                       (const void *)k_deposit_slot<NH, 2, false, false>,
                       (const void *)k_deposit_slot<NH, 1, true, false>,
                       (const void *)k_deposit_slot<NH, 2, true, false>,
                       (const void *)k_deposit_slot<NH, 1, false, true>,
                       (const void *)k_deposit_slot<NH, 2, false, true>,
                       (const void *)k_deposit_slot<NH, 1, true, true>,
                       (const void *)k_deposit_slot<NH, 2, true, true>};
  for (const void *f : fn) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
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

// Multi-period jump settings of felis_run_surrogate (null for felis_run).
struct JumpCtl {
  felis_jump_opts o;
  int *kind = nullptr;          // per step: 0 fine, 1 jump start, 2 inside a jump
  cpx<float> *tab = nullptr;    // ADI scan tables for a field step of J * dz
  int jumps = 0;
  // per-slice moments, network inputs and outputs, per-system emission factor, totals
  double *M = nullptr, *tot = nullptr;
  float *X = nullptr, *Y = nullptr;
  cpx<float> *gain = nullptr;
  bool use_net = false;  // this jump is corrected by the network
  bool use_fno = false;  // ... by the field-to-field network
  float *fno_work = nullptr;
};

// the lattice of steps k..k+J-1 is one undulator section with the field
// solve's deposit on (a jump takes one lattice for all J periods)
bool uniform_section(const felis_lattice *lat_tab, int k, int J) {
  const felis_lattice &a = lat_tab[2 * k + 2];
  if (a.drift_flag || !a.current_flag) return false;
  for (int q = 2 * k; q <= 2 * (k + J); ++q) {
    const felis_lattice &b = lat_tab[q];
    if (b.K0 != a.K0 || b.ue != a.ue || b.K0_alpha != a.K0_alpha || b.ku != a.ku || b.g != a.g ||
        b.und_type != a.und_type || b.drift_flag || !b.current_flag)
      return false;
  }
  return true;
}

// slot chunks of the jump's window passes
int jump_chunks(const felis_ctx *ctx) { return (ctx->d.slice_n + kJumpChunk - 1) / kJumpChunk; }

// buffers of the jump's window passes, allocated on first use and kept; box
// holds the beam box, the push box, then the per-chunk windows of the
// source pass and of the push-field pass
int jump_buffers(felis_ctx *ctx) {
  if (ctx->Sw) return FELIS_OK;
  const size_t fsz = (size_t)2 * ctx->d.n_harm * ctx->S * ctx->N * ctx->esize();
  void *sw = nullptr, *ub = nullptr;
  int4 *w2 = nullptr, *bx = nullptr;
  if (hipMalloc(&sw, fsz) != hipSuccess || hipMalloc(&ub, fsz) != hipSuccess ||
      hipMalloc((void **)&w2, (size_t)ctx->S * sizeof(int4)) != hipSuccess ||
      hipMalloc((void **)&bx, (2 + 2 * (size_t)jump_chunks(ctx)) * sizeof(int4)) != hipSuccess) {
    (void)hipFree(sw);
    (void)hipFree(ub);
    (void)hipFree(w2);
    (void)hipFree(bx);
    ctx->err = "jump buffers";
    return FELIS_ERR_HIP;
  }
  ctx->allocs.push_back(sw);
  ctx->allocs.push_back(ub);
  ctx->allocs.push_back(w2);
  ctx->allocs.push_back(bx);
  ctx->Sw = sw;
  ctx->Ub = ub;
  ctx->win2 = w2;
  ctx->box = bx;
  hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)sw,
                     (long long)(fsz / sizeof(unsigned)), 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)ub,
                     (long long)(fsz / sizeof(unsigned)), 0u);
  return FELIS_OK;
}

// Device words of a graph-captured step: the ring position and the step
// counter that picks the diagnostics row.
__global__ void k_set_ring_step(int *ring_dev, int *step_dev, int ring, int step) {
  if (threadIdx.x == 0) {
    *ring_dev = ring;
    *step_dev = step;
  }
}
// slippage of one slot with the ring position in device memory: clear the
// slot that becomes logical 0, then rotate (a second launch, so that no block
// of the first reads a rotated position)
__global__ void k_zero_next(cpx<float> *U, const int *ring_dev, int S, long long N, int nph) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ph = blockIdx.y;
  const int phys = (*ring_dev - 1 + S) % S;
  if (k < N && ph < nph) U[((long long)ph * S + phys) * N + k] = mk(0.f, 0.f);
}
__global__ void k_ring_rotate(int *ring_dev, int *step_dev, int S) {
  if (threadIdx.x == 0) {
    *ring_dev = (*ring_dev - 1 + S) % S;
    *step_dev += 1;
  }
}

// the particle step's arguments for the lattice (it,0) l0 and (it,0.5) l1
int pstep_args(felis_ctx *ctx, const felis_lattice *l0, const felis_lattice *l1, int nred, PStepArgs *out) {
  const felis_desc &d = ctx->d;
  PStepArgs a{};
  if (l0->und_type != 1 && l0->und_type != 2) {
    ctx->err = "undulator_type must be BiPolar or QuadPolar (particlePush.cpp:53-68)";
    return FELIS_ERR_ARG;
  }
  if (!l1->drift_flag) {
    int e = lat_args(ctx, l1, &a.l1);
    if (e) return e;
  } else {
    a.l1.ku = l1->ku;
  }
  a.l0.K0 = l0->K0; a.l0.ue = l0->ue; a.l0.K0_alpha = l0->K0_alpha; a.l0.ku = l0->ku; a.l0.g = l0->g;
  a.l0.und_type = l0->und_type;
  a.l1.K0 = l1->K0; a.l1.ue = l1->ue; a.l1.K0_alpha = l1->K0_alpha; a.l1.ku = l1->ku; a.l1.g = l1->g;
  a.l1.und_type = l1->und_type;
  a.t0 = trans_args(a.l0);
  a.t1 = trans_args(a.l1);
  a.trans_same = a.t0.aa == a.t1.aa && a.t0.bb == a.t1.bb && a.t0.cc == a.t1.cc && a.t0.ddc == a.t1.ddc &&
                 a.t0.ue == a.t1.ue && a.t0.xCoef == a.t1.xCoef && a.t0.yCoef == a.t1.yCoef && a.t0.ka == a.t1.ka;
  if (!l1->drift_flag) {
    const double K0 = a.l1.K0, ue = a.l1.ue, ku = a.l1.ku;
    a.K2a = (1.0 + ue * ue) * K0 * K0 * 0.5;
    a.K2b = ku * ku * 0.5;
    a.xic = d.ks / ku * K0 * K0 * 0.125 * (1.0 - ue * ue);
    a.halfPhi = a.l1.Phi * 0.5;
    a.xi2c = (float)(std::sqrt(2.0) * K0 / (1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue)));
    a.sq = (float)(K0 * std::sqrt(ue * ue + 1.0));
    a.kaf = (float)a.l1.K0_alpha;
    a.rotm = expi_f32(-a.halfPhi);
    a.etaX = mk((float)a.l1.etaXr, (float)a.l1.etaXi);
    a.etaY = mk((float)a.l1.etaYr, (float)a.l1.etaYi);
  }
  a.drift = l1->drift_flag;
  a.nred = nred;
  a.gamR = d.gamR;
  a.inv_dx = 1.0 / d.dx;
  a.inv_dy = 1.0 / d.dy;
  a.inv_dr = 1.0 / d.dr;
  a.inv_drf = (float)a.inv_dr;
  for (int h = 0; h < d.n_harm; ++h) {
    const int H = d.harmony[h];
    a.hinv[h] = (float)(1.0 / H);
    a.hdb[h] = (float)(H / ctx->dB);
    a.hu2[h] = (float)(1.0 / (2.0 * H * H));
    a.hstep[h] = H - (h ? d.harmony[h - 1] : 0);
  }
  *out = a;
  return FELIS_OK;
}

bool same_lattice(const felis_lattice &a, const felis_lattice &b) {
  return a.K0 == b.K0 && a.ue == b.ue && a.K0_alpha == b.K0_alpha && a.ku == b.ku && a.g == b.g &&
         a.und_type == b.und_type && a.current_flag == b.current_flag && a.drift_flag == b.drift_flag;
}

// One step of a uniform undulator section with the ring position and the
// diagnostics row in device memory, captured once and replayed: the
// diagnostics of the state (slippage of one slot before it), the field solve,
// the push, the next deposit and the slippage.
struct StepGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  felis_lattice key{};
  int npol = 0;
  bool synced = false;  // the device ring position and step counter are current
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

// Per-call buffers of the batched loop handed in by the caller instead of
// allocated (felis_run_noalloc): tile partial sums, fixed-order reduction
// parts, the first solve's energy scratch, the species' first tiles and the
// ADI scan tables (adi_tab_elems entries).
struct RunScratch {
  double *tile_red, *red_parts, *scratch_e;
  int *species_tile;
  cpx<float> *adi_tab;
  size_t adi_tab_elems;
};

// The batched FP32/FP64 loop; with jc, iterations whose gate is open advance
// J periods at once (felis_jump.h). With ext the loop allocates nothing: its
// buffers come from ext and the captured-step replay is not used.
int run_mixed(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
              const double *wakeF, double *dev_out, size_t per, JumpCtl *jc, const RunScratch *ext = nullptr) {
  const felis_desc &d = ctx->d;
  if (d.sc_l > kMaxModes) {
    ctx->err = "the batched path models up to 4 longitudinal space-charge modes (number_longitudinal_mode)";
    return FELIS_ERR_UNSUPPORTED;
  }
  // LDS-resident scan ADI when the system fits, else the two-pass ADI with a
  // full clear of Sc after each use
  const bool fused = adi_scan_fits(ctx);
  if (jc && !fused) {
    ctx->err = "jumps need the LDS-resident ADI (grid too wide)";
    return FELIS_ERR_UNSUPPORTED;
  }
  // The batched deposit takes its windows through LDS in harmonic chunks or
  // strips of node rows, and then builds the radial source there; the tuning
  // may give it less LDS than the device allows. Its budget never goes below
  // one node row of the widest window (nx cells of both polarisations) and
  // one radial row per mode (sc_l (sc_f nr + 1) complex doubles), so both always fit
  // (a fixed-point deposit keeps 16 bytes per window cell instead of 8).
  const int cell_bytes = ctx->dep_mode == FELIS_DEPOSIT_FIXED ? 16 : 8;
  const int dep_floor =
      std::max(d.nx * 2 * cell_bytes, d.sc_l * (d.sc_f * d.nr + 1) * 2 * (int)sizeof(double));
  if (dep_floor > ctx->lds_max) {
    ctx->err = "the radial source row (number_azimuthal_mode x radial_grids) exceeds the LDS of one workgroup";
    return FELIS_ERR_UNSUPPORTED;
  }
  int dep_lds = ctx->lds_max;
  if (ctx->tune.deposit_lds > 0) dep_lds = std::max(4096, std::min(ctx->lds_max, ctx->tune.deposit_lds));
  dep_lds = std::max(dep_lds, (dep_floor + 255) / 256 * 256);
  if (deposit_lds_attr(d.n_harm, dep_lds) != hipSuccess) {
    ctx->err = "hipFuncSetAttribute(k_deposit_slot)";
    return FELIS_ERR_HIP;
  }
  const int nh = d.n_harm, ns = d.n_species;
  const int nred = 2 * nh + 1 + 7;
  double *tile_red = nullptr, *scratch_e = nullptr, *red_parts = nullptr;
  int *species_tile = nullptr;
  cpx<float> *adi_tab = nullptr;
  std::vector<cpx<float>> h_tab;
  adi_scan_tables(d.nx, d.ny, nh, ctx->h_adix, ctx->h_adiy, ctx->h_abc, h_tab);
  const bool own = ext == nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(ctx->stream);
    if (own) {
      (void)hipFree(tile_red);
      (void)hipFree(red_parts);
      (void)hipFree(species_tile);
      (void)hipFree(adi_tab);
      (void)hipFree(scratch_e);
    }
    if (jc) {
      (void)hipFree(jc->tab);
      (void)hipFree(jc->M);
      (void)hipFree(jc->fno_work);
      jc->tab = nullptr;
      jc->M = nullptr;
      jc->fno_work = nullptr;
    }
    ctx->dz_mul = 1.0;
  };

  if (!own) {
    if (h_tab.size() > ext->adi_tab_elems) {
      ctx->err = "ADI tables larger than the scratch carved for them";
      return FELIS_ERR_ARG;
    }
    tile_red = ext->tile_red;
    red_parts = ext->red_parts;
    scratch_e = ext->scratch_e;
    species_tile = ext->species_tile;
    adi_tab = ext->adi_tab;
    if (hipMemcpyAsync(adi_tab, h_tab.data(), h_tab.size() * sizeof(cpx<float>), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(species_tile, ctx->species_tile.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess) {
      cleanup();
      ctx->err = "batched step tables";
      return FELIS_ERR_HIP;
    }
  } else if (hipMalloc((void **)&tile_red, (size_t)std::max(1, ctx->ntiles) * nred * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&species_tile, (ns + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&adi_tab, h_tab.size() * sizeof(cpx<float>)) != hipSuccess ||
      hipMalloc((void **)&scratch_e, 2 * nh * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&red_parts, (size_t)(4 * nh + 1 + 7 * ns) * kRedParts * sizeof(double)) != hipSuccess ||
      hipMemcpy(adi_tab, h_tab.data(), h_tab.size() * sizeof(cpx<float>), hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    ctx->err = "batched step buffers";
    return FELIS_ERR_HIP;
  }
  if (own && hipMemcpy(species_tile, ctx->species_tile.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice) !=
                 hipSuccess) {
    cleanup();
    ctx->err = "species_tile upload";
    return FELIS_ERR_HIP;
  }
  // replay one captured step over the uniform stretches of the run when the
  // tuning asks for it, unless kernels are being timed one by one
  const bool use_graph = fused && !ctx->profiling && ctx->tune.step_graph && own;
  StepGraph sg;
  int *ring_dev = nullptr;
  if (use_graph && hipMalloc((void **)&ring_dev, 2 * sizeof(int)) != hipSuccess) {
    cleanup();
    ctx->err = "graph step words";
    return FELIS_ERR_HIP;
  }
  int *step_dev = ring_dev ? ring_dev + 1 : nullptr;
  const int J = jc ? jc->o.periods : 1;
  if (jc) {
    // the field solve of a jump: ABC and Thomas tables for a step of J dz
    std::vector<cpx<double>> abcJ, adixJ, adiyJ;
    std::vector<cpx<float>> tabJ;
    abc_tables(d, d.dz * J, abcJ, adixJ, adiyJ);
    adi_scan_tables(d.nx, d.ny, nh, adixJ, adiyJ, abcJ, tabJ);
    int e = jump_buffers(ctx);
    if (!e && (hipMalloc((void **)&jc->tab, tabJ.size() * sizeof(cpx<float>)) != hipSuccess ||
               hipMemcpy(jc->tab, tabJ.data(), tabJ.size() * sizeof(cpx<float>), hipMemcpyHostToDevice) !=
                   hipSuccess)) {
      ctx->err = "jump tables";
      e = FELIS_ERR_HIP;
    }
    // one buffer: moments, totals, inputs, outputs, gain
    const size_t nm = (size_t)d.slice_n * moments_per_slice(nh), nx_ = (size_t)d.slice_n * features_per_slice(nh),
                 ny_ = (size_t)d.slice_n * outputs_per_slice(nh), ng = (size_t)2 * nh * d.slice_n;
    if (!e && hipMalloc((void **)&jc->M, (nm + 4) * sizeof(double) + ng * sizeof(cpx<float>) +
                                             (nx_ + ny_) * sizeof(float)) != hipSuccess) {
      ctx->err = "jump correction buffers";
      e = FELIS_ERR_HIP;
    }
    if (!e && jc->o.fno && hipMalloc((void **)&jc->fno_work, felis_fno_work_floats(ctx) * sizeof(float)) != hipSuccess) {
      ctx->err = "field-to-field correction buffers";
      e = FELIS_ERR_HIP;
    }
    if (!e) {
      jc->tot = jc->M + nm;
      jc->gain = (cpx<float> *)(jc->tot + 4);
      jc->X = (float *)(jc->gain + ng);
      jc->Y = jc->X + nx_;
    }
    if (e) {
      cleanup();
      (void)hipFree(ring_dev);
      return e;
    }
  }
  const size_t ne = 2 * (size_t)nh, nbf = 2 * (size_t)nh + 1;
  // the two-pass ADI reads all of Sc: it starts clear and is cleared after
  // every solve; the fused ADI reads each slot's deposit window only
  const size_t fsz = (size_t)2 * nh * ctx->S * ctx->N * ctx->esize();
  if (!fused)
    hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)ctx->Sc,
                       (long long)(fsz / sizeof(unsigned)), 0u);
  hipLaunchKernelGGL(k_win_empty, dim3((unsigned)((ctx->S + 255) / 256)), dim3(256), 0, ctx->stream, ctx->win,
                     ctx->S);
  // the fused ADI sums the slice energies of the systems it solves; those of
  // a dead y polarisation stay zero
  hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, ctx->stream, (unsigned *)ctx->slice_e,
                     (long long)(2 * nh * (size_t)d.slice_n * sizeof(double) / sizeof(unsigned)), 0u);
  int shifted = 0;  // slots the last iteration slipped the field by
  bool gate = jc != nullptr;
  int err = FELIS_OK;
  const unsigned ncb_all = (unsigned)((ctx->N + 255) / 256);
  for (int k = 0; k < n_steps && !err;) {
    const int it = it0 + k;
    const felis_lattice *ls = &lat_tab[2 * k], *l0 = &lat_tab[2 * k + 1], *l1 = &lat_tab[2 * k + 2];
    const bool wake_now = d.mode == 2 && wakeF && wake_step > 0 && it % wake_step == 0;
    // a step of a uniform undulator section after a one-slot slippage, with
    // no wake update and the jump gate shut: the captured step
    if (use_graph && k > 0 && shifted == 1 && !wake_now && !gate && d.mode == 2 &&
        same_lattice(*ls, *l0) && same_lattice(*l0, *l1) && ls->current_flag && !l1->drift_flag) {
      ctx->lat_uy(l1);
      if (!sg.exec || !same_lattice(sg.key, *l1) || sg.npol != ctx->npol()) {
        sg.release();
        PStepArgs a;
        DepArgs dep;
        if ((err = pstep_args(ctx, l0, l1, nred, &a)) || (err = dep_args(ctx, l1, &dep))) break;
        if (!sg.synced) {
          hipLaunchKernelGGL(k_set_ring_step, dim3(1), dim3(64), 0, ctx->stream, ring_dev, step_dev,
                             ctx->ring_base, k);
          sg.synced = true;
        }
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
          ctx->err = "hipStreamBeginCapture";
          err = FELIS_ERR_HIP;
          break;
        }
        ctx->ring_dev_active = ring_dev;
        reduce_diag(ctx, 1, tile_red, species_tile, red_parts, dev_out, step_dev, (int)per);
        int ce = ctx->ez_solve(ls);
        if (!ce) ce = launch_adi_scan(ctx, adi_tab, 1, (cpx<float> *)ctx->Sc, ctx->win);
        if (!ce) ce = particles_and_deposit(ctx, a, tile_red, ctx->U, 1.0, l1, &dep, dep_lds);
        if (!ce) {
          hipLaunchKernelGGL(k_zero_next, dim3(ncb_all, 2 * nh), dim3(256), 0, ctx->stream, (cpx<float> *)ctx->U,
                             (const int *)ring_dev, ctx->S, (long long)ctx->N, 2 * nh);
          hipLaunchKernelGGL(k_ring_rotate, dim3(1), dim3(64), 0, ctx->stream, ring_dev, step_dev, ctx->S);
        }
        ctx->ring_dev_active = nullptr;
        hipError_t he = hipStreamEndCapture(ctx->stream, &sg.graph);
        if (ce) {
          err = ce;
          break;
        }
        if (he != hipSuccess || hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0) != hipSuccess) {
          ctx->err = "step graph capture";
          err = FELIS_ERR_HIP;
          break;
        }
        sg.key = *l1;
        sg.npol = ctx->npol();
      }
      if (!sg.synced) {
        hipLaunchKernelGGL(k_set_ring_step, dim3(1), dim3(64), 0, ctx->stream, ring_dev, step_dev, ctx->ring_base,
                           k);
        sg.synced = true;
      }
      if (hipGraphLaunch(sg.exec, ctx->stream) != hipSuccess) {
        ctx->err = "hipGraphLaunch";
        err = FELIS_ERR_HIP;
        break;
      }
      ctx->ring_base = (ctx->ring_base - 1 + ctx->S) % ctx->S;
      if (jc && jc->kind) jc->kind[k] = 0;
      shifted = 1;
      if ((k & 31) == 31 || k + 1 >= n_steps) err = ctx->check_err();
      ++k;
      continue;
    }
    sg.synced = false;  // an eager step moves the ring position on the host
    if (wake_now) {
      if ((err = ctx->update_wake(wakeF, nullptr))) break;
    }
    double *row = dev_out + (size_t)k * per;
    if (k == 0) {
      // diagnostics of the state handed in and, when the first solve needs one, its deposit
      if ((err = ctx->energy(row)) || (err = ctx->bfactor(row + ne)) || (err = ctx->twiss(row + ne + nbf))) break;
      if (ls->current_flag) {
        ctx->lat_uy(ls);
        DepArgs dep;
        if ((err = dep_args(ctx, ls, &dep)) || (err = deposit_slot(ctx, dep, dep_lds))) break;
      }
    } else {
      KTimer kt(ctx, FELIS_K_ENERGY);
      reduce_diag(ctx, shifted, tile_red, species_tile, red_parts, row, nullptr, 0);
    }
    // jump gate: J periods left in one undulator section with no wake update
    // inside, the jump budget not spent, and the fundamental bunching of the
    // state below the threshold (read back: the host decides the launches)
    bool jump = false;
    if (gate) {
      const felis_jump_opts &o = jc->o;
      bool ok = k + J <= n_steps && (o.max_jumps < 0 || jc->jumps < o.max_jumps) &&
                uniform_section(lat_tab, k, J) && d.mode == 2;
      if (ok && wakeF && wake_step > 0) {
        for (int q = 1; q < J && ok; ++q) ok = (it + q) % wake_step != 0;
      }
      if (ok) {
        // the network's inputs (and the energy spread) of the state before the jump
        const bool moments = o.net || o.fno || o.spread_max > 0;
        if (moments && ((err = felis_slice_features_dev(ctx, jc->M, jc->X)) ||
                        (err = felis_moment_totals_dev(ctx, jc->M, jc->tot))))
          break;
        double b[6] = {0, 0, 0, 0, 0, 0};
        double *bsrc = row + ne;
        if (hipMemcpyAsync(b, bsrc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(b + 2, bsrc + 2 * nh, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            (moments &&
             hipMemcpyAsync(b + 3, jc->tot, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
          ctx->err = "jump gate read";
          err = FELIS_ERR_HIP;
          break;
        }
        const double b1 = b[2] > 0 ? std::sqrt(b[0] * b[0] + b[1] * b[1]) / b[2] : 0.0;
        if (!(b1 < o.bunching_max)) gate = false;  // the linear regime has ended: no more jumps
        jc->use_net = !o.fno && o.net && (o.net_bunching_max <= 0 || b1 < o.net_bunching_max);
        jc->use_fno = o.fno && (o.net_bunching_max <= 0 || b1 < o.net_bunching_max);
        if (o.spread_max > 0 && b[3] > 0) {
          const double mu = b[4] / b[3], var = std::max(b[5] / b[3] - mu * mu, 0.0);
          if (!(std::sqrt(var) / d.gamma_ref < o.spread_max)) gate = false;
        }
        ok = gate;
      }
      jump = ok;
    }
    if (jc && jc->kind) {
      jc->kind[k] = jump ? 1 : 0;
      for (int q = 1; jump && q < J; ++q) jc->kind[k + q] = 2;
    }
    // solveField: the radial Ez from the last deposit, then the ADI
    if (ls->current_flag && (err = ctx->ez_solve(ls))) break;
    GridArgs gw = ctx->grid();
    const unsigned ncb = (unsigned)((ctx->N + 255) / 256);
    const unsigned nchunk = (unsigned)jump_chunks(ctx);
    int4 *cwin_src = jump ? ctx->box + 2 : nullptr, *cwin_push = jump ? ctx->box + 2 + nchunk : nullptr;
    if (jump) {
      // source of the J slices each field slot passes, then the ADI over J dz
      hipLaunchKernelGGL(k_win_bbox, dim3(1), dim3(256), 0, ctx->stream, ctx->win, d.slice_n, d.nx, d.ny, 0,
                         ctx->box);
      hipLaunchKernelGGL(k_win_union, dim3((unsigned)((ctx->S + 255) / 256)), dim3(256), 0, ctx->stream, ctx->win,
                         d.slice_n, J, ctx->win2);
      hipLaunchKernelGGL(k_chunk_windows, dim3((nchunk + 63) / 64), dim3(64), 0, ctx->stream, ctx->win2, d.slice_n, 0,
                         d.nx, d.ny, cwin_src);
      hipLaunchKernelGGL(k_chunk_windows, dim3((nchunk + 63) / 64), dim3(64), 0, ctx->stream, ctx->win, d.slice_n,
                         kJumpMargin, d.nx, d.ny, cwin_push);
      {
        KTimer kt(ctx, FELIS_K_JUMP_WINDOW);
        hipLaunchKernelGGL(k_src_window, dim3(ncb, ctx->npol() * nh, nchunk), dim3(256), 0, ctx->stream, d.nx, d.ny, ctx->S,
                           d.slice_n, J, ctx->box, cwin_src, ctx->win, (const cpx<float> *)ctx->Sc,
                           (cpx<float> *)ctx->Sw);
      }
      // the learned correction: per-slice inputs -> network -> emission factor
      if (jc->use_net) {
        if ((err = felis_mlp_forward_dev(jc->o.net, d.slice_n, jc->X, jc->Y, ctx->stream)) ||
            (err = felis_decode_corr_dev(ctx, jc->Y, jc->gain)))
          break;
      }
      // the push field: the J field slots that pass each slice, each at its
      // time in the jump, from the field before the solve and after it
      hipLaunchKernelGGL(k_win_bbox, dim3(1), dim3(256), 0, ctx->stream, ctx->win, d.slice_n, d.nx, d.ny,
                         kJumpMargin, ctx->box + 1);
      {
        KTimer kt(ctx, FELIS_K_JUMP_WINDOW);
        hipLaunchKernelGGL(k_field_window<0>, dim3(ncb, ctx->npol() * nh, nchunk), dim3(256), 0, ctx->stream, d.nx,
                           d.ny, ctx->S, d.slice_n, gw.ring_base, J, ctx->box + 1, cwin_push, (const cpx<float> *)ctx->U,
                           (cpx<float> *)ctx->Ub);
      }
      // the field solve keeps the windowed source: the correction and the
      // training residual read it after the push
      if ((err = launch_adi_scan(ctx, jc->tab, 2, (cpx<float> *)ctx->Sw, ctx->win2))) break;
      {
        KTimer kt(ctx, FELIS_K_JUMP_WINDOW);
        hipLaunchKernelGGL(k_field_window<1>, dim3(ncb, ctx->npol() * nh, nchunk), dim3(256), 0, ctx->stream, d.nx,
                           d.ny, ctx->S, d.slice_n, gw.ring_base, J, ctx->box + 1, cwin_push, (const cpx<float> *)ctx->U,
                           (cpx<float> *)ctx->Ub);
      }
    } else if (fused) {
      if ((err = launch_adi_scan(ctx, adi_tab, ls->current_flag, (cpx<float> *)ctx->Sc, ctx->win))) break;
    } else {
      if ((err = ctx->adi_only(ls->current_flag)) || (err = ctx->energy(scratch_e))) break;
      if (ls->current_flag) {
        hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, ctx->stream, (unsigned *)ctx->Sc,
                           (long long)(fsz / sizeof(unsigned)), 0u);
      }
    }
    // particles: transverse x2, theta/gamma or drift, wrap, diagnostic moments
    PStepArgs a;
    if ((err = pstep_args(ctx, l0, l1, nred, &a))) break;
    const int kn = k + (jump ? J : 1);  // the next iteration
    const felis_lattice *ln = jump ? &lat_tab[2 * kn] : l1;
    // the next solveField's deposit, with the lattice it will run under,
    // after the push (pipelined with it over slice chunks when tuned so);
    // none after the call's last step: no step of this call reads it, and a
    // later call deposits the state it is handed before its first solve
    DepArgs dep{};
    const bool deposit_next = ln->current_flag && kn < n_steps;
    if (deposit_next && (err = dep_args(ctx, ln, &dep))) break;
    if ((err = particles_and_deposit(ctx, a, tile_red, jump ? ctx->Ub : ctx->U, jump ? J : 1.0, ln,
                                     deposit_next ? &dep : nullptr, dep_lds)))
      break;
    if (jump) {
      // the learned emission correction, before the field slips
      if (jc->use_fno) {
        KTimer kt(ctx, FELIS_K_JUMP_WINDOW);
        if ((err = felis_fno_correct_dev(ctx, jc->o.fno, jc->X, jc->fno_work))) break;
      }
      if (jc->use_net) {
        GridArgs gc = ctx->grid();
        KTimer kt(ctx, FELIS_K_JUMP_WINDOW);
        hipLaunchKernelGGL(k_emit_correct, dim3(d.slice_n, ctx->npol() * nh), dim3(256), 0, ctx->stream, gc, ctx->win2,
                           (const cpx<float> *)jc->gain, (const cpx<float> *)ctx->Sw, (cpx<float> *)ctx->U,
                           ctx->slice_e);
      }
      jc->jumps++;
      ctx->last_jump = J;
      ctx->ring_base = ((ctx->ring_base - J) % ctx->S + ctx->S) % ctx->S;
      KTimer kt(ctx, FELIS_K_SHIFT);
      hipLaunchKernelGGL(k_zero_slots, dim3(ncb, 2 * nh, J), dim3(256), 0, ctx->stream, (cpx<float> *)ctx->U,
                         ctx->ring_base, ctx->S, (long long)ctx->N, J);
      shifted = J;
    } else {
      shifted = (!l1->drift_flag && d.mode == 2) ? 1 : 0;
      if (shifted && (err = ctx->shift())) break;
    }
    if ((k & 31) == 31 || kn >= n_steps || jump) err = ctx->check_err();
    k = kn;
  }
  cleanup();
  sg.release();
  (void)hipFree(ring_dev);
  return err;
}

// The plain FP64 loop: the per-phase operators of every step in the
// reference's order, the step's outputs into row k of dev (per doubles).
int run_plain(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
              const double *wakeF, double *dev, size_t per, const double *energy, const double *bfac,
              const double *twiss) {
  const felis_desc &d = ctx->d;
  const size_t ne = 2 * (size_t)d.n_harm, nb = 2 * (size_t)d.n_harm + 1;
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
  return err;
}

// Shared by felis_run and felis_run_surrogate: the loop, then the per-step
// outputs copied out of one device buffer (rows of periods inside a jump
// are NaN).
int run_common(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
               const double *wakeF, double *energy, double *bfac, double *twiss, JumpCtl *jc) {
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
      err = run_mixed(ctx, lat_tab, it0, n_steps, wake_step, wakeF, dev, per, jc);
    }
  } else {
    err = run_plain(ctx, lat_tab, it0, n_steps, wake_step, wakeF, dev, per, energy, bfac, twiss);
  }
  if (!err) {
    std::vector<double> h((size_t)n_steps * per);
    hipError_t he = hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
      ctx->err = std::string("run outputs: ") + hipGetErrorString(he);
      err = FELIS_ERR_HIP;
    } else {
      const double nan = std::nan("");
      for (int k = 0; k < n_steps; ++k) {
        double *row = h.data() + (size_t)k * per;
        if (jc && jc->kind && jc->kind[k] == 2) std::fill(row, row + per, nan);
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

}  // namespace

extern "C" int felis_run(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
                         const double *wakeF, double *energy, double *bfac, double *twiss) {
  return run_common(ctx, lat_tab, it0, n_steps, wake_step, wakeF, energy, bfac, twiss, nullptr);
}

extern "C" int felis_run_surrogate(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps,
                                   int wake_step, const double *wakeF, const felis_jump_opts *opts, double *energy,
                                   double *bfac, double *twiss, int *kind) {
  if (!ctx || !opts || !lat_tab || n_steps < 0) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  if (!ctx->mixed() || d.dimension != 3) {
    ctx->err = "jumps run in the batched FP32/FP64 path (3D)";
    return FELIS_ERR_UNSUPPORTED;
  }
  if (opts->periods < 1 || opts->periods > 64 || opts->periods > d.slice_n) {
    ctx->err = "jump periods must be in 1..min(64, slice_n)";
    return FELIS_ERR_ARG;
  }
  if (opts->net) {
    int in = 0, out = 0, cap = 0;
    felis_mlp_shape(opts->net, &in, &out, &cap);
    if (in != features_per_slice(d.n_harm) || out != outputs_per_slice(d.n_harm) || cap < d.slice_n) {
      ctx->err = "correction network: inputs 5 + 5 n_harm, outputs 2 n_harm, capacity >= slice_n";
      return FELIS_ERR_ARG;
    }
  }
  if (opts->fno && !felis_fno_maps_ok(ctx)) {
    ctx->err = "field-to-field correction: a square field grid of 64 cells or more";
    return FELIS_ERR_UNSUPPORTED;
  }
  JumpCtl jc;
  jc.o = *opts;
  std::vector<int> kv(std::max(1, n_steps), 0);
  jc.kind = kv.data();
  const int err = run_common(ctx, lat_tab, it0, n_steps, wake_step, wakeF, energy, bfac, twiss, &jc);
  if (kind) std::copy(kv.begin(), kv.begin() + n_steps, kind);
  return err;
}

extern "C" int felis_live_polarisations(felis_ctx *ctx, int *npol) {
  if (!ctx || !npol) return FELIS_ERR_ARG;
  *npol = ctx->npol();
  return FELIS_OK;
}

extern "C" int felis_set_tuning(felis_ctx *ctx, const felis_tuning *t) {
  if (!ctx || !t || t->deposit_lds < 0) return FELIS_ERR_ARG;
  if (t->deposit_threads != 0 && t->deposit_threads != 256 && t->deposit_threads != 512 &&
      t->deposit_threads != 1024) {
    ctx->err = "felis_set_tuning: deposit_threads must be 256, 512 or 1024";
    return FELIS_ERR_ARG;
  }
  if (t->adi_threads != 0 && t->adi_threads != 512 && t->adi_threads != 1024) {
    ctx->err = "felis_set_tuning: adi_threads must be 512 or 1024";
    return FELIS_ERR_ARG;
  }
  if (t->deposit_run != 0 && t->deposit_run != 2 && t->deposit_run != 4 && t->deposit_run != 8 &&
      t->deposit_run != 16) {
    ctx->err = "felis_set_tuning: deposit_run must be 2, 4, 8 or 16";
    return FELIS_ERR_ARG;
  }
  if (t->adi_lanes < 0 || t->adi_lanes > 15) {
    ctx->err = "felis_set_tuning: adi_lanes must be 0 .. 15";
    return FELIS_ERR_ARG;
  }
  ctx->tune = *t;
  return FELIS_OK;
}

extern "C" int felis_get_tuning(felis_ctx *ctx, felis_tuning *t) {
  if (!ctx || !t) return FELIS_ERR_ARG;
  *t = ctx->tune;
  return FELIS_OK;
}

extern "C" int felis_set_deposit_mode(felis_ctx *ctx, int mode) {
  if (!ctx) return FELIS_ERR_ARG;
  if (mode != FELIS_DEPOSIT_ATOMIC && mode != FELIS_DEPOSIT_FIXED) {
    ctx->err = "felis_set_deposit_mode: FELIS_DEPOSIT_ATOMIC or FELIS_DEPOSIT_FIXED";
    return FELIS_ERR_ARG;
  }
  ctx->dep_mode = mode;
  return FELIS_OK;
}

extern "C" int felis_get_deposit_mode(felis_ctx *ctx, int *mode) {
  if (!ctx || !mode) return FELIS_ERR_ARG;
  *mode = ctx->dep_mode;
  return FELIS_OK;
}

// ------------------------------------------------------------------------
// felis_run_noalloc: felis_run with its per-call device buffers taken from a
// file-scope device array instead of hipMalloc. The array lives in every
// device the library runs on, for the life of the process; its address on a
// device comes from hipGetSymbolAddress, taken once per device with that
// device current, and the kernels only ever see that address. Every call on a
// device shares the array and leaves its contents behind: the loop clears
// what it sums into with fill kernels, and calls that overlap on one device
// race.
constexpr size_t kNoAllocBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices] = {};

// the device address of `scratch` on the current device
int scratch_address(felis_ctx *ctx, unsigned char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    ctx->err = "felis_run_noalloc: no current device";
    return FELIS_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ctx->err = std::string("felis_run_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return FELIS_ERR_HIP;
    }
    g_scratch_dev[dev] = (unsigned char *)p;
  }
  *out = g_scratch_dev[dev];
  return FELIS_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int felis_run_noalloc(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps,
                                 int wake_step, const double *wakeF, double *energy, double *bfac, double *twiss) {
  if (!ctx || !lat_tab || n_steps < 0) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  const int nh = d.n_harm, ns = d.n_species;
  const size_t ne = 2 * (size_t)nh, nb = 2 * (size_t)nh + 1, nt = 7 * (size_t)ns, per = ne + nb + nt;
  if (ctx->mixed() && !ctx->np) {
    ctx->err = "no particles uploaded";
    return FELIS_ERR_STATE;
  }
  // the carve: output rows, then (batched) tile sums, reduction parts,
  // energy scratch, species tiles and the ADI scan tables
  size_t need = align256(std::max<size_t>(1, (size_t)n_steps * per) * sizeof(double));
  const size_t off_tile = need;
  RunScratch rs{};
  std::vector<cpx<float>> h_tab;
  size_t off_parts = 0, off_e = 0, off_sp = 0, off_tab = 0;
  if (ctx->mixed()) {
    adi_scan_tables(d.nx, d.ny, nh, ctx->h_adix, ctx->h_adiy, ctx->h_abc, h_tab);
    need += align256((size_t)std::max(1, ctx->ntiles) * (2 * nh + 8) * sizeof(double));
    off_parts = need;
    need += align256((size_t)(4 * nh + 1 + 7 * ns) * kRedParts * sizeof(double));
    off_e = need;
    need += align256(2 * (size_t)nh * sizeof(double));
    off_sp = need;
    need += align256((size_t)(ns + 1) * sizeof(int));
    off_tab = need;
    need += align256(std::max<size_t>(1, h_tab.size()) * sizeof(cpx<float>));
  }
  if (need > kNoAllocBytes) {
    ctx->err = "felis_run_noalloc: the run needs " + std::to_string(need) + " bytes of scratch (" +
               std::to_string(n_steps) + " periods, " + std::to_string(ctx->ntiles) +
               " particle tiles); the library's array holds " + std::to_string(kNoAllocBytes) +
               ": use felis_run or fewer periods per call";
    return FELIS_ERR_ARG;
  }
  unsigned char *base = nullptr;
  int err = scratch_address(ctx, &base);
  if (err) return err;
  double *dev = (double *)base;
  if (ctx->mixed()) {
    rs.tile_red = (double *)(base + off_tile);
    rs.red_parts = (double *)(base + off_parts);
    rs.scratch_e = (double *)(base + off_e);
    rs.species_tile = (int *)(base + off_sp);
    rs.adi_tab = (cpx<float> *)(base + off_tab);
    rs.adi_tab_elems = h_tab.size();
    err = run_mixed(ctx, lat_tab, it0, n_steps, wake_step, wakeF, dev, per, nullptr, &rs);
  } else {
    err = run_plain(ctx, lat_tab, it0, n_steps, wake_step, wakeF, dev, per, energy, bfac, twiss);
  }
  // the outputs straight from the rows into the caller's arrays
  const struct {
    double *dst;
    size_t off, n;
  } outs[3] = {{energy, 0, ne}, {bfac, ne, nb}, {twiss, ne + nb, nt}};
  for (int q = 0; q < 3 && !err && n_steps > 0; ++q) {
    if (!outs[q].dst) continue;
    hipError_t he = hipMemcpy2DAsync(outs[q].dst, outs[q].n * sizeof(double), dev + outs[q].off,
                                     per * sizeof(double), outs[q].n * sizeof(double), (size_t)n_steps,
                                     hipMemcpyDeviceToHost, ctx->stream);
    if (he != hipSuccess) {
      ctx->err = std::string("run outputs: ") + hipGetErrorString(he);
      err = FELIS_ERR_HIP;
    }
  }
  const hipError_t he = hipStreamSynchronize(ctx->stream);
  if (!err && he != hipSuccess) {
    ctx->err = std::string("run outputs: ") + hipGetErrorString(he);
    err = FELIS_ERR_HIP;
  }
  return err;
}
