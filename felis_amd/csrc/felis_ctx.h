// Context: the device-resident equivalent of the reference Domain
// (mesh.h:66-145) for one GPU that owns every slice.
//
// HBM layout
//   U   [pol][h][phys slot][ny][nx]  complex T; slots form a ring of
//       slice_n+2 so that slippage (solveField.cpp:19-47) is a rotation.
//       logical slot s lives at phys (s + ring_base) mod (slice_n+2).
//   Sc  [pol][h][slot][ny][nx]       complex T, not rotated.
//   Uc  [pol][h][slice][ny][nx]      complex T, ADI half-step scratch.
//   Ez  [l][slot][f][nr]             complex double.
//   particles SoA: theta double; x, y, px, py, gamma in R (gamma stored as
//       an offset from gamma_ref in the mixed path); segment-major.
#pragma once

#include <string>
#include <vector>

#include "felis_common.h"

struct felis_ctx {
  felis_desc d{};
  int S = 0;       // slots incl. the two ghosts
  size_t N = 0;    // cells per slice
  int bo = 0;      // Bessel orders
  double dB = 0;   // Bessel table step
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void *> allocs, part_allocs;

  // host copies of the setup tables
  std::vector<felis::cpx<double>> h_abc, h_adix, h_adiy;
  std::vector<double> h_bessel, h_den;
  std::vector<long long> h_seg_off;
  std::vector<double> h_seg_w;
  std::vector<int> species_tile;

  // device state
  void *U = nullptr, *Sc = nullptr, *Uc = nullptr;
  felis::cpx<double> *Ez = nullptr, *ScR = nullptr, *ez_d = nullptr;
  double *ez_b = nullptr;
  double ez_A = -1.0;  // A of the coefficients held in ez_b (negative: none)
  // per slot: interior cells [x0..x1] x [y0..y1] the last batched deposit wrote
  int4 *win = nullptr;
  double *beam_sig = nullptr;  // [n_species][2] rms beam sizes (1D source area)
  bool sig_set = false;
  felis::cpx<double> *abc = nullptr, *adix = nullptr, *adiy = nullptr;
  double *bessel = nullptr, *wakeE = nullptr, *wakeF = nullptr, *den = nullptr;
  float *bessel_f = nullptr;  // the Bessel table in FP32, for the batched push
  double *slice_e = nullptr, *red_out = nullptr;
  unsigned *errw = nullptr;
  int ring_base = 0;

  long long np = 0;
  int ntiles = 0;
  double *theta = nullptr;
  void *gamma = nullptr, *x = nullptr, *y = nullptr, *px = nullptr, *py = nullptr;
  long long *seg_off = nullptr;
  double *seg_w = nullptr;
  int *tile_seg = nullptr, *tile_cnt = nullptr;
  long long *tile_start = nullptr;
  double *tile_red = nullptr;
  // first tile of each segment (nseg + 1 entries) and, per tile, the box of
  // in-grid cells the batched push left (min i, max i, min j, max j)
  int *seg_tile = nullptr;
  int4 *tile_bb = nullptr;
  // deposit chunks: <= kChunk particles of one segment per workgroup
  int nchunks = 0;
  int *chunk_seg = nullptr, *chunk_cnt = nullptr;
  long long *chunk_start = nullptr;
  // Em carry (felis_emcarry.h): set when an in-grid particle can lie at
  // idxR >= nr; then emc holds np * sc_l complex doubles and em_last /
  // em_src one int per tile
  bool em_carry = false;
  void *emc = nullptr;
  int *em_last = nullptr, *em_src = nullptr;
  int lds_max = 65536;  // dynamic LDS bytes a workgroup may use
  int n_cu = 256;       // compute units of the device
  int adi_cols = 0;     // interior columns the batched ADI keeps in LDS (adi_scan_fits)
  felis::Tiles chunks() const { return felis::Tiles{chunk_seg, chunk_start, chunk_cnt}; }

  // per-kernel timing (felis_profile)
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_marks;
  size_t ev_used = 0;
  double prof_ms[FELIS_K_COUNT] = {};
  long long prof_n[FELIS_K_COUNT] = {};
  hipEvent_t ev_get();
  void ev_flush();

  // multi-period jumps (felis_run_surrogate): the field step is dz * dz_mul
  // while a jump is enqueued; the window passes' buffers are allocated on
  // the first jump and kept (Sw: windowed source, Ub: windowed push field,
  // both shaped like Sc and U; win2: windows of Sw; box: beam box)
  // Uy may be nonzero. Clear while the y field and its source are exactly
  // zero (zero at creation, a seed or upload with no y part, every lattice so
  // far with eta_Y = 0): the batched step then solves, deposits and reads the
  // x polarisation only. Set for good by a y upload or seed, or by a lattice
  // with eta_Y != 0 (lat_uy).
  bool uy_live = false;
  int npol() const { return uy_live ? 2 : 1; }
  // eta_Y != 0 couples the y polarisation (particlePush.cpp:228-238)
  void lat_uy(const felis_lattice *lat) {
    if (lat->K0_alpha != 1 || lat->ue != 0.0) uy_live = true;
  }
  double dz_mul = 1.0;
  void *Sw = nullptr, *Ub = nullptr;
  int4 *win2 = nullptr, *box = nullptr;
  int last_jump = 0;  // periods of the last jump (its windowed source is kept in Sw)
  // device-side copies of the evolving state (felis_state_save/restore):
  // field with its ring position, particles, per-slice wake
  struct Snap {
    void *U = nullptr, *part = nullptr;
    double *wake = nullptr;
    int ring_base = 0;
    long long np = -1;
    bool valid = false;
  } snap[2];

  felis_tuning tune{};  // launch shapes (felis_set_tuning); zero: the defaults
  int dep_mode = FELIS_DEPOSIT_ATOMIC;  // batched deposit accumulation (felis_set_deposit_mode)

  // while a step is captured as a graph: the device word holding the ring
  // position, which grid() hands to the kernels (null otherwise)
  const int *ring_dev_active = nullptr;

  // batched step (felis_run) resources
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  void *run_scratch = nullptr;

  bool mixed() const { return d.precision == FELIS_MIXED_FP32; }
  size_t esize() const { return mixed() ? 2 * sizeof(float) : 2 * sizeof(double); }
  size_t rsize() const { return mixed() ? sizeof(float) : sizeof(double); }
  size_t pol_bytes() const { return (size_t)d.n_harm * S * N * esize(); }
  felis::Tiles tiles() const { return felis::Tiles{tile_seg, tile_start, tile_cnt}; }
  felis::GridArgs grid() const;

  int check_err();
  void free_particles();
  int alloc_particles(const int64_t *seg_off, const double *seg_weight);
  void release_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  // per-phase operators (enqueue only; the C-ABI wrappers check errors)
  int solve_field(const felis_lattice *lat);
  int adi_only(int current_flag);  // the two ADI half steps on U with Sc as source
  int transverse(const felis_lattice *lat);
  int push_theta_gamma(const felis_lattice *lat);
  int drift(const felis_lattice *lat);
  int periodic();
  int shift();
  int energy(double *dev_out);
  int bfactor(double *dev_out);
  int twiss(double *dev_out);
  int update_wake(const double *wakeF_host, double *total_cnt);
  int deposit_only(const felis_lattice *lat);
  int ez_solve(const felis_lattice *lat);
};

// Records HIP events around one kernel launch while profiling is on.
struct KTimer {
  felis_ctx *c;
  int id;
  hipEvent_t a = nullptr;
  KTimer(felis_ctx *ctx, int kid) : c(ctx), id(kid) {
    if (c->profiling) {
      a = c->ev_get();
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~KTimer() {
    if (c->profiling && a) {
      hipEvent_t b = c->ev_get();
      (void)hipEventRecord(b, c->stream);
      c->ev_marks.push_back({id, {a, b}});
      if (c->ev_marks.size() >= 4096) c->ev_flush();
    }
  }
};

int lat_args(felis_ctx *ctx, const felis_lattice *lat, felis::LatArgs *out);
void abc_tables(const felis_desc &d, double dz, std::vector<felis::cpx<double>> &abc,
                std::vector<felis::cpx<double>> &adix, std::vector<felis::cpx<double>> &adiy);
int dep_args(felis_ctx *ctx, const felis_lattice *lat, felis::DepArgs *out);
// learned jump correction (felis_sur.hip), enqueued on ctx->stream:
// moments M and network inputs X of every slice; network outputs Y into the
// per-system emission factor; sum w, sum w dg, sum w dg^2 over the slices
int felis_slice_features_dev(felis_ctx *ctx, double *M, float *X);
int felis_decode_corr_dev(felis_ctx *ctx, const float *Y, void *gain);
int felis_moment_totals_dev(felis_ctx *ctx, const double *M, double *out3);
