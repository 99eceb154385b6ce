/*
 * felis.h -- C-ABI of libfelis_hip.so, the MI355X (gfx950) implementation of
 * the FELIS per-undulator-step slice kernel.
 *
 * Each entry point replaces one free function of the reference's operator
 * surface (mesh.h:150-178), called by the reference's main loop in the same
 * order (main.cpp:92-156). The reference functions mutate a Domain; here a context
 * owns the equivalent device state, and the caller passes the lattice state
 * that the reference reads from the Domain (K0, ue, K0_alpha, ku, g,
 * undType, currentFlag, driftFlag: updateK_quadG.cpp:8-94).
 *
 * Conventions
 *   - Every function returns int: FELIS_OK (0) or a negative felis_status.
 *     No exception crosses the ABI; felis_last_error() gives the message.
 *   - Host arrays are caller-owned and copied. Complex arrays are interleaved
 *     (re, im) doubles.
 *   - Field arrays use the reference layout U[h][slot*N + j*nx + i] with
 *     slots 0..slice_n+1 (boundary.cpp:49-60, one rank owns every slice).
 *   - Particles are SoA, segment-major: segment = species*slice_n + slice,
 *     seg_off[n_species*slice_n + 1], one weight per segment (particle.h:32).
 *   - All work is ordered on the context's single HIP stream; a context is
 *     used by one host thread at a time.
 */
#ifndef FELIS_H
#define FELIS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FELIS_MAX_HARM 8

enum felis_status {
  FELIS_OK = 0,
  FELIS_ERR_ARG = -1,         /* bad argument or shape */
  FELIS_ERR_HIP = -2,         /* a HIP runtime call failed */
  FELIS_ERR_GUARD = -3,       /* |dTheta| > 2*pi*numSlice (particlePush.cpp:370-374) */
  FELIS_ERR_RANGE = -4,       /* theta out of range after the wrap (rearrangeParticles.cpp:105-108) */
  FELIS_ERR_NAN = -5,         /* NaN theta met by the wrap */
  FELIS_ERR_UNSUPPORTED = -6, /* input outside what the kernels model (see felis_last_error) */
  FELIS_ERR_STATE = -7,       /* call out of order (e.g. no particles uploaded) */
  FELIS_ERR_IO = -8           /* deck or file error */
};

enum felis_precision {
  FELIS_PLAIN_FP64 = 0, /* every quantity in FP64, as the reference */
  FELIS_MIXED_FP32 = 1  /* FP32 push/ADI/deposit, FP64 theta, slippage phase and reductions */
};

enum felis_field { FELIS_UX = 0, FELIS_UY = 1, FELIS_SCUX = 2, FELIS_SCUY = 3, FELIS_EZ = 4 };

/* Grid, beam and solver description: the Domain scalars of mesh.h:66-145
 * that the hot path reads, as derived by parameterSetting.cpp:42-285. */
typedef struct {
  int dimension;   /* 3 (the 1D path, dimension=1, is host-only) */
  int mode;        /* 1 Static, 2 Time_Dependent (mesh.h:16-21) */
  int precision;   /* felis_precision */
  int nx, ny;      /* array sizes: deck nx+1, ny+1 (parameterSetting.cpp:108-111) */
  int slice_n;
  int n_harm;
  int harmony[FELIS_MAX_HARM];
  int num_slice;   /* slices_in_bucket */
  int n_species;
  int nr, sc_l, sc_f; /* radial grids, longitudinal and azimuthal space-charge modes */
  int bessel_n;    /* Bessel table rows (num_grids, default 2001) */
  int abc_n;       /* absorbing layer width in cells (ABC_N) */
  int wake_on;
  int pad0;
  double abc_sig;  /* ABC_coef */
  double dx, dy, dz, min_x, min_y, min_z, max_z;
  double ks, lambda0, gamR, dr;
  double gamma_ref; /* FELIS_MIXED_FP32: gamma is stored as an FP32 offset from this */
} felis_desc;

/* Lattice state at one half step (updateK_quadG.cpp:8-94 outputs). */
typedef struct {
  double K0, ue, K0_alpha, ku, g, lambdaU;
  int und_type;    /* 1 BiPolar, 2 QuadPolar */
  int current_flag;
  int drift_flag;
  int pad;
} felis_lattice;

typedef struct felis_ctx felis_ctx;

/* -------- lifetime (replaces boundary.cpp:20-130 allocation and tables) -- */
int felis_create(const felis_desc *desc, felis_ctx **out);
int felis_destroy(felis_ctx *ctx);
const char *felis_last_error(const felis_ctx *ctx);
int felis_sync(felis_ctx *ctx);
/* The description the context was created with. */
int felis_get_desc(const felis_ctx *ctx, felis_desc *out);
/* ABC coefficients alpha, alphaP (nx), beta, betaP (ny) as complex and the
 * Bessel table [bessel_n][maxH/2+2] (boundary.cpp:73-123; one order more than
 * the reference for an even maxH, whose last order it reads past its rows). */
int felis_get_tables(felis_ctx *ctx, double *alpha, double *alphaP, double *beta,
                     double *betaP, double *bessel);

/* -------- state transfer (replaces loadBeam/loadSeed/restore*HDF) ------- */
int felis_upload_particles(felis_ctx *ctx, const int64_t *seg_off, const double *seg_weight,
                           const double *x, const double *y, const double *px, const double *py,
                           const double *theta, const double *gamma);
int felis_download_particles(felis_ctx *ctx, double *x, double *y, double *px, double *py,
                             double *theta, double *gamma);
/* The segment layout of the last upload: seg_off [n_species*slice_n + 1]
 * (segment = species*slice_n + slice), seg_weight [n_species*slice_n], and
 * the particle count; any pointer may be NULL. */
int felis_particle_layout(felis_ctx *ctx, int64_t *seg_off, double *seg_weight, int64_t *n_particles);
/* which = felis_field. Ux/Uy/ScUx/ScUy: [n_harm][(slice_n+2)*nx*ny] complex;
 * Ez: [sc_l][(slice_n+2)*sc_f*nr] complex. */
int felis_upload_field(felis_ctx *ctx, int which, const double *data);
int felis_download_field(felis_ctx *ctx, int which, double *data);
/* Per-slice slice power sum_{cells,h,pol} |U|^2 of slots 1..slice_n. */
int felis_slice_energy(felis_ctx *ctx, double *out /* [2][n_harm][slice_n] */);

/* -------- field diagnostics (diag/felField.cpp on the live field) --------- */
/* saveFieldHDF.cpp:66-73: powerNorm and intensityNorm of the field dump
 * (spot_sig_r = Domain::spotSigR, used when dimension = 1). */
int felis_field_norms(const felis_ctx *ctx, double spot_sig_r, double *power_norm, double *intensity_norm);
/* felField.cpp:100-160 without the dump: power[pol][h][slice] = powerNorm *
 * sum over cells of |U|^2 (W) and fluence[pol][h][j*nx + i] = intensityNorm *
 * sum over slices of |U|^2 (J/m^2), for slices 1..slice_n; either may be NULL. */
int felis_field_diag(felis_ctx *ctx, double spot_sig_r, double *power /* [2][n_harm][slice_n] */,
                     double *fluence /* [2][n_harm][nx*ny] */);

/* -------- the per-step operator surface (mesh.h:153-176) --------------- */
/* solveField.cpp:49-78: deposition (if current_flag) + radial Ez + ADI. */
int felis_solve_field(felis_ctx *ctx, const felis_lattice *lat);
/* particlePush.cpp:32-117 (one dz/2 RK4). */
int felis_transverse_push(felis_ctx *ctx, const felis_lattice *lat);
/* particlePush.cpp:198-386 (3D RK4 of theta, gamma). */
int felis_push_theta_gamma(felis_ctx *ctx, const felis_lattice *lat, int iteration);
/* particlePush.cpp:119-195. */
int felis_drift_theta_gamma(felis_ctx *ctx, const felis_lattice *lat, int iteration);
/* rearrangeParticles.cpp:77-114. */
int felis_periodic_particles(felis_ctx *ctx, int iteration);
/* solveField.cpp:19-47 + fieldShareZ.cpp:39-78: slippage by one slice. */
int felis_shift_field(felis_ctx *ctx, int iteration);
/* wakeField.cpp:9-50 + 161-223. wakeF[slice_n] in; wakeE, den out (may be NULL). */
int felis_update_wake(felis_ctx *ctx, const double *wakeF, double *wakeE, double *den,
                      double *total_cnt);
/* Sets the per-slice wake field directly (restart path). */
int felis_set_wake(felis_ctx *ctx, const double *wakeE);
/* dimension = 1 only: sig[2s], sig[2s+1] = rms beam sizes of species s, the
 * source area of solve_Sc_1D (solveField.cpp:463-468). */
int felis_set_beam_sigma(felis_ctx *ctx, const double *sig);

/* -------- per-step diagnostics (main.cpp:123-125) ----------------------- */
/* updateTotalEnergy.cpp:25-36: raw sums of |Ux|^2, |Uy|^2 per harmonic. */
int felis_update_total_energy(felis_ctx *ctx, double *ex, double *ey);
/* saveFile.cpp:124-154: b[2*h] + i b[2*h+1] = sum w exp(i H theta); total weight. */
int felis_update_bfactor(felis_ctx *ctx, double *b, double *total_n);
/* twiss.cpp:34-57: per species x^2, y^2, x'^2, y'^2, xx', yy' sums and count. */
int felis_calculate_twiss(felis_ctx *ctx, double *sums /* [n_species][7] */);

/* -------- whole loop (main.cpp:92-156 with I/O off) --------------------- */
/* lat_tab: 2*n_steps+1 entries, [0] = lattice before step it0, [2k+1] =
 * (it0+k, 0), [2k+2] = (it0+k, 0.5). Outputs per step (any may be NULL):
 * energy[k][2*n_harm] (x sums then y sums), bfac[k][2*n_harm+1] (complex b,
 * total weight), twiss[k][n_species*7]. wakeF (slice_n) may be NULL. */
int felis_run(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
              const double *wakeF, double *energy, double *bfac, double *twiss);
/* felis_run that allocates no device memory. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The same
 * loop and outputs as felis_run; its per-call device buffers (output rows,
 * tile partial sums, reduction parts, ADI scan tables) come from one 16 MiB
 * array in the library that each device holds for the life of the process.
 * A run needing more (about n_steps * (32 n_harm + 56 n_species + 8) bytes
 * plus, batched, particle tiles * (16 n_harm + 64) bytes) returns
 * FELIS_ERR_ARG. The captured-step replay is not used. Calls that overlap on
 * one device share the array and race. */
int felis_run_noalloc(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
                      const double *wakeF, double *energy, double *bfac, double *twiss);

/* -------- multi-period jumps (no counterpart in the reference) ---------- */
/* While the gate is open, one iteration advances every slice `periods`
 * undulator periods: one field solve and one particle push of step
 * periods * dz, with the slippage of the fine steps kept by window passes
 * (each field slot takes the source of the slices it slips past; each slice
 * is pushed in the field slots that slip past it, each weighted in time
 * between the field before and after the jump's solve by the period in which
 * it meets the slice), then a slippage of `periods` slots. The gate needs `periods` steps left in one
 * undulator section with no wake update inside, and the fundamental
 * bunching |b1| of the state below bunching_max; once it fails on the
 * bunching, the rest of the run takes fine steps (the plain batched step). */
typedef struct felis_mlp felis_mlp;
typedef struct felis_fno felis_fno;
typedef struct {
  int periods;          /* J, 1..min(64, slice_n) */
  int max_jumps;        /* jumps allowed in this call (< 0: no limit) */
  double bunching_max;  /* open while |b1| < bunching_max */
  double spread_max;    /* ... and the rms relative energy spread < spread_max (<= 0: not checked) */
  felis_mlp *net;       /* learned correction (NULL: the jump alone): maps the
                           felis_slice_features of each slice before the jump to
                           (log|e_h|, arg e_h) per harmonic (2*n_harm outputs), the
                           complex factor on the jump's emission E = 2 Sw / H: after the
                           push the field gains (e_h - 1) E */
  double net_bunching_max; /* the network corrects jumps while |b1| < this (<= 0: always) */
  felis_fno *fno;       /* field-to-field correction (NULL: none; in place of net when set): maps
                           each slice's emission and field on 64 x 64 cells, with its moments, to
                           factor maps o_h on the emission: the field gains o_h E_h (felis_fno_jump_maps) */
} felis_jump_opts;
/* felis_run with jumps (FELIS_MIXED_FP32, Time_Dependent). kind[n_steps]
 * (may be NULL): 0 a fine step, 1 the first period of a jump, 2 a period
 * inside a jump, whose output rows are NaN. */
int felis_run_surrogate(felis_ctx *ctx, const felis_lattice *lat_tab, int it0, int n_steps, int wake_step,
                        const double *wakeF, const felis_jump_opts *opts, double *energy, double *bfac,
                        double *twiss, int *kind);

/* Device-side snapshots of the evolving state (field with its slippage
 * position, particles, per-slice wake) in one of two slots, batched 3D path:
 * the training pipeline of the jump correction runs the fine steps and a
 * jump from one state. */
int felis_state_save(felis_ctx *ctx, int slot);
int felis_state_restore(felis_ctx *ctx, int slot);
/* Per-slice moments of the current state, out[slice_n][9 + 8*n_harm]:
 * sum w, sum w dg, sum w dg^2, sum w dg e^{i theta} (re, im), then per
 * harmonic sum w e^{iH theta}, sum w Ux e^{iH theta}, sum w Uy e^{iH theta}
 * (re, im each), sum w |Ux|^2, sum w |Uy|^2, then sum w Ux, sum w Uy of the
 * first harmonic (re, im each); dg = gamma - gamma_ref, U bilinear at the
 * particle. */
int felis_slice_moments(felis_ctx *ctx, double *out);
/* The correction network's inputs per slice, out[slice_n][5 + 5*n_harm]
 * (FP32): log10 slice current [kA]; per harmonic log10 mean |U|^2 at the
 * particles, log10 |b_h|, the coherent fraction |sum w U e^{iH theta}| /
 * sqrt(sum w sum w |U|^2) and the cos, sin of its phase; log10 rms relative
 * energy spread, the energy modulation fraction and the cos, sin of its
 * phase against b_1. */
int felis_slice_features(felis_ctx *ctx, float *out);
/* The current state (a) against snapshot `slot` (b), same slippage:
 * field_out[2*n_harm*slice_n][4] = sum conj(a) b (re, im), sum |a|^2,
 * sum |b|^2 per (pol x harmonic, slice); part_out[slice_n][12] = sums of w
 * times 1, c, s, c^2, s^2, cs, dth, dth c, dth s, dgm, dgm c, dgm s with c, s
 * of theta_a, dth = theta_b - theta_a wrapped, dgm = gamma_b - gamma_a. */
int felis_jump_residual(felis_ctx *ctx, int slot, double *field_out, double *part_out);
/* After a jump of J periods (the current state a) against snapshot `slot`
 * (b), same slippage: out[2*n_harm*slice_n][3] = sum conj(E) (b - a) (re,
 * im) and sum |E|^2 per (pol x harmonic, slice), E = 2 Sw / H the jump's
 * emission into the field now at the slice's slot (solved J slots back). */
int felis_jump_emission_residual(felis_ctx *ctx, int slot, double *out);

/* Launch shapes of the batched step, picked per workload by the project's
 * autotuner (felis_amd/tune.py) from measured step times. */
typedef struct {
  int step_graph;   /* 1: replay one captured step over uniform undulator stretches; 0: launch each kernel */
  int deposit_lds;  /* LDS bytes per deposit workgroup (0: all a workgroup may use) */
  int deposit_threads; /* threads per deposit workgroup: 256, 512 or 1024 (0: 1024) */
  int adi_threads;  /* threads per field-solve workgroup: 512 or 1024 (0: 512) */
  int deposit_run;  /* consecutive particles one deposit thread combines: 2, 4, 8 or 16 (0: 8) */
  int adi_lanes;    /* 1..15: the field solve keeps the y right-hand side of that many lanes' columns in LDS
                       and passes the rest through scratch (less LDS per system); 0: all columns that fit */
} felis_tuning;
int felis_set_tuning(felis_ctx *ctx, const felis_tuning *t);
int felis_get_tuning(felis_ctx *ctx, felis_tuning *t);

/* How the batched step's deposit (solveField.cpp:246-359, whose serial loops
 * add in particle order) accumulates: FELIS_DEPOSIT_ATOMIC adds with float
 * and double atomics, whose order varies from run to run (results agree to
 * rounding); FELIS_DEPOSIT_FIXED adds in 64-bit fixed point, which is exact
 * and order-free, so the step is bitwise reproducible from run to run and
 * across deposit launch shapes (for diagnosing against the atomic mode; it
 * reads the particles once more and keeps 16 bytes per window cell). The
 * per-phase FP64 deposit is not affected. FELIS_ERR_ARG for another mode. */
enum { FELIS_DEPOSIT_ATOMIC = 0, FELIS_DEPOSIT_FIXED = 1 };
int felis_set_deposit_mode(felis_ctx *ctx, int mode);
int felis_get_deposit_mode(felis_ctx *ctx, int *mode);

/* Polarisations the batched step solves (1: the y field and its source are
 * exactly zero so far and are skipped; 2: both). No counterpart in the
 * reference, which always solves both. */
int felis_live_polarisations(felis_ctx *ctx, int *npol);

/* -------- device ------------------------------------------------------------ */
/* info[8]: compute units, max LDS bytes per workgroup, LDS bytes per CU,
 * wavefront size, L2 bytes, clock kHz, memory clock kHz, memory bus width. */
int felis_device_info(int *info);

/* -------- per-kernel timing ------------------------------------------------ */
/* Kernels timed with HIP events on the context's stream while profiling is
 * on (the reference has wall-clock minutes only, main.cpp:95-100). */
enum felis_kernel_id {
  FELIS_K_TRANSVERSE = 0, FELIS_K_PUSH, FELIS_K_DRIFT, FELIS_K_PERIODIC, FELIS_K_DEPOSIT,
  FELIS_K_EZ_DEPOSIT, FELIS_K_EZ_SOLVE, FELIS_K_ADI_X, FELIS_K_ADI_Y, FELIS_K_SHIFT,
  FELIS_K_ENERGY, FELIS_K_BFACTOR, FELIS_K_TWISS, FELIS_K_WAKE, FELIS_K_ZERO,
  FELIS_K_PARTICLE_STEP, FELIS_K_ADI_FUSED, FELIS_K_JUMP_WINDOW, FELIS_K_COUNT
};
/* enable = 1 clears the counters and starts timing, 0 stops it. */
int felis_profile(felis_ctx *ctx, int enable);
/* Summed milliseconds and launch counts per felis_kernel_id (FELIS_K_COUNT each). */
int felis_profile_read(felis_ctx *ctx, double *ms, int64_t *launches);

/* -------- input deck and initial state (host side of the reference) ----- */
typedef struct felis_deck felis_deck;
/* findparam.cpp:30-116: value of key in the rank-th [block] of file into out.
 * Returns 1 found, 0 not found, negative on error. */
int felis_find_parameter(const char *block, int rank, const char *key, const char *file, char *out, int cap);
/* parameterSetting.cpp:42-569: parses a deck and derives the Domain scalars.
 * *out is set even on failure so that felis_deck_error can explain. */
int felis_deck_open(const char *path, felis_deck **out);
int felis_deck_close(felis_deck *deck);
const char *felis_deck_error(const felis_deck *deck);
int felis_deck_desc(const felis_deck *deck, int precision, felis_desc *out);
/* info[8]: maxStep, wake update_step, save_step, save_start, field_save,
 * particle_save, mode, sliceN */
int felis_deck_steps(const felis_deck *deck, int *info);
/* lattice table for felis_run from the deck's initial Domain state
 * (updateK_quadG.cpp:8-94 at (it,0) and (it,0.5) for it = it0..it0+n-1). */
int felis_deck_lattice(const felis_deck *deck, int it0, int n_steps, felis_lattice *tab);
/* wakeField.cpp:53-159: wake function wakeF[sliceN]; impZ[1001] may be NULL. */
int felis_deck_wake_function(const felis_deck *deck, double *wakeF, double *impZ);
/* the same table computed on the device (the flat-pipe impedance integral
 * is 5e7 complex evaluations). */
int felis_deck_wake_function_device(const felis_deck *deck, double *wakeF, double *impZ);
/* wakeField.cpp:53-159 on the device for explicit settings: shape 1 flat,
 * 2 circular pipe; acdc 1 AC, else DC conductivity; radius [m], cond [S/m],
 * ctau [m]; dZ = numSlice * lambda0. wakeF[slice_n], impZ[1001] (may be
 * NULL). FELIS_ERR_NAN where the reference prints the NaN and exits. */
int felis_wake_function(int shape, int acdc, double radius, double cond, double ctau, int slice_n, double dZ,
                        double *wakeF, double *impZ);
/* loadBeam.cpp:94-157: particle counts and weights per (species, slice). */
int felis_deck_beam_layout(const felis_deck *deck, int64_t *seg_off, double *seg_weight);
/* loadBeam.cpp:28-269 generated on the device; seed keys the shot noise. */
int felis_deck_load_beam(const felis_deck *deck, felis_ctx *ctx, uint64_t seed);
/* loadSeed.cpp:25-80 generated on the device. */
int felis_deck_load_seed(const felis_deck *deck, felis_ctx *ctx);
/* solveField.cpp:463-468: rms beam sizes sig[2s], sig[2s+1] of each species
 * from the deck's emittance and Twiss parameters. */
int felis_deck_beam_sigma(const felis_deck *deck, double *sig);
/* Savers' settings from the deck: gamma0[n_species] = loadList[s].energy /
 * mc2 + 1 (saveParticleHDF.cpp:145-148) and spot_sig_r = Domain::spotSigR
 * (parameterSetting.cpp:210) for felis_hdf.h. */
int felis_deck_save_settings(const felis_deck *deck, double *gamma0, double *spot_sig_r);

/* -------- learned surrogate: a matrix-core perceptron (no counterpart in
 * the reference; BF16 operands, FP32 accumulation and FP32 master weights).
 * dims[0..layers] are the widths; hidden layers use GELU, the output is
 * linear. Host arrays, row-major, rows <= capacity. */
int felis_mlp_create(int layers, const int *dims, int capacity, uint64_t seed, felis_mlp **out);
int felis_mlp_destroy(felis_mlp *net);
/* in/out widths and row capacity */
int felis_mlp_shape(const felis_mlp *net, int *in, int *out, int *capacity);
const char *felis_mlp_last_error(const felis_mlp *net);
int felis_mlp_get_layer(felis_mlp *net, int layer, float *W /* [in][out] */, float *b /* [out] */);
int felis_mlp_set_layer(felis_mlp *net, int layer, const float *W, const float *b);
/* the network reads (x - mu) * inv_sd per input column, in FP32 before the
 * BF16 rounding (default mu = 0, inv_sd = 1) */
int felis_mlp_set_input_affine(felis_mlp *net, const float *mu, const float *inv_sd);
int felis_mlp_get_input_affine(felis_mlp *net, float *mu, float *inv_sd);
int felis_mlp_forward(felis_mlp *net, int rows, const float *X, float *Y);
/* one Adam step on the (row-weighted, wrow may be NULL) mean squared error;
 * lr = 0 only computes the gradients (felis_mlp_get_grad) */
int felis_mlp_train_step(felis_mlp *net, int rows, const float *X, const float *T, const float *wrow, float lr,
                         double *loss);
int felis_mlp_get_grad(felis_mlp *net, int layer, float *gW, float *gb);
/* C[M][N] = A[M][K] B[K][N] with BF16-rounded operands on the matrix cores */
int felis_gemm_bf16(int M, int N, int K, const float *A, const float *B, float *C);

/* -------- learned surrogate: a Fourier neural operator on 64 x 64 maps (no
 * counterpart in the reference; the field-to-field network of the jumps) ----
 * Channel-last maps [map][y][x][channel], 64 x 64 positions. lift (cin ->
 * width), `layers` Fourier layers (a = h W + b + spectral part over `modes`
 * kept x and y modes, GELU after all but the last), projection (width ->
 * proj -> cout, GELU between). BF16 operands on the matrix cores, FP32
 * accumulation and storage, FP32 Adam. width and modes powers of two. */
int felis_fno_create(int cin, int cout, int width, int modes, int layers, int proj, int capacity, uint64_t seed,
                     felis_fno **out);
int felis_fno_destroy(felis_fno *net);
const char *felis_fno_last_error(const felis_fno *net);
/* dims[7] = cin, cout, width, modes, layers, proj, capacity */
int felis_fno_shape(const felis_fno *net, int *dims);
int64_t felis_fno_param_count(const felis_fno *net);
/* all parameters in one vector: lift W [cin][w], b [w]; per layer W [w][w],
 * b [w], mode weights [2 modes][modes][2w][2w] (ky in 0..m-1 then -m..-1, kx
 * in 0..m-1; rows (re, im) x channel in, columns out); Q1 [w][proj], c1,
 * Q2 [proj][cout], c2 */
int felis_fno_get_params(felis_fno *net, float *P);
int felis_fno_set_params(felis_fno *net, const float *P);
int felis_fno_get_grad(felis_fno *net, float *G);
/* (x - mu) * inv_sd per input channel, applied before the lift */
int felis_fno_set_input_affine(felis_fno *net, const float *mu, const float *inv_sd);
/* X [maps][4096][cin] -> Y [maps][4096][cout], host buffers */
int felis_fno_forward(felis_fno *net, int maps, const float *X, float *Y);
/* one step on `maps` maps (dev != 0: device pointers): X [maps][4096][cin];
 * E, R [maps][4096][cout] as cout/2 complex channels; wb [maps] weights.
 * Loss sum_b wb sum_{p,h} |o_h E_h - R_h|^2 / (maps 4096) of the output o
 * (cout/2 complex factors), its gradients (felis_fno_get_grad) and, for lr >
 * 0, an Adam step. */
int felis_fno_train_step(felis_fno *net, int maps, const float *X, const float *E, const float *R, const float *wb,
                         int dev, float lr, double *loss);
/* After a jump (the current state a) against snapshot `slot` (b, the fine
 * steps from the same state, same slippage): for each slice listed in
 * rows[nrows], the network's input maps X [4096][4 n_harm + 12] (feat: the
 * slices' felis_slice_features before the jump, host [slice_n][...]), the
 * jump's emission E and the residual R = b - a [4096][2 n_harm], both
 * divided by the rms of E per harmonic, where the field solved at the
 * slice's slot is now (periods slots on; zero once it has left the slices),
 * and info [3][n_harm]: sum over the map cells of |E|^2 and |U|^2 (before
 * the division) and over the field cells of the fine field's |b|^2. */
int felis_fno_jump_maps(felis_ctx *ctx, int slot, const float *feat, int nrows, const int *rows, float *X, float *E,
                        float *R, float *info);

#ifdef __cplusplus
}
#endif
#endif /* FELIS_H */
