/*
 * FP64 CPU oracle for the FELIS per-step slice kernel.
 *
 * TEST INFRASTRUCTURE ONLY. This restatement is the checker the HIP path is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Every function restates one reference routine (file:line given at each
 * definition in felis_oracle.c) for a single MPI rank that owns all slices:
 * slots 0 and slice_n+1 are the ghost slots of boundary.cpp:49-60.
 *
 * Layouts (identical to the reference's, mesh.h:74-75, particle.h:24-35):
 *   field   U[h][slot*N + j*nx + i]   complex double, slots 0..slice_n+1
 *   Ez      Ez[l][slot*F*nr + f*nr + j]
 *   particles: SoA, segment-major, segment = species*slice_n + (slot-1),
 *   seg_off[n_species*slice_n + 1] offsets, seg_weight per segment
 *   (the reference keeps one weight per slice list, particle.h:32).
 */
#ifndef FELIS_ORACLE_H
#define FELIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_HARM 8

enum {
  ORC_OK = 0,
  ORC_ERR_ARG = -1,
  ORC_ERR_GUARD = -3,   /* dTheta guard, particlePush.cpp:370-374 */
  ORC_ERR_RANGE = -4,   /* theta wrap range, rearrangeParticles.cpp:105-108 */
  ORC_ERR_UNDTYPE = -6, /* unknown undulator type, particlePush.cpp:64-68 */
  ORC_ERR_NAN = -5      /* NaN in the Flat-wake impedance, wakeField.cpp:97-100 */
};

typedef struct {
  int nx, ny, slice_n, n_harm;
  int harmony[ORC_MAX_HARM];
  int num_slice;     /* slices_in_bucket */
  int n_species;
  int nr, sc_l, sc_f;
  int bn, b_orders;  /* Bessel table rows / orders */
  int wake_on;
  int dimension;
  double dx, dy, dz, min_x, min_y;
  double ks, lambda0, gamR, dr, dBessel;
} orc_params;

typedef struct {
  const orc_params *p;
  const double *alpha, *alphaP, *beta, *betaP; /* complex, interleaved re/im */
  const double *bessel;                        /* [bn][b_orders] */
  const int64_t *seg_off;
  const double *seg_weight;
  double *x, *y, *px, *py, *theta, *gamma;
  double *Ux, *Uy, *ScUx, *ScUy;               /* complex interleaved */
  double *Ez;                                  /* complex interleaved */
  const double *wakeE;                         /* slice_n */
  double *sc_scratch;                          /* complex, sc_l*(sc_f*nr+1) */
  const double *beam_sig;                      /* 1D only: sigX, sigY per species */
} orc_state;

/* Lattice state written by updateK_quadG (updateK_quadG.cpp:8-94). */
typedef struct {
  double K0, ue, K0_alpha, ku, g, lambdaU;
  int und_type;      /* 1 BiPolar, 2 QuadPolar (mesh.h:23-28) */
  int current_flag;
  int drift_flag;
  int pad;
} orc_lattice;

/* One undulator unit as expanded by findUndulatorLoadParameters
 * (parameterSetting.cpp:345-429); units are listed block by block. */
typedef struct {
  double unit_start, unit_end, und_start, und_end, K0, slopeK, ue, lambdaU;
  int K0_alpha, und_type, air, pad;
} orc_und_unit;

/* One quad as expanded by findQuadLoadParameters (parameterSetting.cpp:287-338). */
typedef struct {
  double q_start, q_end, g;
  int list_id, pad;
} orc_quad;

void orc_set_threads(int n);

/* setup: boundary.cpp:73-106 and 109-123 */
void orc_abc_setup(int nx, int ny, int abcN, double abcSig, double dx, double dy,
                   double dz, double ks, double *alpha, double *alphaP,
                   double *beta, double *betaP);
void orc_bessel_table(int bn, int max_h, double *table);

/* lattice: updateK_quadG.cpp:8-94 */
void orc_update_k_quadg(orc_lattice *lat, double dz, int iteration, double half,
                        const orc_und_unit *units, int n_units,
                        const orc_quad *quads, int n_quads);

/* particles: particlePush.cpp */
int orc_transverse_push(orc_state *st, const orc_lattice *lat);
int orc_push_theta_gamma_3d(orc_state *st, const orc_lattice *lat, int iteration);
int orc_drift_theta_gamma(orc_state *st, const orc_lattice *lat);
int orc_periodic_particles(orc_state *st);

/* fields: solveField.cpp */
int orc_solve_field(orc_state *st, const orc_lattice *lat);
int orc_solve_sc_3d(orc_state *st, const orc_lattice *lat);
int orc_solve_field_u_3d(orc_state *st, double *U, const double *Sc, int current_flag);
void orc_shift_field(orc_state *st);

/* 1D path (dimension=1): particlePush.cpp:388-524, solveField.cpp:409-549 */
int orc_push_theta_gamma_1d(orc_state *st, const orc_lattice *lat, int iteration);
int orc_solve_sc_1d(orc_state *st, const orc_lattice *lat);
void orc_solve_field_u_1d(orc_state *st);

/* diagnostics */
void orc_total_energy(const orc_state *st, double *ex, double *ey);
void orc_bfactor(const orc_state *st, double *b, double *total_n);
void orc_twiss(const orc_state *st, double *sums, double *derived);

/* wake: wakeField.cpp */
int orc_wake_function(int shape, int acdc, double radius, double cond, double ctau,
                      int slice_n, double dZ, double *impZ, double *wakeF);
void orc_update_wake(const orc_state *st, const double *wakeF, double *den,
                     double *wakeE, double *total_cnt);

/* seed: loadSeed.cpp:25-80 */
void orc_load_seed_3d(const orc_params *p, double *Ux, double *Uy, int h_pick,
                      double a0, double duration, double min_z, double zR,
                      double focus, double spotSigR, double polarity,
                      double laserAlpha, double laserPsi);

/* whole step loop of main.cpp:92-156 with I/O off. lat_tab holds 2*n_steps+1
 * entries: [0] = state before the loop, [2n+1] = (n, 0), [2n+2] = (n, 0.5).
 * energy_out[n_steps][2*n_harm], bfac_out[n_steps][n_harm+1],
 * twiss_out[n_steps][n_species*7] (any may be NULL). */
int orc_run(orc_state *st, const orc_lattice *lat_tab, int it0, int n_steps,
            int time_dependent, int wake_step, const double *wakeF, double *wakeE_buf,
            double *energy_out, double *bfac_out, double *twiss_out);

#ifdef __cplusplus
}
#endif
#endif
