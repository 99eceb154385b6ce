/*
 * FP64 CPU oracle for the FELIS per-step slice kernel -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference routines, loop for loop and quirk for
 * quirk, for one MPI rank owning every slice. It is the checker for the HIP
 * path and the CPU baseline; the product never links it.
 *
 * How each function is pinned (the checks live in tests/):
 *   orc_periodic_particles, orc_total_energy, orc_twiss, orc_bfactor,
 *   orc_update_wake, orc_wake_function, orc_update_k_quadg, orc_load_seed_3d,
 *   orc_shift_field (its MPI half)   -> golden vectors written by the
 *       reference's own compiled sources (tests/golden/, oracle/_ref;
 *       test_oracle_golden.py).
 *   orc_bessel_table                 -> scipy.special.jv.
 *   orc_abc_setup                    -> the profile's closed form with and
 *       without the layer; a tilted beam is absorbed, never amplified
 *       (test_oracle_pinning.py).
 *   orc_solve_field_u_3d             -> dense complex solves of the row
 *       equations with the special last row; Gaussian-beam diffraction
 *       (width, Gouy phase) at second order in the cell size; sum |U|^2
 *       conserved without the layer.
 *   orc_transverse_push              -> the step map is the fourth-order
 *       Taylor polynomial of exp(hM) for the linear system, coupled (ue != 0)
 *       and uncoupled; symplectic to O(h^5) for K0_alpha = 1.
 *   orc_push_theta_gamma_3d          -> zero-field closed form (K2 with r and
 *       p, wake loss, grid truncation); one RK4 step against an adaptive
 *       integrator of the same ODE with O(dz^5) local error; the Madey
 *       low-gain curve of a frozen field.
 *   orc_drift_theta_gamma            -> zero- and uniform-field closed forms.
 *   orc_solve_sc_3d                  -> single-particle corner writes with
 *       scipy's Bessel functions, quiet-start cancellation; the radial Ez
 *       solve against a dense solve with the slot-0 read; the field energy it
 *       adds equals the work the push's field does on the beam.
 *   test_oracle_mutations.py plants mistakes in copies of this file and
 *       shows that the checks above catch each of them.
 *   orc_push_theta_gamma_1d, orc_solve_sc_1d, orc_solve_field_u_1d
 *                                    -> zero-field slippage, quiet-start
 *       cancellation and U += Sc closed forms; the push's field coupling by
 *       the Madey low-gain curve, the radiation source by the energy the
 *       push's field work puts into the field, the space-charge modes by the
 *       1D plasma wavenumber of a modulated cold beam (test_oracle_1d.py,
 *       with planted mistakes in test_oracle_mutations.py).
 */
#include "felis_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Physical constants exactly as constants.h:1-10 defines them. */
#define K_MU0 1.2566370614e-6
#define K_C 299792458.0
#define K_EMASS 9.10938291e-31
#define K_ECHARGE 1.60217662e-19
#define K_MC2 0.510998918
#define K_Z0 376.730313668

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

#define C(ptr) ((cplx *)(ptr))
#define CC_(ptr) ((const cplx *)(ptr))

static size_t field_stride(const orc_params *p) {
  return (size_t)(p->slice_n + 2) * (size_t)p->nx * (size_t)p->ny;
}

/* ------------------------------------------------------------------ setup */

/* boundary.cpp:73-106: absorbing-layer profiles and ADI coefficients. */
void orc_abc_setup(int nx, int ny, int abcN, double abcSig, double dx, double dy,
                   double dz, double ks, double *alpha, double *alphaP,
                   double *beta, double *betaP) {
  double *sx = calloc((size_t)nx, sizeof(double));
  double *sy = calloc((size_t)ny, sizeof(double));
  cplx *cx = calloc((size_t)nx, sizeof(cplx));
  cplx *cy = calloc((size_t)ny, sizeof(cplx));
  int L = abcN;
  for (int i = 0; i < L; ++i) sx[i] = abcSig * (L - i) * (L - i) / (1.0 * L * L);
  for (int i = nx - L; i < nx; ++i) sx[i] = abcSig * (i - nx + L) * (i - nx + L) / (1.0 * L * L);
  for (int j = 0; j < L; ++j) sy[j] = abcSig * (L - j) * (L - j) / (1.0 * L * L);
  for (int j = ny - L; j < ny; ++j) sy[j] = abcSig * (j - ny + L) * (j - ny + L) / (1.0 * L * L);
  for (int i = 0; i < nx; ++i) cx[i] = 1.0 + I * sx[i];
  for (int j = 0; j < ny; ++j) cy[j] = 1.0 + I * sy[j];
  cplx *a = C(alpha), *ap = C(alphaP), *b = C(beta), *bp = C(betaP);
  for (int i = 0; i < nx; ++i) { a[i] = 0; ap[i] = 0; }
  for (int j = 0; j < ny; ++j) { b[j] = 0; bp[j] = 0; }
  for (int i = 1; i < nx - 1; ++i) {
    a[i] = -I * dz / (2.0 * ks * dx * dx * cx[i] * (cx[i] + cx[i - 1]));
    ap[i] = -I * dz / (2.0 * ks * dx * dx * cx[i] * (cx[i] + cx[i + 1]));
  }
  for (int j = 1; j < ny - 1; ++j) {
    b[j] = -I * dz / (2.0 * ks * dy * dy * cy[j] * (cy[j] + cy[j - 1]));
    bp[j] = -I * dz / (2.0 * ks * dy * dy * cy[j] * (cy[j] + cy[j + 1]));
  }
  free(sx); free(sy); free(cx); free(cy);
}

/* boundary.cpp:109-123: J_n table on [0, 2*pi*maxH), bn rows. The reference
 * calls gsl_sf_bessel_Jn; glibc's jn() evaluates the same function. The
 * reference sizes the rows (maxH+1)/2+1, one order short of what an even
 * maxH reads (particlePush.cpp:356, undefined behaviour there); the rows here
 * hold maxH/2+2 orders, the same count for odd maxH. */
void orc_bessel_table(int bn, int max_h, double *table) {
  double bmax = max_h * 2.0 * M_PI;
  double db = bmax / (double)bn;
  int orders = max_h / 2 + 2;
  for (int i = 0; i < bn; ++i) {
    double xi = db * i;
    for (int j = 0; j < orders; ++j) table[(size_t)i * orders + j] = jn(j, xi);
  }
}

/* ---------------------------------------------------------------- lattice */

/* updateK_quadG.cpp:8-94. lat carries D->K0/ue/K0_alpha/lambdaU/und_type in
 * from the previous call, exactly as the Domain does. */
void orc_update_k_quadg(orc_lattice *lat, double dz, int iteration, double half,
                        const orc_und_unit *units, int n_units,
                        const orc_quad *quads, int n_quads) {
  int inUnd = 0, inInter = 0, air = 0;
  double z = (iteration + half) * dz;
  double K0 = lat->K0, K0_alpha = lat->K0_alpha, ue = lat->ue, lambdaU = lat->lambdaU;
  int undType = lat->und_type;
  lat->current_flag = 0;
  lat->drift_flag = 0;
  for (int n = 0; n < n_units; ++n) {
    const orc_und_unit *u = &units[n];
    if (z >= u->unit_start && z < u->unit_end) {
      if (z >= u->und_start && z < u->und_end) {
        inUnd = 1;
        K0 = u->K0 * (1 + u->slopeK * (z - u->und_start));
        K0_alpha = u->K0_alpha;
        ue = u->ue;
        undType = u->und_type;
        lambdaU = u->lambdaU;
      } else {
        inInter = 1;
        if (u->air) air = 1;
      }
    }
  }
  lat->K0 = K0;
  lat->ue = ue;
  lat->K0_alpha = K0_alpha;
  lat->lambdaU = lambdaU;
  lat->ku = 2 * M_PI / lambdaU;
  lat->und_type = undType;
  if (K0 == 0.0) air = 1;
  if (!inUnd && !inInter) air = 1;
  if (inUnd) lat->current_flag = 1;
  if (air) {
    lat->ku = 0.0;
    lat->drift_flag = 1;
    lat->current_flag = 0;
  }
  /* quads: first overlapping quad of each list wins, later lists override */
  double g = 0, x0 = z, x1 = z + dz * 0.5;
  int skip_list = -1;
  for (int n = 0; n < n_quads; ++n) {
    const orc_quad *q = &quads[n];
    if (q->list_id == skip_list) continue;
    double q0 = q->q_start, q1 = q->q_end;
    if (q1 <= x0 || q0 >= x1) continue;
    if (x0 < q0 && q1 < x1) g = (q1 - q0) / dz * 2.0 * q->g;
    else if (x0 < q0 && q0 < x1) g = (x1 - q0) / dz * 2.0 * q->g;
    else if (x0 < q1 && q1 < x1) g = (q1 - x0) / dz * 2.0 * q->g;
    else g = 1.0 * q->g;
    skip_list = q->list_id;
  }
  lat->g = g;
}

/* -------------------------------------------------------------- particles */

static int64_t n_particles(const orc_state *st) {
  const orc_params *p = st->p;
  return st->seg_off[(size_t)p->n_species * p->slice_n];
}

/* particlePush.cpp:32-117: RK4 over dz/2 of the transverse motion. */
int orc_transverse_push(orc_state *st, const orc_lattice *lat) {
  const orc_params *p = st->p;
  double dz = p->dz * 0.5, ku = lat->ku, K0 = lat->K0, ue = lat->ue;
  double aa = K0 * K0 * 0.25 * ue * ku;
  double bb = K0 * K0 * 0.25 * (1.0 + ue * ue) * ku * ku;
  double cc = K_ECHARGE * lat->g / K_EMASS / K_C;
  double xCoef = 0.0, yCoef = 0.0;
  int K0_alpha = (int)lat->K0_alpha;
  if (lat->und_type == 1) {
    if (K0_alpha == 1) { xCoef = 1; yCoef = 0; }
    else if (K0_alpha == -1) { xCoef = 0; yCoef = 1; }
  } else if (lat->und_type == 2) {
    xCoef = 1; yCoef = 1;
  } else {
    return ORC_ERR_UNDTYPE;
  }
  const double coef[5] = {0, 0, 0.5, 0.5, 1};
  int64_t np = n_particles(st);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t n = 0; n < np; ++n) {
    double kx[5] = {0}, ky[5] = {0}, kpx[5] = {0}, kpy[5] = {0};
    double gam = st->gamma[n];
    double x0 = st->x[n], y0 = st->y[n], px0 = st->px[n], py0 = st->py[n];
    double dd = ku * K0 * K0 / (4.0 * gam * gam) * ue;
    for (int m = 1; m < 5; ++m) {
      double x = x0 + dz * kx[m - 1] * coef[m];
      double y = y0 + dz * ky[m - 1] * coef[m];
      double px = px0 + dz * kpx[m - 1] * coef[m];
      double py = py0 + dz * kpy[m - 1] * coef[m];
      kx[m] = px / gam - aa * (K0_alpha * x + y) / (gam * gam);
      ky[m] = py / gam + aa * (K0_alpha * x - y) / (gam * gam);
      kpx[m] = K0_alpha * (px - py) * dd - (bb / gam * xCoef + cc) * x;
      kpy[m] = K0_alpha * (px + py) * dd - (bb / gam * yCoef - cc) * y;
    }
    st->x[n] += dz / 6.0 * (kx[1] + 2 * kx[2] + 2 * kx[3] + kx[4]);
    st->y[n] += dz / 6.0 * (ky[1] + 2 * ky[2] + 2 * ky[3] + ky[4]);
    st->px[n] += dz / 6.0 * (kpx[1] + 2 * kpx[2] + 2 * kpx[3] + kpx[4]);
    st->py[n] += dz / 6.0 * (kpy[1] + 2 * kpy[2] + 2 * kpy[3] + kpy[4]);
  }
  return ORC_OK;
}

/* Undulator coupling constants shared by push and deposit
 * (particlePush.cpp:222-238, solveField.cpp:216-231). */
static int coupling(double ue, double K0_alpha, double *Phi, cplx *etaX, cplx *etaY) {
  cplx U = (1.0 - ue * ue + I * 2.0 * ue) / (1.0 + ue * ue);
  *Phi = carg(U);
  if (K0_alpha == 1) {
    *etaX = cos(*Phi * 0.5);
    *etaY = I * sin(*Phi * 0.5);
  } else if (K0_alpha == -1) {
    *etaX = I * sin(*Phi * 0.5);
    *etaY = cos(*Phi * 0.5);
  } else {
    return ORC_ERR_ARG;
  }
  return ORC_OK;
}

static inline double bessel_at(const orc_state *st, int idx, int order, const double *w) {
  int bo = st->p->b_orders;
  return st->bessel[(size_t)idx * bo + order] * w[0] + st->bessel[(size_t)(idx + 1) * bo + order] * w[1];
}

/* particlePush.cpp:198-386: RK4 of (theta, gamma) with bilinear field gather.
 * Em is carried from the previous particle when idxR >= nr (line 248). The
 * slices are walked in order because of that carry. */
int orc_push_theta_gamma_3d(orc_state *st, const orc_lattice *lat, int iteration) {
  (void)iteration;
  const orc_params *p = st->p;
  int nx = p->nx, ny = p->ny, nr = p->nr, L = p->sc_l, F = p->sc_f;
  size_t N = (size_t)nx * ny;
  double dz = p->dz, dx = p->dx, dy = p->dy, dr = p->dr;
  double ku = lat->ku, ks = p->ks, K0 = lat->K0, dB = p->dBessel;
  double minX = p->min_x, minY = p->min_y;
  double dPhi = 2 * M_PI * p->num_slice;
  double e_mc2 = K_ECHARGE / K_EMASS / K_C / K_C;
  double ue = lat->ue, Phi;
  cplx etaX, etaY;
  double K0_alpha = lat->K0_alpha;
  if (coupling(ue, K0_alpha, &Phi, &etaX, &etaY) != ORC_OK) return ORC_ERR_ARG;
  const double coef[5] = {0, 0, 0.5, 0.5, 1};
  size_t fs = field_stride(p);
  size_t es = (size_t)(p->slice_n + 2) * F * nr;
  int err = ORC_OK;
  /* Slices are independent unless an in-grid particle has idxR >= nr, where
   * the reference carries Em over from the previous particle: then the walk
   * is serial, otherwise the slices run in parallel. */
  int nseg = p->n_species * p->slice_n;
  int carry = 0;
  if (g_threads > 1) {
    int64_t np = st->seg_off[nseg];
#pragma omp parallel for num_threads(g_threads) reduction(| : carry) schedule(static)
    for (int64_t n = 0; n < np; ++n) {
      double x0 = st->x[n], y0 = st->y[n];
      int idxI = (int)((x0 - minX) / dx), idxJ = (int)((y0 - minY) / dy);
      if (idxI >= 0 && idxI < nx - 1 && idxJ >= 0 && idxJ < ny - 1 &&
          (int)(sqrt(x0 * x0 + y0 * y0) / dr + 0.5) >= nr)
        carry = 1;
    }
  }
  int nt = (g_threads > 1 && !carry) ? g_threads : 1;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1) reduction(min : err)
  for (int seg = 0; seg < nseg; ++seg) {
    int s = seg / p->slice_n, sl = seg % p->slice_n;
    static __thread cplx Em[16];
    if (seg == 0 || nt > 1)
      for (int l = 0; l < L; ++l) Em[l] = 0;
    cplx Ux[ORC_MAX_HARM], Uy[ORC_MAX_HARM];
    double wakeE = 0.0;
    {
      int slot = sl + 1;
      if (p->wake_on) wakeE = st->wakeE[sl] / (K_MC2 * 1.0e6);
      int64_t b = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b; n < e; ++n) {
        double x0 = st->x[n], y0 = st->y[n];
        double r2 = x0 * x0 + y0 * y0;
        double px = st->px[n], py = st->py[n];
        double pr2 = px * px + py * py;
        double th0 = st->theta[n], gam0 = st->gamma[n];
        double K2 = 1.0 + pr2 + (1.0 + ue * ue) * K0 * K0 * 0.5 * (1.0 + ku * ku * 0.5 * r2);
        double xi = ks / ku * K0 * K0 / (8.0 * gam0 * gam0) * (1.0 - ue * ue);
        double xi2 = sqrt(2.0) * K0 / (1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue)) *
                     sqrt((1 + K0_alpha) * (px * px + ue * ue * py * py) +
                          (1 - K0_alpha) * (px * px * ue * ue + py * py));
        double B = atan(2.0 * K0_alpha * ue * (px - py) /
                        ((1.0 + K0_alpha) * (px + ue * ue * py) + (1 - K0_alpha) * (px * ue * ue + py)));
        cplx expP = cexp(I * (B - Phi * 0.5));
        cplx expM = conj(expP);
        int idxI = (int)((x0 - minX) / dx);
        int idxJ = (int)((y0 - minY) / dy);
        double wx[2], wy[2];
        wx[1] = (x0 - minX) / dx - idxI; wx[0] = 1.0 - wx[1];
        wy[1] = (y0 - minY) / dy - idxJ; wy[0] = 1.0 - wy[1];
        if (!(idxI >= 0 && idxI < nx - 1 && idxJ >= 0 && idxJ < ny - 1)) continue;
        double r = sqrt(r2);
        int idxR = (int)(r / dr + 0.5);
        if (idxR > 0 && idxR < nr) {
          double wr[2];
          wr[1] = (r * r / (dr * dr) - idxR * idxR) / (2.0 * idxR); wr[0] = 1.0 - wr[1];
          for (int l = 0; l < L; ++l) {
            Em[l] = 0;
            for (int f = 0; f < F; ++f)
              for (int jj = 0; jj < 2; ++jj)
                Em[l] += CC_(st->Ez)[l * es + (size_t)slot * nr * F + nr * f + idxR + jj] * wr[jj];
          }
        } else if (idxR == 0) {
          for (int l = 0; l < L; ++l) {
            Em[l] = 0;
            for (int f = 0; f < F; ++f)
              for (int jj = 0; jj < 2; ++jj)
                Em[l] += CC_(st->Ez)[l * es + (size_t)slot * nr * F + nr * f + idxR + jj];
          }
        }
        for (int h = 0; h < p->n_harm; ++h) {
          Ux[h] = 0; Uy[h] = 0;
          for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj) {
              size_t c = h * fs + slot * N + (size_t)(idxJ + jj) * nx + (idxI + ii);
              Ux[h] += CC_(st->Ux)[c] * wx[ii] * wy[jj];
              Uy[h] += CC_(st->Uy)[c] * wx[ii] * wy[jj];
            }
        }
        double sumU2 = 0.0;
        for (int h = 0; h < p->n_harm; ++h) {
          int H = p->harmony[h];
          sumU2 += (creal(Ux[h]) * creal(Ux[h]) + cimag(Ux[h]) * cimag(Ux[h]) +
                    creal(Uy[h]) * creal(Uy[h]) + cimag(Uy[h]) * cimag(Uy[h])) / (2.0 * H * H);
        }
        double kth[5] = {0}, kg[5] = {0};
        for (int m = 1; m < 5; ++m) {
          double th = th0 + dz * kth[m - 1] * coef[m];
          double gam = gam0 + dz * kg[m - 1] * coef[m];
          double sumTh = 0.0, sumG = 0.0, sumEz = 0.0;
          for (int l = 0; l < L; ++l) sumEz += 2.0 * creal(Em[l] * cexp(I * (l + 1.0) * th));
          for (int h = 0; h < p->n_harm; ++h) {
            int H = p->harmony[h];
            double w[2];
            int idx = (int)((1.0 * H * xi) / dB);
            w[1] = (H * xi / dB) - idx; w[0] = 1.0 - w[1];
            if (idx < 0 || idx + 1 >= p->bn) { err = ORC_ERR_ARG; goto seg_done; }
            cplx fx, fy;
            if (H % 2 == 1) {
              double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
              double J1 = bessel_at(st, idx, (H - 1) / 2, w);
              double J2 = bessel_at(st, idx, (H + 1) / 2, w);
              fx = sign * (K0_alpha * J1 - J2);
              fy = sign * (K0_alpha * J1 + J2);
            } else {
              double sign = (H / 2 % 2 == 0) ? 1.0 : -1.0;
              double J1 = bessel_at(st, idx, (H - 2) / 2, w);
              double J2 = bessel_at(st, idx, H / 2, w);
              double J3 = bessel_at(st, idx, (H + 2) / 2, w);
              fx = sign * xi2 * H * 0.5 * (expP * (K0_alpha * J1 - J2) + expM * (K0_alpha * J2 - J3));
              fy = sign * xi2 * H * 0.5 * (expP * (K0_alpha * J1 + J2) + expM * (K0_alpha * J2 + J3));
            }
            cplx t = cexp(I * (1.0 * H) * (th + Phi * 0.5)) * (etaX * fx * Ux[h] - etaY * fy * Uy[h]);
            sumTh += creal(I * t) / (1.0 * H);
            sumG += creal(t) / (2.0 * gam);
          }
          kth[m] = ku - ks / (2.0 * gam * gam) * (K2 + sumU2 + K0 * sqrt(ue * ue + 1.0) * sumTh);
          kg[m] = ks * K0 * sqrt(ue * ue + 1.0) * sumG + e_mc2 * sumEz;
        }
        double dth = dz / 6.0 * (kth[1] + 2.0 * kth[2] + 2.0 * kth[3] + kth[4]);
        if (dth > dPhi || dth < -dPhi) { err = ORC_ERR_GUARD; goto seg_done; }
        st->theta[n] += dth;
        double dgam = dz / 6.0 * (kg[1] + 2.0 * kg[2] + 2.0 * kg[3] + kg[4]) - dz * wakeE;
        st->gamma[n] += dgam;
      }
    }
  seg_done:;
  }
  return err;
}

/* particlePush.cpp:119-195: phase drift in gaps; wake loss hits theta. */
int orc_drift_theta_gamma(orc_state *st, const orc_lattice *lat) {
  const orc_params *p = st->p;
  int nx = p->nx, ny = p->ny;
  size_t N = (size_t)nx * ny, fs = field_stride(p);
  double dz = p->dz, ks = p->ks, ku = lat->ku, gamR = p->gamR;
  double dx = p->dx, dy = p->dy, minX = p->min_x, minY = p->min_y;
  for (int s = 0; s < p->n_species; ++s) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      double wakeE = p->wake_on ? st->wakeE[sl] / (K_MC2 * 1.0e6) : 0.0;
      int64_t b = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b; n < e; ++n) {
        double x0 = st->x[n], y0 = st->y[n];
        double px = st->px[n], py = st->py[n];
        double K2 = 1.0 + px * px + py * py;
        int idxI = (int)((x0 - minX) / dx);
        int idxJ = (int)((y0 - minY) / dy);
        double wx[2], wy[2];
        wx[1] = (x0 - minX) / dx - idxI; wx[0] = 1.0 - wx[1];
        wy[1] = (y0 - minY) / dy - idxJ; wy[0] = 1.0 - wy[1];
        if (!(idxI >= 0 && idxI < nx - 1 && idxJ >= 0 && idxJ < ny - 1)) continue;
        double sumU2 = 0.0;
        for (int h = 0; h < p->n_harm; ++h) {
          cplx ux = 0, uy = 0;
          for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj) {
              size_t c = h * fs + slot * N + (size_t)(idxJ + jj) * nx + (idxI + ii);
              ux += CC_(st->Ux)[c] * wx[ii] * wy[jj];
              uy += CC_(st->Uy)[c] * wx[ii] * wy[jj];
            }
          int H = p->harmony[h];
          sumU2 += (creal(ux) * creal(ux) + cimag(ux) * cimag(ux) + creal(uy) * creal(uy) +
                    cimag(uy) * cimag(uy)) / (2.0 * H * H);
        }
        double t = ku - ks / (2.0 * gamR * gamR) * (K2 + sumU2);
        st->theta[n] += t * dz;
        st->theta[n] -= dz * wakeE;
      }
    }
  }
  return ORC_OK;
}

/* rearrangeParticles.cpp:77-114: theta wrap into [0, 2*pi*numSlice]. */
int orc_periodic_particles(orc_state *st) {
  const orc_params *p = st->p;
  double dTh = p->num_slice * 2 * M_PI;
  int64_t np = n_particles(st);
  int err = ORC_OK;
  for (int64_t n = 0; n < np; ++n) {
    double th = st->theta[n], del = 0.0;
    int it;
    if (th >= dTh) {
      it = (int)(th / dTh);
      del = dTh * it;
    } else if (th < 0) {
      it = (int)(th / dTh - 1.0);
      del = dTh * it;
    }
    st->theta[n] -= del;
    if (st->theta[n] > dTh || st->theta[n] < 0) err = ORC_ERR_RANGE;
  }
  return err;
}

/* ------------------------------------------------------------------ fields */

/* solveField.cpp:187-403: bunching deposition and the radial Ez solve. */
int orc_solve_sc_3d(orc_state *st, const orc_lattice *lat) {
  const orc_params *p = st->p;
  int nx = p->nx, ny = p->ny, nr = p->nr, L = p->sc_l, F = p->sc_f;
  size_t N = (size_t)nx * ny, fs = field_stride(p);
  double minX = p->min_x, minY = p->min_y, dx = p->dx, dy = p->dy, dz = p->dz, dr = p->dr;
  double K0 = lat->K0, ks = p->ks, dB = p->dBessel, ue = lat->ue, gamR = p->gamR;
  double Phi;
  cplx etaX, etaY;
  double K0_alpha = lat->K0_alpha;
  if (coupling(ue, K0_alpha, &Phi, &etaX, &etaY) != ORC_OK) return ORC_ERR_ARG;
  cplx *Sx = C(st->ScUx), *Sy = C(st->ScUy);
  for (int h = 0; h < p->n_harm; ++h)
    for (size_t c = 0; c < fs; ++c) { Sx[h * fs + c] = 0; Sy[h * fs + c] = 0; }
  double coef_U = dz * K_ECHARGE * K_ECHARGE * K_MU0 * K0 * sqrt(1.0 + ue * ue) /
                  (4.0 * ks * K_EMASS * (p->lambda0 * p->num_slice) * dx * dy);
  double xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  double coef_Ez = K_ECHARGE * K_C * K_C * K_MU0 * ks * (1.0 + 0.5 * (1.0 + ue * ue) * K0 * K0) /
                   (2.0 * M_PI * p->lambda0 * p->num_slice * gamR * gamR);
  size_t es = (size_t)(p->slice_n + 2) * F * nr;
  size_t scn = (size_t)F * nr + 1; /* one spare: the idxR = nr-1 write, A.4 */
  double A = ks * ks * (1.0 + 0.5 * (1.0 + ue * ue) * K0 * K0) / (gamR * gamR) * dr * dr;
  double *a = malloc(sizeof(double) * nr), *bb = malloc(sizeof(double) * nr), *c = malloc(sizeof(double) * nr);
  cplx *d = malloc(sizeof(cplx) * nr);
  cplx *Sc = C(st->sc_scratch);
  for (int s = 0; s < p->n_species; ++s) {
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      double macro = st->seg_weight[(size_t)s * p->slice_n + sl];
      int64_t b0 = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e0 = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b0; n < e0; ++n) {
        double x = st->x[n], y = st->y[n], theta = st->theta[n], gam = st->gamma[n];
        double px = st->px[n], py = st->py[n];
        double xi2 = sqrt(2.0) * K0 / (1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue)) *
                     sqrt((1 + K0_alpha) * (px * px + ue * ue * py * py) +
                          (1 - K0_alpha) * (px * px * ue * ue + py * py));
        double B = atan(2.0 * K0_alpha * ue * (px - py) /
                        ((1.0 + K0_alpha) * (px + ue * ue * py) + (1 - K0_alpha) * (px * ue * ue + py)));
        cplx expP = cexp(I * (B - Phi * 0.5));
        cplx expM = conj(expP);
        int idxI = (int)((x - minX) / dx);
        int idxJ = (int)((y - minY) / dy);
        double wx[2], wy[2];
        wx[1] = (x - minX) / dx - idxI; wx[0] = 1.0 - wx[1];
        wy[1] = (y - minY) / dy - idxJ; wy[0] = 1.0 - wy[1];
        if (!(idxI >= 0 && idxI < nx && idxJ >= 0 && idxJ < ny)) continue;
        size_t idx = slot * N + (size_t)idxJ * nx + idxI;
        for (int h = 0; h < p->n_harm; ++h) {
          int H = p->harmony[h];
          double dbH = (double)H, w[2];
          int ib = (int)((dbH * xi) / dB);
          w[1] = (dbH * xi / dB) - ib; w[0] = 1.0 - w[1];
          if (ib < 0 || ib + 1 >= p->bn) { free(a); free(bb); free(c); free(d); return ORC_ERR_ARG; }
          cplx fx, fy;
          if (H % 2 == 1) {
            double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0;
            double J1 = bessel_at(st, ib, (H - 1) / 2, w);
            double J2 = bessel_at(st, ib, (H + 1) / 2, w);
            fx = sign * (J1 - K0_alpha * J2);
            fy = sign * (J1 + K0_alpha * J2);
          } else {
            double sign = (H / 2 % 2 == 0) ? 1.0 : -1.0;
            double J1 = bessel_at(st, ib, (H - 2) / 2, w);
            double J2 = bessel_at(st, ib, H / 2, w);
            double J3 = bessel_at(st, ib, (H + 2) / 2, w);
            fx = sign * xi2 * dbH * 0.5 * (expM * (J1 - K0_alpha * J2) + expP * (J2 - K0_alpha * J3));
            fy = sign * xi2 * dbH * 0.5 * (expM * (J1 + K0_alpha * J2) + expP * (J2 + K0_alpha * J3));
          }
          cplx mc = dbH * macro * coef_U * cexp(-I * dbH * (theta + Phi * 0.5)) / gam;
          for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj) {
              Sx[h * fs + idx + jj * nx + ii] -= etaX * wx[ii] * wy[jj] * fx * mc;
              Sy[h * fs + idx + jj * nx + ii] -= etaY * wx[ii] * wy[jj] * fy * mc;
            }
        }
      }
    }
    /* radial space-charge solve, species s overwrites Ez (solveField.cpp:315-398) */
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      for (int l = 0; l < L; ++l)
        for (size_t j = 0; j < scn; ++j) Sc[l * scn + j] = 0;
      double macro = st->seg_weight[(size_t)s * p->slice_n + sl];
      int64_t b0 = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e0 = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b0; n < e0; ++n) {
        double x = st->x[n], y = st->y[n], theta = st->theta[n];
        double phi = (x == 0.0) ? 0.0 : atan2(y, x);
        double r = sqrt(x * x + y * y);
        int idxR = (int)(r / dr + 0.5);
        if (idxR > 0 && idxR < nr) {
          double wr[2];
          wr[1] = (r * r / (dr * dr) - idxR * idxR) / (2.0 * idxR); wr[0] = 1.0 - wr[1];
          for (int l = 0; l < L; ++l)
            for (int f = 0; f < F; ++f) {
              cplx cc = I * coef_Ez * (l + 1.0) * cexp(-I * ((l + 1) * theta + f * phi)) * macro / (1.0 * idxR);
              for (int jj = 0; jj < 2; ++jj) Sc[l * scn + f * nr + idxR + jj] += cc * wr[jj];
            }
        } else if (idxR == 0) {
          for (int l = 0; l < L; ++l) {
            cplx cc = I * coef_Ez * (l + 1.0) * cexp(-I * ((l + 1) * theta)) * macro;
            Sc[l * scn + idxR] += cc;
            for (int f = 1; f < F; ++f) Sc[l * scn + f * nr + idxR] = 0;
          }
        }
      }
      for (int f = 0; f < F; ++f)
        for (int l = 0; l < L; ++l) {
          if (f == 0) {
            a[0] = 1.0; bb[0] = -(1.0 + (l + 1.0) * (l + 1.0) * A / 8.0); c[0] = 0.0; d[0] = Sc[l * scn + 0];
          } else {
            a[0] = 0.0; bb[0] = 1.0; c[0] = 0.0; d[0] = 0.0;
          }
          for (int j = 1; j < nr - 1; ++j) {
            a[j] = j + 0.5;
            bb[j] = -2.0 * j - (l + 1.0) * (l + 1.0) * A * j - f * f * log((j + 0.5) / (j - 0.5));
            c[j] = j - 0.5;
            d[j] = Sc[l * scn + f * nr + j];
          }
          a[nr - 1] = 0.0; bb[nr - 1] = 1.0; c[nr - 1] = 0.0; d[nr - 1] = 0.0;
          for (int j = 1; j < nr; ++j) {
            double w = c[j] / bb[j - 1];
            bb[j] -= w * a[j - 1];
            d[j] -= w * d[j - 1];
          }
          cplx *ez = C(st->Ez) + l * es;
          ez[f * nr + nr - 1] = 0; /* slot 0, A.3 */
          for (int j = nr - 2; j >= 0; --j)
            ez[(size_t)slot * F * nr + f * nr + j] = (d[j] - a[j] * ez[f * nr + j + 1]) / bb[j];
        }
    }
  }
  free(a); free(bb); free(c); free(d);
  return ORC_OK;
}

/* solveField.cpp:80-185: Peaceman-Rachford ADI with the absorbing layer, per
 * harmonic and slice. Keeps the special last-row equations (146-147, 174-175)
 * and leaves the boundary ring of U untouched. */
int orc_solve_field_u_3d(orc_state *st, double *Ud, const double *Scd, int current_flag) {
  const orc_params *p = st->p;
  int nx = p->nx, ny = p->ny;
  size_t N = (size_t)nx * ny, fs = field_stride(p);
  double cf = current_flag ? 1.0 : 0.0;
  const cplx *al = CC_(st->alpha), *alP = CC_(st->alphaP), *be = CC_(st->beta), *beP = CC_(st->betaP);
  cplx *U = C(Ud);
  const cplx *Sc = CC_(Scd);
  int nt = g_threads;
  cplx *work = malloc(sizeof(cplx) * (size_t)nt * (N + 3 * (size_t)(nx > ny ? nx : ny)));
  for (int h = 0; h < p->n_harm; ++h) {
    double H = (double)p->harmony[h];
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      int nm = nx > ny ? nx : ny;
      cplx *Uc = work + (size_t)tid * (N + 3 * (size_t)nm);
      cplx *CCv = Uc + N, *DDv = CCv + nm, *dd = DDv + nm;
      size_t slot = (size_t)sl + 1;
      cplx *Un = U + h * fs;
      const cplx *S = Sc + h * fs;
      for (int j = 0; j < ny; ++j) { Uc[j * nx] = 0; Uc[j * nx + nx - 1] = 0; }
      for (int i = 0; i < nx; ++i) { Uc[i] = 0; Uc[(ny - 1) * nx + i] = 0; }
      for (int j = 1; j < ny - 1; ++j) {
        for (int i = 1; i < nx - 1; ++i) {
          size_t idx = slot * N + (size_t)j * nx + i;
          dd[i] = -beP[j] * Un[idx + nx] + (H + beP[j] + be[j]) * Un[idx] - be[j] * Un[idx - nx] + S[idx] * cf;
        }
        CCv[1] = (H - alP[1] - al[1]) / alP[1];
        DDv[1] = dd[1] / alP[1];
        for (int i = 2; i < nx - 2; ++i) {
          CCv[i] = (H - alP[i] - al[i]) / alP[i] - al[i] / (alP[i] * CCv[i - 1]);
          DDv[i] = dd[i] / alP[i] - al[i] * DDv[i - 1] / (alP[i] * CCv[i - 1]);
        }
        CCv[nx - 2] = H - alP[nx - 2] - al[nx - 2] - 1.0 / CCv[nx - 3];
        DDv[nx - 2] = dd[nx - 2] - DDv[nx - 3] / CCv[nx - 3];
        Uc[j * nx + nx - 2] = DDv[nx - 2] / CCv[nx - 2];
        for (int i = nx - 3; i > 0; --i) Uc[j * nx + i] = (DDv[i] - Uc[j * nx + i + 1]) / CCv[i];
      }
      for (int i = 1; i < nx - 1; ++i) {
        for (int j = 1; j < ny - 1; ++j) {
          size_t idx = (size_t)j * nx + i;
          dd[j] = -alP[i] * Uc[idx + 1] + (H + alP[i] + al[i]) * Uc[idx] - al[i] * Uc[idx - 1] +
                  S[slot * N + (size_t)j * nx + i] * cf;
        }
        CCv[1] = (H - beP[1] - be[1]) / beP[1];
        DDv[1] = dd[1] / beP[1];
        for (int j = 2; j < ny - 1; ++j) {
          CCv[j] = (H - beP[j] - be[j]) / beP[j] - be[j] / (beP[j] * CCv[j - 1]);
          DDv[j] = dd[j] / beP[j] - be[j] * DDv[j - 1] / (beP[j] * CCv[j - 1]);
        }
        CCv[ny - 2] = H - beP[ny - 2] - be[ny - 2] - 1.0 / CCv[ny - 3];
        DDv[ny - 2] = dd[ny - 2] - DDv[ny - 3] / CCv[ny - 3];
        Un[slot * N + (size_t)(ny - 2) * nx + i] = DDv[ny - 2] / CCv[ny - 2];
        for (int j = ny - 3; j > 0; --j)
          Un[slot * N + (size_t)j * nx + i] = (DDv[j] - Un[slot * N + (size_t)(j + 1) * nx + i]) / CCv[j];
      }
    }
  }
  free(work);
  return ORC_OK;
}

/* solveField.cpp:49-78 (3D branch). */
int orc_solve_field(orc_state *st, const orc_lattice *lat) {
  const orc_params *p = st->p;
  size_t fs = field_stride(p);
  for (int h = 0; h < p->n_harm; ++h)
    for (int k = 0; k < p->slice_n + 2; ++k) {
      C(st->ScUx)[h * fs + k] = 0;
      C(st->ScUy)[h * fs + k] = 0;
    }
  if (p->dimension == 1) {
    if (lat->current_flag) {
      int e = orc_solve_sc_1d(st, lat);
      if (e) return e;
    }
    orc_solve_field_u_1d(st);
    return ORC_OK;
  }
  if (lat->current_flag) {
    int e = orc_solve_sc_3d(st, lat);
    if (e) return e;
  }
  orc_solve_field_u_3d(st, st->Ux, st->ScUx, lat->current_flag);
  orc_solve_field_u_3d(st, st->Uy, st->ScUy, lat->current_flag);
  return ORC_OK;
}

/* solveField.cpp:19-47 + fieldShareZ.cpp:39-78 for one rank: slot s <- s-1
 * for s = slice_n+1..1; the exchange with rank+1 is empty. */
void orc_shift_field(orc_state *st) {
  const orc_params *p = st->p;
  size_t N = (size_t)p->nx * p->ny, fs = field_stride(p);
  for (int h = 0; h < p->n_harm; ++h)
    for (int s = p->slice_n + 1; s >= 1; --s) {
      memmove(C(st->Ux) + h * fs + s * N, C(st->Ux) + h * fs + (s - 1) * N, N * sizeof(cplx));
      memmove(C(st->Uy) + h * fs + s * N, C(st->Uy) + h * fs + (s - 1) * N, N * sizeof(cplx));
    }
}

/* ------------------------------------------------------------------ 1D path */

/* particlePush.cpp:388-524 */
int orc_push_theta_gamma_1d(orc_state *st, const orc_lattice *lat, int iteration) {
  (void)iteration;
  const orc_params *p = st->p;
  double ue = lat->ue, ks = p->ks, ku = lat->ku, K0 = lat->K0, dz = p->dz;
  int L = p->sc_l;
  double dPhi = 2.0 * M_PI * p->num_slice, Phi;
  cplx etaX, etaY;
  double K0_alpha = lat->K0_alpha;
  if (coupling(ue, K0_alpha, &Phi, &etaX, &etaY) != ORC_OK) return ORC_ERR_ARG;
  double e_mc2 = K_ECHARGE / K_EMASS / K_C / K_C;
  double K2 = 1.0 + K0 * K0 * 0.5 * (1.0 + ue * ue);
  const double coef[5] = {0, 0, 0.5, 0.5, 1};
  double dB = p->dBessel;
  double xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  size_t fs = (size_t)(p->slice_n + 2);
  size_t es = (size_t)(p->slice_n + 2);
  for (int s = 0; s < p->n_species; ++s)
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      double wakeE = p->wake_on ? st->wakeE[sl] / (K_MC2 * 1.0e6) : 0.0;
      cplx Em[16], Ux[ORC_MAX_HARM], Uy[ORC_MAX_HARM];
      for (int l = 0; l < L; ++l) Em[l] = CC_(st->Ez)[l * es + slot];
      double sumU2 = 0;
      for (int h = 0; h < p->n_harm; ++h) {
        Ux[h] = CC_(st->Ux)[h * fs + slot];
        Uy[h] = CC_(st->Uy)[h * fs + slot];
        int H = p->harmony[h];
        sumU2 += (creal(Ux[h]) * creal(Ux[h]) + cimag(Ux[h]) * cimag(Ux[h]) + creal(Uy[h]) * creal(Uy[h]) +
                  cimag(Uy[h]) * cimag(Uy[h])) / (2.0 * H * H);
      }
      int64_t b0 = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e0 = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b0; n < e0; ++n) {
        double th0 = st->theta[n], gam0 = st->gamma[n];
        double kth[5] = {0}, kg[5] = {0};
        for (int m = 1; m < 5; ++m) {
          double th = th0 + dz * kth[m - 1] * coef[m];
          double gam = gam0 + dz * kg[m - 1] * coef[m];
          double sumTh = 0, sumG = 0, sumEz = 0;
          for (int l = 0; l < L; ++l) sumEz += 2.0 * creal(Em[l] * cexp(I * (l + 1.0) * th));
          for (int h = 0; h < p->n_harm; ++h) {
            int H = p->harmony[h];
            double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0, w[2];
            int idx = (int)((1.0 * H * xi) / dB);
            w[1] = (H * xi / dB) - idx; w[0] = 1.0 - w[1];
            cplx fx = 0, fy = 0;
            if (H % 2 == 1) {
              double J1 = bessel_at(st, idx, (H - 1) / 2, w);
              double J2 = bessel_at(st, idx, (H + 1) / 2, w);
              fx = sign * (K0_alpha * J1 - J2);
              fy = sign * (K0_alpha * J1 + J2);
            }
            cplx t = cexp(I * (1.0 * H) * (th + Phi * 0.5)) * (etaX * fx * Ux[h] - etaY * fy * Uy[h]);
            sumTh += creal(I * t) / (1.0 * H);
            sumG += creal(t) / (2.0 * gam);
          }
          kth[m] = ku - ks / (2.0 * gam * gam) * (K2 + sumU2 + K0 * sqrt(ue * ue + 1.0) * sumTh);
          kg[m] = ks * K0 * sqrt(ue * ue + 1.0) * sumG + e_mc2 * sumEz;
        }
        double dth = dz / 6.0 * (kth[1] + 2.0 * kth[2] + 2.0 * kth[3] + kth[4]);
        if (dth > dPhi || dth <= -dPhi) return ORC_ERR_GUARD;
        st->theta[n] += dth;
        st->gamma[n] += dz / 6.0 * (kg[1] + 2 * kg[2] + 2 * kg[3] + kg[4]) - dz * wakeE;
      }
    }
  return ORC_OK;
}

/* solveField.cpp:422-549. st->beam_sig[2s], [2s+1] = sigX, sigY of species s
 * (sqrt(emit/gamma0 / gammaTwiss), lines 463-468). */
int orc_solve_sc_1d(orc_state *st, const orc_lattice *lat) {
  const double *beam_sig = st->beam_sig;
  const orc_params *p = st->p;
  double ue = lat->ue, Phi;
  cplx etaX, etaY;
  double K0_alpha = lat->K0_alpha;
  if (coupling(ue, K0_alpha, &Phi, &etaX, &etaY) != ORC_OK) return ORC_ERR_ARG;
  double dz = p->dz, ks = p->ks, K0 = lat->K0, dB = p->dBessel;
  double xi = 0.25 * K0 * K0 * (1.0 - ue * ue) / (1.0 + (1.0 + ue * ue) * K0 * K0 * 0.5);
  size_t fs = (size_t)(p->slice_n + 2);
  int L = p->sc_l;
  for (int s = 0; s < p->n_species; ++s) {
    double area = 2.0 * M_PI * beam_sig[2 * s] * beam_sig[2 * s + 1];
    double coef = dz * K_ECHARGE * K_ECHARGE * K_MU0 * K0 * sqrt(1.0 + ue * ue) /
                  (2.0 * ks * K_EMASS * (p->lambda0 * p->num_slice) * area);
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      double macro = st->seg_weight[(size_t)s * p->slice_n + sl];
      int64_t b0 = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e0 = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b0; n < e0; ++n) {
        double th = st->theta[n], gam = st->gamma[n];
        for (int h = 0; h < p->n_harm; ++h) {
          int H = p->harmony[h];
          double sign = ((H - 1) / 2 % 2 == 0) ? 1.0 : -1.0, w[2];
          cplx fx = 0, fy = 0;
          if (H % 2 == 1) {
            int idx = (int)(H * xi / dB);
            w[1] = (H * xi / dB) - idx; w[0] = 1.0 - w[1];
            double J1 = bessel_at(st, idx, (H - 1) / 2, w);
            double J2 = bessel_at(st, idx, (H + 1) / 2, w);
            fx = sign * (J1 - K0_alpha * J2);
            fy = sign * (J1 + K0_alpha * J2);
          }
          cplx mc = macro * coef * cexp((-1.0 * H) * I * (th + Phi * 0.5)) / gam;
          C(st->ScUx)[h * fs + slot] -= etaX * fx * mc;
          C(st->ScUy)[h * fs + slot] -= etaY * fy * mc;
        }
      }
    }
  }
  for (int l = 0; l < L; ++l)
    for (int k = 0; k <= p->slice_n + 1; ++k) C(st->Ez)[l * fs + k] = 0;
  for (int s = 0; s < p->n_species; ++s) {
    double area = 2.0 * M_PI * beam_sig[2 * s] * beam_sig[2 * s + 1];
    double coef = K_ECHARGE * K_C * K_C * K_MU0 / ks / area / (p->lambda0 * p->num_slice);
    for (int sl = 0; sl < p->slice_n; ++sl) {
      int slot = sl + 1;
      double macro = st->seg_weight[(size_t)s * p->slice_n + sl];
      int64_t b0 = st->seg_off[(size_t)s * p->slice_n + sl];
      int64_t e0 = st->seg_off[(size_t)s * p->slice_n + sl + 1];
      for (int64_t n = b0; n < e0; ++n) {
        double th = st->theta[n];
        for (int l = 0; l < L; ++l)
          C(st->Ez)[l * fs + slot] += -1.0 * I * coef / (1.0 + l) * macro * cexp(-I * (l + 1.0) * th);
      }
    }
  }
  return ORC_OK;
}

/* solveField.cpp:409-420 */
void orc_solve_field_u_1d(orc_state *st) {
  const orc_params *p = st->p;
  size_t fs = (size_t)(p->slice_n + 2);
  for (int h = 0; h < p->n_harm; ++h)
    for (int k = 1; k <= p->slice_n; ++k) {
      C(st->Ux)[h * fs + k] += C(st->ScUx)[h * fs + k];
      C(st->Uy)[h * fs + k] += C(st->ScUy)[h * fs + k];
    }
}

/* ------------------------------------------------------------- diagnostics */

/* updateTotalEnergy.cpp:25-36: raw sums of |U|^2 over slots 1..slice_n. */
void orc_total_energy(const orc_state *st, double *ex, double *ey) {
  const orc_params *p = st->p;
  size_t N = (size_t)p->nx * p->ny, fs = field_stride(p);
  for (int h = 0; h < p->n_harm; ++h) {
    double tx = 0.0, ty = 0.0;
    for (size_t c = N; c < (size_t)(p->slice_n + 1) * N; ++c) {
      cplx ux = CC_(st->Ux)[h * fs + c], uy = CC_(st->Uy)[h * fs + c];
      tx += creal(ux) * creal(ux) + cimag(ux) * cimag(ux);
      ty += creal(uy) * creal(uy) + cimag(uy) * cimag(uy);
    }
    ex[h] = tx;
    ey[h] = ty;
  }
}

/* saveFile.cpp:109-172: b[h] = sum w exp(i H theta) (complex, 2*n_harm) and
 * total weight. Slice-major then species, as the reference loops. */
void orc_bfactor(const orc_state *st, double *b, double *total_n) {
  const orc_params *p = st->p;
  double tot = 0.0;
  for (int h = 0; h < p->n_harm; ++h) {
    double H = p->harmony[h];
    cplx acc = 0;
    for (int sl = 0; sl < p->slice_n; ++sl)
      for (int s = 0; s < p->n_species; ++s) {
        size_t seg = (size_t)s * p->slice_n + sl;
        double w = st->seg_weight[seg];
        for (int64_t n = st->seg_off[seg]; n < st->seg_off[seg + 1]; ++n) {
          acc += cexp(I * H * st->theta[n]) * w;
          if (h == 0) tot += w;
        }
      }
    b[2 * h] = creal(acc);
    b[2 * h + 1] = cimag(acc);
  }
  *total_n = tot;
}

/* twiss.cpp:7-99: sums[s*7 + k] = <x^2>,<y^2>,<x'^2>,<y'^2>,<xx'>,<yy'> sums
 * and the count; derived[s*6 + k] = emitX, betaX, alphaX, emitY, betaY, alphaY
 * (emittances not yet multiplied by gamma0). */
void orc_twiss(const orc_state *st, double *sums, double *derived) {
  const orc_params *p = st->p;
  for (int s = 0; s < p->n_species; ++s) {
    double ax = 0, ay = 0, axp = 0, ayp = 0, cx = 0, cy = 0, cnt = 0;
    for (int sl = 0; sl < p->slice_n; ++sl) {
      size_t seg = (size_t)s * p->slice_n + sl;
      for (int64_t n = st->seg_off[seg]; n < st->seg_off[seg + 1]; ++n) {
        double x = st->x[n], y = st->y[n], inv = 1.0 / st->gamma[n];
        double xp = st->px[n] * inv, yp = st->py[n] * inv;
        ax += x * x; ay += y * y; axp += xp * xp; ayp += yp * yp;
        cx += x * xp; cy += y * yp; cnt += 1.0;
      }
    }
    double *o = sums + 7 * s;
    o[0] = ax; o[1] = ay; o[2] = axp; o[3] = ayp; o[4] = cx; o[5] = cy; o[6] = cnt;
    double *d = derived + 6 * s;
    for (int k = 0; k < 6; ++k) d[k] = 0.0;
    if (cnt > 0.0) {
      double x2 = ax / cnt, y2 = ay / cnt, xp2 = axp / cnt, yp2 = ayp / cnt, xx = cx / cnt, yy = cy / cnt;
      double ex = sqrt(x2 * xp2 - xx * xx), ey = sqrt(y2 * yp2 - yy * yy);
      d[0] = ex; d[1] = x2 / ex; d[2] = -xx / ex;
      d[3] = ey; d[4] = y2 / ey; d[5] = -yy / ey;
    }
  }
}

/* -------------------------------------------------------------------- wake */

/* wakeField.cpp:53-159. shape 1 Flat, 2 Circular; acdc 1 AC, 2 DC. */
int orc_wake_function(int shape, int acdc, double radius, double cond, double ctau,
                      int slice_n, double dZ, double *impZ, double *wakeF) {
  double s0 = pow(2.0 * radius * radius / K_Z0 / cond, 1.0 / 3.0);
  const double maxK = 20, maxX = 2000;
  const int numK = 1000, numX = 50000;
  double minK = 0.0, dK = maxK / (numK * 1.0), minX = 0.0, dX = maxX / (numX * 1.0);
  for (int i = 0; i <= numK; ++i) impZ[i] = 0.0;
  int err = ORC_OK;
  if (shape == 1) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int i = 1; i <= numK; ++i) {
      double k = minK + i * dK;
      cplx csum = 0;
      for (int j = 0; j <= numX; ++j) {
        double x = minX + j * dX;
        cplx ccal;
        if (x == 0.0) {
          ccal = sqrt(k) / (2.0 + k * k * k - 2.0 * sqrt(k) * k);
        } else if (x > 320) {
          ccal = 0;
        } else {
          cplx cd = 2.0 / (1.0 - I) / csqrt(k - I * k * k * ctau / s0) * cosh(x) - I * k * sinh(x) / x;
          ccal = 1.0 / cosh(x) / cd;
          if (isnan(creal(ccal))) err = ORC_ERR_NAN;
        }
        csum += ccal * dX;
      }
      impZ[i] = creal(csum);
    }
  } else if (shape == 2) {
    for (int i = 1; i <= numK; ++i) {
      double k = minK + i * dK;
      if (acdc == 1) {
        cplx cd = sqrt(2.0 * s0 / ctau) / k * (I + s0 * 0.5 / k / ctau) - I * 0.5 * k;
        impZ[i] = creal(1.0 / cd);
      } else {
        double dino = 2.0 / k + 0.25 * k * k - sqrt(k);
        impZ[i] = 1.0 / sqrt(k) / dino;
      }
    }
  }
  double coef = K_Z0 * 3e8 / M_PI / radius / radius / M_PI;
  for (int i = 0; i < slice_n; ++i) {
    double z = (i * dZ) / s0, sum = 0.0;
    for (int j = 0; j <= numK; ++j) {
      double k = minK + j * dK;
      sum += impZ[j] * cos(k * z) * dK;
    }
    wakeF[i] = 0.0 + sum * coef;
  }
  return err;
}

/* wakeField.cpp:9-50 and makeDensity 161-223. */
void orc_update_wake(const orc_state *st, const double *wakeF, double *den,
                     double *wakeE, double *total_cnt) {
  const orc_params *p = st->p;
  int sn = p->slice_n;
  double dZ = p->num_slice * p->lambda0;
  for (int i = 0; i < sn; ++i) { den[i] = 0.0; wakeE[i] = 0.0; }
  for (int s = 0; s < p->n_species; ++s)
    for (int sl = 0; sl < sn; ++sl) {
      size_t seg = (size_t)s * sn + sl;
      double cnt = (double)(st->seg_off[seg + 1] - st->seg_off[seg]);
      den[sl] += st->seg_weight[seg] * cnt;
    }
  double tot = 0.0;
  for (int i = 0; i < sn; ++i) tot += den[i];
  if (tot > 0.0) {
    double norm = 1.0 / (tot * dZ);
    for (int i = 0; i < sn; ++i) den[i] *= norm;
  }
  *total_cnt = tot;
  double coef = tot * 1.602e-19;
  for (int i = sn - 1; i >= 0; --i)
    for (int ii = i; ii >= 0; --ii) wakeE[ii] += den[i] * wakeF[i - ii] * dZ * coef;
}

/* -------------------------------------------------------------------- seed */

/* loadSeed.cpp:25-80; min_z is Domain minZ (the routine adds lambda0/2). */
void orc_load_seed_3d(const orc_params *p, double *Uxd, double *Uyd, int h_pick,
                      double a0, double duration, double min_z, double zR,
                      double focus, double spotSigR, double polarity,
                      double laserAlpha, double laserPsi) {
  int nx = p->nx, ny = p->ny;
  size_t N = (size_t)nx * ny, fs = field_stride(p);
  double bucketZ = p->lambda0 * p->num_slice;
  double sigmaZ = duration * K_C;
  double invS2 = 1.0 / sigmaZ / sigmaZ;
  double minZ = min_z + p->lambda0 * 0.5;
  double ks = p->ks, sigR = sqrt(2.0) * spotSigR;
  double factor = sqrt(1.0 + polarity * polarity) * 0.5;
  cplx *Ux = C(Uxd) + h_pick * fs, *Uy = C(Uyd) + h_pick * fs;
  for (int sl = 0; sl < p->slice_n; ++sl) {
    size_t slot = (size_t)sl + 1;
    double z = sl * bucketZ + minZ;
    double delZ = z - focus;
    double ampZ = a0 * exp(-0.25 * z * z * invS2);
    double curv = (delZ == 0.0) ? 1e100 : delZ * (1.0 + zR * zR / (delZ * delZ));
    double gouy = atan(delZ / zR);
    double w = sigR * sqrt(1.0 + delZ * delZ / (zR * zR));
    for (int i = 0; i < nx; ++i) {
      double x = p->min_x + i * p->dx;
      for (int j = 0; j < ny; ++j) {
        double y = p->min_y + j * p->dy;
        double r2 = x * x + y * y;
        double phase = -r2 / (2.0 * w * w);
        cplx cph = I * (ks * delZ + ks * r2 / (2.0 * curv) - gouy);
        cplx amp = factor * (ampZ * sigR / w) * cexp(cph + phase);
        Ux[slot * N + (size_t)j * nx + i] = amp * (1.0 + laserAlpha * cexp(I * laserPsi));
        Uy[slot * N + (size_t)j * nx + i] = amp * (-1.0 + laserAlpha * cexp(I * laserPsi));
      }
    }
  }
}

/* --------------------------------------------------------------- the loop */

int orc_run(orc_state *st, const orc_lattice *lat_tab, int it0, int n_steps,
            int time_dependent, int wake_step, const double *wakeF, double *wakeE_buf,
            double *energy_out, double *bfac_out, double *twiss_out) {
  const orc_params *p = st->p;
  int nh = p->n_harm, ns = p->n_species;
  double *den = wakeF ? malloc(sizeof(double) * (size_t)p->slice_n) : NULL;
  double *tw = malloc(sizeof(double) * 6 * (size_t)ns);
  double *bb = malloc(sizeof(double) * 2 * (size_t)nh);
  int err = ORC_OK;
  for (int k = 0; k < n_steps && !err; ++k) {
    int it = it0 + k;
    if (time_dependent && wakeF && wake_step > 0 && it % wake_step == 0) {
      double tc;
      orc_update_wake(st, wakeF, den, wakeE_buf, &tc);
      st->wakeE = wakeE_buf;
    }
    if (energy_out) orc_total_energy(st, energy_out + (size_t)k * 2 * nh, energy_out + (size_t)k * 2 * nh + nh);
    if (twiss_out) orc_twiss(st, twiss_out + (size_t)k * 7 * ns, tw);
    if (bfac_out) {
      double tot;
      orc_bfactor(st, bb, &tot);
      for (int h = 0; h < nh; ++h)
        bfac_out[(size_t)k * (nh + 1) + h] = cabs(CMPLX(bb[2 * h], bb[2 * h + 1]) / tot);
      bfac_out[(size_t)k * (nh + 1) + nh] = tot;
    }
    const orc_lattice *ls = &lat_tab[2 * k], *l0 = &lat_tab[2 * k + 1], *l1 = &lat_tab[2 * k + 2];
    err = orc_solve_field(st, ls);
    if (err) break;
    if (p->dimension == 3) {
      err = orc_transverse_push(st, l0);
      if (!err) err = orc_transverse_push(st, l1);
      if (err) break;
    }
    if (!l1->drift_flag)
      err = p->dimension == 3 ? orc_push_theta_gamma_3d(st, l1, it) : orc_push_theta_gamma_1d(st, l1, it);
    else
      err = orc_drift_theta_gamma(st, l1);
    if (err) break;
    err = orc_periodic_particles(st);
    if (err) break;
    if (!l1->drift_flag && time_dependent) orc_shift_field(st);
  }
  free(den); free(tw); free(bb);
  return err;
}
