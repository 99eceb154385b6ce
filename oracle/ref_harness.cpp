// Test harness around the reference's own translation units (compiled from
// /root/reference by oracle/Makefile into oracle/_ref/libfelisref.so).
//
// TEST INFRASTRUCTURE ONLY: it fills a reference Domain (mesh.h:66-145) with
// caller-supplied state for one MPI rank, calls one reference routine, and
// copies the results out. tests/golden/make_golden.py uses it to write the
// golden vectors the oracle is checked against. Nothing here is part of the
// product. Reference routines that print to files (totalEnergy, bFactor,
// twissFile, wakeE, impZ, normDen) do so in the directory given by `workdir`;
// the harness parses the line they append.
#include <mpi.h>
#include <unistd.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "constants.h"
#include "mesh.h"

namespace {

void ensure_mpi() {
  int flag = 0;
  MPI_Initialized(&flag);
  if (!flag) {
    int argc = 0;
    char **argv = nullptr;
    MPI_Init(&argc, &argv);
  }
}

struct Cwd {
  std::string old;
  explicit Cwd(const char *dir) {
    char buf[4096];
    if (getcwd(buf, sizeof buf)) old = buf;
    if (dir && chdir(dir) != 0) std::perror("chdir");
  }
  ~Cwd() {
    if (!old.empty() && chdir(old.c_str()) != 0) std::perror("chdir");
  }
};

// Particles: segment-major (species, slice) lists, one weight per list.
void fill_particles(Domain &D, int n_species, const int64_t *seg_off, const double *w,
                    const double *x, const double *y, const double *px, const double *py,
                    const double *th, const double *ga) {
  D.nSpecies = n_species;
  D.particle.resize(D.subSliceN + 2);
  for (auto &p : D.particle) {
    p.head.resize(n_species);
    for (int s = 0; s < n_species; ++s) {
      p.head[s] = new ptclHead{};
      p.head[s]->pt = new ptclList{};
    }
  }
  for (int s = 0; s < n_species; ++s)
    for (int sl = 0; sl < D.subSliceN; ++sl) {
      size_t seg = (size_t)s * D.subSliceN + sl;
      ptclList *p = D.particle[sl + 1].head[s]->pt;
      int64_t b = seg_off[seg], e = seg_off[seg + 1];
      p->weight = w[seg];
      for (int64_t n = b; n < e; ++n) {
        p->x.push_back(x ? x[n] : 0.0);
        p->y.push_back(y ? y[n] : 0.0);
        p->px.push_back(px ? px[n] : 0.0);
        p->py.push_back(py ? py[n] : 0.0);
        p->theta.push_back(th ? th[n] : 0.0);
        p->gamma.push_back(ga ? ga[n] : 1.0);
      }
    }
}

void free_particles(Domain &D) {
  for (auto &p : D.particle)
    for (auto *h : p.head) {
      delete h->pt;
      delete h;
    }
}

std::string last_line(const char *file) {
  FILE *f = std::fopen(file, "r");
  if (!f) return "";
  char buf[65536];
  std::string last;
  while (std::fgets(buf, sizeof buf, f)) {
    if (buf[0] != '#') last = buf;
  }
  std::fclose(f);
  return last;
}

int parse_doubles(const std::string &s, double *out, int maxn) {
  int n = 0;
  const char *p = s.c_str();
  char *end;
  while (n < maxn) {
    double v = std::strtod(p, &end);
    if (end == p) break;
    out[n++] = v;
    p = end;
  }
  return n;
}

}  // namespace

extern "C" {

// rearrangeParticles.cpp:77-114
int ref_periodic(int slice_n, int num_slice, int n_species, const int64_t *seg_off,
                 const double *w, double *theta) {
  ensure_mpi();
  Domain D{};
  D.subSliceN = slice_n;
  D.sliceN = slice_n;
  D.numSlice = num_slice;
  fill_particles(D, n_species, seg_off, w, nullptr, nullptr, nullptr, nullptr, theta, nullptr);
  periodicParticles(D, 0);
  for (int s = 0; s < n_species; ++s)
    for (int sl = 0; sl < slice_n; ++sl) {
      const ptclList *p = D.particle[sl + 1].head[s]->pt;
      int64_t b = seg_off[(size_t)s * slice_n + sl];
      for (size_t k = 0; k < p->theta.size(); ++k) theta[b + k] = p->theta[k];
    }
  free_particles(D);
  return 0;
}

// updateTotalEnergy.cpp:7-87. fields [h][(slice_n+2)*nx*ny] complex.
// out_raw[2*nh] = raw sums (x then y); out_line[1+2*nh] = the appended line.
int ref_total_energy(const char *workdir, int nx, int ny, int slice_n, int nh, const double *Ux,
                     const double *Uy, double dx, double dy, double ks, double lambda0,
                     int num_slice, int mode, double min_z, double max_z, double dz,
                     int iteration, double *out_raw, double *out_line) {
  ensure_mpi();
  Cwd cwd(workdir);
  Domain D{};
  D.nx = nx; D.ny = ny; D.subSliceN = slice_n; D.sliceN = slice_n; D.numHarmony = nh;
  D.dx = dx; D.dy = dy; D.ks = ks; D.lambda0 = lambda0; D.numSlice = num_slice;
  D.mode = mode == 1 ? OperationMode::Static : OperationMode::Time_Dependent;
  D.minZ = min_z; D.maxZ = max_z; D.dz = dz; D.dimension = 3;
  D.maxStep = iteration + 1;
  size_t fs = (size_t)(slice_n + 2) * nx * ny;
  D.Ux.resize(nh); D.Uy.resize(nh);
  for (int h = 0; h < nh; ++h) {
    D.Ux[h].resize(fs); D.Uy[h].resize(fs);
    for (size_t c = 0; c < fs; ++c) {
      D.Ux[h][c] = cplx(Ux[2 * (h * fs + c)], Ux[2 * (h * fs + c) + 1]);
      D.Uy[h][c] = cplx(Uy[2 * (h * fs + c)], Uy[2 * (h * fs + c) + 1]);
    }
  }
  D.totalEnergyX.assign(D.maxStep, std::vector<double>(nh));
  D.totalEnergyY.assign(D.maxStep, std::vector<double>(nh));
  updateTotalEnergy(&D, iteration);
  for (int h = 0; h < nh; ++h) {
    out_raw[h] = D.totalEnergyX[iteration][h];
    out_raw[nh + h] = D.totalEnergyY[iteration][h];
  }
  return parse_doubles(last_line("totalEnergy"), out_line, 1 + 2 * nh);
}

// saveFile.cpp:109-172. out_line[1+nh] = z then |b_h| as printed (%10.4g).
int ref_bfactor(const char *workdir, int slice_n, int n_species, int nh, const int *harmony,
                const int64_t *seg_off, const double *w, const double *theta, double dz,
                double min_z, int iteration, double *out_line) {
  ensure_mpi();
  Cwd cwd(workdir);
  Domain D{};
  D.subSliceN = slice_n; D.sliceN = slice_n; D.numHarmony = nh; D.dz = dz; D.minZ = min_z;
  D.harmony.assign(harmony, harmony + nh);
  fill_particles(D, n_species, seg_off, w, nullptr, nullptr, nullptr, nullptr, theta, nullptr);
  updatebFactor(D, iteration);
  free_particles(D);
  return parse_doubles(last_line("bFactor"), out_line, 1 + nh);
}

// twiss.cpp:7-99. out_line[1+6*n_species] as printed (%12.6g).
int ref_twiss(const char *workdir, int slice_n, int n_species, const int64_t *seg_off,
              const double *w, const double *x, const double *y, const double *px,
              const double *py, const double *gamma, double gamma0, double dz, double min_z,
              int iteration, double *out_line) {
  ensure_mpi();
  Cwd cwd(workdir);
  Domain D{};
  D.subSliceN = slice_n; D.sliceN = slice_n; D.dz = dz; D.minZ = min_z; D.gamma0 = gamma0;
  fill_particles(D, n_species, seg_off, w, x, y, px, py, nullptr, gamma);
  calculate_twiss(D, iteration);
  free_particles(D);
  return parse_doubles(last_line("twissFile"), out_line, 1 + 6 * n_species);
}

// wakeField.cpp:53-159. shape 1 Flat / 2 Circular, acdc 1 AC / 2 DC.
int ref_wake_function(const char *workdir, int shape, int acdc, double radius, double cond,
                      double ctau, int slice_n, int num_slice, double lambda0,
                      double *wakeF) {
  ensure_mpi();
  Cwd cwd(workdir);
  Domain D{};
  D.sliceN = slice_n; D.numSlice = num_slice; D.lambda0 = lambda0;
  D.radius = radius; D.cond = cond; D.ctau = ctau;
  D.wakeShape = shape == 1 ? WakeShapeMode::Flat : WakeShapeMode::Circular;
  D.ac_dc = acdc == 1 ? WakeACDCMode::AC : WakeACDCMode::DC;
  D.wakeF.assign(slice_n, 0.0);
  wakeFunction(&D, 0);
  for (int i = 0; i < slice_n; ++i) wakeF[i] = D.wakeF[i];
  return 0;
}

// wakeField.cpp:9-50 + makeDensity 161-223.
int ref_update_wake(const char *workdir, int slice_n, int n_species, const int64_t *seg_off,
                    const double *w, int num_slice, double lambda0, double min_z, double dz,
                    const double *wakeF, double *den, double *wakeE, double *total_cnt) {
  ensure_mpi();
  Cwd cwd(workdir);
  Domain D{};
  D.subSliceN = slice_n; D.sliceN = slice_n; D.minI = 0; D.maxI = slice_n;
  D.numSlice = num_slice; D.lambda0 = lambda0; D.minZ = min_z; D.dz = dz;
  fill_particles(D, n_species, seg_off, w, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  D.wakeF.assign(wakeF, wakeF + slice_n);
  D.den.assign(slice_n, 0.0);
  D.wakeE.assign(slice_n, 0.0);
  updateWakeField(&D, 0);
  for (int i = 0; i < slice_n; ++i) { den[i] = D.den[i]; wakeE[i] = D.wakeE[i]; }
  *total_cnt = D.totalCnt;
  free_particles(D);
  return 0;
}

// updateK_quadG.cpp:8-94. Units / quads as flat arrays:
//   und[u*12 + k]: unit_start, unit_end, und_start, und_end, K0, slopeK, ue,
//                  lambdaU, K0_alpha, und_type, air, block_id
//   quad[q*4 + k]: q_start, q_end, g, list_id
//   state[8] in/out: K0, ue, K0_alpha, ku, g, lambdaU, und_type, (unused)
//   flags[2] out: current_flag, drift_flag
int ref_update_k_quadg(const double *und, int n_units, const double *quad, int n_quads, double dz,
                       int iteration, double half, double *state, int *flags) {
  ensure_mpi();
  Domain D{};
  D.dz = dz;
  int last_block = -1;
  for (int u = 0; u < n_units; ++u) {
    const double *r = und + 12 * u;
    int blk = (int)r[11];
    if (blk != last_block) {
      D.undList.push_back(UndList{});
      last_block = blk;
    }
    UndList &UL = D.undList.back();
    UL.unitStart.push_back(r[0]); UL.unitEnd.push_back(r[1]);
    UL.undStart.push_back(r[2]); UL.undEnd.push_back(r[3]);
    UL.K0.push_back(r[4]); UL.slopeK = r[5]; UL.ue = r[6]; UL.lambdaU = r[7];
    UL.K0_alpha = (int)r[8];
    UL.type = (int)r[9] == 1 ? UndMode::BiPolar : ((int)r[9] == 2 ? UndMode::QuadPolar : UndMode::Unknown);
    UL.air = r[10] != 0.0;
    UL.numbers = (int)UL.K0.size();
  }
  int last_list = -1;
  for (int q = 0; q < n_quads; ++q) {
    const double *r = quad + 4 * q;
    if ((int)r[3] != last_list) {
      D.quadList.push_back(QuadList{});
      last_list = (int)r[3];
    }
    QuadList &QD = D.quadList.back();
    QD.qdStart.push_back(r[0]); QD.qdEnd.push_back(r[1]); QD.g.push_back(r[2]);
    QD.unitStart.push_back(0.0); QD.unitEnd.push_back(0.0);
    QD.numbers = (int)QD.g.size();
  }
  D.K0 = state[0]; D.ue = state[1]; D.K0_alpha = state[2]; D.ku = state[3]; D.g = state[4];
  D.lambdaU = state[5];
  D.undType = (int)state[6] == 1 ? UndMode::BiPolar : ((int)state[6] == 2 ? UndMode::QuadPolar : UndMode::Unknown);
  updateK_quadG(&D, iteration, half);
  state[0] = D.K0; state[1] = D.ue; state[2] = D.K0_alpha; state[3] = D.ku; state[4] = D.g;
  state[5] = D.lambdaU;
  state[6] = D.undType == UndMode::BiPolar ? 1 : (D.undType == UndMode::QuadPolar ? 2 : 0);
  flags[0] = D.currentFlag ? 1 : 0;
  flags[1] = D.driftFlag ? 1 : 0;
  return 0;
}

// loadSeed.cpp:25-80. Fields out [numHarmony][(slice_n+2)*nx*ny] complex.
int ref_load_seed_3d(int nx, int ny, int slice_n, int nh, const int *harmony, double dx, double dy,
                     double dz, double min_x, double min_y, double min_z, double lambda0,
                     int num_slice, double ks, double a0, double duration, double zR,
                     double focus, double spotSigR, int loadH, double laserAlpha,
                     double laserPsi, double *Ux, double *Uy) {
  ensure_mpi();
  Domain D{};
  D.nx = nx; D.ny = ny; D.subSliceN = slice_n; D.sliceN = slice_n; D.minI = 0;
  D.numHarmony = nh; D.harmony.assign(harmony, harmony + nh);
  D.dx = dx; D.dy = dy; D.dz = dz; D.minX = min_x; D.minY = min_y; D.minZ = min_z;
  D.lambda0 = lambda0; D.numSlice = num_slice; D.ks = ks; D.a0 = a0; D.duration = duration;
  D.zR = zR; D.focus = focus; D.spotSigR = spotSigR; D.loadH = loadH;
  D.laserAlpha = laserAlpha; D.laserPsi = laserPsi; D.dimension = 3;
  size_t fs = (size_t)(slice_n + 2) * nx * ny;
  D.Ux.assign(nh, std::vector<cplx>(fs, cplx(0, 0)));
  D.Uy.assign(nh, std::vector<cplx>(fs, cplx(0, 0)));
  loadSeed(&D, 0);
  for (int h = 0; h < nh; ++h)
    for (size_t c = 0; c < fs; ++c) {
      Ux[2 * (h * fs + c)] = D.Ux[h][c].real(); Ux[2 * (h * fs + c) + 1] = D.Ux[h][c].imag();
      Uy[2 * (h * fs + c)] = D.Uy[h][c].real(); Uy[2 * (h * fs + c) + 1] = D.Uy[h][c].imag();
    }
  return 0;
}

// fieldShareZ.cpp:39-78 on one rank: nothing is exchanged.
int ref_transfer_zplus(int nh, int N, int from_i, int to_i, int n_slots, double *f) {
  ensure_mpi();
  std::vector<std::vector<cplx>> F(nh, std::vector<cplx>((size_t)n_slots * N));
  for (int h = 0; h < nh; ++h)
    for (size_t c = 0; c < (size_t)n_slots * N; ++c)
      F[h][c] = cplx(f[2 * (h * (size_t)n_slots * N + c)], f[2 * (h * (size_t)n_slots * N + c) + 1]);
  MPI_Transfer1F_Zplus(F, nh, N, from_i, to_i);
  for (int h = 0; h < nh; ++h)
    for (size_t c = 0; c < (size_t)n_slots * N; ++c) {
      f[2 * (h * (size_t)n_slots * N + c)] = F[h][c].real();
      f[2 * (h * (size_t)n_slots * N + c) + 1] = F[h][c].imag();
    }
  return 0;
}

// findparam.cpp:30-116
int ref_find_parameters(const char *block, int rank, const char *key, const char *file, char *out) {
  return FindParameters(block, rank, key, file, out);
}

}  // extern "C"
