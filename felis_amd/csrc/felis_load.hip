// Device-side initial state: the quiet-start beam of loadBeam.cpp:28-269 and
// the Gaussian seed of loadSeed.cpp:25-80, generated in HBM.
//
// Quasi-random numbers: a 7-dimensional Halton sequence (bases 2..17), point
// k+1 for the k-th beamlet of a species, as the reference's gsl_qrng_halton
// stream (loadBeam.cpp:75) is consumed one point per beamlet. The Gaussian
// shot-noise amplitudes (loadBeam.cpp:225-229, Fawley's sigma) come from a
// counter-based generator keyed by (seed, species, beamlet, harmonic), so any
// beamlet can be generated by any thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "felis_ctx.h"
#include "felis_host.h"

using namespace felis;

namespace {

struct SliceLoad {
  double gamma0, dGam, ESn0, sigX, sigY, sigXp, sigYp, delTX, delTY, sigNoise;
  long long b0;    // first beamlet index of this slice within the species
  long long out0;  // first particle index
  int nbeam, pad;
};

__device__ inline double radical_inverse(unsigned long long k, unsigned base) {
  const double inv = 1.0 / base;
  double f = inv, r = 0.0;
  while (k) {
    r += (double)(k % base) * f;
    k /= base;
    f *= inv;
  }
  return r;
}

__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// standard normal from two counter-based uniforms (Box-Muller)
__device__ inline double normal01(unsigned long long key) {
  unsigned long long a = mix64(key), b = mix64(key ^ 0xD1B54A32D192ED03ull);
  double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740993.0);
  double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

template <typename R>
__global__ void k_load_beam(const SliceLoad *__restrict__ sl, int nslice, long long nbeam_total, int nib,
                            int maxH, double dPhi, int trans_flat, double noise_on, unsigned long long seed,
                            int species, double gref, double *__restrict__ theta, R *__restrict__ gamma,
                            R *__restrict__ x, R *__restrict__ y, R *__restrict__ px, R *__restrict__ py) {
  long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbeam_total) return;
  int lo = 0, hi = nslice - 1;
  while (lo < hi) {  // last slice with b0 <= b
    int mid = (lo + hi + 1) >> 1;
    if (sl[mid].b0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const SliceLoad s = sl[lo];
  const unsigned primes[7] = {2, 3, 5, 7, 11, 13, 17};
  double v[7];
  unsigned long long k = (unsigned long long)b + 1ull;
  for (int d = 0; d < 7; ++d) v[d] = radical_inverse(k, primes[d]);
  double theta0 = v[4] * dPhi;
  double r1 = v[0] == 0.0 ? 1e-10 : v[0];
  double r2 = v[1];
  double pr1 = v[2] == 0.0 ? 1e-10 : v[2];
  double pr2 = v[3];
  double gu = v[5] == 0.0 ? 1e-10 : v[5];
  double xx, yy, xp, yp;
  if (!trans_flat) {
    double c = sqrt(-2.0 * log(r1));
    xx = c * cos(2.0 * M_PI * r2) * s.sigX;
    yy = c * sin(2.0 * M_PI * r2) * s.sigY;
  } else {
    double c = sqrt(r1);
    xx = c * cos(2.0 * M_PI * r2) * s.sigX;
    yy = c * sin(2.0 * M_PI * r2) * s.sigY;
  }
  double c2 = sqrt(-2.0 * log(pr1));
  xp = c2 * cos(2.0 * M_PI * pr2) * s.sigXp;
  yp = c2 * sin(2.0 * M_PI * pr2) * s.sigYp;
  double tmp = sqrt(-2.0 * log(gu)) * cos(v[6] * 2.0 * M_PI);
  double gam = s.gamma0 + s.dGam * tmp * s.ESn0;
  double pz = sqrt((gam * gam - 1.0) / (1.0 + xp * xp + yp * yp));
  double ppx = xp * pz, ppy = yp * pz;
  xx -= s.delTX * ppx / gam;
  yy -= s.delTY * ppy / gam;
  double an[16], bn[16];
  for (int m = 1; m <= maxH && m < 16; ++m) {
    double sig = s.sigNoise / m;
    unsigned long long key = seed ^ ((unsigned long long)species << 58) ^ ((unsigned long long)b << 6) ^ (unsigned)m;
    an[m] = sig * normal01(key * 2ull);
    bn[m] = sig * normal01(key * 2ull + 1ull);
  }
  const double div = 2.0 * M_PI / (1.0 * nib);
  long long o = s.out0 + (b - s.b0) * nib;
  for (int n = 0; n < nib; ++n) {
    double th = theta0 + n * div;
    double noise = 0.0;
    for (int m = 1; m <= maxH && m < 16; ++m) noise += an[m] * cos(m * th) + bn[m] * sin(m * th);
    theta[o + n] = th + noise * noise_on;
    gamma[o + n] = (R)(gam - gref);
    x[o + n] = (R)xx;
    y[o + n] = (R)yy;
    px[o + n] = (R)ppx;
    py[o + n] = (R)ppy;
  }
}

struct SeedSlice {
  double ampZ, delZ, curv, gouy, w;
};

template <typename T>
__global__ void k_load_seed(GridArgs g, const SeedSlice *__restrict__ ss, int h_pick, double ks, double sigR,
                            double factor, double la, double psi, cpx<T> *__restrict__ Ux,
                            cpx<T> *__restrict__ Uy) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int sl = blockIdx.y;
  if (c >= g.N) return;
  int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const SeedSlice s = ss[sl];
  double xx = g.minX + i * g.dx, yy = g.minY + j * g.dy;
  double r2 = xx * xx + yy * yy;
  double phase = -r2 / (2.0 * s.w * s.w);
  double cph = ks * s.delZ + ks * r2 / (2.0 * s.curv) - s.gouy;
  double mag = factor * (s.ampZ * sigR / s.w) * exp(phase);
  double sn, cs;
  sincos(cph, &sn, &cs);
  cpx<double> amp = mk(mag * cs, mag * sn);
  double sp, cp;
  sincos(psi, &sp, &cp);
  cpx<double> ex = mk(la * cp, la * sp);
  cpx<double> ax = amp * (mk(1.0, 0.0) + ex);
  cpx<double> ay = amp * (mk(-1.0, 0.0) + ex);
  long long off = ((long long)h_pick * g.S + phys_slot(sl + 1, g.ring_base, g.S)) * g.N + c;
  Ux[off] = cvt<T>(ax);
  Uy[off] = cvt<T>(ay);
}

// Per-slice beam quantities of loadBeam.cpp:94-157 for one species.
struct SpeciesLayout {
  std::vector<SliceLoad> slices;
  std::vector<long long> count;
  std::vector<double> weight;
  long long nbeam = 0;
};

SpeciesLayout species_layout(const felis_deck *D, const BeamBlock &LL) {
  SpeciesLayout out;
  const double bucketZ = D->lambda0 * D->num_slice;
  const int ptclCnt = LL.num_in_beamlet * LL.num_beamlet;
  const double macro = ptclCnt > 0 ? LL.peak_current / 1.60217662e-19 / 299792458.0 * bucketZ / (double)ptclCnt : 0.0;
  long long b0 = 0;
  for (int sl = 0; sl < D->slice_n; ++sl) {
    double n0 = 0, En0 = 0, ESn0 = 0;
    double posZ = sl * bucketZ + D->minZ;
    double gamma0 = LL.energy / 0.510998918 + 1;
    if (LL.type == 1) {
      auto lin = [&](const std::vector<double> &zp, const std::vector<double> &zn, double &v) {
        for (size_t l = 0; l + 1 < zp.size(); ++l)
          if (posZ >= zp[l] && posZ < zp[l + 1]) v = (zn[l + 1] - zn[l]) / (zp[l + 1] - zp[l]) * (posZ - zp[l]) + zn[l];
      };
      lin(LL.zpoint, LL.zn, n0);
      lin(LL.Epoint, LL.En, En0);
      gamma0 = LL.energy * En0 / 0.510998918 + 1.0;
      lin(LL.ESpoint, LL.ESn, ESn0);
    } else if (LL.type == 2) {
      double phase = std::pow((posZ - LL.pos_z) / LL.sig_z, LL.gauss_power);
      n0 = std::exp(-phase);
      gamma0 = (LL.energy + LL.e_chirp * (posZ - LL.pos_z)) / 0.510998918 + 1.0;
    }
    double dGam = LL.spread * gamma0 * ESn0;
    double emitX = LL.emit_x / gamma0, emitY = LL.emit_y / gamma0;
    double gammaX = (1 + LL.alpha_x * LL.alpha_x) / LL.beta_x;
    double gammaY = (1 + LL.alpha_y * LL.alpha_y) / LL.beta_y;
    SliceLoad s{};
    s.gamma0 = gamma0;
    s.dGam = dGam;
    s.ESn0 = ESn0;
    s.sigX = std::sqrt(emitX / gammaX);
    s.sigY = std::sqrt(emitY / gammaY);
    s.sigXp = std::sqrt(emitX * gammaX);
    s.sigYp = std::sqrt(emitY * gammaY);
    double distX = std::sqrt(std::fabs((LL.beta_x - 1.0 / gammaX) / gammaX));
    double distY = std::sqrt(std::fabs((LL.beta_y - 1.0 / gammaY) / gammaY));
    double vz = std::sqrt(gamma0 * gamma0 - 1.0) / gamma0;
    s.delTX = vz == 0.0 ? 0.0 : distX / vz;
    s.delTY = vz == 0.0 ? 0.0 : distY / vz;
    int beamlets = (int)(LL.num_beamlet * n0);
    if (LL.min_beamlets > 0 && beamlets < LL.min_beamlets) beamlets = LL.min_beamlets;
    double remacro = beamlets > 0 ? macro * (double)LL.num_beamlet / (double)beamlets * n0 : 0.0;
    double eNumbers = remacro * LL.num_in_beamlet;
    if (eNumbers < 10) eNumbers = 10;
    s.sigNoise = std::sqrt(2.0 / eNumbers);  // /m applied per harmonic
    s.b0 = b0;
    s.nbeam = beamlets;
    b0 += beamlets;
    out.slices.push_back(s);
    out.count.push_back((long long)beamlets * LL.num_in_beamlet);
    out.weight.push_back(remacro);
  }
  out.nbeam = b0;
  return out;
}

}  // namespace

extern "C" int felis_deck_beam_layout(const felis_deck *D, int64_t *seg_off, double *seg_weight) {
  if (!D || !seg_off) return FELIS_ERR_ARG;
  seg_off[0] = 0;
  size_t k = 0;
  for (const BeamBlock &LL : D->beams) {
    SpeciesLayout L = species_layout(D, LL);
    for (int sl = 0; sl < D->slice_n; ++sl, ++k) {
      seg_off[k + 1] = seg_off[k] + L.count[sl];
      if (seg_weight) seg_weight[k] = L.weight[sl];
    }
  }
  return FELIS_OK;
}

extern "C" int felis_deck_beam_sigma(const felis_deck *D, double *sig) {
  if (!D || !sig) return FELIS_ERR_ARG;
  size_t s = 0;
  for (const BeamBlock &LL : D->beams) {
    const double ex = LL.emit_x / D->gamma0, ey = LL.emit_y / D->gamma0;
    const double gx = (1 + LL.alpha_x * LL.alpha_x) / LL.beta_x, gy = (1 + LL.alpha_y * LL.alpha_y) / LL.beta_y;
    sig[2 * s] = std::sqrt(ex / gx);
    sig[2 * s + 1] = std::sqrt(ey / gy);
    ++s;
  }
  return FELIS_OK;
}

extern "C" int felis_deck_load_beam(const felis_deck *D, felis_ctx *ctx, uint64_t seed) {
  if (!D || !ctx) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  if ((int)D->beams.size() != d.n_species || D->slice_n != d.slice_n) {
    ctx->err = "deck and context disagree on species or slices";
    return FELIS_ERR_ARG;
  }
  std::vector<int64_t> seg_off(d.n_species * d.slice_n + 1);
  std::vector<double> seg_w(d.n_species * d.slice_n);
  felis_deck_beam_layout(D, seg_off.data(), seg_w.data());
  int e = ctx->alloc_particles(seg_off.data(), seg_w.data());
  if (e) return e;
  const int maxH = D->harmony.back();
  const double gref = ctx->mixed() ? d.gamma_ref : 0.0;
  SliceLoad *dsl = nullptr;
  if (hipMalloc((void **)&dsl, sizeof(SliceLoad) * d.slice_n) != hipSuccess) {
    ctx->err = "hipMalloc loader table";
    return FELIS_ERR_HIP;
  }
  for (int s = 0; s < d.n_species; ++s) {
    const BeamBlock &LL = D->beams[s];
    SpeciesLayout L = species_layout(D, LL);
    for (int sl = 0; sl < d.slice_n; ++sl) L.slices[sl].out0 = seg_off[(size_t)s * d.slice_n + sl];
    hipError_t he = hipMemcpy(dsl, L.slices.data(), sizeof(SliceLoad) * d.slice_n, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
      (void)hipFree(dsl);
      ctx->err = std::string("loader table: ") + hipGetErrorString(he);
      return FELIS_ERR_HIP;
    }
    if (L.nbeam == 0) continue;
    unsigned nb = (unsigned)((L.nbeam + 127) / 128);
    double noise_on = LL.noise ? 1.0 : 0.0;
    double dPhi = 2.0 * M_PI * D->num_slice;
    if (ctx->mixed())
      hipLaunchKernelGGL((k_load_beam<float>), dim3(nb), dim3(128), 0, ctx->stream, dsl, d.slice_n, L.nbeam,
                         LL.num_in_beamlet, maxH, dPhi, LL.trans_flat ? 1 : 0, noise_on, (unsigned long long)seed, s,
                         gref, ctx->theta, (float *)ctx->gamma, (float *)ctx->x, (float *)ctx->y, (float *)ctx->px,
                         (float *)ctx->py);
    else
      hipLaunchKernelGGL((k_load_beam<double>), dim3(nb), dim3(128), 0, ctx->stream, dsl, d.slice_n, L.nbeam,
                         LL.num_in_beamlet, maxH, dPhi, LL.trans_flat ? 1 : 0, noise_on, (unsigned long long)seed, s,
                         gref, ctx->theta, (double *)ctx->gamma, (double *)ctx->x, (double *)ctx->y,
                         (double *)ctx->px, (double *)ctx->py);
    he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
      (void)hipFree(dsl);
      ctx->err = std::string("beam loader: ") + hipGetErrorString(he);
      return FELIS_ERR_HIP;
    }
  }
  (void)hipFree(dsl);
  return FELIS_OK;
}

extern "C" int felis_deck_load_seed(const felis_deck *D, felis_ctx *ctx) {
  if (!D || !ctx) return FELIS_ERR_ARG;
  const felis_desc &d = ctx->d;
  if (D->dimension != 3) return FELIS_ERR_UNSUPPORTED;
  int h_pick = 0;
  for (int h = 0; h < (int)D->harmony.size(); ++h)
    if (D->harmony[h] == D->load_h) h_pick = h;
  const double bucketZ = D->lambda0 * D->num_slice;
  const double sigmaZ = D->duration * 299792458.0;
  const double invS2 = 1.0 / sigmaZ / sigmaZ;
  const double minZ = D->minZ + D->lambda0 * 0.5;
  const double sigR = std::sqrt(2.0) * D->spot_sig_r;
  const double zR = D->zR;
  std::vector<SeedSlice> ss(d.slice_n);
  for (int sl = 0; sl < d.slice_n; ++sl) {
    double z = sl * bucketZ + minZ;
    double delZ = z - D->focus;
    ss[sl].ampZ = D->a0 * std::exp(-0.25 * z * z * invS2);
    ss[sl].curv = delZ == 0.0 ? 1e100 : delZ * (1.0 + zR * zR / (delZ * delZ));
    ss[sl].gouy = std::atan(delZ / zR);
    ss[sl].w = sigR * std::sqrt(1.0 + delZ * delZ / (zR * zR));
    ss[sl].delZ = delZ;
  }
  SeedSlice *dss = nullptr;
  if (hipMalloc((void **)&dss, sizeof(SeedSlice) * d.slice_n) != hipSuccess) {
    ctx->err = "hipMalloc seed table";
    return FELIS_ERR_HIP;
  }
  hipError_t he = hipMemcpy(dss, ss.data(), sizeof(SeedSlice) * d.slice_n, hipMemcpyHostToDevice);
  // D.polarity is never parsed, so the factor is sqrt(1+0)/2 (loadSeed.cpp:53)
  const double factor = 0.5;
  // Uy = amp (-1 + laser_alpha e^{i laser_psi}) is exactly zero only for
  // laser_alpha = 1, laser_psi = 0
  if (!(D->laser_alpha == 1.0 && D->laser_psi == 0.0)) ctx->uy_live = true;
  GridArgs g = ctx->grid();
  dim3 grid((unsigned)((ctx->N + 255) / 256), (unsigned)d.slice_n);
  if (he == hipSuccess) {
    if (ctx->mixed())
      hipLaunchKernelGGL((k_load_seed<float>), grid, dim3(256), 0, ctx->stream, g, dss, h_pick, D->ks, sigR, factor,
                         D->laser_alpha, D->laser_psi, (cpx<float> *)ctx->U,
                         (cpx<float> *)((char *)ctx->U + ctx->pol_bytes()));
    else
      hipLaunchKernelGGL((k_load_seed<double>), grid, dim3(256), 0, ctx->stream, g, dss, h_pick, D->ks, sigR, factor,
                         D->laser_alpha, D->laser_psi, (cpx<double> *)ctx->U,
                         (cpx<double> *)((char *)ctx->U + ctx->pol_bytes()));
    he = hipStreamSynchronize(ctx->stream);
  }
  (void)hipFree(dss);
  if (he != hipSuccess) {
    ctx->err = std::string("seed loader: ") + hipGetErrorString(he);
    return FELIS_ERR_HIP;
  }
  return FELIS_OK;
}
