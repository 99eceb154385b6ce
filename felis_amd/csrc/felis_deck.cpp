// Input deck: findparam.cpp:30-116 lookup semantics, the derived Domain
// scalars of parameterSetting.cpp:42-569, the lattice of
// updateK_quadG.cpp:8-94 and the wake tables of wakeField.cpp:53-159.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "felis_deck.h"
#include "felis_host.h"

using felis::BeamBlock;
using felis::QuadItem;
using felis::UndUnit;

namespace {

constexpr double kMc2 = 0.510998918;
constexpr double kC = 299792458.0;
constexpr double kECharge = 1.60217662e-19;
constexpr double kEMass = 9.10938291e-31;
constexpr double kZ0 = 376.730313668;
constexpr double kPlankH = 4.135667696E-15;

// Lines of a deck with everything from "//" removed (findparam.cpp:6-11).
std::vector<std::string> read_lines(const char *path, bool *ok) {
  std::vector<std::string> out;
  std::ifstream f(path);
  *ok = f.good();
  std::string line;
  // like fgets, a line keeps its terminator unless a "//" comment cut it
  while (std::getline(f, line)) {
    size_t c = line.find("//");
    if (c != std::string::npos) line.resize(c);
    else if (!f.eof()) line.push_back('\n');
    out.push_back(line);
  }
  return out;
}

// findparam.cpp:30-116: the rank-th [block]; the first line of it whose first
// word is key followed by ' ' or '='; the value is the first word after '='.
// Returns 1 when found, 0 when not, -1 on a '[' without ']'.
int find_param(const std::vector<std::string> &lines, const char *block, int rank, const char *key,
               std::string *val) {
  int count = 0;
  for (size_t ln = 0; ln < lines.size(); ++ln) {
    const std::string &s = lines[ln];
    size_t p = 0;
    while (p < s.size() && s[p] == ' ') ++p;
    bool found = false;
    if (p < s.size() && s[p] == '[') {
      if (s.find(']', p) == std::string::npos) return -1;
      ++p;
      while (p < s.size() && s[p] == ' ') ++p;
      size_t b = 0, blen = std::strlen(block);
      while (b < blen && p < s.size() && s[p] == block[b]) {
        ++p;
        ++b;
      }
      if (b == blen) {
        while (p < s.size() && s[p] == ' ') ++p;
        if (p < s.size() && s[p] == ']') found = true;
      }
    }
    if (found) ++count;
    if (count == rank && found) {
      for (size_t k = ln + 1; k < lines.size(); ++k) {
        const std::string &t = lines[k];
        if (t.find('[') != std::string::npos) return 0;
        size_t q = 0;
        while (q < t.size() && t[q] == ' ') ++q;
        size_t kl = std::strlen(key), m = 0;
        while (m < kl && q < t.size() && t[q] == key[m]) {
          ++q;
          ++m;
        }
        if (m == kl && q < t.size() && (t[q] == ' ' || t[q] == '=')) {
          size_t eq = t.find('=', q);
          if (eq == std::string::npos) continue;
          size_t v = eq + 1;
          while (v < t.size() && t[v] == ' ') ++v;
          size_t e = v;
          while (e < t.size() && t[e] != ' ') ++e;
          *val = t.substr(v, e - v);
          return 1;
        }
      }
      return 0;
    }
  }
  return 0;
}

struct Parser {
  const std::vector<std::string> &lines;
  bool fail = false;
  std::string why;
  bool get(const char *blk, int rank, const char *key, std::string *v) {
    int r = find_param(lines, blk, rank, key, v);
    if (r < 0) {
      fail = true;
      why = "']' is missing";
    }
    return r == 1;
  }
  double num(const char *blk, int rank, const char *key, double dflt, bool required) {
    std::string v;
    if (get(blk, rank, key, &v)) return std::atof(v.c_str());
    if (required) {
      fail = true;
      why = std::string("[") + blk + "] " + key + " is required";
    }
    return dflt;
  }
  int inum(const char *blk, int rank, const char *key, int dflt, bool required) {
    std::string v;
    if (get(blk, rank, key, &v)) return std::atoi(v.c_str());
    if (required) {
      fail = true;
      why = std::string("[") + blk + "] " + key + " is required";
    }
    return dflt;
  }
  bool has(const char *blk, int rank, const char *key, std::string *v) { return get(blk, rank, key, v); }
};

bool on_off(const std::string &s) { return s.find("ON") != std::string::npos; }

}  // namespace

int felis_deck_parse(const char *path, felis_deck *D) {
  bool ok = false;
  std::vector<std::string> lines = read_lines(path, &ok);
  if (!ok) {
    D->err = std::string("cannot open ") + path;
    return FELIS_ERR_IO;
  }
  D->path = path;
  Parser P{lines};
  std::string s;
  D->dimension = P.inum("Domain", 1, "dimension", 3, true);
  if (P.has("Domain", 1, "mode", &s)) {
    if (s.find("Static") != std::string::npos) D->mode = 1;
    else if (s.find("Time_Dependent") != std::string::npos) D->mode = 2;
    else if (s.find("Twiss") != std::string::npos) D->mode = 3;
    else D->mode = 0;
  } else {
    P.fail = true;
    P.why = "[Domain] mode is required";
  }
  D->photon_energy = P.num("Domain", 1, "photon_energy", 0.0, true);
  D->ks = D->photon_energy * 2 * M_PI / (kPlankH * kC);
  D->lambda0 = 2 * M_PI / D->ks;
  D->Lz = P.num("Save", 1, "total_length", 0.0, true);
  D->save_step = P.inum("Save", 1, "save_step", 1, true);
  D->save_start = P.inum("Save", 1, "save_start", 0, true);
  D->field_save = P.has("Save", 1, "field_save", &s) ? on_off(s) : true;
  D->particle_save = P.has("Save", 1, "particle_save", &s) ? on_off(s) : true;
  D->minX = P.num("Domain", 1, "minX", 0, true) * 1e-6;
  D->maxX = P.num("Domain", 1, "maxX", 0, true) * 1e-6;
  D->minY = P.num("Domain", 1, "minY", 0, true) * 1e-6;
  D->maxY = P.num("Domain", 1, "maxY", 0, true) * 1e-6;
  D->minZ = P.num("Domain", 1, "minZ", 0, true) * 1e-6;
  D->maxZ = P.num("Domain", 1, "maxZ", 0, true) * 1e-6;
  D->nx = P.inum("Domain", 1, "nx", 0, true);
  D->ny = P.inum("Domain", 1, "ny", 0, true);
  D->dx = (D->maxX - D->minX) / (D->nx * 1.0);
  D->dy = (D->maxY - D->minY) / (D->ny * 1.0);
  D->nx += 1;
  D->ny += 1;
  if (D->dimension == 1) {
    D->nx = 1; D->ny = 1; D->dx = 0; D->dy = 0;
  }
  D->abc_n = P.inum("Domain", 1, "ABC_N", 10, false);
  D->abc_sig = P.num("Domain", 1, "ABC_coef", 1.0, false);
  int nh = P.inum("Domain", 1, "num_harmony", 1, false);
  D->harmony.assign(nh, 0);
  for (int i = 0; i < nh; ++i) {
    std::string key = "harmony" + std::to_string(i);
    D->harmony[i] = P.inum("Domain", 1, key.c_str(), 0, true);
  }
  D->num_lambda_u = P.inum("Domain", 1, "lambdaUs_in_iteration", 1, false);
  D->num_slice = P.inum("Domain", 1, "slices_in_bucket", 1, true);
  if (D->num_slice < D->num_lambda_u) {
    P.fail = true;
    P.why = "slices_in_bucket must be >= lambdaUs_in_iteration";
  }
  // [EBeam] blocks (parameterSetting.cpp:431-569); most keys come from block 1
  for (int rank = 1;; ++rank) {
    std::string t;
    BeamBlock B;
    if (!P.has("EBeam", rank, "load_type", &t)) break;
    if (t.find("Polygon") != std::string::npos) B.type = 1;
    else if (t.find("Gaussian") != std::string::npos) B.type = 2;
    else break;
    B.energy = P.num("EBeam", 1, "beam_energy", 0, false);
    B.spread = P.num("EBeam", 1, "energy_spread", 0, false) * 0.01;
    B.peak_current = P.num("EBeam", 1, "peak_current", 0, false);
    B.beta_x = P.num("EBeam", 1, "beta_x", 0, false);
    B.beta_y = P.num("EBeam", 1, "beta_y", 0, false);
    B.alpha_x = P.num("EBeam", 1, "alpha_x", 0, false);
    B.alpha_y = P.num("EBeam", 1, "alpha_y", 0, false);
    B.trans_flat = P.has("EBeam", rank, "transverse_flat", &t) ? on_off(t) : false;
    B.emit_x = P.num("EBeam", 1, "norm_emittance_x", 0, false) * 1e-6;
    B.emit_y = P.num("EBeam", 1, "norm_emittance_y", 0, false) * 1e-6;
    B.noise = P.has("EBeam", 1, "noise_ONOFF", &t) ? on_off(t) : false;
    B.num_beamlet = P.inum("EBeam", 1, "beamlets_in_bucket", 0, false);
    B.num_in_beamlet = P.inum("EBeam", 1, "number_in_beamlet", 0, false);
    B.min_beamlets = P.inum("EBeam", rank, "min_beamlets", 0, false);
    auto nodes = [&](const char *cnt, const char *zk, const char *nk, std::vector<double> &zp,
                     std::vector<double> &zn) {
      int n = P.inum("EBeam", rank, cnt, 0, false);
      for (int i = 0; i < n; ++i) {
        std::string a = std::string(zk) + std::to_string(i), b = std::string(nk) + std::to_string(i);
        zp.push_back(P.num("EBeam", rank, a.c_str(), 0, false));
        zn.push_back(P.num("EBeam", rank, b.c_str(), 0, false));
      }
    };
    if (B.type == 1) {
      nodes("z_nodes", "z", "z_n", B.zpoint, B.zn);
      nodes("energy_nodes", "energy_z", "energy_n", B.Epoint, B.En);
      nodes("energySpread_nodes", "energySpread_z", "energySpread_n", B.ESpoint, B.ESn);
      nodes("emit_nodes", "emit_z", "emit_n", B.EmitPoint, B.EmitN);
    } else {
      B.sig_z = P.num("EBeam", rank, "sigma_z", 0, false);
      B.gauss_power = (int)P.num("EBeam", rank, "gaussian_power", 2, false);
      B.pos_z = P.num("EBeam", rank, "position_z", 0, false);
      B.e_chirp = P.num("EBeam", rank, "energy_chirp", 0, false);
    }
    B.gamma0 = B.energy / kMc2 + 1.0;
    D->beams.push_back(B);
  }
  if (D->beams.empty()) {
    D->err = "no [EBeam] block with a load_type";
    return FELIS_ERR_IO;
  }
  D->gamma0 = D->beams[0].gamma0;
  // [Undulator] blocks (parameterSetting.cpp:345-429)
  int blk = 0;
  for (int rank = 1;; ++rank) {
    std::string t;
    int type = 0;
    if (P.has("Undulator", rank, "undulator_type", &t)) {
      if (t.find("BiPolar") != std::string::npos) type = 1;
      else if (t.find("QuadPolar") != std::string::npos) type = 2;
    }
    if (!type) break;
    int numbers = P.inum("Undulator", rank, "numbers", 0, false);
    bool air = P.has("Undulator", rank, "in_air", &t) ? on_off(t) : false;
    double lambdaU = P.num("Undulator", rank, "lambdaU", 0, false) * 0.01;
    if (numbers > 0) {
      double us = P.num("Undulator", rank, "unit_start", 0, false);
      double ue_ = P.num("Undulator", rank, "unit_end", 0, false);
      double ds = P.num("Undulator", rank, "undulator_start", 0, false);
      double de = P.num("Undulator", rank, "undulator_end", 0, false);
      double lin = P.num("Undulator", rank, "linear_taper", 0, false) / std::sqrt(2.0);
      double quad = P.num("Undulator", rank, "quad_taper", 0, false) / std::sqrt(2.0);
      double slope = P.num("Undulator", rank, "slope_K", 0, false) / std::sqrt(2.0);
      int qid = P.inum("Undulator", rank, "quad_taper_start_index", 30000, false);
      double pol = P.num("Undulator", rank, "polarity", 0, false);
      int ka = P.inum("Undulator", rank, "K0_alpha", 1, false);
      double ku = 2 * M_PI / lambdaU;
      double K0 = P.has("Undulator", rank, "K0", &t)
                      ? std::atof(t.c_str())
                      : std::sqrt((4 * D->gamma0 * D->gamma0 * ku / D->ks - 2) / (1.0 + pol * pol));
      double unitLength = ue_ - us, undStart = ds - us, undLength = de - ds;
      UndUnit u0{us, ue_, ds, de, K0, slope, pol, lambdaU, ka, type, air, blk};
      D->units.push_back(u0);
      for (int i = 1; i < numbers; ++i) {
        UndUnit u = D->units.back();
        u.unit_start = u.unit_end;
        u.unit_end = u.unit_start + unitLength;
        u.und_start = u.unit_start + undStart;
        u.und_end = u.und_start + undLength;
        u.K0 = (i >= qid) ? K0 + i * lin + (i - qid) * (i - qid) * quad : K0 + i * lin;
        D->units.push_back(u);
      }
    } else {
      // a block with numbers=0 still sets lambdaU for block 1's derived values
      UndUnit u0{0, 0, 0, 0, 0, 0, P.num("Undulator", rank, "polarity", 0, false), lambdaU,
                 P.inum("Undulator", rank, "K0_alpha", 1, false), type, air, blk};
      if (blk == 0) D->units.push_back(u0);
    }
    ++blk;
  }
  if (D->units.empty()) {
    D->err = "no [Undulator] block with an undulator_type";
    return FELIS_ERR_IO;
  }
  const UndUnit &U0 = D->units[0];
  D->K0 = U0.K0;
  D->ue = U0.ue;
  D->K0_alpha = U0.K0_alpha;
  D->und_type = U0.und_type;
  D->lambdaU = U0.lambdaU;
  D->ku = 2.0 * M_PI / U0.lambdaU;
  D->gamR = std::sqrt(0.5 * D->ks / D->ku * (1 + 0.5 * (1 + D->ue * D->ue) * D->K0 * D->K0));
  // [Quad] blocks (parameterSetting.cpp:287-338)
  for (int rank = 1;; ++rank) {
    int numbers = P.inum("Quad", rank, "numbers", 0, false);
    if (numbers <= 0) break;
    double us = P.num("Quad", rank, "unit_start", 0, false);
    double ue_ = P.num("Quad", rank, "unit_end", 0, false);
    double qs = P.num("Quad", rank, "quad_start", 0, false);
    double qe = P.num("Quad", rank, "quad_end", 0, false);
    double g = P.num("Quad", rank, "g", 0, false);
    double unitLength = ue_ - us, qStart = qs - us, qLength = qe - qs;
    double ustart = us;
    double qsi = qs, qei = qe;
    for (int i = 0; i < numbers; ++i) {
      if (i > 0) {
        ustart = ustart + unitLength;
        qsi = ustart + qStart;
        qei = qsi + qLength;
      }
      D->quads.push_back(QuadItem{qsi, qei, g, rank});
    }
  }
  // space charge, Bessel table, seed (parameterSetting.cpp:188-229)
  D->sc_f = P.inum("Domain", 1, "number_azimuthal_mode", 1, false);
  D->sc_l = P.inum("Domain", 1, "number_longitudinal_mode", 1, false);
  D->nr = P.has("Domain", 1, "radial_grids", &s) ? std::atoi(s.c_str())
                                                 : (int)std::sqrt((double)(D->nx * D->nx + D->ny * D->ny));
  D->dr = std::sqrt((D->maxX - D->minX) * (D->maxX - D->minX) + (D->maxY - D->minY) * (D->maxY - D->minY)) /
          (1.0 * D->nx);
  if (D->dimension == 1) {
    D->nr = 1; D->dr = 0; D->sc_f = 1;
  }
  D->bn = P.inum("Bessel_table", 1, "num_grids", 2001, false);
  D->P0 = P.num("Seed", 1, "power", 0, true);
  D->spot_sig_r = P.num("Seed", 1, "spot_sigma_R", 0, true) * 1e-6;
  D->duration = P.num("Seed", 1, "rms_duration", 0, true) * 1e-15;
  D->focus = P.num("Seed", 1, "focus", 0, true);
  D->load_h = P.inum("Seed", 1, "loading_harmonics", 1, false);
  double sue = P.num("Seed", 1, "polarity", 0.0, false);
  D->laser_alpha = P.num("Seed", 1, "laser_alpha", -1.0, false);
  double area = 2.0 * M_PI * D->spot_sig_r * D->spot_sig_r;
  D->a0 = std::sqrt(D->P0 * 2.0 * kZ0 / area) * kECharge / (kEMass * kC * kC * D->ks * D->load_h);
  D->zR = 2.0 * D->spot_sig_r * D->spot_sig_r * D->ks * (1.0 * D->load_h);
  D->laser_psi = std::arg(std::complex<double>(1.0 - sue * sue, 2.0 * sue) / (1.0 + sue * sue));
  if (D->mode == 1) {
    D->minZ = 0.0;
    D->maxZ = D->lambda0 * D->num_slice;
  }
  D->dz = D->lambdaU * D->num_lambda_u;
  D->max_step = (int)(D->Lz / D->dz + 1);
  D->slice_n = (int)((D->maxZ - D->minZ) / (D->lambda0 * D->num_slice));
  if (D->slice_n > 50000 || D->slice_n < -50000) {
    D->err = "too many slices";
    return FELIS_ERR_IO;
  }
  if (D->mode == 1) D->slice_n = 1;
  // [Wake_field] (parameterSetting.cpp:249-267)
  D->wake_on = P.has("Wake_field", 1, "activate", &s) ? on_off(s) : true;
  D->wake_step = P.inum("Wake_field", 1, "update_step", D->max_step, false);
  D->wake_shape = 1;
  if (P.has("Wake_field", 1, "shape", &s))
    D->wake_shape = s.find("Flat") != std::string::npos ? 1 : (s.find("Circular") != std::string::npos ? 2 : 0);
  D->wake_acdc = 1;
  if (P.has("Wake_field", 1, "ac_dc", &s))
    D->wake_acdc = s.find("AC") != std::string::npos ? 1 : (s.find("DC") != std::string::npos ? 2 : 0);
  D->radius = P.num("Wake_field", 1, "radius", 0.0, true);
  D->cond = P.num("Wake_field", 1, "conductivity", 3.03e7, false);
  D->ctau = D->wake_acdc == 1 ? P.num("Wake_field", 1, "ctau", 2.4e-6, false) : 0.0;
  if (P.fail) {
    D->err = P.why;
    return FELIS_ERR_IO;
  }
  return FELIS_OK;
}

// updateK_quadG.cpp:8-94 on a lattice state that carries K0/ue/K0_alpha/
// lambdaU/und_type from the previous call.
void felis_update_k_quadg(const felis_deck *D, felis_lattice *lat, int iteration, double half) {
  bool inUnd = false, inInter = false, air = false;
  double dz = D->dz;
  double z = (iteration + half) * dz;
  double K0 = lat->K0, K0_alpha = lat->K0_alpha, ue = lat->ue, lambdaU = lat->lambdaU;
  int undType = lat->und_type;
  lat->current_flag = 0;
  lat->drift_flag = 0;
  for (const UndUnit &u : D->units) {
    if (u.unit_end <= u.unit_start) continue;
    if (z >= u.unit_start && z < u.unit_end) {
      if (z >= u.und_start && z < u.und_end) {
        inUnd = true;
        K0 = u.K0 * (1 + u.slopeK * (z - u.und_start));
        K0_alpha = u.K0_alpha;
        ue = u.ue;
        undType = u.und_type;
        lambdaU = u.lambdaU;
      } else {
        inInter = true;
        if (u.air) air = true;
      }
    }
  }
  lat->K0 = K0;
  lat->ue = ue;
  lat->K0_alpha = K0_alpha;
  lat->lambdaU = lambdaU;
  lat->ku = 2 * M_PI / lambdaU;
  lat->und_type = undType;
  if (K0 == 0.0) air = true;
  if (!inUnd && !inInter) air = true;
  if (inUnd) lat->current_flag = 1;
  if (air) {
    lat->ku = 0.0;
    lat->drift_flag = 1;
    lat->current_flag = 0;
  }
  double g = 0, x0 = z, x1 = z + dz * 0.5;
  int done_list = -1;
  for (const QuadItem &q : D->quads) {
    if (q.list == done_list) continue;
    double q0 = q.q_start, q1 = q.q_end;
    if (q1 <= x0 || q0 >= x1) continue;
    if (x0 < q0 && q1 < x1) g = (q1 - q0) / dz * 2.0 * q.g;
    else if (x0 < q0 && q0 < x1) g = (x1 - q0) / dz * 2.0 * q.g;
    else if (x0 < q1 && q1 < x1) g = (q1 - x0) / dz * 2.0 * q.g;
    else g = 1.0 * q.g;
    done_list = q.list;
  }
  lat->g = g;
}

felis_lattice felis_initial_lattice(const felis_deck *D) {
  felis_lattice l{};
  l.K0 = D->K0;
  l.ue = D->ue;
  l.K0_alpha = D->K0_alpha;
  l.ku = D->ku;
  l.g = 0.0;  // Domain D{} zero-init (main.cpp:32)
  l.lambdaU = D->lambdaU;
  l.und_type = D->und_type;
  l.current_flag = 0;
  l.drift_flag = 0;
  return l;
}

// wakeField.cpp:53-159
int felis_wake_tables(const felis_deck *D, double *wakeF, double *impZ_out) {
  const int sn = D->slice_n;
  double dZ = D->num_slice * D->lambda0;
  double s0 = std::pow(2.0 * D->radius * D->radius / kZ0 / D->cond, 1.0 / 3.0);
  const double maxK = 20, maxX = 2000;
  const int numK = 1000, numX = 50000;
  double dK = maxK / (numK * 1.0), dX = maxX / (numX * 1.0);
  std::vector<double> impZ(numK + 1, 0.0);
  using Cx = std::complex<double>;
  const Cx I(0, 1);
  int err = FELIS_OK;
  if (D->wake_shape == 1) {
    for (int i = 1; i <= numK; ++i) {
      double k = i * dK;
      Cx csum = 0;
      Cx root = std::sqrt(Cx(k, -k * k * D->ctau / s0));
      for (int j = 0; j <= numX; ++j) {
        double x = j * dX;
        Cx ccal;
        if (x == 0.0) {
          ccal = std::sqrt(k) / (2.0 + k * k * k - 2.0 * std::sqrt(k) * k);
        } else if (x > 320) {
          ccal = 0;
        } else {
          Cx cd = 2.0 / (1.0 - I) / root * std::cosh(x) - I * k * std::sinh(x) / x;
          ccal = 1.0 / std::cosh(x) / cd;
          if (std::isnan(ccal.real())) err = FELIS_ERR_NAN;
        }
        csum += ccal * dX;
      }
      impZ[i] = csum.real();
    }
  } else if (D->wake_shape == 2) {
    for (int i = 1; i <= numK; ++i) {
      double k = i * dK;
      if (D->wake_acdc == 1) {
        Cx cd = std::sqrt(2.0 * s0 / D->ctau) / k * (I + s0 * 0.5 / k / D->ctau) - I * 0.5 * k;
        impZ[i] = (1.0 / cd).real();
      } else {
        double dino = 2.0 / k + 0.25 * k * k - std::sqrt(k);
        impZ[i] = 1.0 / std::sqrt(k) / dino;
      }
    }
  }
  if (impZ_out)
    for (int i = 0; i <= numK; ++i) impZ_out[i] = impZ[i];
  double coef = kZ0 * 3e8 / M_PI / D->radius / D->radius / M_PI;
  for (int i = 0; i < sn; ++i) {
    double z = (i * dZ) / s0, sum = 0.0;
    for (int j = 0; j <= numK; ++j) sum += impZ[j] * std::cos(j * dK * z) * dK;
    wakeF[i] = 0.0 + sum * coef;
  }
  return err;
}

// ------------------------------------------------------------------ C-ABI

extern "C" int felis_find_parameter(const char *block, int rank, const char *key, const char *file, char *out,
                                    int cap) {
  if (!block || !key || !file || !out || cap <= 0) return FELIS_ERR_ARG;
  bool ok = false;
  std::vector<std::string> lines = read_lines(file, &ok);
  if (!ok) return FELIS_ERR_IO;
  std::string v;
  int r = find_param(lines, block, rank, key, &v);
  if (r < 0) return FELIS_ERR_IO;
  if (r == 0) return 0;
  std::snprintf(out, (size_t)cap, "%s", v.c_str());
  return 1;
}

extern "C" int felis_deck_open(const char *path, felis_deck **out) {
  if (!path || !out) return FELIS_ERR_ARG;
  *out = nullptr;
  felis_deck *D = new felis_deck();
  int e = felis_deck_parse(path, D);
  *out = D;  // handed back even on error so felis_deck_error can explain
  return e;
}

extern "C" int felis_deck_close(felis_deck *deck) {
  delete deck;
  return FELIS_OK;
}

extern "C" const char *felis_deck_error(const felis_deck *deck) { return deck ? deck->err.c_str() : "null deck"; }

extern "C" int felis_deck_desc(const felis_deck *D, int precision, felis_desc *d) {
  if (!D || !d) return FELIS_ERR_ARG;
  std::memset(d, 0, sizeof(*d));
  d->dimension = D->dimension;
  d->mode = D->mode;
  d->precision = precision;
  d->nx = D->nx;
  d->ny = D->ny;
  d->slice_n = D->slice_n;
  d->n_harm = (int)D->harmony.size();
  if (d->n_harm > FELIS_MAX_HARM) return FELIS_ERR_UNSUPPORTED;
  for (int h = 0; h < d->n_harm; ++h) d->harmony[h] = D->harmony[h];
  d->num_slice = D->num_slice;
  d->n_species = (int)D->beams.size();
  d->nr = D->nr;
  d->sc_l = D->sc_l;
  d->sc_f = D->sc_f;
  d->bessel_n = D->bn;
  d->abc_n = D->abc_n;
  d->wake_on = D->wake_on ? 1 : 0;
  d->abc_sig = D->abc_sig;
  d->dx = D->dx;
  d->dy = D->dy;
  d->dz = D->dz;
  d->min_x = D->minX;
  d->min_y = D->minY;
  d->min_z = D->minZ;
  d->max_z = D->maxZ;
  d->ks = D->ks;
  d->lambda0 = D->lambda0;
  d->gamR = D->gamR;
  d->dr = D->dr;
  d->gamma_ref = D->gamma0;
  return FELIS_OK;
}

extern "C" int felis_deck_save_settings(const felis_deck *D, double *gamma0, double *spot_sig_r) {
  if (!D) return FELIS_ERR_ARG;
  if (gamma0)
    for (size_t s = 0; s < D->beams.size(); ++s) gamma0[s] = D->beams[s].gamma0;
  if (spot_sig_r) *spot_sig_r = D->spot_sig_r;
  return FELIS_OK;
}

extern "C" int felis_deck_steps(const felis_deck *D, int *info) {
  if (!D || !info) return FELIS_ERR_ARG;
  info[0] = D->max_step;
  info[1] = D->wake_step;
  info[2] = D->save_step;
  info[3] = D->save_start;
  info[4] = D->field_save ? 1 : 0;
  info[5] = D->particle_save ? 1 : 0;
  info[6] = D->mode;
  info[7] = D->slice_n;
  return FELIS_OK;
}

extern "C" int felis_deck_lattice(const felis_deck *D, int it0, int n_steps, felis_lattice *tab) {
  if (!D || !tab || n_steps < 0) return FELIS_ERR_ARG;
  felis_lattice l = felis_initial_lattice(D);
  tab[0] = l;
  for (int k = 0; k < n_steps; ++k) {
    felis_update_k_quadg(D, &l, it0 + k, 0.0);
    tab[2 * k + 1] = l;
    felis_update_k_quadg(D, &l, it0 + k, 0.5);
    tab[2 * k + 2] = l;
  }
  return FELIS_OK;
}

extern "C" int felis_deck_wake_function(const felis_deck *D, double *wakeF, double *impZ) {
  if (!D || !wakeF) return FELIS_ERR_ARG;
  return felis_wake_tables(D, wakeF, impZ);
}
