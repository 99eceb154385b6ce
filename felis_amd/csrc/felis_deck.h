// Host-side deck state: the Domain scalars, beam, undulator, quad, seed and
// wake blocks that parameterSetting.cpp:42-569 derives from an input deck.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "felis.h"

namespace felis {

struct BeamBlock {       // LoadList, particle.h:49-77
  int type = 0;          // 1 Polygon, 2 Gaussian
  int num_in_beamlet = 0, num_beamlet = 0;
  bool noise = false, trans_flat = false;
  double energy = 0, spread = 0, gamma0 = 1, peak_current = 0;
  double emit_x = 0, emit_y = 0, alpha_x = 0, alpha_y = 0, beta_x = 1, beta_y = 1;
  std::vector<double> zpoint, zn, Epoint, En, ESpoint, ESn, EmitPoint, EmitN;
  double sig_z = 0, pos_z = 0, e_chirp = 0;
  int gauss_power = 2;
  int min_beamlets = 0;  // extension: floor on beamlets per slice (0 = reference behaviour)
};

struct UndUnit {         // one expanded unit of an [Undulator] block
  double unit_start, unit_end, und_start, und_end, K0, slopeK, ue, lambdaU;
  int K0_alpha, und_type;
  bool air;
  int block;
};

struct QuadItem {
  double q_start, q_end, g;
  int list;
};

}  // namespace felis

struct felis_deck {
  std::string path;
  int dimension = 3, mode = 2;
  double photon_energy = 0, ks = 0, lambda0 = 0;
  double Lz = 0;
  int save_step = 1, save_start = 0;
  bool field_save = true, particle_save = true;
  double minX = 0, maxX = 0, minY = 0, maxY = 0, minZ = 0, maxZ = 0;
  int nx = 0, ny = 0;
  double dx = 0, dy = 0;
  int abc_n = 10;
  double abc_sig = 1;
  std::vector<int> harmony;
  int num_lambda_u = 1, num_slice = 1;
  std::vector<felis::BeamBlock> beams;
  double gamma0 = 1;
  std::vector<felis::UndUnit> units;
  std::vector<felis::QuadItem> quads;
  double K0 = 0, ue = 0, K0_alpha = 1, lambdaU = 0, ku = 0, gamR = 0;
  int und_type = 0;
  int sc_f = 1, sc_l = 1, nr = 0;
  double dr = 0;
  int bn = 2001;
  double P0 = 0, spot_sig_r = 0, duration = 0, focus = 0, laser_alpha = -1, a0 = 0, zR = 0, laser_psi = 0;
  int load_h = 1;
  double dz = 0;
  int max_step = 0, slice_n = 0;
  bool wake_on = true;
  int wake_step = 0, wake_shape = 1, wake_acdc = 1;
  double radius = 0, cond = 3.03e7, ctau = 2.4e-6;
  std::string err;
};
