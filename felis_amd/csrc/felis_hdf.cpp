// HDF5 dump and restore of a felis context in the reference's file layout
// (include/felis_hdf.h). Host code on serial HDF5: the state moves through
// the C-ABI of libfelis_hip.so (download/upload), one dataset per harmonic,
// polarisation or species, written and read whole.
#include "felis_hdf.h"

#include <hdf5.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

constexpr int kDataCnt = 8;  // columns of a particle row

// HDF5 handles closed on scope exit
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
};

void quiet_hdf5() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // failures are reported through the return codes
    done = true;
  }
}

template <class T> hid_t native();
template <> hid_t native<int>() { return H5T_NATIVE_INT; }
template <> hid_t native<long>() { return H5T_NATIVE_LONG; }
template <> hid_t native<double>() { return H5T_NATIVE_DOUBLE; }

// saveMeta (saveParticleHDF.cpp:175-219): a 1-D dataset of n values
template <class T>
bool write_meta(hid_t f, const char *name, const T *data, hsize_t n) {
  Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  Hid ds(H5Dcreate2(f, name, native<T>(), sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  return ds.ok() && H5Dwrite(ds.id, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

// restoreMeta (restoreFieldHDF.cpp:157-184): n values of a 1-D dataset
template <class T>
int read_meta(hid_t f, const char *file, const char *name, T *data, hsize_t n) {
  Hid ds(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(FELIS_ERR_IO, std::string(file) + ": no dataset '" + name + "'");
  Hid sp(H5Dget_space(ds.id), H5Sclose);
  if (!sp.ok() || H5Sget_simple_extent_npoints(sp.id) != (hssize_t)n)
    return fail(FELIS_ERR_IO, std::string(file) + ": dataset '" + name + "' does not hold " + std::to_string(n) +
                                  " values");
  if (H5Dread(ds.id, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(FELIS_ERR_IO, std::string(file) + ": cannot read '" + name + "'");
  return FELIS_OK;
}

// save_attr_HDF (saveParticleHDF.cpp:221-268): a 1-D attribute on a dataset
template <class T>
bool write_attr(hid_t ds, const char *name, const T *data, hsize_t n) {
  Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  Hid at(H5Acreate2(ds, name, native<T>(), sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return at.ok() && H5Awrite(at.id, native<T>(), data) >= 0;
}

// a 2-D double dataset [rows][cols] written from (or read into) contiguous memory
bool write_2d(hid_t f, const char *name, const double *data, hsize_t rows, hsize_t cols, hid_t *keep = nullptr) {
  const hsize_t dims[2] = {rows, cols};
  Hid sp(H5Screate_simple(2, dims, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  hid_t ds = H5Dcreate2(f, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) return false;
  const bool ok = rows * cols == 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
  if (keep && ok) {
    *keep = ds;
  } else {
    H5Dclose(ds);
  }
  return ok;
}

int read_2d(hid_t f, const char *file, const char *name, hsize_t rows, hsize_t cols, double *data) {
  Hid ds(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(FELIS_ERR_IO, std::string(file) + ": no dataset '" + name + "'");
  Hid sp(H5Dget_space(ds.id), H5Sclose);
  hsize_t dims[2] = {0, 0};
  if (!sp.ok() || H5Sget_simple_extent_ndims(sp.id) != 2 || H5Sget_simple_extent_dims(sp.id, dims, nullptr) != 2 ||
      dims[0] != rows || dims[1] != cols)
    return fail(FELIS_ERR_IO, std::string(file) + ": dataset '" + name + "' is not [" + std::to_string(rows) + "][" +
                                  std::to_string(cols) + "]");
  if (rows * cols && H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(FELIS_ERR_IO, std::string(file) + ": cannot read '" + name + "'");
  return FELIS_OK;
}

int core(felis_ctx *ctx, int rc, const char *what) {
  return rc == FELIS_OK ? FELIS_OK : fail(rc, std::string(what) + ": " + felis_last_error(ctx));
}

const char *pol_name(int which) { return which == FELIS_UX ? "Ux" : "Uy"; }

}  // namespace

extern "C" const char *felis_hdf_last_error(void) { return g_err.c_str(); }

extern "C" int felis_save_field_hdf(felis_ctx *ctx, const char *file, double spot_sig_r) {
  if (!ctx || !file) return fail(FELIS_ERR_ARG, "felis_save_field_hdf: null argument");
  quiet_hdf5();
  felis_desc d;
  int e;
  if ((e = core(ctx, felis_get_desc(ctx, &d), "felis_get_desc"))) return e;
  const size_t N = (size_t)d.nx * d.ny, S = (size_t)d.slice_n + 2;
  Hid f(H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FELIS_ERR_IO, std::string("cannot create ") + file);
  // slots 1..slice_n of each harmonic are contiguous in the reference layout
  std::vector<double> buf((size_t)d.n_harm * S * N * 2);
  for (int which : {FELIS_UX, FELIS_UY}) {
    if ((e = core(ctx, felis_download_field(ctx, which, buf.data()), "felis_download_field"))) return e;
    for (int h = 0; h < d.n_harm; ++h) {
      const std::string name = pol_name(which) + std::to_string(d.harmony[h]);
      if (!write_2d(f.id, name.c_str(), buf.data() + ((size_t)h * S + 1) * N * 2, (hsize_t)d.slice_n, (hsize_t)N * 2))
        return fail(FELIS_ERR_IO, std::string(file) + ": cannot write '" + name + "'");
    }
  }
  // metadata, saveFieldHDF.cpp:66-88
  double power_norm = 0.0, intensity_norm = 0.0;
  if ((e = core(ctx, felis_field_norms(ctx, spot_sig_r, &power_norm, &intensity_norm), "felis_field_norms"))) return e;
  const double bucket_z = d.num_slice * d.lambda0;
  const int nh = d.n_harm;
  const bool ok = write_meta(f.id, "sliceN", &d.slice_n, 1) && write_meta(f.id, "harmony", d.harmony, (hsize_t)nh) &&
                  write_meta(f.id, "numHarmony", &nh, 1) && write_meta(f.id, "nx", &d.nx, 1) &&
                  write_meta(f.id, "ny", &d.ny, 1) && write_meta(f.id, "minZ", &d.min_z, 1) &&
                  write_meta(f.id, "dz", &d.dz, 1) && write_meta(f.id, "bucketZ", &bucket_z, 1) &&
                  write_meta(f.id, "powerNorm", &power_norm, 1) &&
                  write_meta(f.id, "intensityNorm", &intensity_norm, 1) && write_meta(f.id, "dx", &d.dx, 1) &&
                  write_meta(f.id, "dy", &d.dy, 1) && write_meta(f.id, "minX", &d.min_x, 1) &&
                  write_meta(f.id, "minY", &d.min_y, 1);
  if (!ok) return fail(FELIS_ERR_IO, std::string(file) + ": cannot write the metadata");
  return FELIS_OK;
}

extern "C" int felis_restore_field_hdf(felis_ctx *ctx, const char *file) {
  if (!ctx || !file) return fail(FELIS_ERR_ARG, "felis_restore_field_hdf: null argument");
  quiet_hdf5();
  felis_desc d;
  int e;
  if ((e = core(ctx, felis_get_desc(ctx, &d), "felis_get_desc"))) return e;
  Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FELIS_ERR_IO, std::string("cannot open ") + file);
  // the settings the reference checks before it reads (restoreFieldHDF.cpp:58-77)
  int nx = 0, ny = 0, slice_n = 0, nh = 0;
  double dx = 0, dy = 0, min_x = 0, min_y = 0;
  if ((e = read_meta(f.id, file, "nx", &nx, 1)) || (e = read_meta(f.id, file, "ny", &ny, 1)) ||
      (e = read_meta(f.id, file, "dx", &dx, 1)) || (e = read_meta(f.id, file, "dy", &dy, 1)) ||
      (e = read_meta(f.id, file, "minX", &min_x, 1)) || (e = read_meta(f.id, file, "minY", &min_y, 1)) ||
      (e = read_meta(f.id, file, "sliceN", &slice_n, 1)) || (e = read_meta(f.id, file, "numHarmony", &nh, 1)))
    return e;
  auto mismatch = [&](const char *what, const std::string &got, const std::string &want) {
    return fail(FELIS_ERR_ARG, std::string(file) + ": " + what + " is " + got + ", the context's is " + want);
  };
  if (nx != d.nx) return mismatch("nx", std::to_string(nx), std::to_string(d.nx));
  if (ny != d.ny) return mismatch("ny", std::to_string(ny), std::to_string(d.ny));
  if (dx != d.dx) return mismatch("dx", std::to_string(dx), std::to_string(d.dx));
  if (dy != d.dy) return mismatch("dy", std::to_string(dy), std::to_string(d.dy));
  if (min_x != d.min_x) return mismatch("minX", std::to_string(min_x), std::to_string(d.min_x));
  if (min_y != d.min_y) return mismatch("minY", std::to_string(min_y), std::to_string(d.min_y));
  if (slice_n != d.slice_n) return mismatch("sliceN", std::to_string(slice_n), std::to_string(d.slice_n));
  if (nh != d.n_harm) return mismatch("numHarmony", std::to_string(nh), std::to_string(d.n_harm));
  const size_t N = (size_t)d.nx * d.ny, S = (size_t)d.slice_n + 2;
  std::vector<double> buf((size_t)d.n_harm * S * N * 2);
  for (int which : {FELIS_UX, FELIS_UY}) {
    // the ghost slots keep the context's values
    if ((e = core(ctx, felis_download_field(ctx, which, buf.data()), "felis_download_field"))) return e;
    for (int h = 0; h < d.n_harm; ++h) {
      const std::string name = pol_name(which) + std::to_string(d.harmony[h]);
      if ((e = read_2d(f.id, file, name.c_str(), (hsize_t)d.slice_n, (hsize_t)N * 2,
                       buf.data() + ((size_t)h * S + 1) * N * 2)))
        return e;
    }
    if ((e = core(ctx, felis_upload_field(ctx, which, buf.data()), "felis_upload_field"))) return e;
  }
  return FELIS_OK;
}

extern "C" int felis_save_particles_hdf(felis_ctx *ctx, const char *file, const double *gamma0) {
  if (!ctx || !file || !gamma0) return fail(FELIS_ERR_ARG, "felis_save_particles_hdf: null argument");
  quiet_hdf5();
  felis_desc d;
  int e;
  if ((e = core(ctx, felis_get_desc(ctx, &d), "felis_get_desc"))) return e;
  const int ns = d.n_species, sn = d.slice_n;
  int64_t np = 0;
  std::vector<int64_t> seg_off((size_t)ns * sn + 1);
  std::vector<double> seg_w((size_t)ns * sn);
  if ((e = core(ctx, felis_particle_layout(ctx, seg_off.data(), seg_w.data(), &np), "felis_particle_layout")))
    return e;
  std::vector<double> x(np), y(np), px(np), py(np), th(np), ga(np);
  if ((e = core(ctx, felis_download_particles(ctx, x.data(), y.data(), px.data(), py.data(), th.data(), ga.data()),
                "felis_download_particles")))
    return e;
  Hid f(H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FELIS_ERR_IO, std::string("cannot create ") + file);
  std::vector<double> rows;
  for (int s = 0; s < ns; ++s) {
    // rows of species s, slice by slice (saveParticleHDF.cpp:84-101)
    const int64_t b = seg_off[(size_t)s * sn], total = seg_off[(size_t)(s + 1) * sn] - b;
    rows.assign((size_t)total * kDataCnt, 0.0);
    size_t r = 0;
    for (int sl = 0; sl < sn; ++sl) {
      const size_t seg = (size_t)s * sn + sl;
      for (int64_t n = seg_off[seg]; n < seg_off[seg + 1]; ++n, ++r) {
        double *row = rows.data() + r * kDataCnt;
        row[0] = th[n];
        row[1] = x[n];
        row[2] = y[n];
        row[3] = ga[n];
        row[4] = px[n];
        row[5] = py[n];
        row[6] = sl;
        row[7] = seg_w[seg];
      }
    }
    const std::string name = std::to_string(s);
    hid_t ds = -1;
    if (!write_2d(f.id, name.c_str(), rows.data(), (hsize_t)total, kDataCnt, &ds))
      return fail(FELIS_ERR_IO, std::string(file) + ": cannot write '" + name + "'");
    Hid dsh(ds, H5Dclose);
    // one rank: cntP holds this process's count (saveParticleHDF.cpp:151-154)
    const long cnt = (long)total;
    if (!write_attr(ds, "cntP", &cnt, 1) || !write_attr(ds, "totalCnt", &cnt, 1))
      return fail(FELIS_ERR_IO, std::string(file) + ": cannot write the attributes of '" + name + "'");
  }
  // metadata, saveParticleHDF.cpp:145-168
  const double bucket_z = d.lambda0 * d.num_slice, d_phi = 2 * M_PI * d.num_slice;
  const int data_cnt = kDataCnt;
  const bool ok = write_meta(f.id, "minZ", &d.min_z, 1) && write_meta(f.id, "dz", &d.dz, 1) &&
                  write_meta(f.id, "bucketZ", &bucket_z, 1) && write_meta(f.id, "dPhi", &d_phi, 1) &&
                  write_meta(f.id, "lambda0", &d.lambda0, 1) && write_meta(f.id, "gamma0", gamma0, (hsize_t)ns) &&
                  write_meta(f.id, "numData", &data_cnt, 1) && write_meta(f.id, "nSpecies", &ns, 1) &&
                  write_meta(f.id, "sliceN", &d.slice_n, 1);
  if (!ok) return fail(FELIS_ERR_IO, std::string(file) + ": cannot write the metadata");
  return FELIS_OK;
}

extern "C" int felis_restore_particles_hdf(felis_ctx *ctx, const char *file) {
  if (!ctx || !file) return fail(FELIS_ERR_ARG, "felis_restore_particles_hdf: null argument");
  quiet_hdf5();
  felis_desc d;
  int e;
  if ((e = core(ctx, felis_get_desc(ctx, &d), "felis_get_desc"))) return e;
  Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(FELIS_ERR_IO, std::string("cannot open ") + file);
  int ns = 0, data_cnt = 0;
  if ((e = read_meta(f.id, file, "nSpecies", &ns, 1)) || (e = read_meta(f.id, file, "numData", &data_cnt, 1))) return e;
  if (ns != d.n_species)
    return fail(FELIS_ERR_ARG, std::string(file) + ": nSpecies is " + std::to_string(ns) + ", the context's is " +
                                   std::to_string(d.n_species));
  if (data_cnt < kDataCnt)
    return fail(FELIS_ERR_IO, std::string(file) + ": numData is " + std::to_string(data_cnt) + ", rows need " +
                                  std::to_string(kDataCnt) + " columns");
  const int sn = d.slice_n;
  std::vector<int64_t> seg_off((size_t)ns * sn + 1, 0);
  std::vector<double> seg_w((size_t)ns * sn, 0.0);
  std::vector<double> x, y, px, py, th, ga;
  std::vector<double> rows;
  for (int s = 0; s < ns; ++s) {
    const std::string name = std::to_string(s);
    long total = 0;
    {
      Hid ds(H5Dopen2(f.id, name.c_str(), H5P_DEFAULT), H5Dclose);
      if (!ds.ok()) return fail(FELIS_ERR_IO, std::string(file) + ": no dataset '" + name + "'");
      Hid at(H5Aopen(ds.id, "totalCnt", H5P_DEFAULT), H5Aclose);
      if (!at.ok() || H5Aread(at.id, H5T_NATIVE_LONG, &total) < 0 || total < 0)
        return fail(FELIS_ERR_IO, std::string(file) + ": no attribute totalCnt on '" + name + "'");
    }
    rows.assign((size_t)total * data_cnt, 0.0);
    if ((e = read_2d(f.id, file, name.c_str(), (hsize_t)total, (hsize_t)data_cnt, rows.data()))) return e;
    // rows come in slice order; each slice becomes its segment
    int last = -1;
    for (long r = 0; r < total; ++r) {
      const double *row = rows.data() + (size_t)r * data_cnt;
      const double sv = row[6];
      const int sl = (int)sv;
      if (!(sv >= 0.0 && sv < (double)sn) || (double)sl != sv || sl < last)
        return fail(FELIS_ERR_IO, std::string(file) + ": row " + std::to_string(r) + " of '" + name +
                                      "' has slice " + std::to_string(sv) + " (slices in order, 0.." +
                                      std::to_string(sn - 1) + ")");
      last = sl;
      const size_t seg = (size_t)s * sn + sl;
      seg_off[seg + 1]++;
      seg_w[seg] = row[7];
      th.push_back(row[0]);
      x.push_back(row[1]);
      y.push_back(row[2]);
      ga.push_back(row[3]);
      px.push_back(row[4]);
      py.push_back(row[5]);
    }
  }
  for (size_t k = 0; k + 1 < seg_off.size(); ++k) seg_off[k + 1] += seg_off[k];
  return core(ctx,
              felis_upload_particles(ctx, seg_off.data(), seg_w.data(), x.data(), y.data(), px.data(), py.data(),
                                     th.data(), ga.data()),
              "felis_upload_particles");
}
