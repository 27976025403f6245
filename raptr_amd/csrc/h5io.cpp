// h5io.cpp -- the reference's HDF5 file formats (host code, HDF5 C API).
//
// Measurement files (README.md "Data Formats", H5Reader.hpp): dataset
// "messung", 32-bit float, 5-D (angle, det0 z, det0 y, det1 z, det1 y), so
// the row-major linear index is the channel id of MeasurementSetupLinIndex.
// Density files (H5DensityWriter.hpp:97-231, H5DensityReader.hpp): dataset
// "density", 32-bit float, 3-D (xnbin, ynbin, znbin) in the grid's linear
// order (z fastest), with scalar float attributes xmin/xmax/ymin/ymax/
// zmin/zmax (grid surfaces) and int attributes xnbin/ynbin/znbin.
#include <hdf5.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "raptr_h5.h"

namespace {

thread_local std::string g_err;

int h5fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

// quiet HDF5's own error stack printing; errors come back as statuses
struct Quiet {
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct Handle {
  hid_t id;
  herr_t (*close)(hid_t);
  Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Handle() {
    if (id >= 0) close(id);
  }
  bool ok() const { return id >= 0; }
};

int write_attr_float(hid_t ds, const char* name, float v) {
  Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
  Handle at(H5Acreate2(ds, name, H5T_IEEE_F32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!sp.ok() || !at.ok() || H5Awrite(at.id, H5T_NATIVE_FLOAT, &v) < 0)
    return h5fail("cannot write attribute %s", name);
  return 0;
}

int write_attr_int(hid_t ds, const char* name, int v) {
  Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
  Handle at(H5Acreate2(ds, name, H5T_STD_I32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!sp.ok() || !at.ok() || H5Awrite(at.id, H5T_NATIVE_INT, &v) < 0)
    return h5fail("cannot write attribute %s", name);
  return 0;
}

int read_attr(hid_t ds, const char* name, hid_t mem_type, void* out) {
  Handle at(H5Aopen(ds, name, H5P_DEFAULT), H5Aclose);
  if (!at.ok() || H5Aread(at.id, mem_type, out) < 0)
    return h5fail("cannot read attribute %s", name);
  return 0;
}

int write_dataset(const char* path, const char* name, const float* data, int ndims,
                  const hsize_t* dims) {
  Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return h5fail("cannot create %s", path);
  Handle sp(H5Screate_simple(ndims, dims, nullptr), H5Sclose);
  Handle ds(H5Dcreate2(f.id, name, H5T_IEEE_F32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT,
                       H5P_DEFAULT),
            H5Dclose);
  if (!sp.ok() || !ds.ok()) return h5fail("cannot create dataset %s in %s", name, path);
  if (H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return h5fail("cannot write dataset %s", name);
  return 0;
}

}  // namespace

extern "C" {

const char* raptr_h5_last_error(void) { return g_err.c_str(); }

int raptr_h5_dims(const char* path, const char* dataset, int* ndims, int64_t* dims,
                  int max_dims) {
  Quiet q;
  if (!path || !dataset || !ndims || !dims || max_dims < 1) return h5fail("bad arguments");
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return h5fail("cannot open %s", path);
  Handle ds(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return h5fail("%s has no dataset \"%s\"", path, dataset);
  Handle sp(H5Dget_space(ds.id), H5Sclose);
  const int nd = H5Sget_simple_extent_ndims(sp.id);
  if (nd < 0 || nd > max_dims) return h5fail("dataset %s has %d dimensions", dataset, nd);
  hsize_t d[32];
  H5Sget_simple_extent_dims(sp.id, d, nullptr);
  *ndims = nd;
  for (int k = 0; k < nd; ++k) dims[k] = int64_t(d[k]);
  return 0;
}

int raptr_h5_read(const char* path, const char* dataset, float* out, int64_t n) {
  Quiet q;
  if (!path || !dataset || (!out && n > 0)) return h5fail("bad arguments");
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return h5fail("cannot open %s", path);
  Handle ds(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return h5fail("%s has no dataset \"%s\"", path, dataset);
  Handle sp(H5Dget_space(ds.id), H5Sclose);
  const hssize_t size = H5Sget_simple_extent_npoints(sp.id);
  if (size != n) return h5fail("dataset %s holds %lld values, expected %lld", dataset,
                               (long long)size, (long long)n);
  if (H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return h5fail("cannot read dataset %s", dataset);
  return 0;
}

int raptr_h5_write_measurement(const char* path, const float* counts, const int dims[5]) {
  Quiet q;
  if (!path || !counts || !dims) return h5fail("bad arguments");
  hsize_t d[5];
  for (int k = 0; k < 5; ++k) {
    if (dims[k] <= 0) return h5fail("dimension %d is %d", k, dims[k]);
    d[k] = hsize_t(dims[k]);
  }
  return write_dataset(path, "messung", counts, 5, d);
}

int raptr_h5_write_density(const char* path, const float* density, const float origin[3],
                           const float voxel[3], const int nvox[3]) {
  Quiet q;
  if (!path || !density || !origin || !voxel || !nvox) return h5fail("bad arguments");
  hsize_t d[3] = {hsize_t(nvox[0]), hsize_t(nvox[1]), hsize_t(nvox[2])};
  if (write_dataset(path, "density", density, 3, d)) return -1;
  Handle f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
  Handle ds(H5Dopen2(f.id, "density", H5P_DEFAULT), H5Dclose);
  if (!f.ok() || !ds.ok()) return h5fail("cannot reopen %s", path);
  static const char* lo[3] = {"xmin", "ymin", "zmin"};
  static const char* hi[3] = {"xmax", "ymax", "zmax"};
  static const char* nb[3] = {"xnbin", "ynbin", "znbin"};
  for (int k = 0; k < 3; ++k) {
    // upper surface as the reference computes it: origin + n * voxel, in float
    const float mx = origin[k] + float(nvox[k]) * voxel[k];
    if (write_attr_float(ds.id, lo[k], origin[k]) || write_attr_float(ds.id, hi[k], mx) ||
        write_attr_int(ds.id, nb[k], nvox[k]))
      return -1;
  }
  return 0;
}

int raptr_h5_density_grid(const char* path, float origin[3], float upper[3], int nvox[3]) {
  Quiet q;
  if (!path || !origin || !upper || !nvox) return h5fail("bad arguments");
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return h5fail("cannot open %s", path);
  Handle ds(H5Dopen2(f.id, "density", H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return h5fail("%s has no dataset \"density\"", path);
  static const char* lo[3] = {"xmin", "ymin", "zmin"};
  static const char* hi[3] = {"xmax", "ymax", "zmax"};
  static const char* nb[3] = {"xnbin", "ynbin", "znbin"};
  for (int k = 0; k < 3; ++k) {
    if (read_attr(ds.id, lo[k], H5T_NATIVE_FLOAT, &origin[k]) ||
        read_attr(ds.id, hi[k], H5T_NATIVE_FLOAT, &upper[k]) ||
        read_attr(ds.id, nb[k], H5T_NATIVE_INT, &nvox[k]))
      return -1;
  }
  return 0;
}

}  // extern "C"
