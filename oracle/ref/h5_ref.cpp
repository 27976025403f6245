// h5_ref.cpp -- binds the reference's own HDF5 writer and readers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (oracle/Makefile, target "h5ref")
// against the reference's headers where they lie in the raptr source tree
// and the image's HDF5 C++ library; adds no reference code of its own.
// H5DensityWriter<TGrid> needs a grid type that implements the reference's
// WritableGrid interface (H5DensityWriter.hpp:34-86); RefGrid below is such
// a client type.  tests/golden/make_golden.py records what these functions
// write and read as fixtures for the tests of raptr_amd.h5io.
#include <cstdint>
#include <string>

#include "H5DensityReader.hpp"
#include "H5DensityWriter.hpp"
#include "H5Reader.hpp"

namespace {
struct RefGrid : public WritableGrid<RefGrid, WritableGridTraits<float> > {
  float o[3], d[3];
  int n[3];
  void getOrigin(float* const origin) const {
    for (int k = 0; k < 3; ++k) origin[k] = o[k];
  }
  void getVoxelSize(float* const voxelSize) const {
    for (int k = 0; k < 3; ++k) voxelSize[k] = d[k];
  }
  void getNumberOfVoxels(int* const numberOfVoxels) const {
    for (int k = 0; k < 3; ++k) numberOfVoxels[k] = n[k];
  }
};
}  // namespace

extern "C" {

// H5DensityWriter<RefGrid>(path).write(mem, grid)
int ref_h5_write_density(const char* path, const float* mem, const float* origin,
                         const float* voxel, const int* nvox) {
  try {
    RefGrid g;
    for (int k = 0; k < 3; ++k) g.o[k] = origin[k], g.d[k] = voxel[k], g.n[k] = nvox[k];
    H5DensityWriter<RefGrid>(std::string(path)).write(mem, g);
    return 0;
  } catch (...) {
    return -1;
  }
}

// H5Reader ("messung") / H5DensityReader ("density"): sizeOfFile, read.
// Returns the element count, -1 on error or when it differs from n.
int64_t ref_h5_read(const char* path, int density, float* out, int64_t n) {
  try {
    if (density) {
      H5DensityReader r{std::string(path)};
      if (!r.is_open() || r.sizeOfFile() != n) return -1;
      r.read(out);
    } else {
      H5Reader r{std::string(path)};
      if (!r.is_open() || r.sizeOfFile() != n) return -1;
      r.read(out);
    }
    return n;
  } catch (...) {
    return -1;
  }
}
}
