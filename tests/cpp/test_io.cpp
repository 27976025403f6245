// C++ client of include/raptr/io.hpp: the density writer case of
// test_H5DensityWriter.cpp read back through the C ABI, and a measurement
// written dense and read as the sparse vector / measurement list.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "raptr/io.hpp"

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);     \
      ++fails;                                                    \
    }                                                             \
  } while (0)

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // density: 5 x 6 x 7, origin (0.1, 0.2, 0.3), voxel (1, 2, 3), voxel (1,2,3) = 1
  DefaultVoxelGrid<float> grid(0.1f, 0.2f, 0.3f, 1.0f, 2.0f, 3.0f, 5, 6, 7);
  std::vector<float> dens(5 * 6 * 7, 0.0f);
  dens[(1 * 6 + 2) * 7 + 3] = 1.0f;
  const std::string dfn = dir + "/cpp_density.h5";
  writeHDF5_Density(dens.data(), dfn, grid);
  float lo[3], hi[3];
  int n[3];
  CHECK(raptr_h5_density_grid(dfn.c_str(), lo, hi, n) == 0);
  CHECK(n[0] == 5 && n[1] == 6 && n[2] == 7);
  CHECK(lo[0] == 0.1f && lo[1] == 0.2f && lo[2] == 0.3f);
  CHECK(hi[0] == 0.1f + 5.0f * 1.0f && hi[1] == 0.2f + 6.0f * 2.0f && hi[2] == 0.3f + 7.0f * 3.0f);
  std::vector<float> back = raptr::h5_read_all(dfn, "density");
  CHECK(back == dens);
  // measurement: 2 x 2 x 2 x 2 x 2 channels, three non-zero
  std::vector<float> meas(32, 0.0f);
  meas[3] = 2.0f;
  meas[17] = 5.0f;
  meas[31] = 1.0f;
  const std::string mfn = dir + "/cpp_meas.h5";
  const int dims[5] = {2, 2, 2, 2, 2};
  CHECK(raptr_h5_write_measurement(mfn.c_str(), meas.data(), dims) == 0);
  std::vector<int> id;
  std::vector<float> val;
  int nnz = 0;
  readHDF5_MeasVct(id, val, nnz, mfn);
  CHECK(nnz == 3 && id == (std::vector<int>{3, 17, 31}) && val == (std::vector<float>{2, 5, 1}));
  DefaultMeasurementList<float> list = H5File2DefaultMeasurementList<float>(mfn, 32);
  CHECK(list.size() == 3 && list.cnlId(1) == 17 && list.val(1) == 5.0f);
  bool threw = false;
  try {
    readHDF5_MeasVct(id, val, nnz, dfn);  // no "messung" there
  } catch (const std::runtime_error&) {
    threw = true;
  }
  CHECK(threw);
  std::printf("%s\n", fails ? "FAILED" : "OK");
  return fails ? 1 : 0;
}
