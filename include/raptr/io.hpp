// raptr/io.hpp -- C++ mirror of the reference's HDF5 helpers over
// raptr_h5.h (link raptr_amd/lib/libraptr_h5.so):
//   readHDF5_MeasVct      wrappers.hpp:329-368 (dense "messung" -> sparse)
//   writeHDF5_Density     wrappers.hpp:400-407 (H5DensityWriter)
//   H5File2DefaultMeasurementList   H5File2DefaultMeasurementList.h
// Errors: an unreadable file or dataset throws std::runtime_error (the
// reference lets H5::FileIException escape); a raw size mismatch in
// H5File2DefaultMeasurementList prints and exits, as the reference does.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "raptr/geometry.hpp"
#include "raptr_h5.h"

namespace raptr {

inline void h5_check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + raptr_h5_last_error());
}

inline std::vector<float> h5_read_all(std::string const& fn, const char* dataset) {
  int nd = 0;
  int64_t dims[8];
  h5_check(raptr_h5_dims(fn.c_str(), dataset, &nd, dims, 8), "raptr_h5_dims");
  int64_t n = 1;
  for (int k = 0; k < nd; ++k) n *= dims[k];
  std::vector<float> mem(size_t(n), 0.0f);
  h5_check(raptr_h5_read(fn.c_str(), dataset, mem.data(), n), "raptr_h5_read");
  return mem;
}

}  // namespace raptr

template <typename T>
void readHDF5_MeasVct(std::vector<int>& vctId, std::vector<T>& vctVal, int& vctNnz,
                      std::string const& ifn) {
  const std::vector<float> mem = raptr::h5_read_all(ifn, "messung");
  vctId.clear();
  vctVal.clear();
  for (size_t i = 0; i < mem.size(); ++i)
    if (T(mem[i]) != T(0)) {
      vctId.push_back(int(i));
      vctVal.push_back(T(mem[i]));
    }
  vctNnz = int(vctId.size());
}

template <typename T>
void writeHDF5_Density(T const* const density, std::string const ofn,
                       DefaultVoxelGrid<T> const& grid) {
  const float o[3] = {float(grid.gridox()), float(grid.gridoy()), float(grid.gridoz())};
  const float d[3] = {float(grid.griddx()), float(grid.griddy()), float(grid.griddz())};
  const int n[3] = {grid.gridnx(), grid.gridny(), grid.gridnz()};
  std::vector<float> buf(size_t(n[0]) * n[1] * n[2]);
  for (size_t i = 0; i < buf.size(); ++i) buf[i] = float(density[i]);
  raptr::h5_check(raptr_h5_write_density(ofn.c_str(), buf.data(), o, d, n),
                  "raptr_h5_write_density");
}

template <typename T>
DefaultMeasurementList<T> H5File2DefaultMeasurementList(std::string const fn,
                                                      int const rawSize) {
  int nd = 0;
  int64_t dims[8];
  if (raptr_h5_dims(fn.c_str(), "messung", &nd, dims, 8) != 0) {
    std::cerr << "Could not open file " << fn << std::endl;
    std::exit(EXIT_FAILURE);
  }
  int64_t n = 1;
  for (int k = 0; k < nd; ++k) n *= dims[k];
  if (n != rawSize) {
    std::cerr << "File " << fn << " does not have raw size " << rawSize << std::endl;
    std::exit(EXIT_FAILURE);
  }
  const std::vector<float> mem = raptr::h5_read_all(fn, "messung");
  int nz = 0;
  for (float v : mem) nz += T(v) != T(0);
  DefaultMeasurementList<T> list(nz);
  int id = 0;
  for (int c = 0; c < rawSize; ++c)
    if (T(mem[c]) != T(0)) list.set(id++, c, T(mem[c]));
  return list;
}
