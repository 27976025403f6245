/*
 * raptr_h5.h -- C ABI of the HDF5 measurement and density files used by the
 * reference's tools (library raptr_amd/lib/libraptr_h5.so, host only).
 *
 *   measurement: dataset "messung", float32, 5-D (angle, det0 z, det0 y,
 *                det1 z, det1 y); row-major linear index = channel id
 *                (README.md "Data Formats"; H5Reader.hpp; read into a
 *                sparse vector by readHDF5_MeasVct, wrappers.hpp:329-368)
 *   density:     dataset "density", float32, 3-D (xnbin, ynbin, znbin),
 *                z fastest; scalar attributes xmin, xmax, ymin, ymax, zmin,
 *                zmax (float) and xnbin, ynbin, znbin (int32)
 *                (H5DensityWriter.hpp:97-231, H5DensityReader.hpp)
 *
 * Every function returns 0 or -1; raptr_h5_last_error() gives the message.
 */
#ifndef RAPTR_H5_H
#define RAPTR_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* raptr_h5_last_error(void);
/* Rank and extents of a dataset (H5Reader::sizeOfFile). */
int raptr_h5_dims(const char* path, const char* dataset, int* ndims, int64_t* dims,
                  int max_dims);
/* Read all n values of a dataset as float (H5Reader::read,
 * H5DensityReader::read); fails if the dataset does not hold n values. */
int raptr_h5_read(const char* path, const char* dataset, float* out, int64_t n);
/* Write a dense measurement (file truncated). */
int raptr_h5_write_measurement(const char* path, const float* counts, const int dims[5]);
/* Write a density with the reference's attributes (H5DensityWriter::write;
 * x/y/zmax = origin + n * voxel size in float; file truncated). */
int raptr_h5_write_density(const char* path, const float* density, const float origin[3],
                           const float voxel[3], const int nvox[3]);
/* Grid attributes of a density file. */
int raptr_h5_density_grid(const char* path, float origin[3], float upper[3], int nvox[3]);

#ifdef __cplusplus
}
#endif
#endif
