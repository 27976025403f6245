// wq_ref.cpp -- binds the reference's own host-side workqueue code.
//
// TEST INFRASTRUCTURE ONLY.  This translation unit is compiled (by
// oracle/Makefile, target "ref") against the reference's headers where they
// lie in the raptr source tree; it adds no reference code of its own.  It
// instantiates getWorkqueue (getWorkqueue.hpp:193-243, unbounded, and
// :135-191, bounded) with the reference's Default* geometry types and exposes
// them through a C ABI so that tests/golden/make_golden.py can record the
// reference's exact (channel, voxel) decisions as fixtures.
//
// The reference marks these functions __host__ __device__; the build line
// defines both attributes away (-D__host__= -D__device__=), nothing else.
#include <cstdint>
#include <vector>

#include "VoxelGrid.hpp"
#include "VoxelGridLinIndex.hpp"
#include "MeasurementSetup.hpp"
#include "MeasurementSetupLinIndex.hpp"
#include "MeasurementSetupTrafo2CartCoord.hpp"
#include "MeasurementList.hpp"
#include "getWorkqueue.hpp"

namespace {
typedef float T;
typedef DefaultVoxelGrid<T> VG;
typedef DefaultVoxelGridIdx<VG> Idx;
typedef DefaultVoxelGridIdy<VG> Idy;
typedef DefaultVoxelGridIdz<VG> Idz;
typedef DefaultMeasurementSetup<T> MS;
typedef DefaultMeasurementSetupId0z<MS> Id0z;
typedef DefaultMeasurementSetupId0y<MS> Id0y;
typedef DefaultMeasurementSetupId1z<MS> Id1z;
typedef DefaultMeasurementSetupId1y<MS> Id1y;
typedef DefaultMeasurementSetupIda<MS> Ida;
typedef DefaultMeasurementSetupTrafo2CartCoordFirstPixel<T, MS> Trafo0;
typedef DefaultMeasurementSetupTrafo2CartCoordSecndPixel<T, MS> Trafo1;
typedef DefaultMeasurementList<T> ML;
}  // namespace

extern "C" {

// setup: pos0x, pos1x, da, segx, segy, segz; isetup: na, n0z, n0y, n1z, n1y
// grid: ox, oy, oz, dx, dy, dz; igrid: nx, ny, nz.  n < 0: unbounded.
int64_t ref_workqueue(const float* setup, const int* isetup, const float* grid,
                      const int* igrid, const int* list, int list_size,
                      int* list_id, int* vxl_id, int64_t n, int* cnl_out,
                      int* vxl_out) {
  MS ms(setup[0], setup[1], isetup[0], isetup[1], isetup[2], isetup[3],
        isetup[4], setup[2], setup[3], setup[4], setup[5]);
  VG vg(grid[0], grid[1], grid[2], grid[3], grid[4], grid[5], igrid[0],
        igrid[1], igrid[2]);
  ML ml(list_size);
  for (int i = 0; i < list_size; ++i) ml.set(i, list[i], T(1));
  std::vector<int> cnl, vxl;
  int found;
  if (n < 0) {
    found = getWorkqueue<T, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y,
                         Ida, Trafo0, Trafo1>(cnl, vxl, *list_id, *vxl_id, &ml,
                                              &vg, &ms);
  } else {
    found = getWorkqueue<T, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y,
                         Ida, Trafo0, Trafo1>(cnl, vxl, *list_id, *vxl_id, &ml,
                                              &vg, &ms, int(n));
  }
  for (int i = 0; i < found; ++i) {
    cnl_out[i] = cnl[i];
    vxl_out[i] = vxl[i];
  }
  return found;
}

// Out-of-place pixel trafo of the reference (det 0 or 1).
void ref_trafo(const float* setup, const int* isetup, int det, int idz, int idy,
               int ida, const float* box, float* out) {
  MS ms(setup[0], setup[1], isetup[0], isetup[1], isetup[2], isetup[3],
        isetup[4], setup[2], setup[3], setup[4], setup[5]);
  if (det == 0)
    Trafo0()(out, box, idz, idy, ida, &ms);
  else
    Trafo1()(out, box, idz, idy, ida, &ms);
}

// Reference point-line distance (distancePointLine.h:87-101).
float ref_distance(const float* a, const float* b, const float* p) {
  return distance<float>(a, b, p);
}
}
