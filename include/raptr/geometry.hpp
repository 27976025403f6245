// raptr/geometry.hpp -- host mirror of the reference's geometry "plugin"
// interfaces: the CRTP bases MeasurementSetup<T,C> (MeasurementSetup.hpp:
// 36-175) and VoxelGrid<T,C> (VoxelGrid.hpp:37-142), their Default*
// specialisations, the index functors (MeasurementSetupLinIndex.hpp:62-320,
// VoxelGridLinIndex.hpp:52-221), the pixel trafos
// (MeasurementSetupTrafo2CartCoord.hpp:70-466) and the measurement list
// (MeasurementList.hpp:66-128).  Names, template parameters, constructor
// argument order and functor call shapes follow the reference so that code
// written against it compiles unchanged; the bodies are this project's.
//
// The GPU kernels implement the Default* family; Default geometry objects
// convert to the POD structs of the C ABI (to_c()), which is how they cross
// into raptr_hip.h.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "raptr_hip.h"

namespace raptr_detail {
inline float angle_sin(int ida, float da) {
  float arg = 3.14159265358979323846f * float(ida);
  arg = arg * da;
  arg = arg / 180.0f;
  return ::sinf(arg);  // the host C library's, as the reference evaluates it
}
inline float angle_cos(int ida, float da) {
  float arg = 3.14159265358979323846f * float(ida);
  arg = arg * da;
  arg = arg / 180.0f;
  return ::cosf(arg);
}
}  // namespace raptr_detail

// ------------------------------------------------------ measurement setup
template <typename T, typename ConcreteMeasurementSetup>
class MeasurementSetup {
 public:
  T pos0x() const { return self()->pos0x(); }
  T pos1x() const { return self()->pos1x(); }
  int na() const { return self()->na(); }
  int n0z() const { return self()->n0z(); }
  int n0y() const { return self()->n0y(); }
  int n1z() const { return self()->n1z(); }
  int n1y() const { return self()->n1y(); }
  T da() const { return self()->da(); }
  T segx() const { return self()->segx(); }
  T segy() const { return self()->segy(); }
  T segz() const { return self()->segz(); }

 private:
  const ConcreteMeasurementSetup* self() const {
    return static_cast<const ConcreteMeasurementSetup*>(this);
  }
};

template <typename T>
class DefaultMeasurementSetup : public MeasurementSetup<T, DefaultMeasurementSetup<T> > {
 public:
  DefaultMeasurementSetup() {}
  DefaultMeasurementSetup(T pos0x, T pos1x, int na, int n0z, int n0y, int n1z, int n1y, T da,
                          T segx, T segy, T segz)
      : pos0x_(pos0x), pos1x_(pos1x), na_(na), da_(da), n0z_(n0z), n0y_(n0y), n1z_(n1z),
        n1y_(n1y), segx_(segx), segy_(segy), segz_(segz) {}
  T pos0x() const { return pos0x_; }
  T pos1x() const { return pos1x_; }
  int na() const { return na_; }
  T da() const { return da_; }
  int n0z() const { return n0z_; }
  int n0y() const { return n0y_; }
  int n1z() const { return n1z_; }
  int n1y() const { return n1y_; }
  T segx() const { return segx_; }
  T segy() const { return segy_; }
  T segz() const { return segz_; }
  int nChannels() const { return na_ * n0z_ * n0y_ * n1z_ * n1y_; }

  raptr_setup_t to_c() const {
    raptr_setup_t s;
    s.pos0x = float(pos0x_);
    s.pos1x = float(pos1x_);
    s.na = na_;
    s.n0z = n0z_;
    s.n0y = n0y_;
    s.n1z = n1z_;
    s.n1y = n1y_;
    s.da = float(da_);
    s.segx = float(segx_);
    s.segy = float(segy_);
    s.segz = float(segz_);
    return s;
  }

 private:
  T pos0x_ = 0, pos1x_ = 0;
  int na_ = 0;
  T da_ = 0;
  int n0z_ = 0, n0y_ = 0, n1z_ = 0, n1y_ = 0;
  T segx_ = 0, segy_ = 0, segz_ = 0;
};

// channel index functors: (ida, id0z, id0y, id1z, id1y), id1y fastest
template <typename ConcreteMSLinId, typename MS>
struct MeasurementSetupLinId {
  int operator()(int id0z, int id0y, int id1z, int id1y, int ida, MS const* meas) {
    return static_cast<ConcreteMSLinId*>(this)->operator()(id0z, id0y, id1z, id1y, ida, meas);
  }
};
template <typename MS>
struct DefaultMeasurementSetupLinId
    : public MeasurementSetupLinId<DefaultMeasurementSetupLinId<MS>, MS> {
  int operator()(int id0z, int id0y, int id1z, int id1y, int ida, MS const* m) {
    return (((ida * m->n0z() + id0z) * m->n0y() + id0y) * m->n1z() + id1z) * m->n1y() + id1y;
  }
};

#define RAPTR_MS_SPLIT_FUNCTOR(NAME, EXPR)                                              \
  template <typename C, typename MS>                                                    \
  struct MeasurementSetup##NAME {                                                       \
    int operator()(int linId, MS const* meas) {                                         \
      return static_cast<C*>(this)->operator()(linId, meas);                            \
    }                                                                                   \
  };                                                                                    \
  template <typename MS>                                                                \
  struct DefaultMeasurementSetup##NAME                                                  \
      : public MeasurementSetup##NAME<DefaultMeasurementSetup##NAME<MS>, MS> {          \
    int operator()(int linId, MS const* m) {                                            \
      const int p1 = m->n1y(), p2 = p1 * m->n1z(), p3 = p2 * m->n0y(), p4 = p3 * m->n0z(); \
      (void)p1; (void)p2; (void)p3; (void)p4;                                           \
      return EXPR;                                                                      \
    }                                                                                   \
  };
RAPTR_MS_SPLIT_FUNCTOR(Ida, linId / p4)
RAPTR_MS_SPLIT_FUNCTOR(Id0z, (linId % p4) / p3)
RAPTR_MS_SPLIT_FUNCTOR(Id0y, (linId % p3) / p2)
RAPTR_MS_SPLIT_FUNCTOR(Id1z, (linId % p2) / p1)
RAPTR_MS_SPLIT_FUNCTOR(Id1y, linId % p1)
#undef RAPTR_MS_SPLIT_FUNCTOR

// pixel unit-cube -> world (out of place and in place), detector 0 and 1
template <typename T, typename MS, int DET>
struct DefaultMeasurementSetupTrafo2CartCoordPixel {
  void operator()(T* out, T const* box, int idz, int idy, int ida, MS const* m) const {
    const T e0 = m->segx(), e1 = m->segy(), e2 = m->segz();
    const T p0 = DET == 0 ? m->pos0x() : m->pos1x();
    const int ny = DET == 0 ? m->n0y() : m->n1y();
    const int nz = DET == 0 ? m->n0z() : m->n1z();
    const T p1 = (T(idy) - T(.5) * T(ny - 1)) * e1;
    const T p2 = (T(idz) - T(.5) * T(nz - 1)) * e2;
    const T sn = T(raptr_detail::angle_sin(ida, float(m->da())));
    const T cs = T(raptr_detail::angle_cos(ida, float(m->da())));
    const T t0 = cs * e0, t1 = T(0), t2 = sn * e2;
    const T t3 = cs * (p0 - T(.5) * e0) + sn * (p2 - T(.5) * e2);
    const T t4 = T(0), t5 = e1, t6 = T(0), t7 = p1 - T(.5) * e1;
    const T t8 = -sn * e0, t9 = T(0), t10 = cs * e2;
    const T t11 = -sn * (p0 - T(.5) * e0) + cs * (p2 - T(.5) * e2);
    const T b0 = box[0], b1 = box[1], b2 = box[2];
    out[0] = t0 * b0 + t1 * b1 + t2 * b2 + t3 * T(1);
    out[1] = t4 * b0 + t5 * b1 + t6 * b2 + t7 * T(1);
    out[2] = t8 * b0 + t9 * b1 + t10 * b2 + t11 * T(1);
  }
  void operator()(T* coords, int idz, int idy, int ida, MS const* m) const {
    T tmp[3] = {coords[0], coords[1], coords[2]};
    (*this)(coords, tmp, idz, idy, ida, m);
  }
};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordFirstPixel
    : DefaultMeasurementSetupTrafo2CartCoordPixel<T, MS, 0> {};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordSecndPixel
    : DefaultMeasurementSetupTrafo2CartCoordPixel<T, MS, 1> {};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordFirstPixel_inplace
    : DefaultMeasurementSetupTrafo2CartCoordPixel<T, MS, 0> {};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordSecndPixel_inplace
    : DefaultMeasurementSetupTrafo2CartCoordPixel<T, MS, 1> {};

// ---------------------------------------------------------------- grid
template <typename T, typename ConcreteVoxelGrid>
class VoxelGrid {
 public:
  T gridox() const { return self()->gridox(); }
  T gridoy() const { return self()->gridoy(); }
  T gridoz() const { return self()->gridoz(); }
  T griddx() const { return self()->griddx(); }
  T griddy() const { return self()->griddy(); }
  T griddz() const { return self()->griddz(); }
  int gridnx() const { return self()->gridnx(); }
  int gridny() const { return self()->gridny(); }
  int gridnz() const { return self()->gridnz(); }

 private:
  const ConcreteVoxelGrid* self() const { return static_cast<const ConcreteVoxelGrid*>(this); }
};

template <typename T>
class DefaultVoxelGrid : public VoxelGrid<T, DefaultVoxelGrid<T> > {
 public:
  DefaultVoxelGrid() {}
  DefaultVoxelGrid(T o0, T o1, T o2, T d0, T d1, T d2, int n0, int n1, int n2) {
    o_[0] = o0, o_[1] = o1, o_[2] = o2;
    d_[0] = d0, d_[1] = d1, d_[2] = d2;
    n_[0] = n0, n_[1] = n1, n_[2] = n2;
  }
  T gridox() const { return o_[0]; }
  T gridoy() const { return o_[1]; }
  T gridoz() const { return o_[2]; }
  T griddx() const { return d_[0]; }
  T griddy() const { return d_[1]; }
  T griddz() const { return d_[2]; }
  int gridnx() const { return n_[0]; }
  int gridny() const { return n_[1]; }
  int gridnz() const { return n_[2]; }
  int nVoxels() const { return n_[0] * n_[1] * n_[2]; }

  raptr_grid_t to_c() const {
    raptr_grid_t g;
    for (int k = 0; k < 3; ++k) {
      g.o[k] = float(o_[k]);
      g.d[k] = float(d_[k]);
      g.n[k] = n_[k];
    }
    return g;
  }

 private:
  T o_[3] = {0, 0, 0}, d_[3] = {0, 0, 0};
  int n_[3] = {0, 0, 0};
};

// voxel index functors, z fastest: lin = idx*ny*nz + idy*nz + idz
template <typename C, typename VG>
struct VoxelGridLinId {
  int operator()(int idx, int idy, int idz, VG const* vg) {
    return static_cast<C*>(this)->operator()(idx, idy, idz, vg);
  }
};
template <typename VG>
struct DefaultVoxelGridLinId : public VoxelGridLinId<DefaultVoxelGridLinId<VG>, VG> {
  int operator()(int idx, int idy, int idz, VG const* vg) {
    return (idx * vg->gridny() + idy) * vg->gridnz() + idz;
  }
};
template <typename VG>
struct DefaultVoxelGridIdx {
  int operator()(int linId, VG const* vg) { return linId / (vg->gridnz() * vg->gridny()); }
};
template <typename VG>
struct DefaultVoxelGridIdy {
  int operator()(int linId, VG const* vg) {
    return (linId % (vg->gridnz() * vg->gridny())) / vg->gridnz();
  }
};
template <typename VG>
struct DefaultVoxelGridIdz {
  int operator()(int linId, VG const* vg) { return linId % vg->gridnz(); }
};

// ------------------------------------------------------ measurement list
template <typename T, typename ConcreteMeasurementList>
class MeasurementList {
 public:
  int cnlId(int listId) const { return self()->cnlId(listId); }
  T val(int listId) const { return self()->val(listId); }

 private:
  const ConcreteMeasurementList* self() const {
    return static_cast<const ConcreteMeasurementList*>(this);
  }
};

template <typename T>
class DefaultMeasurementList : public MeasurementList<T, DefaultMeasurementList<T> > {
 public:
  explicit DefaultMeasurementList(int size) : cnl_(size), val_(size) {}
  int cnlId(int listId) const { return cnl_[listId]; }
  T val(int listId) const { return val_[listId]; }
  int size() const { return int(cnl_.size()); }
  // MeasurementList.hpp:113-125: an out-of-range set is reported on
  // stderr and ends the process
  void set(int listId, int cnlId, T val) {
    if (listId >= size()) {
      std::cerr << "raptr: measurement list index " << listId << " >= size " << size()
                << std::endl;
      std::exit(EXIT_FAILURE);
    }
    cnl_[listId] = cnlId;
    val_[listId] = val;
  }
  const int* cnlData() const { return cnl_.data(); }

 private:
  std::vector<int> cnl_;
  std::vector<T> val_;
};
