// common.hpp -- shared device/host helpers of the gfx950 raptr kernels.
//
// All arithmetic here keeps the reference's operation order; the library is
// compiled with -ffp-contract=off and HIP's correctly rounded FP32 divide and
// sqrt, so results are bit-identical to the host restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "raptr_hip.h"

namespace raptr {

// ------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define RAPTR_HIP_TRY(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::raptr::fail(RAPTR_ERR_DEVICE, "%s:%d: %s: %s", __FILE__,       \
                           __LINE__, #expr, hipGetErrorString(_e));           \
  } while (0)

#define RAPTR_LAUNCH_CHECK()                                                  \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess)                                                     \
      return ::raptr::fail(RAPTR_ERR_DEVICE, "%s:%d: launch: %s", __FILE__,   \
                           __LINE__, hipGetErrorString(_e));                  \
  } while (0)

#define RAPTR_TRY(expr)                                                       \
  do {                                                                        \
    int _s = (expr);                                                          \
    if (_s != RAPTR_OK) return _s;                                            \
  } while (0)

inline hipStream_t S(raptr_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------- workspace carving
struct Carver {
  char* base;
  size_t cap;
  size_t off = 0;
  bool ok = true;
  Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <class T>
  T* take(size_t count) {
    size_t o = (off + 255) & ~size_t(255);
    size_t need = o + count * sizeof(T);
    if (need > cap || (count && !base)) {
      ok = false;
      off = need;
      return nullptr;
    }
    off = need;
    return reinterpret_cast<T*>(base + o);
  }
};
// Size-only counterpart of Carver (same alignment rule).
struct Sizer {
  size_t off = 0;
  template <class T>
  void take(size_t count) {
    off = ((off + 255) & ~size_t(255)) + count * sizeof(T);
  }
  size_t bytes() const { return off + 256; }
};

// ---------------------------------------------------- geometry args
// Kernel-argument bundle: the setup, the grid and the per-angle sin/cos
// table (host computed, identical for every kernel and for the oracle).
struct Geom {
  raptr_setup_t s;
  raptr_grid_t g;
  float sumRadii;
  int nvox;
  int nch;  // channel count; ids outside [0, nch) are rejected by the kernels
  float sn[RAPTR_MAX_ANGLES];
  float cs[RAPTR_MAX_ANGLES];
};

int make_geom(const raptr_setup_t* setup, const raptr_grid_t* grid, Geom* out);

// Channel index split, MeasurementSetupLinIndex.hpp:117-320.
__host__ __device__ inline void cnl_split(const raptr_setup_t& s, int cnl,
                                          int& ida, int& id0z, int& id0y,
                                          int& id1z, int& id1y) {
  int p4 = s.n1y * s.n1z * s.n0y * s.n0z;
  int p3 = s.n1y * s.n1z * s.n0y;
  int p2 = s.n1y * s.n1z;
  int p1 = s.n1y;
  int t = cnl;
  ida = t / p4;
  t %= p4;
  id0z = t / p3;
  t %= p3;
  id0y = t / p2;
  t %= p2;
  id1z = t / p1;
  id1y = t % p1;
}

// Voxel index split, VoxelGridLinIndex.hpp:114-205 (z fastest).
__host__ __device__ inline void vxl_split(const raptr_grid_t& g, int vxl,
                                          int& ix, int& iy, int& iz) {
  int pyz = g.n[1] * g.n[2];
  ix = vxl / pyz;
  int t = vxl % pyz;
  iy = t / g.n[2];
  iz = t % g.n[2];
}

// Pixel unit-cube -> world, MeasurementSetupTrafo2CartCoord.hpp:94-148
// (det 0) / :305-359 (det 1); products summed left to right.
__host__ __device__ inline void trafo(const raptr_setup_t& s, int det, float sn,
                                      float cs, int idz, int idy, float b0,
                                      float b1, float b2, float* out) {
  float e0 = s.segx, e1 = s.segy, e2 = s.segz;
  float p0, p1, p2;
  if (det == 0) {
    p0 = s.pos0x;
    p1 = (float(idy) - 0.5f * float(s.n0y - 1)) * e1;
    p2 = (float(idz) - 0.5f * float(s.n0z - 1)) * e2;
  } else {
    p0 = s.pos1x;
    p1 = (float(idy) - 0.5f * float(s.n1y - 1)) * e1;
    p2 = (float(idz) - 0.5f * float(s.n1z - 1)) * e2;
  }
  float t0 = cs * e0, t1 = 0.0f, t2 = sn * e2;
  float t3 = cs * (p0 - 0.5f * e0) + sn * (p2 - 0.5f * e2);
  float t4 = 0.0f, t5 = e1, t6 = 0.0f, t7 = p1 - 0.5f * e1;
  float t8 = -sn * e0, t9 = 0.0f, t10 = cs * e2;
  float t11 = -sn * (p0 - 0.5f * e0) + cs * (p2 - 0.5f * e2);
  out[0] = t0 * b0 + t1 * b1 + t2 * b2 + t3 * 1.0f;
  out[1] = t4 * b0 + t5 * b1 + t6 * b2 + t7 * 1.0f;
  out[2] = t8 * b0 + t9 * b1 + t10 * b2 + t11 * 1.0f;
}

// |v|, distancePointLine.h:38-43.
__host__ __device__ inline float absolute3(float a, float b, float c) {
  return sqrtf(a * a + b * b + c * c);
}

// Distance of p from the line through a and b, distancePointLine.h:87-101.
__host__ __device__ inline float distance_point_line(const float* a,
                                                     const float* b,
                                                     const float* p) {
  float ap0 = p[0] - a[0], ap1 = p[1] - a[1], ap2 = p[2] - a[2];
  float ab0 = b[0] - a[0], ab1 = b[1] - a[1], ab2 = b[2] - a[2];
  float c0 = ab1 * ap2 - ab2 * ap1;
  float c1 = ab2 * ap0 - ab0 * ap2;
  float c2 = ab0 * ap1 - ab1 * ap0;
  return absolute3(c0, c1, c2) / absolute3(ab0, ab1, ab2);
}

// Pixel centres of a channel (unit-cube point (0.5, 0.5, 0.5)).
__host__ __device__ inline void channel_centres(const Geom& G, int cnl,
                                                float* p0, float* p1) {
  int ida, id0z, id0y, id1z, id1y;
  cnl_split(G.s, cnl, ida, id0z, id0y, id1z, id1y);
  trafo(G.s, 0, G.sn[ida], G.cs[ida], id0z, id0y, 0.5f, 0.5f, 0.5f, p0);
  trafo(G.s, 1, G.sn[ida], G.cs[ida], id1z, id1y, 0.5f, 0.5f, 0.5f, p1);
}

// Voxel centre, getWorkqueue.hpp:91-94.
__host__ __device__ inline void voxel_centre(const raptr_grid_t& g, int ix,
                                             int iy, int iz, float* c) {
  c[0] = g.o[0] + (float(ix) + 0.5f) * g.d[0];
  c[1] = g.o[1] + (float(iy) + 0.5f) * g.d[1];
  c[2] = g.o[2] + (float(iz) + 0.5f) * g.d[2];
}

// Grid plane coordinate o + k*d (getSystemMatrixFromWorkqueue.cu:181-186).
__host__ __device__ inline float plane(const raptr_grid_t& g, int dim, int k) {
  return g.o[dim] + float(k) * g.d[dim];
}

// ---------------------------------------------------- Philox4x32-10
// Salmon et al., SC'11: two 32x32->64 multiplies per round, Weyl key bump.
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1,
                                              uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1,
                                              uint32_t* out) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

__host__ __device__ inline float u01(uint32_t x) {
  return float((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// RandRayGen / ReguRayGen (RayGenerators.hpp:36-99): six uniforms placed in
// the two pixel volumes by the trafos.
// element: the voxel of a RAND_ELEMENT ray (counter word 3 = element + 1);
// ignored by the other generators.
__host__ __device__ inline void make_ray(const Geom& G, int raygen,
                                         uint32_t seed, int cnl, int ray,
                                         float* out, int element = -1) {
  float u[6];
  if (raygen == RAPTR_RAYGEN_REGU) {
    for (int i = 0; i < 6; ++i) u[i] = 0.5f;
  } else {
    const uint32_t w3 = raygen == RAPTR_RAYGEN_RAND_ELEMENT ? uint32_t(element) + 1u : 0u;
    uint32_t r0[4], r1[4];
    philox4x32_10(uint32_t(cnl), uint32_t(ray), 0u, w3, seed, 0u, r0);
    philox4x32_10(uint32_t(cnl), uint32_t(ray), 1u, w3, seed, 0u, r1);
    u[0] = u01(r0[0]);
    u[1] = u01(r0[1]);
    u[2] = u01(r0[2]);
    u[3] = u01(r0[3]);
    u[4] = u01(r1[0]);
    u[5] = u01(r1[1]);
  }
  int ida, id0z, id0y, id1z, id1y;
  cnl_split(G.s, cnl, ida, id0z, id0y, id1z, id1y);
  trafo(G.s, 0, G.sn[ida], G.cs[ida], id0z, id0y, u[0], u[1], u[2], out);
  trafo(G.s, 1, G.sn[ida], G.cs[ida], id1z, id1y, u[3], u[4], u[5], out + 3);
}

// Slab intersection length, getSystemMatrixFromWorkqueue.cu:44-89 with the
// masked max/min of ChordsCalc_lowlevel.hpp:49-175.
__host__ __device__ inline float slab_length(const float* v, const float* r) {
  bool sx0 = r[0] != r[3], sx1 = r[1] != r[4], sx2 = r[2] != r[5];
  if (!(sx0 || sx1 || sx2)) return 0.0f;
  bool sx[3] = {sx0, sx1, sx2};
  float amin = 0.0f, amax = 0.0f;
  bool good = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!sx[d]) continue;
    float lo = (v[d] - r[d]) / (r[3 + d] - r[d]);
    float hi = (v[3 + d] - r[d]) / (r[3 + d] - r[d]);
    if (hi < lo) {
      float t = lo;
      lo = hi;
      hi = t;
    }
    amin = good ? (amin < lo ? lo : amin) : lo;
    amax = good ? (hi < amax ? hi : amax) : hi;
    good = true;
  }
  if (!(amin < amax)) return 0.0f;
  float dx = r[3] - r[0], dy = r[4] - r[1], dz = r[5] - r[2];
  return (amax - amin) * sqrtf(dx * dx + dy * dy + dz * dz);
}

// ------------------------------------------------------ device utils
constexpr int kWave = 64;

__device__ inline uint64_t lanemask_lt() {
  uint32_t lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return unsigned(g);
}

}  // namespace raptr
