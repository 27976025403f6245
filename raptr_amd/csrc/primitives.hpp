// primitives.hpp -- wave64/block scan helpers, invariant-divisor division,
// device-wide exclusive scan (own implementation, no library scan).
#pragma once

#include "common.hpp"

namespace raptr {

// Unsigned division by an invariant divisor (Granlund & Montgomery, "Division
// by invariant integers using multiplication", PLDI'94, fig. 4.1):
// q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(m, n),
// m = floor(2^32 * (2^l - d) / d) + 1, l = ceil(log2 d).
struct FastDiv {
  uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  f.s1 = l ? 1 : 0;
  f.s2 = l ? l - 1 : 0;
  return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t = __umulhi(f.m, n);
#else
  uint32_t t = uint32_t((uint64_t(f.m) * n) >> 32);
#endif
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

template <class T>
__device__ inline T wave_inclusive_scan(T v) {
  int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T n = __shfl_up(v, off, kWave);
    if (lane >= off) v += n;
  }
  return v;
}

template <class T>
__device__ inline T wave_reduce_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Exclusive block scan (BLOCK threads); lds needs BLOCK/64 slots.  Returns
// the exclusive prefix, the block total in *total.  Contains barriers.
template <class T, int BLOCK>
__device__ inline T block_exclusive_scan(T v, T* lds, T* total) {
  constexpr int NW = BLOCK / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int w = threadIdx.x / kWave;
  T inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) lds[w] = inc;
  __syncthreads();
  T pre = T(0), tot = T(0);
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    T t = lds[i];
    if (i < w) pre += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return pre + inc - v;
}

// Device-wide exclusive scan of n values (TIn summed into TOut); optional
// device pointer receives the grand total.  Workspace from
// scan_workspace_bytes(n).
size_t scan_workspace_bytes(int64_t n);
int scan_exclusive_i32_i64(const int* in, int64_t* out, int64_t n,
                           int64_t* d_total, void* ws, size_t ws_bytes,
                           hipStream_t st);
int scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n,
                       int64_t* d_total, void* ws, size_t ws_bytes,
                       hipStream_t st);
int scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n,
                       uint32_t* d_total, void* ws, size_t ws_bytes,
                       hipStream_t st);

// Stable LSD radix sort of (key, value) pairs by bits [0, bits) of the key;
// ping-pong buffers k0/k1, v0/v1; histogram and scan scratch come from cv.
// *result = 0 or 1: which buffer pair holds the sorted data.
int radix_sort_pairs(uint64_t* k0, uint64_t* k1, uint32_t* v0, uint32_t* v1, int64_t n,
                     int bits, hipStream_t st, int* result, Carver& cv);

// Native vector types (accepted by the non-temporal load builtin).
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef int vi4 __attribute__((ext_vector_type(4)));
typedef unsigned int vu2 __attribute__((ext_vector_type(2)));

}  // namespace raptr
