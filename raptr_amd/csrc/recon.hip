// recon.hip -- device-resident MLEM reconstruction plan on gfx950.
//
// A plan holds the system matrix of the measured channels in two layouts
// tuned for the two passes of an MLEM iteration (reco.cu:246-303):
//  * forward: CSR with every row padded to a multiple of 4 entries and
//    16-byte aligned, so each lane streams 4 values and 4 column indices per
//    load (row sums are short, ~360 entries, and scalar traversal loses a
//    quarter of the bandwidth to row overheads);
//  * backward: the transpose cut into row blocks (channel ranges) whose ratio
//    vector e fits in one XCD's L2.  Each block is one launch, so at any time
//    all waves gather from the same L2-sized slice of e; the column sums are
//    carried across blocks in a fixed order (deterministic).
// Values are FP32, or 16-bit with a per-row FP32 scale: FP16 or BF16 of
// value/rowmax, or unsigned 16-bit fixed point round(65535*value/rowmax)
// (absolute error <= rowmax/131070, uniform over the row).  The scale is
// applied once per row in the forward pass and folded into the ratio
// (e'_i = scale_i * e_i) for the backward pass.  Vectors hold F frames
// per index ([index][frame]); one launch handles all frames, so the matrix
// stream is shared by the batch.  The normalisation sums are FP64.
#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>
#include <vector>

#include "recon.hpp"

namespace raptr {
namespace {

typedef unsigned short vus4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------- layout construction
__global__ void k_padded_len(const int64_t* rp, int m, int64_t* plen) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    plen[i] = (rp[i + 1] - rp[i] + 3) & ~int64_t(3);
}

__global__ void k_set_last(int64_t* p, int m, const int64_t* total) { p[m] = *total; }

// One wave per row: row maximum -> scale (mixed), copy with zero padding.
// interleave chunks: forward rows for 32 lanes x 4 entries per step, column
// segments for 64 lanes x 4 entries per step
constexpr int kIlvFwd = 128;
constexpr int kIlvBwd = 256;

// Single-frame plans store every padded row (forward) and column segment
// (transpose) interleaved in chunks of C entries: the chunk's L entries
// (L = C, or the remainder, a multiple of 4) are laid out as L/4 groups of
// 4 with group g holding sorted entries g, g + L/4, g + 2L/4, g + 3L/4.  A
// wave that vector-loads one group per lane then gathers sorted-consecutive
// entries in each of its 4 gather instructions; consecutive entries of a
// row (voxels along z) or of a column (neighbouring channels) share cache
// lines, so the gathers touch about a quarter of the lines.  C = 0: no
// interleave.  Returns the position of sorted entry o in a segment of
// padded length P.
__device__ __forceinline__ int64_t ilv(int64_t o, int64_t P, int C) {
  if (C == 0) return o;
  const int64_t c = o / C, base = c * C;
  const int64_t L = P - base < C ? P - base : C;
  const int64_t q = L / 4, r = o - base;
  return base + 4 * (r % q) + r / q;
}

__global__ __launch_bounds__(256) void k_pad_copy(const int64_t* rp, const int* col,
                                                  const float* val, int m, int vt,
                                                  const int64_t* prp, int* pcol, void* pval,
                                                  float* scale, int C) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    int64_t b = rp[i], e = rp[i + 1], pb = prp[i], pe = prp[i + 1];
    float mx = 0.0f;
    if (vt != kF32) {
      for (int64_t k = b + lane; k < e; k += kWave) mx = fmaxf(mx, fabsf(val[k]));
      for (int off = kWave / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
    }
    float sc = mx > 0.0f ? mx : 1.0f;
    if (vt != kF32 && lane == 0) scale[i] = vt == kU16 ? sc * (1.0f / 65535.0f) : sc;
    int lastc = e > b ? col[e - 1] : 0;
    for (int64_t k = lane; k < pe - pb; k += kWave) {
      bool real = b + k < e;
      int c = real ? col[b + k] : lastc;
      float v = real ? val[b + k] : 0.0f;
      const int64_t d = pb + ilv(k, pe - pb, C);
      pcol[d] = c;
      if (vt == kF32)
        static_cast<float*>(pval)[d] = v;
      else
        static_cast<uint16_t*>(pval)[d] = f2h(v / sc, vt);
    }
  }
}

// per-row scales of 16-bit plans (as k_pad_copy computes them) without the
// padded copy: one wave per row
__global__ __launch_bounds__(256) void k_row_scales(const int64_t* rp, const float* val, int m,
                                                    int vt, float* scale) {
  const int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    float mx = 0.0f;
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += kWave) mx = fmaxf(mx, fabsf(val[k]));
    for (int off = kWave / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
    const float sc = mx > 0.0f ? mx : 1.0f;
    if (lane == 0) scale[i] = vt == kU16 ? sc * (1.0f / 65535.0f) : sc;
  }
}

// transpose keys: (row block << colbits) | col, payload = entry index
__global__ void k_bt_keys(const int64_t* rp, const int* col, int m, int blk_rows, int colbits,
                          uint64_t* keys, uint32_t* idx) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    uint64_t hi = uint64_t(i / blk_rows) << colbits;
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += kWave) {
      keys[k] = hi | uint64_t(uint32_t(col[k]));
      idx[k] = uint32_t(k);
    }
  }
}

__global__ void k_expand_rows32(const int64_t* rp, int m, uint32_t* rowid) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw)
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += kWave) rowid[k] = uint32_t(i);
}

// padded segment lengths of the blocked transpose: segment (b, j) of
// column j in row block b padded to a multiple of 4 entries; the (b, n)
// slot is empty so an exclusive scan gives every block's start
__global__ void k_bt_padlen(const int64_t* uptr, int64_t total, int n, int64_t* plen) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t j = t % (n + 1);
    plen[t] = j < n ? ((uptr[t + 1] - uptr[t] + 3) & ~int64_t(3)) : 0;
  }
}

// sorted entry j -> its place in the padded segment of its (block, column)
__global__ void k_bt_scatter(const uint64_t* keys, const uint32_t* idx, int64_t nnz, int n,
                             int colbits, const int64_t* uptr, const int64_t* pptr,
                             const uint32_t* rowid, const float* val, const float* scale,
                             int vt, int* brow, void* bval, int C) {
  const uint64_t cmask = (uint64_t(1) << colbits) - 1;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nnz;
       j += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t key = keys[j];
    const int64_t t = int64_t(key >> colbits) * (n + 1) + int64_t(key & cmask);
    const int64_t d = pptr[t] + ilv(j - uptr[t], pptr[t + 1] - pptr[t], C);
    const uint32_t k = idx[j];
    const uint32_t r = rowid[k];
    brow[d] = int(r);
    const float v = val[k];
    if (vt == kF32)
      static_cast<float*>(bval)[d] = v;
    else
      static_cast<uint16_t*>(bval)[d] =
          f2h(vt == kU16 ? v / (scale[r] * 65535.0f) : v / scale[r], vt);
  }
}

// padding entries: value 0, row = the block's first row (inside the block)
__global__ void k_bt_pad(const int64_t* uptr, const int64_t* pptr, int64_t total, int n,
                         int blk_rows, int* brow, void* bval, int vt, int C) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = t / (n + 1), j = t - b * (n + 1);
    if (j >= n) continue;
    const int64_t len = uptr[t + 1] - uptr[t];
    const int64_t plen = (len + 3) & ~int64_t(3);
    for (int64_t o = len; o < plen; ++o) {
      const int64_t p = pptr[t] + ilv(o, plen, C);
      brow[p] = int(b * blk_rows);
      if (vt == kF32)
        static_cast<float*>(bval)[p] = 0.0f;
      else
        static_cast<uint16_t*>(bval)[p] = 0;
    }
  }
}

// segment pointers of the blocked transpose: ptr[b*(n+1) + j] = first entry
// with key >= (b << colbits | j)
__global__ void k_bt_ptr(const uint64_t* keys, int64_t nnz, int nblk, int n, int colbits,
                         int64_t* ptr) {
  int64_t total = int64_t(nblk) * (n + 1);
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    int64_t b = t / (n + 1), j = t - b * (n + 1);
    uint64_t key = (uint64_t(b) << colbits) + uint64_t(j);
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    ptr[t] = lo;
  }
}

// ------------------------------------------------------ MLEM kernels
// Frames per launch FR (1 or a power of two up to 64).  A wave handles one
// row (forward) / column (backward); its 64 lanes form 64/FR entry slots
// x FR frames.  F = 1 uses vector loads of 4 entries per lane (forward).


// Single frame: each lane streams U groups of 4 entries per step (vector
// loads), so 4*U gathers of x are in flight per lane.
template <int VT, int LANES, int U>
__global__ __launch_bounds__(256) void k_fwd1(int64_t m, const int64_t* __restrict__ rp,
                                              const int* __restrict__ col,
                                              const typename Val<VT>::T* __restrict__ val,
                                              const float* __restrict__ rscale,
                                              const float* __restrict__ x,
                                              const float* __restrict__ y,
                                              float* __restrict__ e) {
  const int g = threadIdx.x % LANES;
  const int64_t i = int64_t(blockIdx.x) * (256 / LANES) + threadIdx.x / LANES;
  if (i >= m) return;
  const int64_t b = rp[i], en = rp[i + 1];  // multiples of 4
  float acc = 0.0f;
  int64_t k = b + 4 * g;
  for (; k + int64_t(4) * LANES * (U - 1) < en; k += int64_t(4) * LANES * U) {
    vi4 c[U];
    float v[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(col + k + 4 * LANES * u));
      load4<VT>(val, k + 4 * LANES * u, v[u][0], v[u][1], v[u][2], v[u][3]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      acc = fmaf(v[u][0], x[c[u].x], acc);
      acc = fmaf(v[u][1], x[c[u].y], acc);
      acc = fmaf(v[u][2], x[c[u].z], acc);
      acc = fmaf(v[u][3], x[c[u].w], acc);
    }
  }
  for (; k < en; k += int64_t(4) * LANES) {
    vi4 c = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(col + k));
    float v0, v1, v2, v3;
    load4<VT>(val, k, v0, v1, v2, v3);
    acc = fmaf(v0, x[c.x], acc);
    acc = fmaf(v1, x[c.y], acc);
    acc = fmaf(v2, x[c.z], acc);
    acc = fmaf(v3, x[c.w], acc);
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  if (g == 0) {
    float sc = VT == kF32 ? 1.0f : rscale[i];
    float yt = sc * acc;
    float r = yt != 0.0f ? y[i] / yt : 0.0f;
    e[i] = VT == kF32 ? r : sc * r;  // ratio folded with the row scale
  }
}

// Single frame with 16-bit column low parts: as k_fwd1, with the 4 columns
// of a lane's 4 stored entries rebuilt from their low parts and the row's
// split points.  Stored entry d of a padded row of length P sits in chunk
// cb = d - d % C; with L = min(C, P - cb) and q = L/4 its sorted position
// is cb + (d % 4)*q + (d - cb)/4 (the inverse of the interleave).
template <int VT, int LANES, int U>
__global__ __launch_bounds__(256) void k_fwd1c(int64_t m, const int64_t* __restrict__ rp,
                                               const uint16_t* __restrict__ lo,
                                               const uint64_t* __restrict__ split,
                                               const typename Val<VT>::T* __restrict__ val,
                                               const float* __restrict__ rscale,
                                               const float* __restrict__ x,
                                               const float* __restrict__ y,
                                               float* __restrict__ e) {
  const int g = threadIdx.x % LANES;
  const int64_t i = int64_t(blockIdx.x) * (256 / LANES) + threadIdx.x / LANES;
  if (i >= m) return;
  const int64_t b = rp[i], en = rp[i + 1];  // multiples of 4
  const int P = int(en - b);
  const uint64_t sp = split[i];
  const int s1 = int(sp & 0xffff), s2 = int((sp >> 16) & 0xffff), s3 = int((sp >> 32) & 0xffff);
  float acc = 0.0f;
  auto cols4 = [&](int64_t k, int (&c)[4]) {
    const vus4 h = __builtin_nontemporal_load(reinterpret_cast<const vus4*>(lo + k));
    const int d = int(k - b);
    const int cb = d - d % kIlvFwd;
    const int L = P - cb < kIlvFwd ? P - cb : kIlvFwd;
    const int q = L >> 2;
    const int o0 = cb + ((d - cb) >> 2);
    const unsigned short hl[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = o0 + j * q;
      const int hi = int(o >= s1) + int(o >= s2) + int(o >= s3);
      c[j] = (hi << 16) | int(hl[j]);
    }
  };
  int64_t k = b + 4 * g;
  for (; k + int64_t(4) * LANES * (U - 1) < en; k += int64_t(4) * LANES * U) {
    int c[U][4];
    float v[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cols4(k + 4 * LANES * u, c[u]);
      load4<VT>(val, k + 4 * LANES * u, v[u][0], v[u][1], v[u][2], v[u][3]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc = fmaf(v[u][j], x[c[u][j]], acc);
  }
  for (; k < en; k += int64_t(4) * LANES) {
    int c[4];
    float v[4];
    cols4(k, c);
    load4<VT>(val, k, v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc = fmaf(v[j], x[c[j]], acc);
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  if (g == 0) {
    float sc = VT == kF32 ? 1.0f : rscale[i];
    float yt = sc * acc;
    float r = yt != 0.0f ? y[i] / yt : 0.0f;
    e[i] = VT == kF32 ? r : sc * r;  // ratio folded with the row scale
  }
}

// Single frame with x staged in LDS: a workgroup takes `rows` (the plan's
// union group size) consecutive plan rows (neighbouring lines through the same voxels in the
// locality row order), copies x over the sorted union of their columns into
// LDS, then runs k_fwd1c's loop with each entry's 16-bit index into that
// union.  The matrix stream stays 6 bytes per FP32 entry; the gathers read
// LDS instead of the L1 cache, whose tag lookups bound k_fwd1/k_fwd1c.  The
// stream uses ordinary cached loads: rows share the cache lines at their
// boundaries, which non-temporal loads fetched twice (12.2 -> 10.3 GB of
// HBM reads per pass at full size; x lives in LDS, so nothing is evicted).
constexpr int kLocMaxUnion = 16384;  // 64 KB of LDS per workgroup
// union group sizes a plan may take (plan rows per group), largest first
constexpr int kLocRowsChoices[] = {256, 128, 64, 32, 16};

template <int VT, int LANES, int U>
__global__ __launch_bounds__(256) void k_fwd1l(int64_t m, int rows,
                                               const int64_t* __restrict__ rp,
                                               const uint16_t* __restrict__ lcol,
                                               const int64_t* __restrict__ up,
                                               const int* __restrict__ ucol,
                                               const typename Val<VT>::T* __restrict__ val,
                                               const float* __restrict__ rscale,
                                               const float* __restrict__ x,
                                               const float* __restrict__ y,
                                               float* __restrict__ e) {
  extern __shared__ float xs[];
  const int64_t grp = blockIdx.x;
  const int64_t ub = up[grp];
  const int nu = int(up[grp + 1] - ub);
  for (int j = threadIdx.x; j < nu; j += 256) xs[j] = x[ucol[ub + j]];
  __syncthreads();
  const int g = threadIdx.x % LANES;
  const int64_t r0 = grp * rows, r1 = r0 + rows < m ? r0 + rows : m;
  for (int64_t i = r0 + threadIdx.x / LANES; i < r1; i += 256 / LANES) {
    const int64_t b = rp[i], en = rp[i + 1];  // multiples of 4
    float acc = 0.0f;
    int64_t k = b + 4 * g;
    for (; k + int64_t(4) * LANES * (U - 1) < en; k += int64_t(4) * LANES * U) {
      vus4 c[U];
      float v[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = *reinterpret_cast<const vus4*>(lcol + k + 4 * LANES * u);
        load4c<VT>(val, k + 4 * LANES * u, v[u][0], v[u][1], v[u][2], v[u][3]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        acc = fmaf(v[u][0], xs[c[u].x], acc);
        acc = fmaf(v[u][1], xs[c[u].y], acc);
        acc = fmaf(v[u][2], xs[c[u].z], acc);
        acc = fmaf(v[u][3], xs[c[u].w], acc);
      }
    }
    for (; k < en; k += int64_t(4) * LANES) {
      const vus4 c = *reinterpret_cast<const vus4*>(lcol + k);
      float v0, v1, v2, v3;
      load4c<VT>(val, k, v0, v1, v2, v3);
      acc = fmaf(v0, xs[c.x], acc);
      acc = fmaf(v1, xs[c.y], acc);
      acc = fmaf(v2, xs[c.z], acc);
      acc = fmaf(v3, xs[c.w], acc);
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    if (g == 0) {
      float sc = VT == kF32 ? 1.0f : rscale[i];
      float yt = sc * acc;
      float r = yt != 0.0f ? y[i] / yt : 0.0f;
      e[i] = VT == kF32 ? r : sc * r;  // ratio folded with the row scale
    }
  }
}

// Frame vectors: a lane carries V = min(F, 4) consecutive frames of one
// matrix entry (one 4/8/16-byte gather), a wave's 64 lanes are SLOTS entry
// slots x Q frame vectors.  The value and index of an entry are loaded by
// the Q lanes of its slot.
template <int FR>
struct FrameVec {
  static constexpr int V = FR >= 4 ? 4 : FR;
  static constexpr int Q = FR / V;
  static constexpr int SLOTS = kWave / Q;
};

template <int V>
__device__ __forceinline__ void load_frames(const float* __restrict__ p, float (&o)[V]) {
  if constexpr (V == 4) {
    vf4 t = *reinterpret_cast<const vf4*>(p);
    o[0] = t.x, o[1] = t.y, o[2] = t.z, o[3] = t.w;
  } else if constexpr (V == 2) {
    typedef float vf2 __attribute__((ext_vector_type(2)));
    vf2 t = *reinterpret_cast<const vf2*>(p);
    o[0] = t.x, o[1] = t.y;
  } else {
    o[0] = *p;
  }
}

template <int V>
__device__ __forceinline__ void store_frames(float* __restrict__ p, const float (&o)[V]) {
  if constexpr (V == 4) {
    *reinterpret_cast<vf4*>(p) = vf4{o[0], o[1], o[2], o[3]};
  } else if constexpr (V == 2) {
    typedef float vf2 __attribute__((ext_vector_type(2)));
    *reinterpret_cast<vf2*>(p) = vf2{o[0], o[1]};
  } else {
    *p = o[0];
  }
}

// XCD-contiguous block index: consecutive workgroups of one XCD (dispatch
// is round-robin over the 8 XCDs) take consecutive logical blocks, so the
// rows an XCD works on at any moment are neighbours and share its L2.
__device__ __forceinline__ int64_t xcd_block(int64_t bx, int64_t G) {
  const int64_t xcd = bx % 8, q = G / 8, r8 = G % 8;
  return xcd * q + (xcd < r8 ? xcd : r8) + bx / 8;
}

// F frames forward + ratio: one wave per row.  Each entry slot takes 4
// consecutive entries per step (one 16-byte value load and one 16-byte
// index load for the 4, rows are padded to multiples of 4), then gathers
// the 4 entries' frame vectors; U groups of 4 are in flight per step.  This
// halves the vector-memory instructions per entry against one entry per
// slot, and those instructions (not bytes) bound the batched passes.
template <int VT, int FR, int U>
__global__ __launch_bounds__(256) void k_fwdF(int64_t m, const int64_t* __restrict__ rp,
                                              const int* __restrict__ col,
                                              const typename Val<VT>::T* __restrict__ val,
                                              const float* __restrict__ rscale,
                                              const float* __restrict__ x,
                                              const float* __restrict__ y,
                                              float* __restrict__ e) {
  using FV = FrameVec<FR>;
  constexpr int V = FV::V, Q = FV::Q, SLOTS = FV::SLOTS;
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane % Q, slot = lane / Q;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * (256 / kWave) + threadIdx.x / kWave;
  if (i >= m) return;
  const int64_t b = rp[i], en = rp[i + 1];  // multiples of 4
  float acc[V];
#pragma unroll
  for (int v = 0; v < V; ++v) acc[v] = 0.0f;
  const float* xq = x + V * q;
  constexpr int64_t STEP = 4 * SLOTS;
  int64_t k = b + 4 * slot;
  for (; k + STEP * (U - 1) < en; k += STEP * U) {
    float w[U][4];
    vi4 c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      load4<VT>(val, k + STEP * u, w[u][0], w[u][1], w[u][2], w[u][3]);
      c[u] = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(col + k + STEP * u));
    }
    float g[U][4][V];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      load_frames<V>(xq + int64_t(c[u].x) * FR, g[u][0]);
      load_frames<V>(xq + int64_t(c[u].y) * FR, g[u][1]);
      load_frames<V>(xq + int64_t(c[u].z) * FR, g[u][2]);
      load_frames<V>(xq + int64_t(c[u].w) * FR, g[u][3]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = fmaf(w[u][j], g[u][j][v], acc[v]);
  }
  for (; k < en; k += STEP) {
    float w[4];
    load4<VT>(val, k, w[0], w[1], w[2], w[3]);
    const vi4 c = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(col + k));
    float g[4][V];
    load_frames<V>(xq + int64_t(c.x) * FR, g[0]);
    load_frames<V>(xq + int64_t(c.y) * FR, g[1]);
    load_frames<V>(xq + int64_t(c.z) * FR, g[2]);
    load_frames<V>(xq + int64_t(c.w) * FR, g[3]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] = fmaf(w[j], g[j][v], acc[v]);
  }
#pragma unroll
  for (int off = kWave / 2; off >= Q; off >>= 1)
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  if (slot == 0) {
    const float sc = VT == kF32 ? 1.0f : rscale[i];
    float yy[V], r[V];
    load_frames<V>(y + i * FR + V * q, yy);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      float yt = sc * acc[v];
      float t = (yt != 0.0f && yy[v] != 0.0f) ? yy[v] / yt : 0.0f;
      r[v] = VT == kF32 ? t : sc * t;  // ratio folded with the row scale
    }
    store_frames<V>(e + i * FR + V * q, r);
  }
}

// End of a backward launch for column j: reduce over entry slots (done by
// the caller), carry c across launches, and in the last launch apply
// x' = x*c/(s + (s==0)) and write the workgroup's FP64 partial sums.
template <int FR>
__device__ __forceinline__ void bwd_finish(float (&acc)[FrameVec<FR>::V], int j, int n, int q,
                                           int slot, int w, int lb, int first, int last,
                                           float* __restrict__ c, const float* __restrict__ x,
                                           const float* __restrict__ s,
                                           float* __restrict__ xx,
                                           double* __restrict__ partials) {
  constexpr int V = FrameVec<FR>::V;
  double part[V];
#pragma unroll
  for (int v = 0; v < V; ++v) part[v] = 0.0;
  if (j < n && slot == 0) {
    const int64_t o = int64_t(j) * FR + V * q;
    float cv[V];
    if (first) {
#pragma unroll
      for (int v = 0; v < V; ++v) cv[v] = acc[v];
    } else {
      load_frames<V>(c + o, cv);
#pragma unroll
      for (int v = 0; v < V; ++v) cv[v] += acc[v];
    }
    if (!last) {
      store_frames<V>(c + o, cv);
    } else {
      float d = s[j];
      d += float(d == 0.0f);
      float xv[V], nv[V];
      load_frames<V>(x + o, xv);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        nv[v] = xv[v] * cv[v] / d;
        part[v] = double(nv[v]);
      }
      store_frames<V>(xx + o, nv);
    }
  }
  if (last) {
    // per-frame block sums: lanes with slot 0 of each wave hold frames
    // V*q .. V*q+V-1
    __shared__ double red[256 / kWave][64];
    if (slot == 0) {
#pragma unroll
      for (int v = 0; v < V; ++v) red[w][V * q + v] = part[v];
    }
    __syncthreads();
    if (threadIdx.x < FR) {
      double t = 0.0;
      for (int ww = 0; ww < 256 / kWave; ++ww) t += red[ww][threadIdx.x];
      partials[int64_t(lb) * FR + threadIdx.x] = t;
    }
  }
}

// Backward over nbl consecutive row blocks: c[j][f] (+)= sum over those
// blocks' entries of column j of v * e'[row][f]; the launch that ends with
// the last block applies the update xx = x * c / (s + (s==0)) and writes
// FP64 partial sums per block/frame.  One wave per column, lanes as in the
// forward; column segments are padded to multiples of 4 (value 0, row in
// the block) and read 4 entries per slot and step.
template <int VT, int FR, int UNROLL>
__global__ __launch_bounds__(256) void k_bwd(int n, const int64_t* __restrict__ ptr, int nbl,
                                             const int* __restrict__ brow,
                                             const typename Val<VT>::T* __restrict__ val,
                                             const float* __restrict__ e, float* __restrict__ c,
                                             int first, int last, const float* __restrict__ x,
                                             const float* __restrict__ s,
                                             float* __restrict__ xx,
                                             double* __restrict__ partials) {
  using FV = FrameVec<FR>;
  constexpr int V = FV::V, Q = FV::Q, SLOTS = FV::SLOTS;
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane % Q, slot = lane / Q;
  const int w = threadIdx.x / kWave;
  const int lb = int(xcd_block(blockIdx.x, gridDim.x));  // XCD-contiguous column ranges
  const int j = lb * (256 / kWave) + w;
  float acc[V];
#pragma unroll
  for (int v = 0; v < V; ++v) acc[v] = 0.0f;
  // segment bounds of column j in the launch's nbl row blocks, one block
  // per lane, loaded together and broadcast block by block
  int64_t pb = 0, pe = 0;
  if (j < n && lane < nbl) {
    pb = ptr[int64_t(lane) * (n + 1) + j];
    pe = ptr[int64_t(lane) * (n + 1) + j + 1];
  }
  const float* eq = e + V * q;
  constexpr int64_t STEP = 4 * SLOTS;
  for (int bb = 0; bb < nbl && j < n; ++bb) {
    const int64_t b = __shfl(pb, bb, kWave), en = __shfl(pe, bb, kWave);
    int64_t k = b + 4 * slot;
    for (; k + STEP * (UNROLL - 1) < en; k += STEP * UNROLL) {
      float wv[UNROLL][4];
      vi4 rr[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        load4<VT>(val, k + STEP * u, wv[u][0], wv[u][1], wv[u][2], wv[u][3]);
        rr[u] = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(brow + k + STEP * u));
      }
      float g[UNROLL][4][V];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        load_frames<V>(eq + int64_t(rr[u].x) * FR, g[u][0]);
        load_frames<V>(eq + int64_t(rr[u].y) * FR, g[u][1]);
        load_frames<V>(eq + int64_t(rr[u].z) * FR, g[u][2]);
        load_frames<V>(eq + int64_t(rr[u].w) * FR, g[u][3]);
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
          for (int v = 0; v < V; ++v) acc[v] = fmaf(wv[u][jj], g[u][jj][v], acc[v]);
    }
    for (; k < en; k += STEP) {
      float wv[4];
      load4<VT>(val, k, wv[0], wv[1], wv[2], wv[3]);
      const vi4 rr = __builtin_nontemporal_load(reinterpret_cast<const vi4*>(brow + k));
      float g[4][V];
      load_frames<V>(eq + int64_t(rr.x) * FR, g[0]);
      load_frames<V>(eq + int64_t(rr.y) * FR, g[1]);
      load_frames<V>(eq + int64_t(rr.z) * FR, g[2]);
      load_frames<V>(eq + int64_t(rr.w) * FR, g[3]);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = fmaf(wv[jj], g[jj][v], acc[v]);
    }
  }
  if (j < n) {
#pragma unroll
    for (int off = kWave / 2; off >= Q; off >>= 1)
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  }
  bwd_finish<FR>(acc, j, n, q, slot, w, lb, first, last, c, x, s, xx, partials);
}

// 64-frame batches (16-lane frame groups): the 16 lanes of an entry slot
// form one DPP row.  A wave loads the values and indices of 64 consecutive
// entries with one 4-byte load each (lane (slot, q) holds entry
// 16*(q/4) + 4*slot + q%4 of the chunk) and broadcasts them inside each row
// with row_newbcast, so a chunk costs 2 index/value loads instead of 8; the
// 16 gathers of 1 KiB each (4 entries x 256 B per instruction) remain.
__device__ __forceinline__ int row_bcast_i(int v, int lane16) {
  switch (lane16) {
#define RAPTR_RB(L) \
  case L:           \
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + L, 0xf, 0xf, false);
    RAPTR_RB(0) RAPTR_RB(1) RAPTR_RB(2) RAPTR_RB(3) RAPTR_RB(4) RAPTR_RB(5) RAPTR_RB(6)
    RAPTR_RB(7) RAPTR_RB(8) RAPTR_RB(9) RAPTR_RB(10) RAPTR_RB(11) RAPTR_RB(12) RAPTR_RB(13)
    RAPTR_RB(14) RAPTR_RB(15)
#undef RAPTR_RB
  }
  return v;
}

// one 64-entry chunk [k0, k0 + 64) of a padded segment ending at en: the
// lane's frame vector gathers from base + index*64 are accumulated
template <int VT>
__device__ __forceinline__ void chunk64(const typename Val<VT>::T* __restrict__ val,
                                        const int* __restrict__ idx, int64_t k0, int64_t en,
                                        int slot, int q, const float* __restrict__ base,
                                        float (&acc)[4]) {
  const int64_t kk = k0 + 16 * (q >> 2) + 4 * slot + (q & 3);
  const bool in = kk < en;
  const float wv = in ? Val<VT>::get(val, kk) : 0.0f;
  const int cv = in ? __builtin_nontemporal_load(idx + kk) : 0;
  float g[4][4][4];
  float w[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      w[t][jj] = __int_as_float(row_bcast_i(__float_as_int(wv), 4 * t + jj));
      const int c = row_bcast_i(cv, 4 * t + jj);
      if (k0 + 16 * t + 4 * slot < en) {
        load_frames<4>(base + int64_t(c) * 64, g[t][jj]);
      } else {
        g[t][jj][0] = g[t][jj][1] = g[t][jj][2] = g[t][jj][3] = 0.0f;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = fmaf(w[t][jj], g[t][jj][v], acc[v]);
}

template <int VT>
__global__ __launch_bounds__(256) void k_fwd64(int64_t m, const int64_t* __restrict__ rp,
                                               const int* __restrict__ col,
                                               const typename Val<VT>::T* __restrict__ val,
                                               const float* __restrict__ rscale,
                                               const float* __restrict__ x,
                                               const float* __restrict__ y,
                                               float* __restrict__ e) {
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane & 15, slot = lane >> 4;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * (256 / kWave) + threadIdx.x / kWave;
  if (i >= m) return;
  const int64_t b = rp[i], en = rp[i + 1];
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const float* xq = x + 4 * q;
  for (int64_t k0 = b; k0 < en; k0 += 64) chunk64<VT>(val, col, k0, en, slot, q, xq, acc);
#pragma unroll
  for (int off = 32; off >= 16; off >>= 1)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  if (slot == 0) {
    const float sc = VT == kF32 ? 1.0f : rscale[i];
    float yy[4], r[4];
    load_frames<4>(y + i * 64 + 4 * q, yy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float yt = sc * acc[v];
      float t = (yt != 0.0f && yy[v] != 0.0f) ? yy[v] / yt : 0.0f;
      r[v] = VT == kF32 ? t : sc * t;
    }
    store_frames<4>(e + i * 64 + 4 * q, r);
  }
}

template <int VT>
__global__ __launch_bounds__(256) void k_bwd64(int n, const int64_t* __restrict__ ptr, int nbl,
                                               const int* __restrict__ brow,
                                               const typename Val<VT>::T* __restrict__ val,
                                               const float* __restrict__ e,
                                               float* __restrict__ c, int first, int last,
                                               const float* __restrict__ x,
                                               const float* __restrict__ s,
                                               float* __restrict__ xx,
                                               double* __restrict__ partials) {
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane & 15, slot = lane >> 4;
  const int w = threadIdx.x / kWave;
  const int lb = int(xcd_block(blockIdx.x, gridDim.x));
  const int j = lb * (256 / kWave) + w;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int64_t pb = 0, pe = 0;
  if (j < n && lane < nbl) {
    pb = ptr[int64_t(lane) * (n + 1) + j];
    pe = ptr[int64_t(lane) * (n + 1) + j + 1];
  }
  const float* eq = e + 4 * q;
  for (int bb = 0; bb < nbl && j < n; ++bb) {
    const int64_t b = __shfl(pb, bb, kWave), en = __shfl(pe, bb, kWave);
    for (int64_t k0 = b; k0 < en; k0 += 64) chunk64<VT>(val, brow, k0, en, slot, q, eq, acc);
  }
  if (j < n) {
#pragma unroll
    for (int off = 32; off >= 16; off >>= 1)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  }
  bwd_finish<64>(acc, j, n, q, slot, w, lb, first, last, c, x, s, xx, partials);
}

// ------------------------------------------- grouped 64-frame layouts
// A 64-frame pass gathers a 256-byte frame vector per matrix entry, and
// those gathers (the cache lines they pull from L2), not the matrix bytes,
// bound it.  Neighbouring plan rows (lines with one midpoint and
// neighbouring slopes) and neighbouring voxels share most of their entries,
// so the grouped layouts store groups of G rows (forward) or of G voxels of
// a brick (backward) as one ascending union list with G values per union
// entry (0 where a member has no entry).  One gathered frame vector then
// serves all G members.  Every member's sum still runs over its own entries
// in ascending order, with exact zero terms between them.
constexpr int kGrpPad = 64;    // union entries of tail padding (>= 2 x the largest unroll)
// union lists are padded (value 0, the list's last key) to multiples of
// kAlignFwd entries (row groups) / kAlignBwd entries (block x brick lists,
// which are short)
constexpr int kAlignFwd = 32;
constexpr int kAlignBwd = 8;

// forward groups: list s of group g is CSR row G*g + s
template <int G>
struct RowLists {
  const int64_t* rp;
  const int* col;
  const float* val;
  int64_t m;
  __device__ void bounds(int64_t g, int64_t (&b)[G], int64_t (&e)[G]) const {
#pragma unroll
    for (int s = 0; s < G; ++s) {
      const int64_t r = int64_t(G) * g + s;
      b[s] = r < m ? rp[r] : 0;
      e[s] = r < m ? rp[r + 1] : 0;
    }
  }
  __device__ int key(int64_t k) const { return col[k]; }
  __device__ float value(int64_t k) const { return val[k]; }
  __device__ int64_t row(int64_t g, int s, int) const { return int64_t(G) * g + s; }
};

// backward groups: list s of group t = (block b, brick h) is the block's
// sorted transpose segment (rows ascending) of column cols[h*G + s]
// (-1: the brick has no such member)
template <int G>
struct ColLists {
  const int64_t* uptr;
  const uint32_t* idx;
  const uint32_t* rowid;
  const float* val;
  const int* cols;
  int n, ngc;
  __device__ void bounds(int64_t t, int64_t (&b)[G], int64_t (&e)[G]) const {
    const int64_t blk = t / ngc, h = t - blk * ngc;
#pragma unroll
    for (int s = 0; s < G; ++s) {
      const int j = cols[h * G + s];
      const int64_t o = blk * (n + 1) + j;
      b[s] = j >= 0 ? uptr[o] : 0;
      e[s] = j >= 0 ? uptr[o + 1] : 0;
    }
  }
  __device__ int key(int64_t k) const { return int(rowid[idx[k]]); }
  __device__ float value(int64_t k) const { return val[idx[k]]; }
  __device__ int64_t row(int64_t, int, int key) const { return key; }
};

// Merge of the G ascending lists of a group: emit(key, at) once per
// distinct key in ascending order, at[s] = position in list s or -1.
template <int G, class L, class Emit>
__device__ __forceinline__ void merge_lists(const L& lists, int64_t g, Emit emit) {
  int64_t p[G], e[G];
  lists.bounds(g, p, e);
  int h[G];
#pragma unroll
  for (int s = 0; s < G; ++s) h[s] = p[s] < e[s] ? lists.key(p[s]) : INT_MAX;
  for (;;) {
    int mn = INT_MAX;
#pragma unroll
    for (int s = 0; s < G; ++s) mn = min(mn, h[s]);
    if (mn == INT_MAX) break;
    int64_t at[G];
#pragma unroll
    for (int s = 0; s < G; ++s) {
      if (h[s] == mn) {
        at[s] = p[s]++;
        h[s] = p[s] < e[s] ? lists.key(p[s]) : INT_MAX;
      } else {
        at[s] = -1;
      }
    }
    emit(mn, at);
  }
}

template <int G, class L>
__global__ __launch_bounds__(256) void k_grp_count(L lists, int64_t ng, int align,
                                                   int64_t* __restrict__ cnt) {
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < ng;
       g += int64_t(gridDim.x) * blockDim.x) {
    int64_t c = 0;
    merge_lists<G>(lists, g, [&](int, const int64_t (&)[G]) { ++c; });
    cnt[g] = (c + align - 1) / align * align;
  }
}

// union keys and G values per union entry; 16-bit values are quantized
// against the member row's scale as in the transpose (k_bt_scatter)
template <int G, class L>
__global__ __launch_bounds__(256) void k_grp_fill(L lists, int64_t ng,
                                                  const int64_t* __restrict__ gptr,
                           const float* __restrict__ scale, int vt, int* __restrict__ ukey,
                           void* __restrict__ uval) {
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < ng;
       g += int64_t(gridDim.x) * blockDim.x) {
    int64_t o = gptr[g];
    int last = 0;
    merge_lists<G>(lists, g, [&](int key, const int64_t (&at)[G]) {
      ukey[o] = key;
      last = key;
#pragma unroll
      for (int s = 0; s < G; ++s) {
        const float v = at[s] >= 0 ? lists.value(at[s]) : 0.0f;
        if (vt == kF32) {
          static_cast<float*>(uval)[G * o + s] = v;
        } else {
          const int64_t r = lists.row(g, s, key);
          static_cast<uint16_t*>(uval)[G * o + s] =
              at[s] >= 0 ? f2h(vt == kU16 ? v / (scale[r] * 65535.0f) : v / scale[r], vt)
                         : uint16_t(0);
        }
      }
      ++o;
    });
    // alignment padding: the list's last key (a valid index), values 0
    for (const int64_t e = gptr[g + 1]; o < e; ++o) {
      ukey[o] = last;
#pragma unroll
      for (int s = 0; s < G; ++s) {
        if (vt == kF32)
          static_cast<float*>(uval)[G * o + s] = 0.0f;
        else
          static_cast<uint16_t*>(uval)[G * o + s] = 0;
      }
    }
  }
}

// the wave's index inside its workgroup, as a wave-uniform value (so that
// per-wave pointers and loop bounds live in scalar registers)
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
}

// ---- 16- and 32-member groups on the matrix cores
// v_mfma_f32_16x16x4f32 computes D[16][16] += A[16][4] B[4][16]: lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and holds D[4*(l/16)+r][l%16]
// in register r.  v_mfma_f32_32x32x2f32 computes D[32][32] += A[32][2]
// B[2][32]: lane l supplies A[l%32][l/32] and B[l/32][l%32] and holds
// D[8*(r/4)+4*(l/32)+r%4][l%32] in register r (16 registers).  Here A = the
// members' values of K union entries (rows of A = group members), B = a
// block of frames of those entries' gathered frame vectors (16 or 32
// frames), D = the members' sums for that frame block; 64/W instructions
// per K entries cover the 64 frames.  Frame block fb, column c holds frame
// FB*c + fb (FB = 64/W blocks): lane l gathers frames FB*(l%W) .. FB*(l%W) +
// FB-1 of its entry with one FB-float vector load, so the W lanes of an
// entry read its 256-byte frame vector with one instruction.
// The values and keys arrive with coalesced vector loads (lane l loads
// member l%W of entry l/W), so nothing streams through the scalar cache.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

template <int G>
struct MF;
template <>
struct MF<16> {
  static constexpr int W = 16, K = 4, NR = 4, FB = 4;  // width, entries, registers, frame blocks
  using acc_t = f4v;
  __device__ static acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int member(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct MF<32> {
  static constexpr int W = 32, K = 2, NR = 16, FB = 2;
  using acc_t = f16v;
  __device__ static acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  __device__ static int member(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
};

// streamed once per pass: non-temporal, so they do not push the gathered
// frame vectors out of L2
template <int VT>
__device__ __forceinline__ float load_val1(const typename Val<VT>::T* __restrict__ v,
                                           int64_t i) {
  if constexpr (VT == kF32)
    return __builtin_nontemporal_load(v + i);
  else
    return h2f(__builtin_nontemporal_load(v + i), VT);
}

// a lane's FB gathered frames: one FB-float vector load (FB = 1: a float)
template <int FB>
struct GV {
  typedef float type __attribute__((ext_vector_type(FB)));
  __device__ static float at(const type& v, int i) { return v[i]; }
};
template <>
struct GV<1> {
  typedef float type;
  __device__ static float at(const type& v, int) { return v; }
};

// keys, values and gathered frame vectors of one step (K*U union entries;
// lane l takes entry K*u + l/W and member l%W).  FB frame blocks: the
// first W*FB frames of each 64-frame vector (FB < MF<G>::FB: passes over
// the first 16 or 32 frames only)
template <int G, int U, int FB>
struct MSet {
  using gvec = typename GV<FB>::type;
  int c[U];
  float a[U];
  gvec g[U];
};

template <int VT, int G, int U, int FB>
__device__ __forceinline__ void mload(MSet<G, U, FB>& t, const int* __restrict__ key,
                                      const typename Val<VT>::T* __restrict__ val, int64_t k0,
                                      int kq, int rw) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t k = k0 + MF<G>::K * u + kq;
    t.c[u] = __builtin_nontemporal_load(key + k);
    t.a[u] = load_val1<VT>(val, k * G + rw);
  }
}

// the frame vectors of a loaded set (lane l: frames FB*(l%W) .. + FB-1)
template <int G, int U, int FB>
__device__ __forceinline__ void mgather(MSet<G, U, FB>& t, const float* __restrict__ base) {
  using gvec = typename MSet<G, U, FB>::gvec;
#pragma unroll
  for (int u = 0; u < U; ++u) t.g[u] = *reinterpret_cast<const gvec*>(base + int64_t(t.c[u]) * 64);
}

// one step: the loads of T (two steps ahead), the gathers of N (one step
// ahead), then the matrix products of S.  Vector loads complete in order,
// so S's products wait for S's gathers (issued a step earlier) but neither
// for N's gathers nor for T's loads: each gather has a whole step of other
// waves' work to hide behind.
template <int VT, int G, int U, int FB>
__device__ __forceinline__ void mstep(const MSet<G, U, FB>& S, MSet<G, U, FB>& N,
                                      MSet<G, U, FB>& T, bool load_t, bool gather_n,
                                      const int* __restrict__ key,
                                      const typename Val<VT>::T* __restrict__ val, int64_t kt,
                                      int kq, int rw, const float* __restrict__ base,
                                      typename MF<G>::acc_t (&acc)[FB]) {
  using M = MF<G>;
  if (load_t) mload<VT, G, U, FB>(T, key, val, kt, kq, rw);
  __builtin_amdgcn_sched_barrier(0);
  if (gather_n) mgather<G, U, FB>(N, base);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) acc[fb] = M::op(S.a[u], GV<FB>::at(S.g[u], fb), acc[fb]);
}

// one step with the gathers in it: the gathers of S, the loads of T (two
// steps ahead), then the products of S (waiting for S's gathers only)
template <int VT, int G, int U, int FB>
__device__ __forceinline__ void mstep_in(MSet<G, U, FB>& S, MSet<G, U, FB>& T, bool load_t,
                                         const int* __restrict__ key,
                                         const typename Val<VT>::T* __restrict__ val,
                                         int64_t kt, int kq, int rw,
                                         const float* __restrict__ base,
                                         typename MF<G>::acc_t (&acc)[FB]) {
  using M = MF<G>;
  mgather<G, U, FB>(S, base);
  __builtin_amdgcn_sched_barrier(0);
  if (load_t) mload<VT, G, U, FB>(T, key, val, kt, kq, rw);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) acc[fb] = M::op(S.a[u], GV<FB>::at(S.g[u], fb), acc[fb]);
}

// acc[fb] += the group's sums over the union entries [b, en) (a multiple
// of K*U) for every frame, in steps of K*U entries with three register sets
// (keys and values two steps ahead; frame gathers in the step, or one step
// ahead with AHEAD); reads keys and values up to 2*K*U entries past en and
// gathers frame vectors of keys up to K*U entries past en (inside the next
// list or the tail padding, whose keys are valid columns).  Gathering ahead
// gives each gather a step of other waves' work to hide behind (the 64-frame
// backward, 12.0 -> 11.6 ms at full size) but costs the forward, whose
// higher occupancy already hides it (9.2 -> 9.9 ms).
template <int VT, int G, int U, bool AHEAD, int FB>
__device__ __forceinline__ void mfma_sum(const int* __restrict__ key,
                                         const typename Val<VT>::T* __restrict__ val,
                                         int64_t b, int64_t en, const float* __restrict__ base,
                                         int kq, int rw, typename MF<G>::acc_t (&acc)[FB]) {
  constexpr int E = MF<G>::K * U;
  const int64_t steps = (en - b) / E;
  if (steps <= 0) return;
  MSet<G, U, FB> A, B, C;
  mload<VT, G, U, FB>(A, key, val, b, kq, rw);
  mload<VT, G, U, FB>(B, key, val, b + E, kq, rw);
  int64_t kt = b + 2 * E;
  int64_t s = 0;
  if constexpr (AHEAD) {
    mgather<G, U, FB>(A, base);
    for (; s + 3 <= steps; s += 3) {
      mstep<VT, G, U, FB>(A, B, C, true, true, key, val, kt, kq, rw, base, acc);
      mstep<VT, G, U, FB>(B, C, A, true, true, key, val, kt + E, kq, rw, base, acc);
      mstep<VT, G, U, FB>(C, A, B, true, true, key, val, kt + 2 * E, kq, rw, base, acc);
      kt += 3 * E;
    }
    if (s < steps)
      mstep<VT, G, U, FB>(A, B, C, false, s + 1 < steps, key, val, kt, kq, rw, base, acc);
    if (s + 1 < steps)
      mstep<VT, G, U, FB>(B, C, A, false, false, key, val, kt, kq, rw, base, acc);
  } else {
    for (; s + 3 <= steps; s += 3) {
      mstep_in<VT, G, U, FB>(A, C, true, key, val, kt, kq, rw, base, acc);
      mstep_in<VT, G, U, FB>(B, A, true, key, val, kt + E, kq, rw, base, acc);
      mstep_in<VT, G, U, FB>(C, B, true, key, val, kt + 2 * E, kq, rw, base, acc);
      kt += 3 * E;
    }
    if (s < steps) mstep_in<VT, G, U, FB>(A, C, false, key, val, kt, kq, rw, base, acc);
    if (s + 1 < steps) mstep_in<VT, G, U, FB>(B, C, false, key, val, kt, kq, rw, base, acc);
  }
}

template <int G, int FB>
__device__ __forceinline__ void acc_zero(typename MF<G>::acc_t (&acc)[FB]) {
#pragma unroll
  for (int fb = 0; fb < FB; ++fb)
#pragma unroll
    for (int r = 0; r < MF<G>::NR; ++r) acc[fb][r] = 0.0f;
}

// FB frame blocks: frames 0 .. W*FB - 1 of each 64-frame row (FB =
// MF<G>::FB: all 64); ratios of the other frames are not written.
template <int VT, int G, int U, int FB>
__global__ __launch_bounds__(256) void k_fwdm64(int64_t ng, int64_t m,
                                                const int64_t* __restrict__ gp,
                                                const int* __restrict__ gcol,
                                                const typename Val<VT>::T* __restrict__ gval,
                                                const float* __restrict__ rscale,
                                                const float* __restrict__ x,
                                                const float* __restrict__ y,
                                                float* __restrict__ e) {
  using M = MF<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int kq = lane / M::W, rw = lane % M::W;
  const int64_t g = xcd_block(blockIdx.x, gridDim.x) * (256 / kWave) + wave_id();
  if (g >= ng) return;
  typename M::acc_t acc[FB];
  acc_zero<G, FB>(acc);
  mfma_sum<VT, G, U, false, FB>(gcol, gval, gp[g], gp[g + 1], x + FB * rw, kq, rw, acc);
#pragma unroll
  for (int r = 0; r < M::NR; ++r) {
    const int64_t row = int64_t(G) * g + M::member(lane, r);
    if (row >= m) continue;
    const float sc = VT == kF32 ? 1.0f : rscale[row];
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) {
      const int64_t o = row * 64 + FB * rw + fb;
      const float yy = y[o];
      const float yt = sc * acc[fb][r];
      const float t = (yt != 0.0f && yy != 0.0f) ? yy / yt : 0.0f;
      e[o] = VT == kF32 ? t : sc * t;  // ratio folded with the row scale
    }
  }
}

template <int VT, int U>
__global__ __launch_bounds__(256) void k_bwdm64(int ngc, const int64_t* __restrict__ gp,
                                                int nbl, const int* __restrict__ grow,
                                                const typename Val<VT>::T* __restrict__ gval,
                                                const int* __restrict__ cols,
                                                const float* __restrict__ e,
                                                float* __restrict__ c, int first, int last,
                                                const float* __restrict__ x,
                                                const float* __restrict__ s,
                                                float* __restrict__ xx,
                                                double* __restrict__ partials) {
  using M = MF<16>;
  const int lane = threadIdx.x & (kWave - 1);
  const int kq = lane >> 4, r16 = lane & 15;
  const int w = wave_id();
  const int lb = int(xcd_block(blockIdx.x, gridDim.x));
  const int h = lb * (256 / kWave) + w;
  M::acc_t acc[4];
  acc_zero<16, 4>(acc);
  if (h < ngc) {
    for (int bb = 0; bb < nbl; ++bb)
      mfma_sum<VT, 16, 2, true, 4>(grow, gval, gp[int64_t(bb) * ngc + h],
                                gp[int64_t(bb) * ngc + h + 1],
                          e + 4 * r16, kq, r16, acc);
  }
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  if (h < ngc) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = cols[int64_t(h) * 16 + 4 * kq + i];
      if (j < 0) continue;
      float d = 0.0f;
      if (last) {
        d = s[j];
        d += float(d == 0.0f);
      }
#pragma unroll
      for (int fb = 0; fb < 4; ++fb) {
        const int64_t o = int64_t(j) * 64 + 4 * r16 + fb;
        const float cv = first ? acc[fb][i] : c[o] + acc[fb][i];
        if (!last) {
          c[o] = cv;
        } else {
          const float nv = x[o] * cv / d;
          xx[o] = nv;
          part[fb] += double(nv);
        }
      }
    }
  }
  if (last) {
    __shared__ double red[256 / kWave][4][64];
#pragma unroll
    for (int fb = 0; fb < 4; ++fb) red[w][kq][4 * r16 + fb] = part[fb];
    __syncthreads();
    if (threadIdx.x < 64) {
      double t = 0.0;
      for (int ww = 0; ww < 256 / kWave; ++ww)
        for (int q = 0; q < 4; ++q) t += red[ww][q][threadIdx.x];
      partials[int64_t(lb) * 64 + threadIdx.x] = t;
    }
  }
}

// Backward with per-XCD row ranges (64 frames, bricks of G = 16 or 32):
// workgroup i runs on XCD i % 8 (the dispatcher's round-robin), which takes
// row blocks [nblk*x/8, nblk*(x+1)/8).  Each wave walks its brick through
// all of its XCD's blocks; the bricks resident on one XCD at a time are
// consecutive (neighbouring voxels) and walk the same small ratio slices
// together, so their ratio gathers hit that XCD's L2.  Each XCD writes its
// partial image sums to its own slice of c8; k_bwd_xreduce adds the 8
// slices in a fixed order and applies the update.  On a device that assigns
// workgroups differently the result is the same (only the cache behaviour
// changes).
// boff / bcnt: this launch covers row blocks b0 + boff .. b0 + boff + bcnt
// - 1 of the XCD's range (several launches keep the bricks resident on an
// XCD within a few blocks of each other, so the ratio rows they gather fit
// its L2); accum: add to the partial images instead of writing them.
// FB frame blocks as in k_fwdm64: the partial images of frames W*FB .. 63
// are not written.
template <int VT, int G, int FB>
__global__ __launch_bounds__(256) void k_bwdm64x(int ngc, int nblk, int n,
                                                 const int64_t* __restrict__ gp,
                                                 const int* __restrict__ grow,
                                                 const typename Val<VT>::T* __restrict__ gval,
                                                 const int* __restrict__ cols,
                                                 const float* __restrict__ e,
                                                 float* __restrict__ c8, int boff, int bcnt,
                                                 int accum) {
  using M = MF<G>;
  constexpr int U = 8 / M::K;  // 8-entry steps (backward lists are aligned to 8)
  const int lane = threadIdx.x & (kWave - 1);
  const int kq = lane / M::W, rw = lane % M::W;
  const int xcd = int(blockIdx.x % 8);
  const int h = int(blockIdx.x / 8) * (256 / kWave) + wave_id();
  if (h >= ngc) return;
  const int b0 = int(int64_t(nblk) * xcd / 8), b1 = int(int64_t(nblk) * (xcd + 1) / 8);
  const int s0 = b0 + boff < b1 ? b0 + boff : b1;
  const int s1 = s0 + bcnt < b1 ? s0 + bcnt : b1;
  typename M::acc_t acc[FB];
  acc_zero<G, FB>(acc);
  for (int b = s0; b < s1; ++b)
    mfma_sum<VT, G, U, true, FB>(grow, gval, gp[int64_t(b) * ngc + h],
                                 gp[int64_t(b) * ngc + h + 1], e + FB * rw, kq, rw, acc);
  float* out = c8 + int64_t(xcd) * n * 64;
#pragma unroll
  for (int r = 0; r < M::NR; ++r) {
    const int j = cols[int64_t(h) * G + M::member(lane, r)];
    if (j < 0) continue;
#pragma unroll
    for (int fb = 0; fb < FB; ++fb) {
      const int64_t o = int64_t(j) * 64 + FB * rw + fb;
      out[o] = accum ? out[o] + acc[fb][r] : acc[fb][r];
    }
  }
}

// c = sum of the 8 partial images (XCD order), x' = x*c/(s + (s==0)), and
// FP64 partial sums per workgroup (4 voxels) and frame; frames fa .. 63 are
// skipped (x' and their partial sums not written)
__global__ __launch_bounds__(256) void k_bwd_xreduce(int n, int fa, const float* __restrict__ c8,
                                                     const float* __restrict__ x,
                                                     const float* __restrict__ s,
                                                     float* __restrict__ xx,
                                                     double* __restrict__ partials) {
  const int f = threadIdx.x & 63, jj = threadIdx.x >> 6;
  const int64_t j = int64_t(blockIdx.x) * 4 + jj;
  double part = 0.0;
  if (j < n && f < fa) {
    float cv = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) cv += c8[(int64_t(q) * n + j) * 64 + f];
    float d = s[j];
    d += float(d == 0.0f);
    const float nv = x[j * 64 + f] * cv / d;
    xx[j * 64 + f] = nv;
    part = double(nv);
  }
  __shared__ double red[4][64];
  red[jj][f] = part;
  __syncthreads();
  if (threadIdx.x < fa)
    partials[int64_t(blockIdx.x) * 64 + f] = red[0][f] + red[1][f] + red[2][f] + red[3][f];
}

// Single-frame backward with per-XCD row ranges: workgroup i runs on XCD
// i % 8 and takes row blocks [nblk*x/8, nblk*(x+1)/8), so each XCD's ratio
// slices stay in its own L2.  A wave walks one column through those blocks
// (4 entries per lane and step, UNROLL steps in flight, as k_bwd) and writes
// the XCD's partial column sum to c8[x][j]; k_bwd1_xreduce adds the 8 in a
// fixed order and applies the update.
// the rows of 4 stored entries: 32-bit rows, or 16-bit offsets in the block
// (non-temporal: the stream must not evict the ratio slices from L2)
template <bool R16>
__device__ __forceinline__ vi4 rows4(const void* __restrict__ brow, int64_t k, int r0) {
  if constexpr (R16) {
    const vus4 h =
        __builtin_nontemporal_load(reinterpret_cast<const vus4*>(static_cast<const uint16_t*>(brow) + k));
    return vi4{r0 + int(h.x), r0 + int(h.y), r0 + int(h.z), r0 + int(h.w)};
  } else {
    return __builtin_nontemporal_load(reinterpret_cast<const vi4*>(static_cast<const int*>(brow) + k));
  }
}

// LPC lanes per column (16, 32 or 64), 256 / LPC columns per workgroup.
template <int VT, int UNROLL, bool R16, int LPC>
__global__ __launch_bounds__(256) void k_bwd1x(int n, int nblk, int blk_rows,
                                               const int64_t* __restrict__ ptr,
                                               const void* __restrict__ brow,
                                               const typename Val<VT>::T* __restrict__ val,
                                               const float* __restrict__ e,
                                               float* __restrict__ c8) {
  const int lane = threadIdx.x % LPC;
  const int xcd = int(blockIdx.x % 8);
  const int j = int(blockIdx.x / 8) * (256 / LPC) + int(threadIdx.x) / LPC;
  if (j >= n) return;  // the column's LPC lanes leave together
  const int b0 = int(int64_t(nblk) * xcd / 8), b1 = int(int64_t(nblk) * (xcd + 1) / 8);
  constexpr int64_t STEP = 4 * LPC;
  float acc = 0.0f;
  // the segment bounds of up to LPC/2 blocks in one load (lane 2t, 2t+1:
  // block bc + t), handed to the block loop by shuffles
  int64_t pv = 0;
  for (int b = b0; b < b1; ++b) {
    const int t = (b - b0) % (LPC / 2);
    if (t == 0) {
      const int nbc = b1 - b < LPC / 2 ? b1 - b : LPC / 2;
      pv = lane < 2 * nbc ? ptr[int64_t(b + lane / 2) * (n + 1) + j + (lane & 1)] : 0;
    }
    const int64_t bb = __shfl(pv, 2 * t, LPC), en = __shfl(pv, 2 * t + 1, LPC);
    const int r0 = b * blk_rows;
    int64_t k = bb + 4 * lane;
    for (; k + STEP * (UNROLL - 1) < en; k += STEP * UNROLL) {
      float wv[UNROLL][4];
      vi4 rr[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        load4<VT>(val, k + STEP * u, wv[u][0], wv[u][1], wv[u][2], wv[u][3]);
        rr[u] = rows4<R16>(brow, k + STEP * u, r0);
      }
      float g[UNROLL][4];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        g[u][0] = e[rr[u].x];
        g[u][1] = e[rr[u].y];
        g[u][2] = e[rr[u].z];
        g[u][3] = e[rr[u].w];
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = fmaf(wv[u][q], g[u][q], acc);
    }
    for (; k < en; k += STEP) {
      float wv[4];
      load4<VT>(val, k, wv[0], wv[1], wv[2], wv[3]);
      const vi4 rr = rows4<R16>(brow, k, r0);
      acc = fmaf(wv[0], e[rr.x], acc);
      acc = fmaf(wv[1], e[rr.y], acc);
      acc = fmaf(wv[2], e[rr.z], acc);
      acc = fmaf(wv[3], e[rr.w], acc);
    }
  }
#pragma unroll
  for (int off = LPC / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, LPC);
  if (lane == 0) c8[int64_t(xcd) * n + j] = acc;
}

// c = sum of the 8 partial column sums (XCD order), x' = x*c/(s + (s==0)),
// FP64 partial sum per workgroup of 256 voxels
__global__ __launch_bounds__(256) void k_bwd1_xreduce(int n, const float* __restrict__ c8,
                                                      const float* __restrict__ x,
                                                      const float* __restrict__ s,
                                                      float* __restrict__ xx,
                                                      double* __restrict__ partials) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double part = 0.0;
  if (j < n) {
    float cv = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) cv += c8[int64_t(q) * n + j];
    float d = s[j];
    d += float(d == 0.0f);
    const float nv = x[j] * cv / d;
    xx[j] = nv;
    part = double(nv);
  }
  __shared__ double red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------------
// Single frame, both products in one pass over the matrix.  A workgroup of
// W waves takes one group of `rows` consecutive plan rows (the groups of
// the LDS-staged forward) and copies x over the group's column union into
// LDS.  The group's rows go out in batches: in each, every wave takes
// 64 / LANES rows, LANES lanes per row, with S steps of each row's entries
// held in registers and the next batch's already in flight:
//   yt_i = sum_j A_ij x_j (x from LDS), e_i = y_i / yt_i,
// then A_ij e_i is added into column sums over the union, also in LDS:
// one region per wave, or (SH) one region that the waves add to in turn,
// a barrier between turns; inside one wave the adds follow its
// instruction order.  At the end the regions are added in wave order and
// written over the union (cpart[up[g] + rank]); a column reduce
// (k_fb1_treduce + k_fb1_tcombine, k_fb1_reduce8 + k_bwd1_xreduce or
// k_fb1_reduce) adds each column's group partials in a fixed order.  The
// matrix is streamed once per iteration instead of once per product.
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
// 4 entries at padded position k of a row [b, en); lanes past the row get
// value 0 and rank 0 from a load of the row's first entries (position b is
// inside the arrays' padding even for an empty last row), so the loads are
// not branched around and the next row's stay in flight
template <int VT>
__device__ __forceinline__ void fb_load(const uint16_t* __restrict__ lcol,
                                        const typename Val<VT>::T* __restrict__ val, int64_t k,
                                        int64_t b, int64_t en, vus4& c, float (&v)[4]) {
  const bool ok = k < en;
  const int64_t kk = ok ? k : b;
  const vus4 cc = *reinterpret_cast<const vus4*>(lcol + kk);
  float t[4];
  load4c<VT>(val, kk, t[0], t[1], t[2], t[3]);
  c = ok ? cc : vus4{0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = ok ? t[q] : 0.0f;
}

// sum over the LANES lanes of each row slot, the same bits in all of them
// (DPP quad swaps and row mirrors inside 16-lane rows, then swaps across
// rows); every lane must be active
template <int LANES>
__device__ __forceinline__ float slot_sum(float v) {
  v += dppf<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dppf<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dppf<0x141>(v);  // row_half_mirror
  v += dppf<0x140>(v);  // row_mirror
  if constexpr (LANES >= 32) v += __shfl_xor(v, 16, 64);
  if constexpr (LANES >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

// W waves per group; LANES lanes per row (64 / LANES rows per wave and
// batch); S steps of 4 x LANES entries per row held in registers (longer
// rows continue from memory); the next batch in flight while one is
// summed.  SH: one column-sum region for the workgroup, the waves adding
// their rows in turn (wave order, a barrier between turns) -- a third of
// the LDS of W regions, so about three times the workgroups per CU
template <int VT, int W, int LANES, int S, bool SH>
__global__ __launch_bounds__(64 * W) void k_fb1l(int64_t m, int rows,
                                                 const int64_t* __restrict__ rp,
                                                 const uint16_t* __restrict__ lcol,
                                                 const int64_t* __restrict__ up,
                                                 const int* __restrict__ ucol,
                                                 const typename Val<VT>::T* __restrict__ val,
                                                 const float* __restrict__ rscale,
                                                 const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 float* __restrict__ e,
                                                 float* __restrict__ cpart) {
  constexpr int RPW = 64 / LANES;  // rows per wave and batch
  constexpr int STEP = 4 * LANES;  // entries of a row per step
  constexpr int NT = 64 * W;
  const int kBatches = (rows + W * RPW - 1) / (W * RPW);
  constexpr int NR = SH ? 1 : W;  // column-sum regions
  // the group's rows + 1 row bounds, x over the union (nus floats),
  // then NR column-sum regions of nus + 1 floats each (the last a spare
  // slot)
  extern __shared__ int64_t fb_lds[];
  int64_t* rps = fb_lds;
  float* xs = reinterpret_cast<float*>(fb_lds + rows + 1);
  const int tid = int(threadIdx.x);
  const int64_t grp = blockIdx.x;
  const int64_t ub = up[grp];
  const int nu = int(up[grp + 1] - ub);
  const int nus = nu > 0 ? nu : 1;
  const int64_t r0 = grp * rows;
  const int nr = int(r0 + rows < m ? rows : m - r0);  // >= 1
  for (int i = tid; i <= rows; i += NT) rps[i] = rp[r0 + (i < nr ? i : nr)];
  // x over the union, 8 gathers in flight per thread
  for (int j0 = 0; j0 < nu; j0 += NT * 8) {
    int cc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + tid + u * NT;
      cc[u] = ucol[ub + (j < nu ? j : 0)];
    }
    float xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = x[cc[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + tid + u * NT;
      if (j < nu) xs[j] = xv[u];
    }
  }
  for (int j = tid; j < NR * (nus + 1); j += NT) xs[nus + j] = 0.0f;
  if (nu == 0 && tid == 0) xs[0] = 0.0f;  // lanes past a row read rank 0
  __syncthreads();
  const int w = tid / 64, lane = tid % 64, q = lane / LANES, g = lane % LANES;
  float* cw = xs + nus + (SH ? 0 : int64_t(nus + 1) * w);
  // row of this lane's slot in batch bt (rows past the group are empty)
  auto row_of = [&](int bt) { return bt * (W * RPW) + w * RPW + q; };
  int t = row_of(0);
  int64_t b = rps[t < nr ? t : nr], en = t < nr ? rps[t + 1] : b;
  float yv = y[r0 + (t < nr ? t : nr - 1)];
  float sv = VT != kF32 ? rscale[r0 + (t < nr ? t : nr - 1)] : 1.0f;
  vus4 c[S];
  float v[S][4];
#pragma unroll
  for (int s = 0; s < S; ++s) fb_load<VT>(lcol, val, b + 4 * g + STEP * s, b, en, c[s], v[s]);
  for (int bt = 0; bt < kBatches; ++bt) {
    // the next batch's rows in flight
    const int tn = row_of(bt + 1);
    const bool has_next = bt + 1 < kBatches && tn < nr;
    const int64_t bn = rps[tn < nr ? tn : nr], enn = has_next ? rps[tn + 1] : bn;
    const float yn = y[r0 + (tn < nr ? tn : nr - 1)];
    const float sn = VT != kF32 ? rscale[r0 + (tn < nr ? tn : nr - 1)] : 1.0f;
    vus4 cn[S];
    float vn[S][4];
#pragma unroll
    for (int s = 0; s < S; ++s)
      fb_load<VT>(lcol, val, bn + 4 * g + STEP * s, bn, enn, cn[s], vn[s]);
    float acc = 0.0f;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      acc = fmaf(v[s][0], xs[c[s].x], acc);
      acc = fmaf(v[s][1], xs[c[s].y], acc);
      acc = fmaf(v[s][2], xs[c[s].z], acc);
      acc = fmaf(v[s][3], xs[c[s].w], acc);
    }
    const int64_t kx = b + 4 * g + int64_t(STEP) * S;  // rows past S steps
    for (int64_t kk = kx; kk < en; kk += STEP) {
      const vus4 cc = *reinterpret_cast<const vus4*>(lcol + kk);
      float vv[4];
      load4c<VT>(val, kk, vv[0], vv[1], vv[2], vv[3]);
      acc = fmaf(vv[0], xs[cc.x], acc);
      acc = fmaf(vv[1], xs[cc.y], acc);
      acc = fmaf(vv[2], xs[cc.z], acc);
      acc = fmaf(vv[3], xs[cc.w], acc);
    }
    const float yt = sv * slot_sum<LANES>(acc);
    const float r = yt != 0.0f ? yv / yt : 0.0f;
    const float ev = VT == kF32 ? r : sv * r;  // ratio folded with the row scale
    if (t < nr && g == 0) e[r0 + t] = ev;
    // Column sums, one row slot at a time: the lanes of one LDS instruction
    // then hold distinct columns of one row (the plan has no repeated
    // column in a row), so a plain read-add-write per entry is exact and
    // ordered; entries of value 0 (padding, lanes past the row) go to the
    // region's spare slot nus.  LDS float atomics are far slower.
    int si[S][4];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const bool ok = b + 4 * g + STEP * s < en;
      si[s][0] = ok && v[s][0] != 0.0f ? int(c[s].x) : nus;
      si[s][1] = ok && v[s][1] != 0.0f ? int(c[s].y) : nus;
      si[s][2] = ok && v[s][2] != 0.0f ? int(c[s].z) : nus;
      si[s][3] = ok && v[s][3] != 0.0f ? int(c[s].w) : nus;
    }
    for (int turn = 0; turn < (SH ? W : 1); ++turn) {
      if (!SH || w == turn) {
#pragma unroll
        for (int qq = 0; qq < RPW; ++qq) {
          if (q == qq) {
            float old[S][4];
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
              for (int j = 0; j < 4; ++j) old[s][j] = cw[si[s][j]];
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
              for (int j = 0; j < 4; ++j) cw[si[s][j]] = fmaf(v[s][j], ev, old[s][j]);
            for (int64_t kk = kx; kk < en; kk += STEP) {
              const vus4 cc = *reinterpret_cast<const vus4*>(lcol + kk);
              float vv[4];
              load4c<VT>(val, kk, vv[0], vv[1], vv[2], vv[3]);
              const int i0 = vv[0] != 0.0f ? int(cc.x) : nus;
              const int i1 = vv[1] != 0.0f ? int(cc.y) : nus;
              const int i2 = vv[2] != 0.0f ? int(cc.z) : nus;
              const int i3 = vv[3] != 0.0f ? int(cc.w) : nus;
              const float o0 = cw[i0], o1 = cw[i1], o2 = cw[i2], o3 = cw[i3];
              cw[i0] = fmaf(vv[0], ev, o0);
              cw[i1] = fmaf(vv[1], ev, o1);
              cw[i2] = fmaf(vv[2], ev, o2);
              cw[i3] = fmaf(vv[3], ev, o3);
            }
          }
          // The row slots of a wave add in turn.  A lane runs only its own
          // slot's block, so the compiler, which orders memory per lane,
          // may otherwise move the next slot's LDS reads above this slot's
          // writes (rows of one group share columns): a compiler barrier
          // keeps the program order, which the wave's LDS accesses follow.
          asm volatile("" ::: "memory");
        }
      }
      if (SH) __syncthreads();
    }
    t = tn;
    b = bn;
    en = enn;
    yv = yn;
    sv = sn;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      c[s] = cn[s];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[s][j] = vn[s][j];
    }
  }
  __syncthreads();
  for (int j = tid; j < nu; j += NT) {
    float sum = xs[nus + j];
#pragma unroll
    for (int k = 1; k < NR; ++k) sum += xs[nus + int64_t(nus + 1) * k + j];
    cpart[ub + j] = sum;
  }
}

// c_j = the sum of column j's group partials in group order (vpos: the
// union positions holding column j, ascending; LPC lanes per column, each
// lane's terms in order, then a fixed tree over the lanes), then
// x'_j = x_j c_j / (s_j + (s_j == 0)) and an FP64 partial sum per
// workgroup.  Workgroups of one XCD take consecutive columns, whose
// partials lie side by side in each group's sorted union.
template <int LPC, int U>
__global__ __launch_bounds__(256) void k_fb1_reduce(int n, const int64_t* __restrict__ vptr,
                                                    const int* __restrict__ vpos,
                                                    const float* __restrict__ cpart,
                                                    const float* __restrict__ x,
                                                    const float* __restrict__ s,
                                                    float* __restrict__ xx,
                                                    double* __restrict__ partials) {
  const int lane = int(threadIdx.x) % LPC;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t j = blk * (256 / LPC) + int(threadIdx.x) / LPC;
  float acc = 0.0f;
  if (j < n) {
    const int64_t k1 = vptr[j + 1];
    int64_t k = vptr[j] + lane;
    for (; k + int64_t(LPC) * (U - 1) < k1; k += int64_t(LPC) * U) {
      int p[U];
#pragma unroll
      for (int u = 0; u < U; ++u) p[u] = vpos[k + LPC * u];
      float g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) g[u] = cpart[p[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += g[u];
    }
    for (; k < k1; k += LPC) acc += cpart[vpos[k]];
  }
#pragma unroll
  for (int off = LPC / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, LPC);
  double part = 0.0;
  if (j < n && lane == 0) {
    float d = s[j];
    d += float(d == 0.0f);
    const float nv = x[j] * acc / d;
    xx[j] = nv;
    part = double(nv);
  }
  __shared__ double red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// bits of the keys 0 .. count - 1
inline int bits_for_count(int64_t count) {
  int b = 1;
  while ((int64_t(1) << b) < count) ++b;
  return b;
}

// Per-XCD form of the same sums: workgroup i runs on XCD x = i % 8 and
// adds, for its columns, the partials of groups [ng x / 8, ng (x+1) / 8)
// only (the positions of column j in [vb[x-1][j], vb[x][j]), with vb[-1] =
// vptr[j] and vb[7] = vptr[j+1]), so an XCD's L2 holds the lines of an
// eighth of the groups; c8[x][j] takes the XCD's sum, and k_bwd1_xreduce
// adds the 8 in XCD order and applies the update.  The positions stream
// past L2 (non-temporal) so that they do not evict the partial sums' lines.
template <int LPC, int U>
__global__ __launch_bounds__(256) void k_fb1_reduce8(int n, const int64_t* __restrict__ vptr,
                                                     const int64_t* __restrict__ vb,
                                                     const int* __restrict__ vpos,
                                                     const float* __restrict__ cpart,
                                                     float* __restrict__ c8) {
  const int lane = int(threadIdx.x) % LPC;
  const int xcd = int(blockIdx.x % 8);
  const int64_t j = int64_t(blockIdx.x / 8) * (256 / LPC) + int(threadIdx.x) / LPC;
  if (j >= n) return;  // the column's LPC lanes leave together
  const int64_t k1 = xcd == 7 ? vptr[j + 1] : vb[int64_t(xcd) * n + j];
  int64_t k = (xcd == 0 ? vptr[j] : vb[int64_t(xcd - 1) * n + j]) + lane;
  float acc = 0.0f;
  for (; k + int64_t(LPC) * (U - 1) < k1; k += int64_t(LPC) * U) {
    int p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) p[u] = __builtin_nontemporal_load(vpos + k + LPC * u);
    float g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) g[u] = cpart[p[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += g[u];
  }
  for (; k < k1; k += LPC) acc += cpart[__builtin_nontemporal_load(vpos + k)];
#pragma unroll
  for (int off = LPC / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, LPC);
  if (lane == 0) c8[int64_t(xcd) * n + j] = acc;
}

// vb[x-1][j] = first position of column j (in [vptr[j], vptr[j+1])) that
// belongs to a group of XCD range x or later, x = 1 .. 7
__global__ void k_vbounds(int n, const int64_t* __restrict__ vptr, const int* __restrict__ vpos,
                          const int64_t* __restrict__ up, int64_t ng, int64_t* __restrict__ vb) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < int64_t(7) * n;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int x = int(t / n) + 1;
    const int64_t j = t % n;
    const int64_t pstart = up[ng * x / 8];
    int64_t lo = vptr[j], hi = vptr[j + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (int64_t(vpos[mid]) < pstart) lo = mid + 1; else hi = mid;
    }
    vb[t] = lo;
  }
}

// (column, position) pairs of the unions, for the column-major index
__global__ void k_upos_keys(const int* __restrict__ ucol, int64_t nu, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ idx) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nu;
       p += int64_t(gridDim.x) * blockDim.x) {
    keys[p] = uint64_t(uint32_t(ucol[p]));
    idx[p] = uint32_t(p);
  }
}

__global__ void k_u32_to_i32_64(const uint32_t* __restrict__ a, int64_t m, int* __restrict__ b) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = int(a[i]);
}

// vptr[j] = first sorted position whose column is >= j, j = 0 .. n
__global__ void k_vptr(const uint64_t* __restrict__ keys, int64_t nu, int n,
                       int64_t* __restrict__ vptr) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j <= n;
       j += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = nu;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (int64_t(keys[mid]) < j) lo = mid + 1; else hi = mid;
    }
    vptr[j] = lo;
  }
}

// Column-tile form of the sums: workgroup w takes part k of tile t
// (twg[w] = {t, k, first segment, end segment}: a run of the tile's
// segments in group order) and adds its partials for the 2^tshift columns
// of the tile.  The segments (a group's union positions inside the tile:
// distinct columns, ascending; tcol holds each position's column inside
// its tile) are split into NREG consecutive runs, one
// per slot of LPS lanes, each with its own column sums in LDS, so the
// lanes of one read-add-write never share a column.  A slot reads its
// segments' partials and columns in the order k_fb1l wrote them (no
// gather): it holds the bounds of LPS segments in its lanes (the next LPS
// already in flight), cuts them into chunks of LPS positions and loads U
// chunks at a time before adding them in order.  Then each column's NREG
// sums are added in slot order into ck[k][j]; k_fb1_tcombine adds a
// column's parts in order.  Fixed order, deterministic.
template <int LPS, int W>
__global__ __launch_bounds__(64 * W) void k_fb1_treduce(int n, int tshift,
                                                        const int4* __restrict__ twg,
                                                        const int2* __restrict__ tseg,
                                                        const uint16_t* __restrict__ tcol,
                                                        const float* __restrict__ cpart,
                                                        float* __restrict__ ck) {
  constexpr int SUB = 64 / LPS;
  constexpr int NREG = W * SUB;
  constexpr int U = 8;
  extern __shared__ float tr_lds[];
  const int T = 1 << tshift;
  const int tid = int(threadIdx.x);
  const int4 wg = twg[blockIdx.x];
  const int t = wg.x, kp = wg.y;
  const int j0 = t << tshift;
  for (int i = tid; i < NREG * T; i += 64 * W) tr_lds[i] = 0.0f;
  __syncthreads();
  const int s0 = wg.z, s1 = wg.w;
  const int lane = tid % 64, sl = lane % LPS, slot = lane / LPS;
  const int r = (tid / 64) * SUB + slot;
  const int64_t cnt = s1 - s0;
  const int a = s0 + int(cnt * r / NREG), b = s0 + int(cnt * (r + 1) / NREG);
  float* sums = tr_lds + r * T;
  const uint64_t smask = LPS == 64 ? ~uint64_t(0) : ((uint64_t(1) << LPS) - 1) << (slot * LPS);
  // the first window's bounds; lanes past the slot's segments load segment
  // a (the loop runs only if a < b) and hold an empty range
  int2 nxt;
  {
    const bool in = a + sl < b;
    const int2 v = tseg[in ? a + sl : a];
    nxt = in ? v : int2{0, 0};
  }
  for (int s = a; s < b; s += LPS) {
    const int2 se = nxt;
    {
      const bool in = s + LPS + sl < b;
      const int2 v = tseg[in ? s + LPS + sl : a];
      nxt = in ? v : int2{0, 0};
    }
    // chunks per segment and their inclusive prefix over the slot's lanes
    const int nch = (se.y - se.x + LPS - 1) / LPS;
    int incl = nch;
#pragma unroll
    for (int off = 1; off < LPS; off <<= 1) {
      const int o = __shfl_up(incl, off, LPS);
      if (sl >= off) incl += o;
    }
    const int excl = incl - nch;
    const int total = __shfl(incl, LPS - 1, LPS);
    for (int q0 = 0; q0 < total; q0 += U) {
      int c[U];
      float v[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = q0 + u;
        // segment of chunk q: the lanes whose chunks end at or before q
        const int i = min(__popcll(__ballot(incl <= q) & smask), LPS - 1);
        const int p0 = __shfl(se.x, i, LPS), p1 = __shfl(se.y, i, LPS);
        const int p = p0 + (q - __shfl(excl, i, LPS)) * LPS + sl;
        ok[u] = q < total && p < p1;
        const int pp = ok[u] ? p : 0;  // position 0 exists (the unions are not empty)
        c[u] = tcol[pp];
        v[u] = cpart[pp];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ok[u]) sums[c[u]] += v[u];
    }
  }
  __syncthreads();
  for (int i = tid; i < T; i += 64 * W) {
    const int j = j0 + i;
    if (j >= n) break;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < NREG; ++k) acc += tr_lds[k * T + i];
    ck[int64_t(kp) * n + j] = acc;
  }
}

// c_j = the sum of column j's tk[tile] part sums in part order, then the
// update x'_j = x_j c_j / (s_j + (s_j == 0)) and an FP64 partial sum per
// workgroup (as k_bwd1_xreduce)
__global__ __launch_bounds__(256) void k_fb1_tcombine(int n, int tshift, const int* __restrict__ tk,
                                                      const float* __restrict__ ck,
                                                      const float* __restrict__ x,
                                                      const float* __restrict__ s,
                                                      float* __restrict__ xx,
                                                      double* __restrict__ partials) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double part = 0.0;
  if (j < n) {
    const int K = tk[j >> tshift];
    float cv = 0.0f;
    for (int q = 0; q < K; ++q) cv += ck[int64_t(q) * n + j];
    float d = s[j];
    d += float(d == 0.0f);
    const float nv = x[j] * cv / d;
    xx[j] = nv;
    part = double(nv);
  }
  __shared__ double red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// column of every union position inside its tile of 2^tshift columns
__global__ void k_tile_cols(const int* __restrict__ ucol, int64_t nu, int tshift,
                            uint16_t* __restrict__ tcol) {
  const int mask = (1 << tshift) - 1;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nu;
       p += int64_t(gridDim.x) * blockDim.x)
    tcol[p] = uint16_t(ucol[p] & mask);
}

// segment starts: the first position of every group and every position
// whose column lies in another tile than the one before it
__global__ void k_seg_flags(const int* __restrict__ ucol, int64_t nu, int tshift,
                            uint32_t* __restrict__ flag) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nu;
       p += int64_t(gridDim.x) * blockDim.x)
    flag[p] = (p == 0 || (ucol[p] >> tshift) != (ucol[p - 1] >> tshift)) ? 1u : 0u;
}
__global__ void k_seg_gstarts(const int64_t* __restrict__ up, int64_t ng, int64_t nu,
                              uint32_t* __restrict__ flag) {
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < ng;
       g += int64_t(gridDim.x) * blockDim.x)
    if (up[g] < nu) flag[up[g]] = 1u;
}
// segment s (in position order) starts at ps[s]; ps[nseg] = nu; key: its tile
__global__ void k_seg_emit(const int* __restrict__ ucol, int64_t nu, int tshift,
                           const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                           int* __restrict__ ps) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < nu;
       p += int64_t(gridDim.x) * blockDim.x) {
    if (flag[p]) {
      const uint32_t s = idx[p];
      keys[s] = uint64_t(uint32_t(ucol[p] >> tshift));
      vals[s] = s;
      ps[s] = int(p);
    }
    if (p == nu - 1) ps[idx[p] + flag[p]] = int(nu);
  }
}
__global__ void k_seg_pairs(const uint32_t* __restrict__ sorted, int64_t nseg,
                            const int* __restrict__ ps, int2* __restrict__ tseg) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nseg;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t s = sorted[i];
    tseg[i] = int2{ps[s], ps[s + 1]};
  }
}
// tsb[t] = first segment (sorted by tile) of tile t, t = 0 .. ntiles
__global__ void k_seg_bounds(const uint64_t* __restrict__ keys, int64_t nseg, int ntiles,
                             int* __restrict__ tsb) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t <= ntiles;
       t += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = nseg;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (keys[mid] < uint64_t(t)) lo = mid + 1; else hi = mid;
    }
    tsb[t] = int(lo);
  }
}

// *dup = 1 if a row holds one union rank twice among its entries of
// nonzero stored value (the fused pass adds each such entry to its column
// sum with a plain read-add-write, exact only for distinct columns).  One
// workgroup per group of `rows` rows, a bitmap of `words` words per wave
template <int VT>
__global__ __launch_bounds__(256) void k_fb_distinct(int64_t m, int rows,
                                                     const int64_t* __restrict__ rp,
                                                     const uint16_t* __restrict__ lcol,
                                                     const typename Val<VT>::T* __restrict__ val,
                                                     int words, int* __restrict__ dup) {
  extern __shared__ uint32_t fb_bm[];
  const int w = int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
  uint32_t* bm = fb_bm + int64_t(w) * words;
  const int64_t r0 = int64_t(blockIdx.x) * rows;
  const int nr = int(r0 + rows < m ? rows : m - r0);
  int found = 0;
  for (int t = w; t < nr; t += 4) {
    for (int k = lane; k < words; k += 64) bm[k] = 0u;
    const int64_t b = rp[r0 + t], en = rp[r0 + t + 1];
    for (int64_t k = b + lane; k < en; k += 64) {
      const bool nz = VT == kF32 ? reinterpret_cast<const float*>(val)[k] != 0.0f
                                 : reinterpret_cast<const uint16_t*>(val)[k] != 0;
      if (nz) {
        const uint32_t r = lcol[k], bit = 1u << (r & 31u);
        if (atomicOr(&bm[r >> 5], bit) & bit) found = 1;
      }
    }
  }
  if (found) atomicOr(dup, 1);
}


// brick of the backward groups: G voxels (x, y, z extents), wide in x and z
// and thin in y.  The lines of a two-head setup rotated about y run nearly
// parallel to the x-z planes, so a line that crosses a 4 x 1 x 4 brick meets
// more of its voxels than one crossing a 2 x 2 x 4 brick: on the 52^3 grid
// of 13 x 13-pixel heads 6.1 of 16 members per union entry against 5.4
// (10 % fewer entries), and 8.6 of 32 for 4 x 2 x 4 against 6.8 for
// 2 x 4 x 4.
__host__ __device__ inline void brick_dims(int G, int& bx, int& by, int& bz) {
  bx = 4;
  by = G >= 32 ? 2 : 1;
  bz = 4;
}

// cols[h*G + s]: member s of brick h; -1 outside the grid.  Bricks in y,
// x, z order (z fastest), so consecutive bricks fill x-z planes: the lines
// of a two-head setup rotated about y run nearly parallel to those planes,
// so the bricks one XCD holds at a time (consecutive ids) are crossed by the
// same lines over many bricks, and the ratio rows they gather stay in its
// L2.  Members in x, y, z order, z fastest.
__global__ void k_brick_cols(int nx, int ny, int nz, int G, int64_t ngc, int* __restrict__ cols) {
  int bx, by, bz;
  brick_dims(G, bx, by, bz);
  const int nbz = (nz + bz - 1) / bz, nbx = (nx + bx - 1) / bx;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < ngc * G;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t h = t / G;
    const int sidx = int(t - h * G);
    const int hz = int(h % nbz), hx = int((h / nbz) % nbx), hy = int(h / (int64_t(nbz) * nbx));
    const int mz = sidx % bz, my = (sidx / bz) % by, mx = sidx / (bz * by);
    const int ix = hx * bx + mx, iy = hy * by + my, iz = hz * bz + mz;
    cols[t] = (ix < nx && iy < ny && iz < nz) ? (ix * ny + iy) * nz + iz : -1;
  }
}

// groups of G consecutive column ids (no grid known)
__global__ void k_seq_cols(int n, int64_t total, int* __restrict__ cols) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x)
    cols[t] = t < n ? int(t) : -1;
}

// rows of the transpose as 16-bit offsets in their row block (blocks of at
// most 65,536 rows): 6 bytes per FP32 entry instead of 8
__global__ void k_bt_rows16(int64_t nb, int64_t bnnz, const int64_t* pptr, int n, int blk_rows,
                            const int* brow, uint16_t* brow16) {
  // one thread per (block, column) segment
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nb * (n + 1);
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = t / (n + 1);
    const int64_t r0 = b * blk_rows;
    for (int64_t p = pptr[t]; p < pptr[t + 1]; ++p) brow16[p] = uint16_t(brow[p] - r0);
  }
}

// forward columns as 16-bit low parts (grids of <= 4 x 65,536 voxels): one
// wave per row; split[i] packs the row's first positions with a column >=
// 65,536*k (k = 1, 2, 3; 0xffff: none) so col = 65,536*hi + lo with hi =
// number of splits <= the position.  Padding entries repeat the row's last
// column, so they decode to it.  *bad = 1 if a padded row is longer than
// 65,535 entries (positions must stay below the 0xffff "none") or a row's
// high parts decrease (columns not ascending in 65,536-blocks).
__global__ __launch_bounds__(256) void k_fwd_cols16(const int64_t* __restrict__ rp,
                                                    const int* __restrict__ col, int m,
                                                    const int64_t* __restrict__ prp,
                                                    const int* __restrict__ pcol,
                                                    uint16_t* __restrict__ lo,
                                                    uint64_t* __restrict__ split,
                                                    int* __restrict__ bad) {
  const int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    const int64_t b = rp[i], len = rp[i + 1] - b;
    if (prp[i + 1] - prp[i] > 65535) {
      if (lane == 0) bad[0] = 1;
      continue;
    }
    int c1 = 0, c2 = 0, c3 = 0;
    bool down = false;
    for (int64_t k = lane; k < len; k += kWave) {
      const int c = col[b + k];
      c1 += c < 65536;
      c2 += c < 2 * 65536;
      c3 += c < 3 * 65536;
      down |= k > 0 && (col[b + k - 1] >> 16) > (c >> 16);
    }
    if (down) bad[0] = 1;
    for (int off = kWave / 2; off > 0; off >>= 1) {
      c1 += __shfl_xor(c1, off, kWave);
      c2 += __shfl_xor(c2, off, kWave);
      c3 += __shfl_xor(c3, off, kWave);
    }
    if (lane == 0) {
      const uint64_t s1 = c1 < len ? uint64_t(c1) : 0xffffu;
      const uint64_t s2 = c2 < len ? uint64_t(c2) : 0xffffu;
      const uint64_t s3 = c3 < len ? uint64_t(c3) : 0xffffu;
      split[i] = s1 | (s2 << 16) | (s3 << 32);
    }
    for (int64_t d = prp[i] + lane; d < prp[i + 1]; d += kWave) lo[d] = uint16_t(pcol[d] & 0xffff);
  }
}

// Column unions of the LDS-staged forward (k_fwd1l), one workgroup per
// group of `rows` plan rows.  The voxels are taken in windows of
// kUnionWin (256,000: a window's bitmap and popcount prefix, 8,000 words
// each, and 257 partials fit 64 KB of LDS); per window a bitmap of the group's columns in it
// and the words' exclusive popcount prefix, offset by the union entries of
// the windows before.  pass 0 writes the union size cnt[group]; pass 1 the
// sorted union ucol[up[group] + r] and every padded entry's rank lcol[d] in
// it.  Each window re-reads the group's entries (grids of more than one
// window: a build-time cost only).
constexpr int kUnionWinWords = 8000;
constexpr int kUnionWin = kUnionWinWords * 32;
__global__ __launch_bounds__(256) void k_loc_union(const int64_t* __restrict__ frp,
                                                   const int* __restrict__ fcol, int64_t m,
                                                   int rows, int n, int pass,
                                                   int64_t* __restrict__ cnt,
                                                   const int64_t* __restrict__ up,
                                                   int* __restrict__ ucol,
                                                   uint16_t* __restrict__ lcol) {
  // bitmap, prefix (nwmax words each), 256 partials and the window's total
  extern __shared__ float ldsf[];
  uint32_t* lw = reinterpret_cast<uint32_t*>(ldsf);
  const int nwmax = (min(n, kUnionWin) + 31) >> 5;
  uint32_t* bm = lw;
  uint32_t* pre = lw + nwmax;
  uint32_t* part = lw + 2 * nwmax;
  uint32_t& win_total = part[256];
  const int64_t grp = blockIdx.x;
  const int64_t r0 = grp * rows, r1 = r0 + rows < m ? r0 + rows : m;
  const int64_t d0 = frp[r0], d1 = frp[r1];
  const int64_t ub = pass ? up[grp] : 0;
  uint32_t base = 0;  // union entries of the windows before (same in every thread)
  for (int c0 = 0; c0 < n; c0 += kUnionWin) {
    const int c1 = n - c0 < kUnionWin ? n : c0 + kUnionWin;
    const int nw = (c1 - c0 + 31) >> 5;
    for (int w = threadIdx.x; w < nw; w += 256) bm[w] = 0;
    __syncthreads();
    for (int64_t d = d0 + threadIdx.x; d < d1; d += 256) {
      const int c = fcol[d];
      if (c >= c0 && c < c1) atomicOr(&bm[(c - c0) >> 5], 1u << ((c - c0) & 31));
    }
    __syncthreads();
    // thread t sums words [t*per, (t+1)*per); partial sums scanned by thread 0
    const int per = (nw + 255) / 256;
    const int w0 = threadIdx.x * per, w1 = w0 + per < nw ? w0 + per : nw;
    uint32_t t = 0;
    for (int w = w0; w < w1; ++w) t += __popc(bm[w]);
    part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int j = 0; j < 256; ++j) {
        const uint32_t v = part[j];
        part[j] = run;
        run += v;
      }
      win_total = run;
    }
    __syncthreads();
    if (pass == 1) {
      uint32_t run = base + part[threadIdx.x];
      for (int w = w0; w < w1; ++w) {
        pre[w] = run;
        run += __popc(bm[w]);
      }
      __syncthreads();
      for (int w = threadIdx.x; w < nw; w += 256) {
        uint32_t bits = bm[w];
        uint32_t r = pre[w];
        while (bits) {
          const int bb = __ffs(bits) - 1;
          ucol[ub + r] = c0 + ((w << 5) | bb);
          ++r;
          bits &= bits - 1;
        }
      }
      for (int64_t d = d0 + threadIdx.x; d < d1; d += 256) {
        const int c = fcol[d];
        if (c >= c0 && c < c1) {
          const uint32_t w = uint32_t(c - c0) >> 5;
          lcol[d] = uint16_t(pre[w] + __popc(bm[w] & ((1u << ((c - c0) & 31)) - 1)));
        }
      }
    }
    base += win_total;
    __syncthreads();  // the next window clears the bitmap and rewrites win_total
  }
  if (pass == 0 && threadIdx.x == 0) cnt[grp] = int64_t(base);
}

// scale_f = 1/sum_b partials[b][f] (FP64, fixed order); 0 for an empty frame
// one workgroup per frame f = blockIdx.x (each thread sums blocks t, t+256,
// ... in order, then a fixed tree): the same order for every launch shape
__global__ __launch_bounds__(256) void k_norms(const double* partials, int nb, int FR,
                                               float* scale, double* norm) {
  __shared__ double red[256];
  const int f = blockIdx.x;
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) t += partials[int64_t(b) * FR + f];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tot = red[0];
    norm[f] = tot;
    scale[f] = tot != 0.0 ? float(1.0 / tot) : 0.0f;
  }
}

// x = scale * x' for frames 0 .. fa - 1 of each FR-frame row (the others
// keep their values)
__global__ void k_scale_frames(int64_t n, int FR, int fa, const float* xx, const float* scale,
                               float* x) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n * FR;
       t += int64_t(gridDim.x) * blockDim.x)
    if (t % FR < fa) x[t] = scale[t % FR] * xx[t];
}

__global__ void k_sens_norm32(const double* sens, int n, float* s) {
  __shared__ double red[256];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t += sens[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  double tot = red[0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = float(sens[i] / tot);
}

}  // namespace
}  // namespace raptr

using namespace raptr;


namespace raptr {
namespace {


// ------------------------------------------------ grouped layout builds
// union lists of ng groups: count pass, scan, fill pass
template <int G, class L>
int build_groups(raptr_recon* R, const L& lists, int64_t ng, int align, int64_t** gptr,
                 int** gkey, void** gval, int64_t* gnnz, hipStream_t st) {
  int rc;
  if ((rc = ralloc(R, gptr, size_t(ng) + 1))) return rc;
  const size_t sb = scan_workspace_bytes(ng > 0 ? ng : 1);
  Sizer z;
  z.take<int64_t>(size_t(ng) + 1);
  z.take<int64_t>(4);
  z.take<char>(sb);
  void* tmp = nullptr;
  hipError_t ae = hipMalloc(&tmp, z.bytes());
  if (ae != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "grouped layout scratch: %s", hipGetErrorString(ae));
  }
  Carver cv(tmp, z.bytes());
  int64_t* cnt = cv.take<int64_t>(size_t(ng) + 1);
  int64_t* tot = cv.take<int64_t>(4);
  char* sws = cv.take<char>(sb);
  int64_t h = 0;
  if (ng > 0) {
    hipLaunchKernelGGL((k_grp_count<G, L>), dim3(grid_for(ng, 256, 1 << 16)), dim3(256), 0, st,
                       lists, ng, align, cnt);
    rc = scan_exclusive_i64(cnt, *gptr, ng, tot, sws, sb, st);
    if (!rc) {
      hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, st, *gptr, int(ng),
                         (const int64_t*)tot);
      (void)hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st);
    }
  } else {
    (void)hipMemsetAsync(*gptr, 0, 8, st);
  }
  hipError_t se = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (rc) return rc;
  if (se != hipSuccess) return fail(RAPTR_ERR_DEVICE, "grouped layout: %s", hipGetErrorString(se));
  *gnnz = h;
  // tail padding: pipelined kernels read up to kGrpPad union entries past
  // a list's end (values unused)
  if ((rc = ralloc(R, gkey, size_t(h) + kGrpPad))) return rc;
  if (R->vt == kF32) {
    float* p;
    if ((rc = ralloc(R, &p, G * (size_t(h) + kGrpPad)))) return rc;
    *gval = p;
  } else {
    uint16_t* p;
    if ((rc = ralloc(R, &p, G * (size_t(h) + kGrpPad)))) return rc;
    *gval = p;
  }
  // keep the padding defined (column / row 0, value 0)
  const size_t vb = R->vt == kF32 ? 4 : 2;
  (void)hipMemsetAsync(*gkey + h, 0, kGrpPad * sizeof(int), st);
  (void)hipMemsetAsync(static_cast<char*>(*gval) + G * size_t(h) * vb, 0, G * kGrpPad * vb, st);
  if (ng > 0) {
    hipLaunchKernelGGL((k_grp_fill<G, L>), dim3(grid_for(ng, 256, 1 << 16)), dim3(256), 0, st,
                       lists, ng, (const int64_t*)*gptr, (const float*)R->rscale, R->vt, *gkey,
                       *gval);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(RAPTR_ERR_DEVICE, "grouped fill: %s", hipGetErrorString(le));
  }
  return RAPTR_OK;
}

// grouped kernel shapes compiled per group size: (unroll, pipelined)
// kind: 0 plain, 1 pipelined (scalar loads, vector FMAs), 2 matrix cores
bool grp_shape_ok(int G, int U, int kind) {
  switch (G) {
    case 16: return (U == 8 || U == 16 || U == 32) && (kind == 2 || kind == 3);
    case 32: return (U == 8 || U == 16) && kind == 3;
  }
  return false;
}

// ---------------------------------------------------------- launches
int launch_fwd_any(const raptr_recon* R, int F, const float* x, const float* y, float* e,
                   hipStream_t st) {
  if (R->m == 0) return RAPTR_OK;
  if (R->grouped_only && !(F == 64 && R->grp64))
    return fail(RAPTR_ERR_INVALID_ARG, "a grouped-only plan runs 64-frame grouped passes only");
  const int64_t m = R->m;
  const int U = R->fwd_unroll;
  if (F == 1 && R->fwd_kind == 2 && R->lcol) {
    const int L = R->fwd_lanes;
    const unsigned g = unsigned((int64_t(m) + R->lrows - 1) / R->lrows);
    const size_t shm = size_t(R->lmax > 0 ? R->lmax : 1) * sizeof(float);
#define RAPTR_FL(VV, LL, UU)                                                                 \
  if (R->vt == VV && L == LL && U == UU) {                                                   \
    hipLaunchKernelGGL((k_fwd1l<VV, LL, UU>), dim3(g), dim3(256), shm, st, m, R->lrows,       \
                       R->frp,                                                               \
                       (const uint16_t*)R->lcol, (const int64_t*)R->lup,                     \
                       (const int*)R->lucol, (const typename Val<VV>::T*)R->fval, R->rscale, \
                       x, y, e);                                                             \
    RAPTR_LAUNCH_CHECK();                                                                    \
    return RAPTR_OK;                                                                         \
  }
#define RAPTR_FLV(VV)                                                                        \
  RAPTR_FL(VV, 16, 1) RAPTR_FL(VV, 32, 1) RAPTR_FL(VV, 64, 1) RAPTR_FL(VV, 16, 2)            \
  RAPTR_FL(VV, 32, 2) RAPTR_FL(VV, 64, 2) RAPTR_FL(VV, 16, 4) RAPTR_FL(VV, 32, 4)            \
  RAPTR_FL(VV, 64, 4)
    RAPTR_FLV(kF32) RAPTR_FLV(kF16) RAPTR_FLV(kBF16) RAPTR_FLV(kU16)
#undef RAPTR_FLV
#undef RAPTR_FL
    return fail(RAPTR_ERR_INVALID_ARG, "forward shape lanes %d unroll %d unsupported", L, U);
  }
  if (F == 1 && R->fcol16 && R->fwd_c16) {
    const int L = R->fwd_lanes;
    const unsigned g = unsigned((m + (256 / L) - 1) / (256 / L));
#define RAPTR_F1C(VV, LL, UU)                                                                \
  if (R->vt == VV && L == LL && U == UU) {                                                   \
    hipLaunchKernelGGL((k_fwd1c<VV, LL, UU>), dim3(g), dim3(256), 0, st, m, R->frp,           \
                       (const uint16_t*)R->fcol16, (const uint64_t*)R->fsplit,               \
                       (const typename Val<VV>::T*)R->fval, R->rscale, x, y, e);             \
    RAPTR_LAUNCH_CHECK();                                                                    \
    return RAPTR_OK;                                                                         \
  }
#define RAPTR_F1CV(VV)                                                                       \
  RAPTR_F1C(VV, 16, 1) RAPTR_F1C(VV, 32, 1) RAPTR_F1C(VV, 64, 1) RAPTR_F1C(VV, 16, 2)        \
  RAPTR_F1C(VV, 32, 2) RAPTR_F1C(VV, 64, 2) RAPTR_F1C(VV, 16, 4) RAPTR_F1C(VV, 32, 4)        \
  RAPTR_F1C(VV, 64, 4)
    RAPTR_F1CV(kF32) RAPTR_F1CV(kF16) RAPTR_F1CV(kBF16) RAPTR_F1CV(kU16)
#undef RAPTR_F1CV
#undef RAPTR_F1C
    return fail(RAPTR_ERR_INVALID_ARG, "forward shape lanes %d unroll %d unsupported", L, U);
  }
  if (F == 1) {
    const int L = R->fwd_lanes;
    unsigned g = unsigned((m + (256 / L) - 1) / (256 / L));
#define RAPTR_F1(VV, LL, UU)                                                                 \
  if (R->vt == VV && L == LL && U == UU) {                                                   \
    hipLaunchKernelGGL((k_fwd1<VV, LL, UU>), dim3(g), dim3(256), 0, st, m, R->frp, R->fcol,   \
                       (const typename Val<VV>::T*)R->fval, R->rscale, x, y, e);             \
    RAPTR_LAUNCH_CHECK();                                                                    \
    return RAPTR_OK;                                                                         \
  }
#define RAPTR_F1V(VV)                                                                        \
  RAPTR_F1(VV, 16, 1) RAPTR_F1(VV, 32, 1) RAPTR_F1(VV, 64, 1) RAPTR_F1(VV, 16, 2)            \
  RAPTR_F1(VV, 32, 2) RAPTR_F1(VV, 64, 2) RAPTR_F1(VV, 16, 4) RAPTR_F1(VV, 32, 4)            \
  RAPTR_F1(VV, 64, 4)
    RAPTR_F1V(kF32) RAPTR_F1V(kF16) RAPTR_F1V(kBF16) RAPTR_F1V(kU16)
#undef RAPTR_F1V
#undef RAPTR_F1
    return fail(RAPTR_ERR_INVALID_ARG, "forward shape lanes %d unroll %d unsupported", L, U);
  }
  unsigned g = unsigned((m + 3) / 4);
  if (F == 64 && R->grp64) {
    // kinds 2 and 3 share the forward; frame blocks: the plan's active
    // frames (R->fa) over the group's W columns
    const unsigned gg = unsigned((R->ngf + 3) / 4);
    const int GU = R->grp_unroll, GG = R->grp_g;
    const int FBA = R->fa / GG;
#define RAPTR_M64(VV, G_, UU, FB_)                                                           \
    if (R->vt == VV && GG == G_ && GU == MF<G_>::K * UU && FBA == FB_) {                     \
      hipLaunchKernelGGL((k_fwdm64<VV, G_, UU, FB_>), dim3(gg), dim3(256), 0, st, R->ngf, m,  \
                         R->gfp, R->gfcol, (const typename Val<VV>::T*)R->gfval, R->rscale,  \
                         x, y, e);                                                           \
      RAPTR_LAUNCH_CHECK();                                                                  \
      return RAPTR_OK;                                                                       \
    }
#define RAPTR_M64V(VV)                                                                       \
  RAPTR_M64(VV, 16, 2, 4) RAPTR_M64(VV, 16, 4, 4) RAPTR_M64(VV, 16, 8, 4)                     \
  RAPTR_M64(VV, 16, 2, 2) RAPTR_M64(VV, 16, 4, 2) RAPTR_M64(VV, 16, 8, 2)                     \
  RAPTR_M64(VV, 16, 2, 1) RAPTR_M64(VV, 16, 4, 1) RAPTR_M64(VV, 16, 8, 1)                     \
  RAPTR_M64(VV, 32, 4, 2) RAPTR_M64(VV, 32, 8, 2)
    RAPTR_M64V(kF32) RAPTR_M64V(kF16) RAPTR_M64V(kBF16) RAPTR_M64V(kU16)
#undef RAPTR_M64V
#undef RAPTR_M64
    return fail(RAPTR_ERR_INVALID_ARG, "matrix-core forward shape unsupported");
  }
  if (F == 64 && R->bcast64) {
#define RAPTR_F64(VV)                                                                        \
    if (R->vt == VV) {                                                                       \
      hipLaunchKernelGGL((k_fwd64<VV>), dim3(g), dim3(256), 0, st, m, R->frp, R->fcol,        \
                         (const typename Val<VV>::T*)R->fval, R->rscale, x, y, e);            \
      RAPTR_LAUNCH_CHECK();                                                                  \
      return RAPTR_OK;                                                                       \
    }
    RAPTR_F64(kF32) RAPTR_F64(kF16) RAPTR_F64(kBF16) RAPTR_F64(kU16)
#undef RAPTR_F64
  }
#define RAPTR_FF(VV, FF, UU)                                                               \
  if (R->vt == VV && F == FF && U == UU) {                                                 \
    hipLaunchKernelGGL((k_fwdF<VV, FF, UU>), dim3(g), dim3(256), 0, st, m, R->frp, R->fcol, \
                       (const typename Val<VV>::T*)R->fval, R->rscale, x, y, e);           \
    RAPTR_LAUNCH_CHECK();                                                                  \
    return RAPTR_OK;                                                                       \
  }
#define RAPTR_FFU(VV, FF) RAPTR_FF(VV, FF, 1) RAPTR_FF(VV, FF, 2) RAPTR_FF(VV, FF, 4)
#define RAPTR_FFV(VV)                                                                      \
  RAPTR_FFU(VV, 2) RAPTR_FFU(VV, 4) RAPTR_FFU(VV, 8) RAPTR_FFU(VV, 16) RAPTR_FFU(VV, 32)    \
  RAPTR_FFU(VV, 64)
  RAPTR_FFV(kF32) RAPTR_FFV(kF16) RAPTR_FFV(kBF16) RAPTR_FFV(kU16)
#undef RAPTR_FFV
#undef RAPTR_FFU
#undef RAPTR_FF
  return fail(RAPTR_ERR_INVALID_ARG, "frame count %d / unroll %d unsupported", F, U);
}

// one launch over row blocks [blk, blk + nbl)
int launch_bwd_any(const raptr_recon* R, int F, int blk, int nbl, const float* x,
                   hipStream_t st, int* nb_out) {
  if (R->grouped_only && !(F == 64 && R->grp64))
    return fail(RAPTR_ERR_INVALID_ARG, "a grouped-only plan runs 64-frame grouped passes only");
  const int n = R->n;
  unsigned g = unsigned((n + 3) / 4);
  if (nb_out) *nb_out = int(g);
  const int64_t* ptr = R->bptr + int64_t(blk) * (n + 1);
  int first = blk == 0, last = blk + nbl == R->nblk;
  if (F == 64 && R->grp64) {
    // kind 2: one launch per group of row blocks into the shared image sums
    // (kind 3 runs launch_bwd_xcd instead); all 64 frames
    if (R->fa != 64)
      return fail(RAPTR_ERR_INVALID_ARG, "passes over fewer than 64 frames need kind 3");
    const unsigned gg = unsigned((R->ngc + 3) / 4);
    if (nb_out) *nb_out = int(gg);
    const int64_t* gp = R->gbp + int64_t(blk) * R->ngc;
    // backward lists are aligned to 8 entries: steps of 8 (U = 2)
#define RAPTR_MB64(VV)                                                                       \
    if (R->vt == VV && R->grp_g == 16) {                                                     \
      hipLaunchKernelGGL((k_bwdm64<VV, 2>), dim3(gg), dim3(256), 0, st, int(R->ngc), gp, nbl, \
                         R->gbrow, (const typename Val<VV>::T*)R->gbval,                     \
                         (const int*)R->gcols, R->e, R->c, first, last, x, R->s, R->xx,      \
                         R->partials);                                                       \
      RAPTR_LAUNCH_CHECK();                                                                  \
      return RAPTR_OK;                                                                       \
    }
    RAPTR_MB64(kF32) RAPTR_MB64(kF16) RAPTR_MB64(kBF16) RAPTR_MB64(kU16)
#undef RAPTR_MB64
    return fail(RAPTR_ERR_INVALID_ARG, "matrix-core backward shape unsupported");
  }
  if (F == 64 && R->bcast64) {
#define RAPTR_B64(VV)                                                                        \
    if (R->vt == VV) {                                                                       \
      hipLaunchKernelGGL((k_bwd64<VV>), dim3(g), dim3(256), 0, st, n, ptr, nbl, R->brow,      \
                         (const typename Val<VV>::T*)R->bval, R->e, R->c, first, last, x,     \
                         R->s, R->xx, R->partials);                                          \
      RAPTR_LAUNCH_CHECK();                                                                  \
      return RAPTR_OK;                                                                       \
    }
    RAPTR_B64(kF32) RAPTR_B64(kF16) RAPTR_B64(kBF16) RAPTR_B64(kU16)
#undef RAPTR_B64
  }
  int U = R->bwd_unroll;
#define RAPTR_B(VV, FF, UU)                                                                 \
  if (R->vt == VV && F == FF && U == UU) {                                                  \
    hipLaunchKernelGGL((k_bwd<VV, FF, UU>), dim3(g), dim3(256), 0, st, n, ptr, nbl, R->brow, \
                       (const typename Val<VV>::T*)R->bval, R->e, R->c, first, last, x,     \
                       R->s, R->xx, R->partials);                                           \
    RAPTR_LAUNCH_CHECK();                                                                   \
    return RAPTR_OK;                                                                        \
  }
#define RAPTR_BU(VV, FF) RAPTR_B(VV, FF, 1) RAPTR_B(VV, FF, 2) RAPTR_B(VV, FF, 4)
  RAPTR_BU(kF32, 1) RAPTR_BU(kF32, 2) RAPTR_BU(kF32, 4) RAPTR_BU(kF32, 8) RAPTR_BU(kF32, 16)
  RAPTR_BU(kF32, 32) RAPTR_BU(kF32, 64)
  RAPTR_BU(kF16, 1) RAPTR_BU(kF16, 2) RAPTR_BU(kF16, 4) RAPTR_BU(kF16, 8) RAPTR_BU(kF16, 16)
  RAPTR_BU(kF16, 32) RAPTR_BU(kF16, 64)
  RAPTR_BU(kBF16, 1) RAPTR_BU(kBF16, 2) RAPTR_BU(kBF16, 4) RAPTR_BU(kBF16, 8)
  RAPTR_BU(kBF16, 16) RAPTR_BU(kBF16, 32) RAPTR_BU(kBF16, 64)
  RAPTR_BU(kU16, 1) RAPTR_BU(kU16, 2) RAPTR_BU(kU16, 4) RAPTR_BU(kU16, 8) RAPTR_BU(kU16, 16)
  RAPTR_BU(kU16, 32) RAPTR_BU(kU16, 64)
#undef RAPTR_BU
#undef RAPTR_B
  return fail(RAPTR_ERR_INVALID_ARG, "backward configuration unsupported");
}

// single frame, per-XCD row ranges (backward mode 3)
int launch_bwd1_xcd(const raptr_recon* R, const float* x, hipStream_t st, int* nb_out) {
  if (!R->c8 || !R->bptr) return fail(RAPTR_ERR_INVALID_ARG, "plan has no per-XCD partial sums");
  const int n = R->n;
  // lanes per column: 64 unless set to 16 or 32 (set_backward)
  const int L = R->bwd_lpc == 16 || R->bwd_lpc == 32 ? R->bwd_lpc : 64;
  const unsigned gg = unsigned(8 * ((int64_t(n) + (256 / L) - 1) / (256 / L)));
  const int U = R->bwd_unroll;
  const bool r16 = R->brow16 && R->bwd_r16;
  bool launched = false;
#define RAPTR_B1X(VV, UU, LL)                                                               \
  if (R->vt == VV && U == UU && L == LL) {                                                  \
    if (r16)                                                                                \
      hipLaunchKernelGGL((k_bwd1x<VV, UU, true, LL>), dim3(gg), dim3(256), 0, st, n,         \
                         R->nblk, R->blk_rows, (const int64_t*)R->bptr,                     \
                         (const void*)R->brow16, (const typename Val<VV>::T*)R->bval,       \
                         (const float*)R->e, R->c8);                                        \
    else                                                                                    \
      hipLaunchKernelGGL((k_bwd1x<VV, UU, false, LL>), dim3(gg), dim3(256), 0, st, n,        \
                         R->nblk, R->blk_rows, (const int64_t*)R->bptr,                     \
                         (const void*)R->brow, (const typename Val<VV>::T*)R->bval,         \
                         (const float*)R->e, R->c8);                                        \
    launched = true;                                                                        \
  }
#define RAPTR_B1XV(VV)                                                                      \
  RAPTR_B1X(VV, 1, 16) RAPTR_B1X(VV, 2, 16) RAPTR_B1X(VV, 4, 16) RAPTR_B1X(VV, 1, 32)       \
  RAPTR_B1X(VV, 2, 32) RAPTR_B1X(VV, 4, 32) RAPTR_B1X(VV, 1, 64) RAPTR_B1X(VV, 2, 64)       \
  RAPTR_B1X(VV, 4, 64)
  RAPTR_B1XV(kF32) RAPTR_B1XV(kF16) RAPTR_B1XV(kBF16) RAPTR_B1XV(kU16)
#undef RAPTR_B1XV
#undef RAPTR_B1X
  if (!launched)
    return fail(RAPTR_ERR_INVALID_ARG, "per-XCD backward unroll %d lanes %d unsupported", U, L);
  RAPTR_LAUNCH_CHECK();
  const unsigned nr = unsigned((int64_t(n) + 255) / 256);
  hipLaunchKernelGGL(k_bwd1_xreduce, dim3(nr), dim3(256), 0, st, n, (const float*)R->c8, x,
                     (const float*)R->s, R->xx, R->partials);
  RAPTR_LAUNCH_CHECK();
  if (nb_out) *nb_out = int(nr);
  return RAPTR_OK;
}

// 64-frame backward with per-XCD row ranges and partial images, then the
// reduce + update (grouped layouts of 16, kind 3)
int launch_bwd_xcd(const raptr_recon* R, const float* x, hipStream_t st, int* nb_out) {
  if (!R->c8 || R->grp_g < 16)
    return fail(RAPTR_ERR_INVALID_ARG, "plan has no per-XCD partial images");
  const unsigned gg = unsigned(8 * ((R->ngc + 3) / 4));
  const int per_xcd = (R->nblk + 7) / 8;  // most blocks of any XCD's range
  const int bpl = R->xcd_bpl > 0 ? R->xcd_bpl : per_xcd;
  const int FBA = R->fa / R->grp_g;
  for (int boff = 0; boff < per_xcd; boff += bpl) {
    const int accum = boff > 0;
    bool launched = false;
#define RAPTR_X64(VV, G_, FB_)                                                               \
  if (R->vt == VV && R->grp_g == G_ && FBA == FB_) {                                         \
    hipLaunchKernelGGL((k_bwdm64x<VV, G_, FB_>), dim3(gg), dim3(256), 0, st, int(R->ngc),     \
                       R->nblk, R->n, (const int64_t*)R->gbp, (const int*)R->gbrow,          \
                       (const typename Val<VV>::T*)R->gbval, (const int*)R->gcols,           \
                       (const float*)R->e, R->c8, boff, bpl, accum);                         \
    launched = true;                                                                         \
  }
#define RAPTR_X64V(VV)                                                                       \
  RAPTR_X64(VV, 16, 4) RAPTR_X64(VV, 16, 2) RAPTR_X64(VV, 16, 1) RAPTR_X64(VV, 32, 2)
    RAPTR_X64V(kF32) RAPTR_X64V(kF16) RAPTR_X64V(kBF16) RAPTR_X64V(kU16)
#undef RAPTR_X64V
#undef RAPTR_X64
    if (!launched) return fail(RAPTR_ERR_INVALID_ARG, "per-XCD backward shape unsupported");
    RAPTR_LAUNCH_CHECK();
  }
  const unsigned nr = unsigned((int64_t(R->n) + 3) / 4);
  hipLaunchKernelGGL(k_bwd_xreduce, dim3(nr), dim3(256), 0, st, R->n, R->fa,
                     (const float*)R->c8, x, (const float*)R->s, R->xx, R->partials);
  RAPTR_LAUNCH_CHECK();
  if (nb_out) *nb_out = int(nr);
  return RAPTR_OK;
}

// LDS bytes of the fused pass with W waves per group: the group's row
// bounds, x over the largest union and W column-sum regions, or one when
// the waves share it (one spare slot each)
size_t fused_lds_bytes(const raptr_recon* R, int W, bool shared) {
  const size_t u = size_t(std::max(R->lmax, 1));
  return size_t(R->lrows + 1) * sizeof(int64_t) +
         (u + size_t(shared ? 1 : W) * (u + 1)) * sizeof(float);
}
constexpr size_t kLdsMax = 160 * 1024;

// single frame: forward, ratio and backward in one pass (k_fb1l), then the
// fixed-order column reduce + update
// the dynamic-LDS limit of every fused products kernel raised to the CU's
// 160 KiB, once per process (at plan build, outside any graph capture);
// not fatal where the default already suffices
void fused_lds_attributes() {
  static std::once_flag once;
  std::call_once(once, [] {
#define RAPTR_FA(VV, WW, LL, SS, SH)                                                         \
  if (hipFuncSetAttribute((const void*)k_fb1l<VV, WW, LL, SS, SH>,                           \
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsMax)) !=       \
      hipSuccess)                                                                            \
    (void)hipGetLastError();
#define RAPTR_FAS(VV, WW, SH)                                                                \
  RAPTR_FA(VV, WW, 16, 6, SH) RAPTR_FA(VV, WW, 32, 3, SH) RAPTR_FA(VV, WW, 64, 2, SH)        \
  RAPTR_FA(VV, WW, 64, 3, SH) RAPTR_FA(VV, WW, 64, 4, SH)
#define RAPTR_FAV(VV)                                                                        \
  RAPTR_FAS(VV, 2, false) RAPTR_FAS(VV, 2, true) RAPTR_FAS(VV, 4, false)                     \
  RAPTR_FAS(VV, 4, true) RAPTR_FAS(VV, 8, true)
    RAPTR_FAV(kF32) RAPTR_FAV(kF16) RAPTR_FAV(kBF16) RAPTR_FAV(kU16)
#undef RAPTR_FAV
#undef RAPTR_FAS
#undef RAPTR_FA
  });
}

// part: 0 both kernels, 1 the products (k_fb1l) only, 2 the column reduce
// and update only
// waves per workgroup of the column-tile reduce: as many as keep its LDS
// (one region of 2^tshift floats per slot of lpc lanes) within 32 KiB, 1-8
inline int treduce_waves(int tshift, int lpc) {
  const int64_t per_wave = (int64_t(64) / lpc) << tshift;  // floats
  int w = 1;
  while (w < 8 && per_wave * 2 * w * int64_t(sizeof(float)) <= 32 * 1024) w *= 2;
  return w;
}

int launch_fused1(const raptr_recon* R, const float* x, const float* y, hipStream_t st,
                  int* nb_out, int part = 0) {
  if (!R->cpart || !R->vpos) return fail(RAPTR_ERR_INVALID_ARG, "plan has no fused pass");
  const int64_t m = R->m;
  const int n = R->n;
  const int W = R->fb_waves, LR = R->fb_lanes;
  const bool SHR = R->fused == 2;
  const size_t shm = fused_lds_bytes(R, W, SHR);
  if (shm > kLdsMax) return fail(RAPTR_ERR_INVALID_ARG, "fused pass: %zu B of LDS", shm);
  if (m > 0 && part != 2) {
    const unsigned g = unsigned((m + R->lrows - 1) / R->lrows);
    bool launched = false;
#define RAPTR_FB(VV, WW, LL, SS, SH)                                                         \
  if (R->vt == VV && W == WW && LR == LL && SHR == SH && (LL != 64 || SS == R->fb_steps)) {  \
    hipLaunchKernelGGL((k_fb1l<VV, WW, LL, SS, SH>), dim3(g), dim3(64 * WW), shm, st, m,      \
                       R->lrows, R->frp,                                                     \
                       (const uint16_t*)R->lcol, (const int64_t*)R->lup,                     \
                       (const int*)R->lucol, (const typename Val<VV>::T*)R->fval, R->rscale, \
                       x, y, R->e, R->cpart);                                                \
    launched = true;                                                                         \
  }
#define RAPTR_FBS(VV, WW, SH)                                                                \
  RAPTR_FB(VV, WW, 16, 6, SH) RAPTR_FB(VV, WW, 32, 3, SH) RAPTR_FB(VV, WW, 64, 2, SH)        \
  RAPTR_FB(VV, WW, 64, 3, SH) RAPTR_FB(VV, WW, 64, 4, SH)
#define RAPTR_FBW(VV, WW) RAPTR_FBS(VV, WW, false) RAPTR_FBS(VV, WW, true)
#define RAPTR_FBV(VV) RAPTR_FBW(VV, 2) RAPTR_FBW(VV, 4) RAPTR_FBS(VV, 8, true)
    RAPTR_FBV(kF32) RAPTR_FBV(kF16) RAPTR_FBV(kBF16) RAPTR_FBV(kU16)
#undef RAPTR_FBV
#undef RAPTR_FBW
#undef RAPTR_FBS
#undef RAPTR_FB
    if (!launched)
      return fail(RAPTR_ERR_INVALID_ARG, "fused pass: %d waves, %d lanes per row unsupported", W,
                  LR);
    RAPTR_LAUNCH_CHECK();
  }
  if (part == 1) return RAPTR_OK;
  const int L = R->fb_lpc;
  if (R->fb_xcd == 2) {
    if (!R->tseg) return fail(RAPTR_ERR_INVALID_ARG, "fused pass: no column-tile segments");
    const int TW = treduce_waves(R->tshift, L);
    const size_t tshm = size_t(TW) * size_t(64 / L) * (size_t(1) << R->tshift) * sizeof(float);
    bool launched = false;
#define RAPTR_FT(LL, WW)                                                                       \
  if (L == LL && TW == WW) {                                                                   \
    hipLaunchKernelGGL((k_fb1_treduce<LL, WW>), dim3(R->tnwg), dim3(64 * WW), tshm, st, n,     \
                       R->tshift, (const int4*)R->twg, (const int2*)R->tseg,                   \
                       (const uint16_t*)R->tcol, (const float*)R->cpart, R->ck);               \
    launched = true;                                                                           \
  }
#define RAPTR_FTL(LL) RAPTR_FT(LL, 1) RAPTR_FT(LL, 2) RAPTR_FT(LL, 4) RAPTR_FT(LL, 8)
    RAPTR_FTL(16) RAPTR_FTL(32) RAPTR_FTL(64)
#undef RAPTR_FTL
#undef RAPTR_FT
    if (!launched) return fail(RAPTR_ERR_INVALID_ARG, "reduce lanes %d", L);
    RAPTR_LAUNCH_CHECK();
    const unsigned nr = unsigned((int64_t(n) + 255) / 256);
    hipLaunchKernelGGL(k_fb1_tcombine, dim3(nr), dim3(256), 0, st, n, R->tshift,
                       (const int*)R->tk, (const float*)R->ck, x, (const float*)R->s, R->xx,
                       R->partials);
    RAPTR_LAUNCH_CHECK();
    if (nb_out) *nb_out = int(nr);
    return RAPTR_OK;
  }
  if (R->fb_xcd) {
    const unsigned gg = unsigned(8 * ((int64_t(n) + (256 / L) - 1) / (256 / L)));
#define RAPTR_FR8(LL)                                                                          \
  if (L == LL)                                                                                 \
    hipLaunchKernelGGL((k_fb1_reduce8<LL, 4>), dim3(gg), dim3(256), 0, st, n,                   \
                       (const int64_t*)R->vptr, (const int64_t*)R->vb, (const int*)R->vpos,    \
                       (const float*)R->cpart, R->c8);
    RAPTR_FR8(16) else RAPTR_FR8(32) else RAPTR_FR8(64) else return fail(RAPTR_ERR_INVALID_ARG,
                                                                      "reduce lanes %d", L);
#undef RAPTR_FR8
    RAPTR_LAUNCH_CHECK();
    const unsigned nr = unsigned((int64_t(n) + 255) / 256);
    hipLaunchKernelGGL(k_bwd1_xreduce, dim3(nr), dim3(256), 0, st, n, (const float*)R->c8, x,
                       (const float*)R->s, R->xx, R->partials);
    RAPTR_LAUNCH_CHECK();
    if (nb_out) *nb_out = int(nr);
    return RAPTR_OK;
  }
  const unsigned nr = unsigned((int64_t(n) + (256 / L) - 1) / (256 / L));
#define RAPTR_FR(LL)                                                                           \
  if (L == LL)                                                                                 \
    hipLaunchKernelGGL((k_fb1_reduce<LL, 4>), dim3(nr), dim3(256), 0, st, n,                    \
                       (const int64_t*)R->vptr, (const int*)R->vpos, (const float*)R->cpart,   \
                       x, (const float*)R->s, R->xx, R->partials);
  RAPTR_FR(16) else RAPTR_FR(32) else RAPTR_FR(64) else return fail(RAPTR_ERR_INVALID_ARG,
                                                                    "reduce lanes %d", L);
#undef RAPTR_FR
  RAPTR_LAUNCH_CHECK();
  if (nb_out) *nb_out = int(nr);
  return RAPTR_OK;
}

int iteration(raptr_recon* R, int F, const float* y, float* x, hipStream_t st) {
  int nb = 0;
  if (F == 1 && R->fused && R->cpart) {
    RAPTR_TRY(launch_fused1(R, x, y, st, &nb));
  } else if (F == 1 && R->bwd_lds == 3) {
    RAPTR_TRY(launch_fwd_any(R, F, x, y, R->e, st));
    RAPTR_TRY(launch_bwd1_xcd(R, x, st, &nb));
  } else if (F == 64 && R->grp64 && R->grp_pipe == 3) {
    RAPTR_TRY(launch_fwd_any(R, F, x, y, R->e, st));
    RAPTR_TRY(launch_bwd_xcd(R, x, st, &nb));
  } else {
    RAPTR_TRY(launch_fwd_any(R, F, x, y, R->e, st));
    for (int b = 0; b < R->nblk; b += R->bwd_bpl)
      RAPTR_TRY(launch_bwd_any(R, F, b, std::min(R->bwd_bpl, R->nblk - b), x, st, &nb));
  }
  // active frames: all F, or the plan's fa of a 64-frame grouped pass
  const int FA = F == 64 && R->grp64 ? R->fa : F;
  hipLaunchKernelGGL(k_norms, dim3(FA), dim3(256), 0, st, (const double*)R->partials, nb, F,
                     R->scale, R->norm);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scale_frames, dim3(grid_for(int64_t(R->n) * F, 256, 8192)), dim3(256),
                     0, st, int64_t(R->n), F, FA, (const float*)R->xx,
                     (const float*)R->scale, x);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

bool valid_frames(int F) { return F >= 1 && F <= 64 && (F & (F - 1)) == 0; }

// frame-slot plumbing of the frames-fastest [rows][F] layout (serving)
__global__ void k_frame_fill(float* a, int64_t rows, int F, int slot, float v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < rows;
       i += int64_t(gridDim.x) * blockDim.x)
    a[i * F + slot] = v;
}

__global__ void k_frame_scatter(float* a, int64_t rows, int F, int slot, const int* idx,
                                const float* val, int64_t n, int* bad) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int r = idx[i];
    if (r >= 0 && r < rows)
      a[int64_t(r) * F + slot] = val[i];
    else
      bad[0] = 1;
  }
}

__global__ void k_frame_gather(const float* a, int64_t rows, int F, int slot, float* out) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < rows;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = a[i * F + slot];
}

__global__ void k_frame_copy(float* a, int64_t rows, int F, int src, int dst) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < rows;
       i += int64_t(gridDim.x) * blockDim.x)
    a[i * F + dst] = a[i * F + src];
}

// ---------------------------------------------------- plan row order
// A plan may keep its rows in an order of its own (plan row p holds the
// caller's row order[p]).  Row order changes which rows share cache lines in
// the backward's ratio gather; it changes no row's sum.

// seen[] counts how often each caller row occurs; *bad = 1 on a miss
__global__ void k_perm_check(const int* __restrict__ order, int m, int* __restrict__ seen,
                             int* __restrict__ bad) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < m;
       p += int64_t(gridDim.x) * blockDim.x) {
    const int o = order[p];
    if (o < 0 || o >= m || atomicAdd(&seen[o], 1) != 0) bad[0] = 1;
  }
}

__global__ void k_perm_len(const int64_t* __restrict__ rp, const int* __restrict__ order,
                           int m, int64_t* __restrict__ plen) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < m;
       p += int64_t(gridDim.x) * blockDim.x) {
    const int o = order[p];
    plen[p] = rp[o + 1] - rp[o];
  }
}

// one wave per plan row copies the caller row's entries in their order
__global__ __launch_bounds__(256) void k_perm_copy(const int64_t* __restrict__ rp,
                                                   const int* __restrict__ col,
                                                   const float* __restrict__ val,
                                                   const int* __restrict__ order, int m,
                                                   const int64_t* __restrict__ prp,
                                                   int* __restrict__ pcol,
                                                   float* __restrict__ pval) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t waves = int64_t(gridDim.x) * (blockDim.x / kWave);
  for (int64_t p = int64_t(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; p < m;
       p += waves) {
    const int o = order[p];
    const int64_t b = rp[o], len = rp[o + 1] - b, d = prp[p];
    for (int64_t k = lane; k < len; k += kWave) {
      pcol[d + k] = col[b + k];
      pval[d + k] = val[b + k];
    }
  }
}

// dst[p][f] = src[order[p]][f]  (caller rows -> plan rows)
__global__ void k_rows_gather(const float* __restrict__ src, float* __restrict__ dst,
                              const int* __restrict__ order, int64_t m, int F) {
  const int64_t total = m * F;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = i / F;
    dst[i] = src[int64_t(order[p]) * F + (i - p * F)];
  }
}

// dst[order[p]][f] = src[p][f]  (plan rows -> caller rows)
__global__ void k_rows_scatter(const float* __restrict__ src, float* __restrict__ dst,
                               const int* __restrict__ order, int64_t m, int F) {
  const int64_t total = m * F;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = i / F;
    dst[int64_t(order[p]) * F + (i - p * F)] = src[i];
  }
}

// Locality key of a two-head channel: angle, then the line's midpoint
// between the heads (z0 + z1, y0 + y1) in tiles of T x T midpoints (tile
// first, then the midpoint inside it), then its slope (z1 - z0, y1 - y0).
// Lines through one voxel then sit in a few runs of consecutive keys, and a
// run of consecutive rows covers a compact bundle of lines (T > 1).
// kind 0: (angle, midpoint tile, midpoint in tile, slope); kind 1 (y-major):
// (midpoint y, angle, midpoint z tile, midpoint z in tile, slope); kind 2:
// as kind 0 with the slopes in 4 x 4 tiles (slope tile, slope in tile);
// kind 3: as kind 2 with angle / 4 in place of the angle and angle % 4 last
__global__ void k_locality_keys(raptr_setup_t s, const int* __restrict__ cnl, int m, int T,
                                int kind, uint64_t* __restrict__ keys,
                                uint32_t* __restrict__ idx) {
  const uint64_t zs = uint64_t(s.n0z + s.n1z - 1), ys = uint64_t(s.n0y + s.n1y - 1);
  const uint64_t nzt = (zs + T - 1) / T, nyt = (ys + T - 1) / T;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += int64_t(gridDim.x) * blockDim.x) {
    int ida, id0z, id0y, id1z, id1y;
    cnl_split(s, cnl[i], ida, id0z, id0y, id1z, id1y);
    const uint64_t zm = uint64_t(id0z + id1z), ym = uint64_t(id0y + id1y);
    uint64_t k;
    if (kind == 1) {
      k = ym;
      k = k * uint64_t(s.na) + uint64_t(ida);
      k = k * nzt + zm / T;
      k = k * T + zm % T;
    } else {
      k = uint64_t(kind == 3 ? ida / 4 : ida);
      k = k * nzt + zm / T;
      k = k * nyt + ym / T;
      k = k * T + zm % T;
      k = k * T + ym % T;
    }
    const uint64_t dz = uint64_t(id1z - id0z + s.n0z - 1), dy = uint64_t(id1y - id0y + s.n0y - 1);
    if (kind >= 2) {
      constexpr uint64_t S = 4;
      k = k * ((zs + S - 1) / S) + dz / S;
      k = k * ((ys + S - 1) / S) + dy / S;
      k = k * S + dz % S;
      k = k * S + dy % S;
      if (kind == 3) k = k * 4 + uint64_t(ida % 4);
    } else {
      k = k * zs + dz;
      k = k * ys + dy;
    }
    keys[i] = k;
    idx[i] = uint32_t(i);
  }
}

__global__ void k_u32_to_i32(const uint32_t* __restrict__ a, int m, int* __restrict__ b) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = int(a[i]);
}

// frame vectors are loaded with 4*min(F,4)-byte vector accesses
bool frames_aligned(const void* p, int F) {
  const uintptr_t a = uintptr_t(4) * uintptr_t(F >= 4 ? 4 : F);
  return p == nullptr || uintptr_t(p) % a == 0;
}

}  // namespace
}  // namespace raptr

namespace raptr {
namespace {
// column unions of the LDS-staged forward (k_fwd1l): sizes, scan, unions
// and local indices; none when a group's union exceeds kLocMaxUnion
// The column unions of groups of `rows` consecutive plan rows.  rows = 0:
// among the sizes of kLocRowsChoices that give at least kMinGroups groups
// (four workgroups per CU; without one, the smallest size), the one whose
// fused iteration the model below puts fastest.  Larger groups share more
// of their columns (full-size workload: 112M union entries for 64 rows,
// 51M for 256; 96^3 geometry: 441M for 16, 306M for 32, 153M for 128), so
// the staging of x and the partial sums shrink, while fewer workgroups fit
// a CU.  Model: the products kernel moves 6 bytes per padded entry and 12
// per union entry (column, x, partial sum) at 5.2 TB/s with three or more
// 8-wave workgroups per CU, 4.5 with two, 3.0 with one; the reduce reads 8
// bytes per union entry at 3.3 TB/s.  Measured per iteration, full size:
// 64 rows 2.38 ms, 128 2.21, 256 2.04-2.09, 512 2.36 (model 2.40, 2.22,
// 2.11, 2.36); 96^3: 16 rows 4.44, 32 3.58, 64 4.39, 128 3.86 (model
// 3.88, 3.63, 4.50, 4.07); small geometry (20,480 rows): 16 rows 0.044 ms
// (256: 0.094).
constexpr int64_t kMinGroups = 1024;
double union_model_ms(int64_t pnnz, int64_t entries, size_t lds) {
  const int per_cu = lds > 0 ? int(std::min<size_t>(4, (160 * 1024) / lds)) : 0;
  if (per_cu == 0) return 1e30;
  const double rate = per_cu >= 3 ? 5.2e9 : per_cu == 2 ? 4.5e9 : 3.0e9;  // bytes per ms
  return (6.0 * double(pnnz) + 12.0 * double(entries)) / rate + 8.0 * double(entries) / 3.3e9;
}
int build_local_columns(raptr_recon* R, hipStream_t st, int rows_req = 0) {
  const int64_t m = R->m;
  const int n = R->n;
  const size_t shm = (2 * size_t((std::min(n, kUnionWin) + 31) >> 5) + 257) * sizeof(uint32_t);
  int rc;
  int64_t* cnt = nullptr;
  int64_t h = 0, mx = 0;
  // without a request: the group sizes' union counts first (pass 0 each),
  // then the choice
  if (!rows_req) {
    const int nc = int(sizeof(kLocRowsChoices) / sizeof(int));
    std::vector<double> cost(nc, 1e30);
    for (int ci = 0; ci < nc; ++ci) {
      const int rr = kLocRowsChoices[ci];
      const int64_t ng = (m + rr - 1) / rr;
      if ((rc = ralloc(R, &cnt, size_t(ng)))) return rc;
      hipLaunchKernelGGL(k_loc_union, dim3(unsigned(ng)), dim3(256), shm, st,
                         (const int64_t*)R->frp, (const int*)R->fcol, m, rr, n, 0, cnt,
                         (const int64_t*)nullptr, (int*)nullptr, (uint16_t*)nullptr);
      std::vector<int64_t> hc(static_cast<size_t>(ng));
      hipError_t le = hipGetLastError();
      if (le == hipSuccess)
        le = hipMemcpyAsync(hc.data(), cnt, size_t(ng) * 8, hipMemcpyDeviceToHost, st);
      if (le == hipSuccess) le = hipStreamSynchronize(st);
      rfree(R, cnt);
      if (le != hipSuccess)
        return fail(RAPTR_ERR_DEVICE, "column unions: %s", hipGetErrorString(le));
      const int64_t mxc = hc.empty() ? 0 : *std::max_element(hc.begin(), hc.end());
      int64_t tot = 0;
      for (int64_t v : hc) tot += v;
      if (mxc <= kLocMaxUnion)
        cost[ci] = union_model_ms(
            R->pnnz, tot, size_t(rr + 1) * sizeof(int64_t) + size_t(2 * mxc + 1) * sizeof(float));
    }
    int pick = kLocRowsChoices[nc - 1];
    double best = 1e30;
    for (int ci = 0; ci < nc; ++ci)  // largest first: ties keep the larger groups
      if ((m + kLocRowsChoices[ci] - 1) / kLocRowsChoices[ci] >= kMinGroups && cost[ci] < best) {
        best = cost[ci];
        pick = kLocRowsChoices[ci];
      }
    return build_local_columns(R, st, pick);
  }
  const int rows = rows_req;
  {
    const int64_t ng = (m + rows - 1) / rows;
    if ((rc = ralloc(R, &R->lup, size_t(ng) + 1)) || (rc = ralloc(R, &cnt, size_t(ng)))) return rc;
    hipLaunchKernelGGL(k_loc_union, dim3(unsigned(ng)), dim3(256), shm, st,
                       (const int64_t*)R->frp, (const int*)R->fcol, m, rows, n, 0, cnt,
                       (const int64_t*)nullptr, (int*)nullptr, (uint16_t*)nullptr);
    void* tmp = nullptr;
    const size_t sb = scan_workspace_bytes(ng);
    RAPTR_HIP_TRY(hipMalloc(&tmp, 512 + sb));
    int64_t* tot = static_cast<int64_t*>(tmp);
    rc = scan_exclusive_i64(cnt, R->lup, ng, tot, static_cast<char*>(tmp) + 256, sb, st);
    if (!rc) hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, st, R->lup, int(ng),
                                (const int64_t*)tot);
    std::vector<int64_t> hc(static_cast<size_t>(ng));
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st);
    if (le == hipSuccess)
      le = hipMemcpyAsync(hc.data(), cnt, size_t(ng) * 8, hipMemcpyDeviceToHost, st);
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    if (rc) return rc;
    if (le != hipSuccess)
      return fail(RAPTR_ERR_DEVICE, "column unions: %s", hipGetErrorString(le));
    mx = hc.empty() ? 0 : *std::max_element(hc.begin(), hc.end());
    if (mx > kLocMaxUnion) {  // too spread out for LDS: the plan keeps no unions
      rfree(R, R->lup);
      rfree(R, cnt);
      return RAPTR_OK;
    }
  }
  const int64_t ng = (m + rows - 1) / rows;
  if ((rc = ralloc(R, &R->lucol, size_t(h) + 1)) || (rc = ralloc(R, &R->lcol, size_t(R->pnnz) + 8)))
    return rc;
  hipLaunchKernelGGL(k_loc_union, dim3(unsigned(ng)), dim3(256), shm, st,
                     (const int64_t*)R->frp, (const int*)R->fcol, m, rows, n, 1, cnt,
                     (const int64_t*)R->lup, R->lucol, R->lcol);
  hipError_t le = hipGetLastError();
  if (le == hipSuccess) le = hipStreamSynchronize(st);
  if (le != hipSuccess) return fail(RAPTR_ERR_DEVICE, "column unions: %s", hipGetErrorString(le));
  rfree(R, cnt);
  R->lrows = rows;
  R->lmax = int(mx);
  R->lnu = h;
  return RAPTR_OK;
}

// the segment table of the column-tile reduce for tiles of 2^tshift
// columns, each tile's segments cut into parts of about `target` segments
// (replaces the plan's table; the plan has no clones)
int build_tile_segments(raptr_recon* R, int tshift, int target, hipStream_t st) {
  const int64_t nu = R->lnu;
  const int64_t ng = (int64_t(R->m) + R->lrows - 1) / R->lrows;
  const int ntiles = int((int64_t(R->n) + (int64_t(1) << tshift) - 1) >> tshift);
  rfree(R, R->tseg);
  rfree(R, R->tcol);
  rfree(R, R->twg);
  rfree(R, R->tk);
  rfree(R, R->ck);
  R->nseg = 0;
  R->tshift = R->tseg_target = R->tnwg = R->tkmax = 0;
  Sizer z;
  z.take<uint32_t>(size_t(nu));
  z.take<uint32_t>(size_t(nu));
  z.take<uint32_t>(1);
  z.take<char>(scan_workspace_bytes(nu));
  void* tmp = nullptr;
  hipError_t ae = hipMalloc(&tmp, z.bytes() + 4096);
  if (ae != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "tile segments scratch: %s", hipGetErrorString(ae));
  }
  Carver cv(tmp, z.bytes() + 4096);
  uint32_t* flag = cv.take<uint32_t>(size_t(nu));
  uint32_t* idx = cv.take<uint32_t>(size_t(nu));
  uint32_t* d_total = cv.take<uint32_t>(1);
  const size_t wsb = scan_workspace_bytes(nu);
  char* ws = cv.take<char>(wsb);
  const unsigned gp = unsigned(grid_for(nu, 256, 1 << 16));
  hipLaunchKernelGGL(k_seg_flags, dim3(gp), dim3(256), 0, st, (const int*)R->lucol, nu, tshift,
                     flag);
  hipLaunchKernelGGL(k_seg_gstarts, dim3(grid_for(ng, 256, 1 << 16)), dim3(256), 0, st,
                     (const int64_t*)R->lup, ng, nu, flag);
  int rc = scan_exclusive_u32(flag, idx, nu, d_total, ws, wsb, st);
  uint32_t total = 0;
  if (!rc) {
    hipError_t le = hipGetLastError();
    if (le == hipSuccess)
      le = hipMemcpyAsync(&total, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "tile segments: %s", hipGetErrorString(le));
  }
  const int64_t S = total;
  void* tmp2 = nullptr;
  std::vector<int> tsb_h(size_t(ntiles) + 1);
  if (!rc) {
    Sizer z2;
    z2.take<uint64_t>(size_t(S));
    z2.take<uint64_t>(size_t(S));
    z2.take<uint32_t>(size_t(S));
    z2.take<uint32_t>(size_t(S));
    z2.take<int>(size_t(S) + 1);
    z2.take<int>(size_t(ntiles) + 1);
    const int64_t nt = (S + 2047) / 2048;
    z2.take<uint32_t>(size_t(nt) * 256);
    z2.take<uint32_t>(size_t(nt) * 256);
    z2.take<char>(scan_workspace_bytes(nt * 256));
    ae = hipMalloc(&tmp2, z2.bytes() + 4096);
    if (ae != hipSuccess) {
      (void)hipGetLastError();
      tmp2 = nullptr;
      rc = fail(RAPTR_ERR_OUT_OF_MEMORY, "tile segments scratch: %s", hipGetErrorString(ae));
    }
    if (!rc) rc = ralloc(R, &R->tseg, size_t(S));
    if (!rc) rc = ralloc(R, &R->tcol, size_t(nu));
    if (!rc) {
      Carver c2(tmp2, z2.bytes() + 4096);
      hipLaunchKernelGGL(k_tile_cols, dim3(gp), dim3(256), 0, st, (const int*)R->lucol, nu,
                         tshift, R->tcol);
      uint64_t* k0 = c2.take<uint64_t>(size_t(S));
      uint64_t* k1 = c2.take<uint64_t>(size_t(S));
      uint32_t* v0 = c2.take<uint32_t>(size_t(S));
      uint32_t* v1 = c2.take<uint32_t>(size_t(S));
      int* ps = c2.take<int>(size_t(S) + 1);
      int* tsb = c2.take<int>(size_t(ntiles) + 1);
      hipLaunchKernelGGL(k_seg_emit, dim3(gp), dim3(256), 0, st, (const int*)R->lucol, nu,
                         tshift, (const uint32_t*)flag, (const uint32_t*)idx, k0, v0, ps);
      int res = 0;
      rc = radix_sort_pairs(k0, k1, v0, v1, S, bits_for_count(ntiles), st, &res, c2);
      if (!rc) {
        hipLaunchKernelGGL(k_seg_pairs, dim3(grid_for(S, 256, 1 << 16)), dim3(256), 0, st,
                           (const uint32_t*)(res ? v1 : v0), S, (const int*)ps, R->tseg);
        hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for(int64_t(ntiles) + 1, 256, 1 << 16)),
                           dim3(256), 0, st, (const uint64_t*)(res ? k1 : k0), S, ntiles, tsb);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess)
          le = hipMemcpyAsync(tsb_h.data(), tsb, tsb_h.size() * sizeof(int),
                              hipMemcpyDeviceToHost, st);
        if (le == hipSuccess) le = hipStreamSynchronize(st);
        if (le != hipSuccess)
          rc = fail(RAPTR_ERR_DEVICE, "tile segments: %s", hipGetErrorString(le));
      }
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (tmp2) (void)hipFree(tmp2);
  // the workgroups: tile t's segments in ceil(count / target) parts of
  // equal size (one part for a tile with none)
  std::vector<int4> wg;
  std::vector<int> tk(size_t(ntiles), 1);
  int kmax = 1;
  if (!rc) {
    for (int t = 0; t < ntiles; ++t) {
      const int64_t a = tsb_h[size_t(t)], c = tsb_h[size_t(t) + 1] - a;
      const int K = int(std::max<int64_t>(1, (c + target - 1) / target));
      tk[size_t(t)] = K;
      kmax = std::max(kmax, K);
      for (int k = 0; k < K; ++k)
        wg.push_back(int4{t, k, int(a + c * k / K), int(a + c * (k + 1) / K)});
    }
    if (wg.size() > size_t(INT32_MAX))
      rc = fail(RAPTR_ERR_INVALID_ARG, "tile segments: too many workgroups");
  }
  if (!rc) rc = ralloc(R, &R->twg, wg.size());
  if (!rc) rc = ralloc(R, &R->tk, tk.size());
  if (!rc) rc = ralloc(R, &R->ck, size_t(kmax) * size_t(R->n));
  if (!rc) {
    hipError_t le = hipMemcpyAsync(R->twg, wg.data(), wg.size() * sizeof(int4),
                                   hipMemcpyHostToDevice, st);
    if (le == hipSuccess)
      le = hipMemcpyAsync(R->tk, tk.data(), tk.size() * sizeof(int), hipMemcpyHostToDevice, st);
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "tile segments: %s", hipGetErrorString(le));
  }
  if (rc) {
    rfree(R, R->tseg);
    rfree(R, R->tcol);
    rfree(R, R->twg);
    rfree(R, R->tk);
    rfree(R, R->ck);
    return rc;
  }
  R->nseg = S;
  R->tshift = tshift;
  R->tseg_target = target;
  R->tnwg = int(wg.size());
  R->tkmax = kmax;
  return RAPTR_OK;
}

// the fused pass's index: the union positions of every column, ascending
// (a stable radix sort of the positions by column), the groups' partial
// column sums and the column-tile reduce's segments; none where the unions
// do not fit LDS or a row repeats a column
int build_union_index(raptr_recon* R, hipStream_t st) {
  const int64_t nu = R->lnu;
  const int n = R->n;
  if (!R->lucol || nu <= 0 || nu > int64_t(INT32_MAX) || fused_lds_bytes(R, 2, true) > kLdsMax)
    return RAPTR_OK;
  int rc;
  {  // rows must not repeat a column
    int* dup = nullptr;
    if ((rc = ralloc(R, &dup, 1))) return rc;
    (void)hipMemsetAsync(dup, 0, sizeof(int), st);
    const int words = (R->lmax + 31) / 32;
    const unsigned ng = unsigned((int64_t(R->m) + R->lrows - 1) / R->lrows);
    const size_t shm = size_t(4) * size_t(std::max(words, 1)) * sizeof(uint32_t);
    switch (R->vt) {
#define RAPTR_FD(VV)                                                                         \
  case VV:                                                                                   \
    hipLaunchKernelGGL(k_fb_distinct<VV>, dim3(ng), dim3(256), shm, st, int64_t(R->m),        \
                       R->lrows,                                                             \
                       (const int64_t*)R->frp, (const uint16_t*)R->lcol,                     \
                       (const typename Val<VV>::T*)R->fval, words, dup);                     \
    break;
      RAPTR_FD(kF32) RAPTR_FD(kF16) RAPTR_FD(kBF16) RAPTR_FD(kU16)
#undef RAPTR_FD
    }
    int h = 0;
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipMemcpyAsync(&h, dup, sizeof(int), hipMemcpyDeviceToHost, st);
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    rfree(R, dup);
    if (le != hipSuccess) return fail(RAPTR_ERR_DEVICE, "union index: %s", hipGetErrorString(le));
    if (h) return RAPTR_OK;  // a repeated column: no fused pass
  }
  if ((rc = ralloc(R, &R->vptr, size_t(n) + 1)) || (rc = ralloc(R, &R->vpos, size_t(nu))) ||
      (rc = ralloc(R, &R->vb, size_t(7) * n)) || (rc = ralloc(R, &R->cpart, size_t(nu))) ||
      (!R->c8 && (rc = ralloc(R, &R->c8, size_t(8) * n))))
    return rc;
  const int64_t ntiles = (nu + 2047) / 2048;
  Sizer z;
  z.take<uint64_t>(size_t(nu));
  z.take<uint64_t>(size_t(nu));
  z.take<uint32_t>(size_t(nu));
  z.take<uint32_t>(size_t(nu));
  z.take<uint32_t>(size_t(ntiles) * 256);
  z.take<uint32_t>(size_t(ntiles) * 256);
  z.take<char>(scan_workspace_bytes(ntiles * 256));
  void* tmp = nullptr;
  hipError_t ae = hipMalloc(&tmp, z.bytes() + 4096);
  if (ae != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "union index scratch: %s", hipGetErrorString(ae));
  }
  Carver cv(tmp, z.bytes() + 4096);
  uint64_t* k0 = cv.take<uint64_t>(size_t(nu));
  uint64_t* k1 = cv.take<uint64_t>(size_t(nu));
  uint32_t* v0 = cv.take<uint32_t>(size_t(nu));
  uint32_t* v1 = cv.take<uint32_t>(size_t(nu));
  hipLaunchKernelGGL(k_upos_keys, dim3(grid_for(nu, 256, 1 << 16)), dim3(256), 0, st,
                     (const int*)R->lucol, nu, k0, v0);
  int res = 0;
  rc = radix_sort_pairs(k0, k1, v0, v1, nu, bits_for_count(n), st, &res, cv);
  if (!rc) {
    hipLaunchKernelGGL(k_vptr, dim3(grid_for(int64_t(n) + 1, 256, 1 << 16)), dim3(256), 0, st,
                       (const uint64_t*)(res ? k1 : k0), nu, n, R->vptr);
    hipLaunchKernelGGL(k_u32_to_i32_64, dim3(grid_for(nu, 256, 1 << 16)), dim3(256), 0, st,
                       (const uint32_t*)(res ? v1 : v0), nu, R->vpos);
    const int64_t ng = (int64_t(R->m) + R->lrows - 1) / R->lrows;
    hipLaunchKernelGGL(k_vbounds, dim3(grid_for(int64_t(7) * n, 256, 1 << 16)), dim3(256), 0, st,
                       n, (const int64_t*)R->vptr, (const int*)R->vpos, (const int64_t*)R->lup,
                       ng, R->vb);
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "union index: %s", hipGetErrorString(le));
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (rc) return rc;
  // the default single-frame iteration: one shared column-sum region, 64
  // lanes per row, the per-XCD reduce with 16 lanes per column; register-held
  // steps of 256 entries for the mean padded row plus 64 (2 for the
  // full-size workload's 360 entries, 3 for the 96^3 geometry's 642); 4
  // waves per group where 5 or more such workgroups fit a CU's LDS, else 8
  // (the fastest shapes measured on both)
  R->fused = 2;
  R->fb_lanes = 64;
  R->fb_lpc = 16;
  R->fb_xcd = 1;
  const double mean_row = double(R->pnnz) / double(std::max(R->m, 1));
  R->fb_steps = std::min(4, std::max(2, int(std::ceil((mean_row + 64.0) / 256.0))));
  R->fb_waves = fused_lds_bytes(R, 4, true) * 5 <= kLdsMax ? 4 : 8;
  fused_lds_attributes();
  // the column-tile reduce (tiles of 1,024 columns, parts of 512 segments,
  // 64 lanes per slot: the fastest shape measured at full size and at 96^3)
  // where its table builds with at least 512 workgroups, else the per-XCD
  // reduce (small plans: a few dozen tile workgroups leave the GPU idle)
  if (build_tile_segments(R, 10, 512, st) == RAPTR_OK) {
    if (R->tnwg >= 512) {
      R->fb_xcd = 2;
      R->fb_lpc = 64;
    }
  } else {
    (void)hipGetLastError();
  }
  return RAPTR_OK;
}

int create_plan(raptr_recon** out, int m, int n, int64_t nnz, const int64_t* d_rp,
                const int* d_col, const float* d_val, const double* d_sens, int value_type,
                int max_frames, int block_rows, int layout, const int* gdims, int group,
                int grouped_only, raptr_stream_t stream);
}  // namespace
}  // namespace raptr

extern "C" {

int raptr_recon_create(raptr_recon** out, int m, int n, int64_t nnz, const int64_t* d_rp,
                       const int* d_col, const float* d_val, const double* d_sens,
                       int value_type, int max_frames, int block_rows,
                       raptr_stream_t stream) {
  return raptr_recon_create_ex(out, m, n, nnz, d_rp, d_col, d_val, d_sens, value_type,
                               max_frames, block_rows, RAPTR_LAYOUT_BLOCKED, stream);
}

int raptr_recon_create_ex(raptr_recon** out, int m, int n, int64_t nnz, const int64_t* d_rp,
                          const int* d_col, const float* d_val, const double* d_sens,
                          int value_type, int max_frames, int block_rows, int layout,
                          raptr_stream_t stream) {
  return create_plan(out, m, n, nnz, d_rp, d_col, d_val, d_sens, value_type, max_frames,
                     block_rows, layout, nullptr, 0, 0, stream);
}
}  // extern "C"

namespace raptr {
namespace {
// gdims: the columns' voxel grid (nx, ny, nz) for brick-shaped backward
// groups, or null; group: members per group of the 64-frame grouped
// layouts (16 or 32; 0 = 16)
int create_plan(raptr_recon** out, int m, int n, int64_t nnz, const int64_t* d_rp,
                const int* d_col, const float* d_val, const double* d_sens, int value_type,
                int max_frames, int block_rows, int layout, const int* gdims, int group,
                int grouped_only, raptr_stream_t stream) {
  if (!out) return fail(RAPTR_ERR_INVALID_ARG, "null output");
  if (group == 0) group = 16;
  const bool go = grouped_only != 0;
  if (go && (max_frames != 64 || layout != RAPTR_LAYOUT_BLOCKED || nnz <= 0 || m <= 0))
    return fail(RAPTR_ERR_INVALID_ARG,
                "a grouped-only plan needs max_frames 64, the blocked layout and a non-empty "
                "matrix");
  if (group != 16 && group != 32)
    return fail(RAPTR_ERR_INVALID_ARG, "group must be 16 or 32");
  if (gdims && int64_t(gdims[0]) * gdims[1] * gdims[2] != n)
    return fail(RAPTR_ERR_INVALID_ARG, "grid %d x %d x %d does not hold %d voxels", gdims[0],
                gdims[1], gdims[2], n);
  *out = nullptr;
  if (m < 0 || n <= 0 || nnz < 0) return fail(RAPTR_ERR_INVALID_ARG, "bad sizes");
  if (value_type < kF32 || value_type > kU16)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown value type %d", value_type);
  if (!valid_frames(max_frames))
    return fail(RAPTR_ERR_INVALID_ARG, "max_frames must be a power of two in [1, 64]");
  if (!d_rp || !d_sens || (nnz > 0 && (!d_col || !d_val)))
    return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (nnz > int64_t(UINT32_MAX) - 16)
    return fail(RAPTR_ERR_INVALID_ARG, "nnz exceeds 2^32");
  if (layout != RAPTR_LAYOUT_BLOCKED)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown layout %d (only RAPTR_LAYOUT_BLOCKED)", layout);
  hipStream_t st = S(stream);
  raptr_recon* R = new raptr_recon();
  R->m = m;
  R->n = n;
  R->nnz = nnz;
  R->vt = value_type;
  R->fmax = max_frames;
  auto bail = [&](int rc) {
    raptr_recon_destroy(R);
    return rc;
  };
  // row block: ratio slice of <= 2 MiB per launch (one XCD's L2 holds it);
  // single-frame row-blocked plans: 65,536 rows (256 KB slices, 16-bit
  // transpose rows; the per-XCD backward keeps an eighth of the blocks in
  // each XCD's L2)
  if (block_rows <= 0)
    block_rows = max_frames == 1 ? 65536 : std::max(1, (2 << 20) / (4 * max_frames));
  R->blk_rows = block_rows;
  R->nblk = std::max(1, int((int64_t(m) + block_rows - 1) / block_rows));
  int rc;
  {
    // ---- forward: padded CSR (grouped-only plans: the row scales only)
    R->grouped_only = go ? 1 : 0;
    int64_t* plen = nullptr;
    int64_t* tot = nullptr;
    if (go) {
      if ((rc = ralloc(R, &R->rscale, size_t(m) + 1))) return bail(rc);
      if (value_type != kF32)
        hipLaunchKernelGGL(k_row_scales, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                           dim3(256), 0, st, d_rp, d_val, m, value_type, R->rscale);
    }
    if (!go && (rc = ralloc(R, &R->frp, size_t(m) + 1))) return bail(rc);
    if (!go) {
      void* tmp = nullptr;
      size_t sb = scan_workspace_bytes(m);
      RAPTR_HIP_TRY(hipMalloc(&tmp, size_t(m) * 8 + 512 + sb));
      plen = static_cast<int64_t*>(tmp);
      tot = reinterpret_cast<int64_t*>(static_cast<char*>(tmp) + ((size_t(m) * 8 + 255) & ~size_t(255)));
      char* sws = reinterpret_cast<char*>(tot) + 256;
      if (m > 0) {
        hipLaunchKernelGGL(k_padded_len, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, d_rp,
                           m, plen);
        rc = scan_exclusive_i64(plen, R->frp, m, tot, sws, sb, st);
        if (!rc) hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, st, R->frp, m, (const int64_t*)tot);
      } else {
        (void)hipMemsetAsync(R->frp, 0, 8, st);
        (void)hipMemsetAsync(tot, 0, 8, st);
      }
      int64_t h = 0;
      (void)hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st);
      (void)hipStreamSynchronize(st);
      (void)hipFree(tmp);
      if (rc) return bail(rc);
      R->pnnz = h;
    }
    if (!go && (rc = ralloc(R, &R->fcol, size_t(R->pnnz) + 4))) return bail(rc);
    if (go) {
    } else if (value_type == kF32) {
      float* p;
      if ((rc = ralloc(R, &p, size_t(R->pnnz) + 4))) return bail(rc);
      R->fval = p;
    } else {
      uint16_t* p;
      if ((rc = ralloc(R, &p, size_t(R->pnnz) + 8))) return bail(rc);
      R->fval = p;
    }
    if (!go && (rc = ralloc(R, &R->rscale, size_t(m) + 1))) return bail(rc);
    if (!go && m > 0) {
      hipLaunchKernelGGL(k_pad_copy, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)), dim3(256),
                         0, st, d_rp, d_col, d_val, m, value_type, (const int64_t*)R->frp,
                         R->fcol, R->fval, R->rscale, max_frames == 1 ? kIlvFwd : 0);
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) return bail(fail(RAPTR_ERR_DEVICE, "pad: %s", hipGetErrorString(le)));
    }
    if (max_frames == 64) {
      R->grp_g = group;
      R->ngf = (int64_t(m) + group - 1) / group;
      switch (group) {
#define RAPTR_RG(G_)                                                                       \
  case G_: {                                                                               \
    RowLists<G_> L{d_rp, d_col, d_val, int64_t(m)};                                        \
    rc = build_groups<G_>(R, L, R->ngf, kAlignFwd, &R->gfp, &R->gfcol, &R->gfval, &R->gfnnz, \
                          st);                                                             \
    break;                                                                                 \
  }
        RAPTR_RG(16) RAPTR_RG(32)
#undef RAPTR_RG
      }
      if (rc) return bail(rc);
      // default shape: matrix cores with per-XCD row ranges
      R->grp_unroll = group == 16 ? 8 : 16;
      R->grp_pipe = 3;
      // kind 3's backward walks windows of about 131,072 rows per XCD and
      // launch (the best measured at full size for 16,384- and 32,768-row
      // blocks)
      R->xcd_bpl = block_rows > 0 && block_rows < 131072 ? 131072 / block_rows : 1;
    }
    // ---- backward: row-blocked transpose via a stable radix sort, every
    // (block, column) segment padded to a multiple of 4 entries
    const int64_t T = int64_t(R->nblk) * (n + 1);
    if (!go && (rc = ralloc(R, &R->bptr, size_t(T) + 1))) return bail(rc);
    if (nnz > 0) {
      int colbits = 1;
      while ((int64_t(1) << colbits) <= n) ++colbits;
      int blkbits = 0;
      while ((int64_t(1) << blkbits) < R->nblk) ++blkbits;
      const size_t sort_ws = raptr_csr_transpose_workspace_size(nnz, m, n);  // radix buffers
      const size_t sb = scan_workspace_bytes(T);
      const size_t tb = ((size_t(T) + 1) * 8 + 255) & ~size_t(255);
      void* tmp = nullptr;
      RAPTR_HIP_TRY(hipMalloc(&tmp, sort_ws + 2 * tb + 256 + sb));
      // carve: rowid | key32 (unused) | sort buffers, in the order the
      // transpose helper lays them out, then the unpadded pointers, padded
      // lengths, scan total and scan scratch
      Carver cv(tmp, sort_ws);
      uint32_t* rowid = cv.take<uint32_t>(size_t(nnz));
      (void)cv.take<int>(size_t(nnz));
      rc = RAPTR_OK;
      uint64_t* k0 = cv.take<uint64_t>(size_t(nnz));
      uint64_t* k1 = cv.take<uint64_t>(size_t(nnz));
      uint32_t* v0 = cv.take<uint32_t>(size_t(nnz));
      uint32_t* v1 = cv.take<uint32_t>(size_t(nnz));
      char* extra = static_cast<char*>(tmp) + sort_ws;
      int64_t* uptr = reinterpret_cast<int64_t*>(extra);
      int64_t* plen = reinterpret_cast<int64_t*>(extra + tb);
      int64_t* tot = reinterpret_cast<int64_t*>(extra + 2 * tb);
      char* sws = extra + 2 * tb + 256;
      if (!cv.ok) rc = fail(RAPTR_ERR_WORKSPACE, "transpose scratch carve failed");
      uint64_t* keys = nullptr;
      uint32_t* idx = nullptr;
      int64_t h = 0;
      if (!rc) {
        hipLaunchKernelGGL(k_bt_keys, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)), dim3(256),
                           0, st, d_rp, d_col, m, block_rows, colbits, k0, v0);
        hipLaunchKernelGGL(k_expand_rows32, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                           dim3(256), 0, st, d_rp, m, rowid);
        int res = 0;
        rc = radix_sort_pairs(k0, k1, v0, v1, nnz, colbits + blkbits, st, &res, cv);
        if (!rc) {
          keys = res ? k1 : k0;
          idx = res ? v1 : v0;
          hipLaunchKernelGGL(k_bt_ptr, dim3(grid_for(T, 256, 1 << 16)), dim3(256), 0, st,
                             (const uint64_t*)keys, nnz, R->nblk, n, colbits, uptr);
          if (!go) {
            hipLaunchKernelGGL(k_bt_padlen, dim3(grid_for(T, 256, 1 << 16)), dim3(256), 0, st,
                               (const int64_t*)uptr, T, n, plen);
            rc = scan_exclusive_i64(plen, R->bptr, T, tot, sws, sb, st);
          }
        }
        if (!rc && !go) {
          hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, st, R->bptr, int(T),
                             (const int64_t*)tot);
          (void)hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st);
          hipError_t se = hipStreamSynchronize(st);
          if (se != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "transpose: %s", hipGetErrorString(se));
        }
      }
      if (!rc && !go) {
        R->bnnz = h;
        if (!(rc = ralloc(R, &R->brow, size_t(h) + 4))) {
          if (value_type == kF32) {
            float* p;
            if (!(rc = ralloc(R, &p, size_t(h) + 4))) R->bval = p;
          } else {
            uint16_t* p;
            if (!(rc = ralloc(R, &p, size_t(h) + 8))) R->bval = p;
          }
        }
      }
      if (!rc && !go) {
        hipLaunchKernelGGL(k_bt_scatter, dim3(grid_for(nnz, 256, 1 << 16)), dim3(256), 0, st,
                           (const uint64_t*)keys, (const uint32_t*)idx, nnz, n, colbits,
                           (const int64_t*)uptr, (const int64_t*)R->bptr,
                           (const uint32_t*)rowid, d_val, (const float*)R->rscale, value_type,
                           R->brow, R->bval, max_frames == 1 ? kIlvBwd : 0);
        hipLaunchKernelGGL(k_bt_pad, dim3(grid_for(T, 256, 1 << 16)), dim3(256), 0, st,
                           (const int64_t*)uptr, (const int64_t*)R->bptr, T, n, block_rows,
                           R->brow, R->bval, value_type, max_frames == 1 ? kIlvBwd : 0);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "transpose: %s", hipGetErrorString(le));
      }
      if (!rc && max_frames == 64) {
        // member columns of every group: bricks of the grid, or runs of
        // consecutive ids
        const int G = group;
        if (gdims) {
          int bx, by, bz;
          brick_dims(G, bx, by, bz);
          R->ngc = int64_t((gdims[0] + bx - 1) / bx) * ((gdims[1] + by - 1) / by) *
                   ((gdims[2] + bz - 1) / bz);
        } else {
          R->ngc = (int64_t(n) + G - 1) / G;
        }
        if (!(rc = ralloc(R, &R->gcols, size_t(R->ngc) * G))) {
          const int64_t tot = R->ngc * G;
          if (gdims)
            hipLaunchKernelGGL(k_brick_cols, dim3(grid_for(tot, 256, 1 << 16)), dim3(256), 0, st,
                               gdims[0], gdims[1], gdims[2], G, R->ngc, R->gcols);
          else
            hipLaunchKernelGGL(k_seq_cols, dim3(grid_for(tot, 256, 1 << 16)), dim3(256), 0, st,
                               n, tot, R->gcols);
          switch (G) {
#define RAPTR_CG(G_)                                                                       \
  case G_: {                                                                               \
    ColLists<G_> L{(const int64_t*)uptr, (const uint32_t*)idx, (const uint32_t*)rowid,     \
                   d_val, (const int*)R->gcols, n, int(R->ngc)};                           \
    rc = build_groups<G_>(R, L, int64_t(R->nblk) * R->ngc, kAlignBwd, &R->gbp, &R->gbrow,  \
                          &R->gbval, &R->gbnnz, st);                                       \
    break;                                                                                 \
  }
            RAPTR_CG(16) RAPTR_CG(32)
#undef RAPTR_CG
          }
        }
        // per-XCD partial images of the kind-3 backward
        if (!rc) rc = ralloc(R, &R->c8, size_t(8) * n * 64);
        if (!rc) R->grp64 = 1;
      }
      (void)hipStreamSynchronize(st);
      (void)hipFree(tmp);
      if (rc) return bail(rc);
    } else {
      (void)hipMemsetAsync(R->bptr, 0, sizeof(int64_t) * (size_t(T) + 1), st);
      float* p;
      if ((rc = ralloc(R, &R->brow, 4))) return bail(rc);
      if ((rc = ralloc(R, &p, 4))) return bail(rc);
      R->bval = p;
    }
  }
  // ---- sensitivity and work vectors
  const int F = max_frames;
  if ((rc = ralloc(R, &R->s, size_t(n)))) return bail(rc);
  if ((rc = ralloc(R, &R->e, (size_t(m) + 1) * F))) return bail(rc);
  if ((rc = ralloc(R, &R->c, size_t(n) * F))) return bail(rc);
  if ((rc = ralloc(R, &R->xx, size_t(n) * F))) return bail(rc);
  if ((rc = ralloc(R, &R->partials, (size_t(n) / 4 + 8) * F))) return bail(rc);
  if ((rc = ralloc(R, &R->scale, size_t(F)))) return bail(rc);
  if ((rc = ralloc(R, &R->norm, size_t(F)))) return bail(rc);
  hipLaunchKernelGGL(k_sens_norm32, dim3(1), dim3(256), 0, st, d_sens, n, R->s);
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) return bail(fail(RAPTR_ERR_DEVICE, "sens: %s", hipGetErrorString(le)));
  hipError_t se = hipStreamSynchronize(st);
  if (se != hipSuccess) return bail(fail(RAPTR_ERR_DEVICE, "create: %s", hipGetErrorString(se)));
  R->fwd_lanes = double(nnz) / std::max(m, 1) >= 192 ? 32 : 16;
  R->fwd_unroll = max_frames == 1 ? 1 : 2;
  R->bwd_lds = 0;
  // single-frame blocked plans: per-XCD partial column sums (backward mode 3)
  if (max_frames == 1 && layout == RAPTR_LAYOUT_BLOCKED && R->bptr &&
      (rc = ralloc(R, &R->c8, size_t(8) * n)))
    return bail(rc);
  // ... and the transpose's rows as 16-bit offsets in blocks of <= 65,536 rows
  if (max_frames == 1 && layout == RAPTR_LAYOUT_BLOCKED && R->bptr && R->brow &&
      block_rows <= 65536 && R->bnnz > 0) {
    if ((rc = ralloc(R, &R->brow16, size_t(R->bnnz) + 8))) return bail(rc);
    hipLaunchKernelGGL(k_bt_rows16, dim3(grid_for(int64_t(R->nblk) * (n + 1), 256, 1 << 16)),
                       dim3(256), 0, st, int64_t(R->nblk), R->bnnz, (const int64_t*)R->bptr, n,
                       block_rows, (const int*)R->brow, R->brow16);
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    if (le != hipSuccess) return bail(fail(RAPTR_ERR_DEVICE, "rows16: %s", hipGetErrorString(le)));
    R->bwd_r16 = 1;
  }
  // ... and the forward's columns as 16-bit low parts (grids of <= 4 x 65,536
  // voxels, rows of <= 65,535 entries)
  if (max_frames == 1 && layout == RAPTR_LAYOUT_BLOCKED && R->fcol && int64_t(n) <= 4 * 65536 &&
      m > 0 && R->pnnz > 0) {
    int* bad = nullptr;
    if ((rc = ralloc(R, &R->fcol16, size_t(R->pnnz) + 8)) ||
        (rc = ralloc(R, &R->fsplit, size_t(m) + 1)) || (rc = ralloc(R, &bad, 4)))
      return bail(rc);
    (void)hipMemsetAsync(bad, 0, sizeof(int), st);
    hipLaunchKernelGGL(k_fwd_cols16, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)), dim3(256),
                       0, st, d_rp, d_col, m, (const int64_t*)R->frp, (const int*)R->fcol,
                       R->fcol16, R->fsplit, bad);
    int hbad = 0;
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (le == hipSuccess) le = hipStreamSynchronize(st);
    if (le != hipSuccess) return bail(fail(RAPTR_ERR_DEVICE, "cols16: %s", hipGetErrorString(le)));
    rfree(R, bad);
    if (hbad) {  // the plan keeps 32-bit columns only
      rfree(R, R->fcol16);
      rfree(R, R->fsplit);
    } else {
      R->fwd_c16 = 1;
    }
  }
  // ... and the column unions of the LDS-staged forward
  if (max_frames == 1 && layout == RAPTR_LAYOUT_BLOCKED && R->fcol && m > 0 && R->pnnz > 0) {
    if ((rc = build_local_columns(R, st))) return bail(rc);
    if (R->lcol) {
      R->fwd_kind = 2;
      R->fwd_lanes = 16;  // the fastest shape of the full-size workload
      R->fwd_unroll = 4;
      if ((rc = build_union_index(R, st))) return bail(rc);
    }
  }
  // single-frame row-blocked plans of >= 8 row blocks (one or more per
  // XCD): the per-XCD backward (32 lanes per column, 2 steps in flight) is
  // the default
  if (R->c8 && max_frames == 1 && layout == RAPTR_LAYOUT_BLOCKED && R->nblk >= 8) {
    R->bwd_lds = 3;
    R->bwd_lpc = 32;
    R->bwd_unroll = 2;
  }
  *out = R;
  return RAPTR_OK;
}
}  // namespace
}  // namespace raptr

extern "C" {

int raptr_recon_create_ordered(raptr_recon** out, int m, int n, int64_t nnz,
                               const int64_t* d_rp, const int* d_col, const float* d_val,
                               const double* d_sens, int value_type, int max_frames,
                               int block_rows, int layout, const int* d_row_order,
                               raptr_stream_t stream) {
  raptr_recon_options o = {layout, d_row_order, nullptr, 0, 0};
  return raptr_recon_create_opts(out, m, n, nnz, d_rp, d_col, d_val, d_sens, value_type,
                                 max_frames, block_rows, &o, stream);
}

int raptr_recon_create_opts(raptr_recon** out, int m, int n, int64_t nnz, const int64_t* d_rp,
                            const int* d_col, const float* d_val, const double* d_sens,
                            int value_type, int max_frames, int block_rows,
                            const raptr_recon_options* opts, raptr_stream_t stream) {
  const raptr_recon_options o = opts ? *opts : raptr_recon_options{RAPTR_LAYOUT_BLOCKED, nullptr,
                                                                   nullptr, 0, 0};
  const int* d_row_order = o.d_row_order;
  int gd[3] = {0, 0, 0};
  if (o.grid) {
    for (int k = 0; k < 3; ++k) gd[k] = o.grid->n[k];
  }
  const int* gdims = o.grid ? gd : nullptr;
  if (!d_row_order)
    return create_plan(out, m, n, nnz, d_rp, d_col, d_val, d_sens, value_type, max_frames,
                       block_rows, o.layout, gdims, o.group, o.grouped_only, stream);
  if (!out) return fail(RAPTR_ERR_INVALID_ARG, "null output");
  *out = nullptr;
  if (m < 0 || n <= 0 || nnz < 0) return fail(RAPTR_ERR_INVALID_ARG, "bad sizes");
  if (!d_rp || (nnz > 0 && (!d_col || !d_val))) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipStream_t st = S(stream);
  // scratch: seen[m] | bad | permuted row pointers | lengths | scan total |
  // scan scratch | permuted columns | permuted values
  Sizer z;
  z.take<int>(size_t(m) + 1);
  z.take<int>(4);
  z.take<int64_t>(size_t(m) + 1);
  z.take<int64_t>(size_t(m) + 1);
  z.take<int64_t>(4);
  const size_t sb = scan_workspace_bytes(m > 0 ? m : 1);
  z.take<char>(sb);
  z.take<int>(size_t(nnz) + 1);
  z.take<float>(size_t(nnz) + 1);
  void* tmp = nullptr;
  hipError_t ae = hipMalloc(&tmp, z.bytes());
  if (ae != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "row order scratch (%zu B): %s", z.bytes(),
                hipGetErrorString(ae));
  }
  Carver cv(tmp, z.bytes());
  int* seen = cv.take<int>(size_t(m) + 1);
  int* bad = cv.take<int>(4);
  int64_t* prp = cv.take<int64_t>(size_t(m) + 1);
  int64_t* plen = cv.take<int64_t>(size_t(m) + 1);
  int64_t* tot = cv.take<int64_t>(4);
  char* sws = cv.take<char>(sb);
  int* pcol = cv.take<int>(size_t(nnz) + 1);
  float* pval = cv.take<float>(size_t(nnz) + 1);
  int rc = RAPTR_OK;
  int hbad = 0;
  (void)hipMemsetAsync(seen, 0, (size_t(m) + 1) * sizeof(int), st);
  (void)hipMemsetAsync(bad, 0, 4 * sizeof(int), st);
  (void)hipMemsetAsync(prp, 0, sizeof(int64_t), st);
  if (m > 0) {
    hipLaunchKernelGGL(k_perm_check, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st,
                       d_row_order, m, seen, bad);
  }
  hipError_t e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    rc = fail(RAPTR_ERR_DEVICE, "row order check: %s", hipGetErrorString(e));
  } else if (hbad) {
    rc = fail(RAPTR_ERR_INVALID_ARG, "row order is not a permutation of 0..%d", m - 1);
  } else if (m > 0) {
    hipLaunchKernelGGL(k_perm_len, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, d_rp,
                       d_row_order, m, plen);
    rc = scan_exclusive_i64(plen, prp, m, tot, sws, sb, st);
    if (!rc) {
      hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, st, prp, m, (const int64_t*)tot);
      hipLaunchKernelGGL(k_perm_copy, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                         dim3(256), 0, st, d_rp, d_col, d_val, d_row_order, m,
                         (const int64_t*)prp, pcol, pval);
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "row order: %s", hipGetErrorString(le));
    }
  }
  if (!rc)
    rc = create_plan(out, m, n, nnz, prp, pcol, pval, d_sens, value_type, max_frames,
                     block_rows, o.layout, gdims, o.group, o.grouped_only, stream);
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (rc) return rc;
  raptr_recon* R = *out;
  if ((rc = ralloc(R, &R->rperm, size_t(m) + 1)) ||
      (rc = ralloc(R, &R->yp, size_t(m) * size_t(R->fmax) + 4))) {
    raptr_recon_destroy(R);
    *out = nullptr;
    return rc;
  }
  (void)hipMemcpyAsync(R->rperm, d_row_order, size_t(m) * sizeof(int), hipMemcpyDeviceToDevice,
                       st);
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    raptr_recon_destroy(R);
    *out = nullptr;
    return fail(RAPTR_ERR_DEVICE, "row order copy: %s", hipGetErrorString(e));
  }
  return RAPTR_OK;
}

int raptr_recon_row_ordered(const raptr_recon* R) { return R && R->rperm ? 1 : 0; }

int raptr_locality_order(const raptr_setup_t* setup, const int* d_channels, int m, int tile,
                         int* d_order, raptr_stream_t stream) {
  return raptr_locality_order_ex(setup, d_channels, m, tile, RAPTR_ORDER_ANGLE_MAJOR, d_order,
                                 stream);
}

int raptr_locality_order_ex(const raptr_setup_t* setup, const int* d_channels, int m, int tile,
                            int kind, int* d_order, raptr_stream_t stream) {
  if (!setup || m < 0 || tile < 1 || tile > 64 || (m > 0 && (!d_channels || !d_order)))
    return fail(RAPTR_ERR_INVALID_ARG, "bad locality order arguments");
  if (kind != RAPTR_ORDER_ANGLE_MAJOR && kind != RAPTR_ORDER_Y_MAJOR &&
      kind != RAPTR_ORDER_SLOPE_TILES && kind != RAPTR_ORDER_ANGLE_SLOPE_TILES)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown locality order kind %d", kind);
  if (m == 0) return RAPTR_OK;
  const raptr_setup_t s = *setup;
  if (s.na < 1 || s.n0z < 1 || s.n0y < 1 || s.n1z < 1 || s.n1y < 1)
    return fail(RAPTR_ERR_INVALID_ARG, "bad setup");
  const int zs = s.n0z + s.n1z - 1, ys = s.n0y + s.n1y - 1;
  // largest key + 1 (kind 1 orders the same factors differently; kinds 2
  // and 3 pad the slope ranges to whole 4 x 4 tiles; kind 3 pads the angles
  // to a multiple of 4)
  const bool stiles = kind == RAPTR_ORDER_SLOPE_TILES || kind == RAPTR_ORDER_ANGLE_SLOPE_TILES;
  const double slopes =
      stiles ? double((zs + 3) / 4 * 4) * double((ys + 3) / 4 * 4) : double(zs) * ys;
  const double angles = kind == RAPTR_ORDER_ANGLE_SLOPE_TILES ? double((s.na + 3) / 4 * 4)
                                                              : double(s.na);
  const double kmax = angles * ((zs + tile - 1) / tile) * ((ys + tile - 1) / tile) * tile *
                      tile * slopes;
  if (kmax >= 1.8e19) return fail(RAPTR_ERR_INVALID_ARG, "setup too large for 64-bit keys");
  int bits = 1;
  while (bits < 64 && double(uint64_t(1) << bits) < kmax) ++bits;
  hipStream_t st = S(stream);
  const int64_t ntiles = (int64_t(m) + 2047) / 2048;
  Sizer z;
  z.take<uint64_t>(size_t(m));
  z.take<uint64_t>(size_t(m));
  z.take<uint32_t>(size_t(m));
  z.take<uint32_t>(size_t(m));
  z.take<uint32_t>(size_t(ntiles) * 256);
  z.take<uint32_t>(size_t(ntiles) * 256);
  z.take<char>(scan_workspace_bytes(ntiles * 256));
  void* tmp = nullptr;
  hipError_t ae = hipMalloc(&tmp, z.bytes() + 4096);
  if (ae != hipSuccess) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "locality order scratch: %s", hipGetErrorString(ae));
  }
  Carver cv(tmp, z.bytes() + 4096);
  uint64_t* k0 = cv.take<uint64_t>(size_t(m));
  uint64_t* k1 = cv.take<uint64_t>(size_t(m));
  uint32_t* v0 = cv.take<uint32_t>(size_t(m));
  uint32_t* v1 = cv.take<uint32_t>(size_t(m));
  hipLaunchKernelGGL(k_locality_keys, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, s,
                     d_channels, m, tile, kind, k0, v0);
  int res = 0;
  int rc = radix_sort_pairs(k0, k1, v0, v1, m, bits, st, &res, cv);
  if (!rc) {
    hipLaunchKernelGGL(k_u32_to_i32, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st,
                       (const uint32_t*)(res ? v1 : v0), m, d_order);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) rc = fail(RAPTR_ERR_DEVICE, "locality order: %s", hipGetErrorString(le));
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  return rc;
}

int raptr_recon_destroy(raptr_recon* R) {
  if (!R) return RAPTR_OK;
  if (R->nclones > 0)
    return fail(RAPTR_ERR_INVALID_ARG, "plan has %d live clone(s): destroy them first",
                R->nclones);
  clear_graphs(R);
  for (void* p : R->allocs) (void)hipFree(p);
  if (R->parent) --R->parent->nclones;
  delete R;
  return RAPTR_OK;
}

int raptr_recon_clone(raptr_recon* P, raptr_recon** out, raptr_stream_t stream) {
  if (!P || !out) return fail(RAPTR_ERR_INVALID_ARG, "null plan or output");
  *out = nullptr;
  if (P->parent) P = P->parent;  // clones of a clone share the same layouts
  raptr_recon* R = new raptr_recon(*P);
  // the copy's graphs, allocations and clone count are the parent's: drop
  // them (not destroyed: they stay the parent's)
  R->graphs.clear();
  R->graph_tick = 0;
  R->allocs.clear();
  R->asizes.clear();
  R->bytes = 0;
  R->nclones = 0;
  R->parent = P;
  ++P->nclones;
  // the work vectors, each the size of the parent's
  auto own = [&](auto*& p) -> int {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(p)>>;
    if (!p) return RAPTR_OK;
    size_t bytes = 0;
    for (size_t i = 0; i < P->allocs.size(); ++i)
      if (P->allocs[i] == static_cast<const void*>(p)) bytes = P->asizes[i];
    p = nullptr;
    return ralloc(R, &p, bytes / sizeof(T) + 1);
  };
  int rc = RAPTR_OK;
  if (!rc) rc = own(R->e);
  if (!rc) rc = own(R->c);
  if (!rc) rc = own(R->xx);
  if (!rc) rc = own(R->partials);
  if (!rc) rc = own(R->scale);
  if (!rc) rc = own(R->norm);
  if (!rc) rc = own(R->c8);
  if (!rc) rc = own(R->yp);
  if (!rc) rc = own(R->cpart);
  if (!rc) rc = own(R->ck);
  if (rc) {
    raptr_recon_destroy(R);
    return rc;
  }
  (void)stream;
  *out = R;
  return RAPTR_OK;
}

int raptr_recon_clone_count(const raptr_recon* R) { return R ? R->nclones : -1; }

int raptr_recon_info(const raptr_recon* R, int64_t* device_bytes, int64_t* padded_nnz,
                     int* row_blocks, int* fwd_lanes, int* fwd_unroll, int* bwd_unroll) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (device_bytes) *device_bytes = int64_t(R->bytes);
  if (padded_nnz) *padded_nnz = R->pnnz;
  if (row_blocks) *row_blocks = R->nblk;
  if (fwd_lanes) *fwd_lanes = R->fwd_lanes;
  if (fwd_unroll) *fwd_unroll = R->fwd_unroll;
  if (bwd_unroll) *bwd_unroll = R->bwd_unroll;
  return RAPTR_OK;
}

int raptr_recon_set_config(raptr_recon* R, int fwd_lanes, int fwd_unroll, int bwd_unroll) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (fwd_lanes != 16 && fwd_lanes != 32 && fwd_lanes != 64)
    return fail(RAPTR_ERR_INVALID_ARG, "forward lanes must be 16, 32 or 64");
  // unroll counts groups of 4 entries in flight per lane (entry slot)
  if (fwd_unroll != 1 && fwd_unroll != 2 && fwd_unroll != 4)
    return fail(RAPTR_ERR_INVALID_ARG, "forward unroll must be 1, 2 or 4");
  if (bwd_unroll != 1 && bwd_unroll != 2 && bwd_unroll != 4)
    return fail(RAPTR_ERR_INVALID_ARG, "backward unroll must be 1, 2 or 4");
  if (R->fwd_lanes != fwd_lanes || R->fwd_unroll != fwd_unroll ||
      R->bwd_unroll != bwd_unroll) {
    clear_graphs(R);
  }
  R->fwd_lanes = fwd_lanes;
  R->fwd_unroll = fwd_unroll;
  R->bwd_unroll = bwd_unroll;
  return RAPTR_OK;
}

int raptr_recon_set_backward(raptr_recon* R, int staged, int lanes_per_column, int workgroups) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (staged != 0 && staged != 3)
    return fail(RAPTR_ERR_INVALID_ARG, "staged must be 0 (off) or 3 (per-XCD row ranges)");
  if (staged == 3 && (R->fmax != 1 || !R->c8))
    return fail(RAPTR_ERR_INVALID_ARG,
                "the per-XCD backward needs a single-frame plan with partial sums");
  if (lanes_per_column != 8 && lanes_per_column != 16 && lanes_per_column != 32 &&
      lanes_per_column != 64)
    return fail(RAPTR_ERR_INVALID_ARG, "lanes per column must be 8, 16, 32 or 64");
  if (workgroups < 1) return fail(RAPTR_ERR_INVALID_ARG, "workgroups must be positive");
  if (R->bwd_lds != staged || R->bwd_lpc != lanes_per_column || R->bwd_wgs != workgroups) {
    clear_graphs(R);
  }
  R->bwd_lds = staged;
  R->bwd_lpc = lanes_per_column;
  R->bwd_wgs = workgroups;
  return RAPTR_OK;
}

int raptr_recon_set_compact_rows(raptr_recon* R, int on) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (on && !R->brow16 && !R->fcol16)
    return fail(RAPTR_ERR_INVALID_ARG,
                "no 16-bit indices (single-frame blocked plans: transpose rows with block_rows "
                "<= 65536, forward columns with <= 262144 voxels)");
  if (R->bwd_r16 != (on && R->brow16 ? 1 : 0) || R->fwd_c16 != (on && R->fcol16 ? 1 : 0)) {
    clear_graphs(R);
  }
  R->bwd_r16 = on && R->brow16 ? 1 : 0;
  R->fwd_c16 = on && R->fcol16 ? 1 : 0;
  return RAPTR_OK;
}

int raptr_recon_set_forward_kind(raptr_recon* R, int kind) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (kind != 0 && kind != 2)
    return fail(RAPTR_ERR_INVALID_ARG, "forward kind must be 0 (rows) or 2 (LDS-staged x)");
  if (kind == 2 && !R->lcol)
    return fail(RAPTR_ERR_INVALID_ARG, "no LDS column unions on this plan");
  if (R->fwd_kind != kind) {
    clear_graphs(R);
  }
  R->fwd_kind = kind;
  return RAPTR_OK;
}

int raptr_recon_get_forward_kind(const raptr_recon* R, int* kind, int* available,
                                 int64_t* union_entries) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (kind) *kind = R->fwd_kind;
  if (available) *available = 1 | (R->lcol ? 4 : 0);
  if (union_entries) *union_entries = R->lcol ? R->lnu : 0;
  return RAPTR_OK;
}

// bit 0: transpose rows, bit 1: forward columns (16-bit and on); -1: none
int raptr_recon_get_compact_rows(const raptr_recon* R) {
  if (!R || (!R->brow16 && !R->fcol16)) return -1;
  return (R->brow16 && R->bwd_r16 ? 1 : 0) | (R->fcol16 && R->fwd_c16 ? 2 : 0);
}

int raptr_recon_layout(const raptr_recon* R) { return R ? R->layout : -1; }

int raptr_recon_set_blocks_per_launch(raptr_recon* R, int blocks) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (blocks < 1 || blocks > 64)
    return fail(RAPTR_ERR_INVALID_ARG, "blocks per launch must be in [1, 64]");
  if (R->bwd_bpl != blocks) {
    clear_graphs(R);
  }
  R->bwd_bpl = blocks;
  return RAPTR_OK;
}

int raptr_recon_get_blocks_per_launch(const raptr_recon* R) { return R ? R->bwd_bpl : -1; }

int raptr_recon_set_xcd_blocks(raptr_recon* R, int blocks) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (blocks < 0) return fail(RAPTR_ERR_INVALID_ARG, "blocks per launch must be >= 0");
  if (R->xcd_bpl != blocks) clear_graphs(R);
  R->xcd_bpl = blocks;
  return RAPTR_OK;
}

int raptr_recon_get_xcd_blocks(const raptr_recon* R) { return R ? R->xcd_bpl : -1; }

int64_t raptr_recon_transpose_nnz(const raptr_recon* R) { return R ? R->bnnz : -1; }

int raptr_recon_set_broadcast(raptr_recon* R, int on) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (R->bcast64 != (on ? 1 : 0)) {
    clear_graphs(R);
  }
  R->bcast64 = on ? 1 : 0;
  return RAPTR_OK;
}

int raptr_recon_get_broadcast(const raptr_recon* R) { return R ? R->bcast64 : -1; }

int raptr_recon_graph_count(const raptr_recon* R) { return R ? int(R->graphs.size()) : -1; }

int raptr_recon_set_grouped(raptr_recon* R, int on, int unroll, int kind) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (!on && R->grouped_only)
    return fail(RAPTR_ERR_INVALID_ARG, "a grouped-only plan has no other 64-frame layout");
  if (on && !R->gbp) return fail(RAPTR_ERR_INVALID_ARG, "plan has no grouped layouts");
  if (R->gbp && !grp_shape_ok(R->grp_g, unroll, kind))
    return fail(RAPTR_ERR_INVALID_ARG,
                "groups of %d take unroll %s with kind %s; got unroll %d kind %d", R->grp_g,
                R->grp_g == 16 ? "8, 16 or 32" : "8 or 16", R->grp_g == 16 ? "2 or 3" : "3",
                unroll, kind);
  if (R->grp64 != (on ? 1 : 0) || R->grp_unroll != unroll || R->grp_pipe != kind) {
    clear_graphs(R);
  }
  R->grp64 = on ? 1 : 0;
  R->grp_unroll = unroll;
  R->grp_pipe = kind;
  return RAPTR_OK;
}

int raptr_recon_get_grouped(const raptr_recon* R, int* group, int* on, int* unroll,
                            int* kind, int64_t* fwd_entries, int64_t* bwd_entries) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (group) *group = R->gbp ? R->grp_g : 0;
  if (on) *on = R->grp64;
  if (unroll) *unroll = R->grp_unroll;
  if (kind) *kind = R->grp_pipe;
  if (fwd_entries) *fwd_entries = R->gbp ? R->gfnnz : 0;
  if (bwd_entries) *bwd_entries = R->gbp ? R->gbnnz : 0;
  return RAPTR_OK;
}


int raptr_recon_get_backward(const raptr_recon* R, int* staged, int* lanes_per_column,
                             int* workgroups) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (staged) *staged = R->bwd_lds;
  if (lanes_per_column) *lanes_per_column = R->bwd_lpc;
  if (workgroups) *workgroups = R->bwd_wgs;
  return RAPTR_OK;
}

int raptr_recon_set_fused(raptr_recon* R, int on, int waves, int lanes_per_row,
                          int lanes_per_column) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (on < 0 || on > 2) return fail(RAPTR_ERR_INVALID_ARG, "on must be 0, 1 or 2");
  if (on && !R->cpart)
    return fail(RAPTR_ERR_INVALID_ARG,
                "no fused pass (single-frame blocked plans with column unions that fit LDS)");
  if ((waves != 2 && waves != 4 && waves != 8) || (waves == 8 && on == 1))
    return fail(RAPTR_ERR_INVALID_ARG, "waves must be 2 or 4 (or 8 with a shared region)");
  if (lanes_per_row != 16 && lanes_per_row != 32 && lanes_per_row != 64)
    return fail(RAPTR_ERR_INVALID_ARG, "lanes per row must be 16, 32 or 64");
  if (lanes_per_column != 16 && lanes_per_column != 32 && lanes_per_column != 64)
    return fail(RAPTR_ERR_INVALID_ARG, "lanes per column must be 16, 32 or 64");
  if (R->cpart && on && fused_lds_bytes(R, waves, on == 2) > kLdsMax)
    return fail(RAPTR_ERR_INVALID_ARG, "%d waves need %zu B of LDS (> 160 KiB)", waves,
                fused_lds_bytes(R, waves, on == 2));
  const int nf = on;
  if (R->fused != nf || R->fb_waves != waves || R->fb_lanes != lanes_per_row ||
      R->fb_lpc != lanes_per_column)
    clear_graphs(R);
  R->fused = nf;
  R->fb_waves = waves;
  R->fb_lanes = lanes_per_row;
  R->fb_lpc = lanes_per_column;
  return RAPTR_OK;
}

int raptr_recon_set_fused_steps(raptr_recon* R, int steps) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (steps < 2 || steps > 4) return fail(RAPTR_ERR_INVALID_ARG, "steps must be 2, 3 or 4");
  if (R->fb_steps != steps) clear_graphs(R);
  R->fb_steps = steps;
  return RAPTR_OK;
}

int raptr_recon_get_fused_steps(const raptr_recon* R) { return R ? R->fb_steps : -1; }

int raptr_recon_set_fused_reduce(raptr_recon* R, int mode) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (mode < 0 || mode > 2) return fail(RAPTR_ERR_INVALID_ARG, "reduce mode must be 0, 1 or 2");
  if (mode == 2 && !R->tseg) return fail(RAPTR_ERR_INVALID_ARG, "plan has no column-tile segments");
  if (R->fb_xcd != mode) clear_graphs(R);
  R->fb_xcd = mode;
  return RAPTR_OK;
}

int raptr_recon_get_fused_reduce(const raptr_recon* R) { return R ? R->fb_xcd : -1; }

int raptr_recon_set_fused_tiles(raptr_recon* R, int columns, int segments_per_workgroup,
                                raptr_stream_t stream) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (!R->cpart) return fail(RAPTR_ERR_INVALID_ARG, "plan has no fused pass");
  if (R->parent || R->nclones)
    return fail(RAPTR_ERR_INVALID_ARG, "the segment table is shared with clones: set it before cloning");
  int sh = 7;
  while (sh < 11 && (1 << sh) != columns) ++sh;
  if ((1 << sh) != columns)
    return fail(RAPTR_ERR_INVALID_ARG, "tile columns must be a power of two in 128..2048");
  if (segments_per_workgroup < 64 || segments_per_workgroup > (1 << 20))
    return fail(RAPTR_ERR_INVALID_ARG, "segments per workgroup must be in 64..1048576");
  if (R->tseg && R->tshift == sh && R->tseg_target == segments_per_workgroup) return RAPTR_OK;
  clear_graphs(R);
  (void)hipDeviceSynchronize();  // no queued pass may still read the old table
  const int rc = build_tile_segments(R, sh, segments_per_workgroup,
                                     static_cast<hipStream_t>(stream));
  if (rc && R->fb_xcd == 2) R->fb_xcd = 1;
  return rc;
}

int raptr_recon_set_union_rows(raptr_recon* R, int rows, raptr_stream_t stream) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (R->parent || R->nclones)
    return fail(RAPTR_ERR_INVALID_ARG, "the unions are shared with clones: set them before cloning");
  if (rows != 0 && (rows < 16 || rows > 512 || (rows & (rows - 1))))
    return fail(RAPTR_ERR_INVALID_ARG, "union rows must be 0 or a power of two in 16..512");
  if (R->fmax != 1 || R->layout != RAPTR_LAYOUT_BLOCKED || !R->fcol || R->m == 0 || R->pnnz == 0)
    return fail(RAPTR_ERR_INVALID_ARG,
                "column unions need a single-frame blocked plan with a padded CSR");
  hipStream_t st = static_cast<hipStream_t>(stream);
  clear_graphs(R);
  (void)hipDeviceSynchronize();  // no queued pass may still read the old unions
  rfree(R, R->lup);
  rfree(R, R->lucol);
  rfree(R, R->lcol);
  rfree(R, R->cpart);
  rfree(R, R->vptr);
  rfree(R, R->vpos);
  rfree(R, R->vb);
  rfree(R, R->tseg);
  rfree(R, R->tcol);
  rfree(R, R->twg);
  rfree(R, R->tk);
  rfree(R, R->ck);
  R->lmax = 0;
  R->lnu = 0;
  R->nseg = 0;
  R->tshift = R->tseg_target = R->tnwg = R->tkmax = 0;
  R->fused = 0;
  int rc = build_local_columns(R, st, rows);
  if (!rc && R->lcol) rc = build_union_index(R, st);
  if (!R->lcol && R->fwd_kind == 2) R->fwd_kind = 0;
  return rc;
}

int raptr_recon_get_unions(const raptr_recon* R, int* rows, int* largest, int64_t* entries) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (rows) *rows = R->lcol ? R->lrows : 0;
  if (largest) *largest = R->lcol ? R->lmax : 0;
  if (entries) *entries = R->lcol ? R->lnu : 0;
  return RAPTR_OK;
}

int raptr_recon_get_fused_tiles(const raptr_recon* R, int* columns, int* segments_per_workgroup,
                                int64_t* segments, int* workgroups, int* max_parts) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (columns) *columns = R->tseg ? (1 << R->tshift) : 0;
  if (segments_per_workgroup) *segments_per_workgroup = R->tseg ? R->tseg_target : 0;
  if (segments) *segments = R->tseg ? R->nseg : 0;
  if (workgroups) *workgroups = R->tseg ? R->tnwg : 0;
  if (max_parts) *max_parts = R->tseg ? R->tkmax : 0;
  return RAPTR_OK;
}

int raptr_recon_get_fused(const raptr_recon* R, int* available, int* on, int* waves,
                          int* lanes_per_row, int* lanes_per_column) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (available) *available = R->cpart ? 1 : 0;
  if (on) *on = R->fused;
  if (waves) *waves = R->fb_waves;
  if (lanes_per_row) *lanes_per_row = R->fb_lanes;
  if (lanes_per_column) *lanes_per_column = R->fb_lpc;
  return RAPTR_OK;
}

static int mlem_impl(raptr_recon* R, int nframes, const float* d_y, float* d_x, int niter,
                     int use_graph, hipStream_t st);

int raptr_recon_mlem(raptr_recon* R, int nframes, const float* d_y, float* d_x, int niter,
                     int use_graph, raptr_stream_t stream) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  R->fa = 64;
  return mlem_impl(R, nframes, d_y, d_x, niter, use_graph, S(stream));
}

int raptr_recon_mlem_active(raptr_recon* R, int nframes, int active, const float* d_y,
                            float* d_x, int niter, int use_graph, raptr_stream_t stream) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (active == nframes) return raptr_recon_mlem(R, nframes, d_y, d_x, niter, use_graph, stream);
  if (nframes != 64 || (active != 16 && active != 32))
    return fail(RAPTR_ERR_INVALID_ARG,
                "active frames must be 16, 32 or 64 of a 64-frame pass (got %d of %d)", active,
                nframes);
  if (!(R->fmax == 64 && R->grp64 && R->grp_pipe == 3 && R->grp_g == 16))
    return fail(RAPTR_ERR_INVALID_ARG,
                "passes over fewer than 64 frames need the grouped layouts of 16 on, kind 3");
  R->fa = active;
  const int rc = mlem_impl(R, nframes, d_y, d_x, niter, use_graph, S(stream));
  R->fa = 64;
  return rc;
}

static int mlem_impl(raptr_recon* R, int nframes, const float* d_y, float* d_x, int niter,
                     int use_graph, hipStream_t st) {
  if (!valid_frames(nframes) || nframes > R->fmax)
    return fail(RAPTR_ERR_INVALID_ARG, "nframes %d not a power of two <= %d", nframes, R->fmax);
  if (niter < 0 || !d_y || !d_x) return fail(RAPTR_ERR_INVALID_ARG, "bad arguments");
  if (!frames_aligned(d_y, nframes) || !frames_aligned(d_x, nframes))
    return fail(RAPTR_ERR_INVALID_ARG, "y and x must be aligned to %d bytes for %d frames",
                4 * (nframes >= 4 ? 4 : nframes), nframes);
  if (R->rperm && R->m > 0) {
    // counts into plan row order once per call (outside any captured graph)
    hipLaunchKernelGGL(k_rows_gather, dim3(grid_for(int64_t(R->m) * nframes, 256, 1 << 15)),
                       dim3(256), 0, st, d_y, R->yp, (const int*)R->rperm, int64_t(R->m),
                       nframes);
    RAPTR_LAUNCH_CHECK();
    d_y = R->yp;
  }
  if (!use_graph || niter < 2 || !st) {
    for (int it = 0; it < niter; ++it) RAPTR_TRY(iteration(R, nframes, d_y, d_x, st));
    return RAPTR_OK;
  }
  const auto key = std::make_tuple(nframes, R->fa, uintptr_t(d_y), uintptr_t(d_x));
  auto it = R->graphs.find(key);
  hipGraphExec_t exec = nullptr;
  if (it == R->graphs.end()) {
    if (R->graphs.size() >= kMaxGraphs) {
      auto lru = R->graphs.begin();
      for (auto jt = R->graphs.begin(); jt != R->graphs.end(); ++jt)
        if (jt->second.used < lru->second.used) lru = jt;
      // a replay of the evicted graph may still be queued on some stream
      RAPTR_HIP_TRY(hipDeviceSynchronize());
      destroy_graph(lru->second);
      R->graphs.erase(lru);
    }
    hipGraph_t graph = nullptr;
    RAPTR_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int irc = iteration(R, nframes, d_y, d_x, st);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (irc) {
      if (graph) (void)hipGraphDestroy(graph);
      return irc;
    }
    RAPTR_HIP_TRY(ce);
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(RAPTR_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
    }
    R->graphs[key] = {graph, exec, ++R->graph_tick};
  } else {
    exec = it->second.exec;
    it->second.used = ++R->graph_tick;
  }
  for (int i = 0; i < niter; ++i) RAPTR_HIP_TRY(hipGraphLaunch(exec, st));
  return RAPTR_OK;
}

// One forward (+ratio) or one full backward (all row blocks, with update)
// of the plan, for timing and tests.
int raptr_recon_forward(raptr_recon* R, int nframes, const float* d_x, const float* d_y,
                        float* d_e, raptr_stream_t stream) {
  if (!R || !valid_frames(nframes) || nframes > R->fmax)
    return fail(RAPTR_ERR_INVALID_ARG, "bad plan or frame count");
  if (!frames_aligned(d_x, nframes) || !frames_aligned(d_y, nframes) ||
      !frames_aligned(d_e, nframes))
    return fail(RAPTR_ERR_INVALID_ARG, "x, y and e must be aligned to the frame vector");
  hipStream_t st = S(stream);
  if (R->rperm && R->m > 0) {
    // caller rows -> plan rows, the pass into the plan's ratio buffer, back
    const unsigned g = grid_for(int64_t(R->m) * nframes, 256, 1 << 15);
    hipLaunchKernelGGL(k_rows_gather, dim3(g), dim3(256), 0, st, d_y, R->yp,
                       (const int*)R->rperm, int64_t(R->m), nframes);
    RAPTR_LAUNCH_CHECK();
    RAPTR_TRY(launch_fwd_any(R, nframes, d_x, R->yp, R->e, st));
    hipLaunchKernelGGL(k_rows_scatter, dim3(g), dim3(256), 0, st, (const float*)R->e, d_e,
                       (const int*)R->rperm, int64_t(R->m), nframes);
    RAPTR_LAUNCH_CHECK();
    return RAPTR_OK;
  }
  return launch_fwd_any(R, nframes, d_x, d_y, d_e, st);
}

int raptr_recon_fused_pass(raptr_recon* R, const float* d_x, const float* d_y, int part,
                           raptr_stream_t stream) {
  if (!R) return fail(RAPTR_ERR_INVALID_ARG, "null plan");
  if (!R->cpart || R->fused == 0)
    return fail(RAPTR_ERR_INVALID_ARG, "the fused iteration is off or absent on this plan");
  if (part < 0 || part > 2) return fail(RAPTR_ERR_INVALID_ARG, "part must be 0, 1 or 2");
  if (!d_x || (part != 2 && !d_y)) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipStream_t st = S(stream);
  if (R->rperm && R->m > 0 && part == 0) {
    const unsigned g = grid_for(int64_t(R->m), 256, 1 << 15);
    hipLaunchKernelGGL(k_rows_gather, dim3(g), dim3(256), 0, st, d_y, R->yp,
                       (const int*)R->rperm, int64_t(R->m), 1);
    RAPTR_LAUNCH_CHECK();
    d_y = R->yp;
  }
  return launch_fused1(R, d_x, d_y, st, nullptr, part);
}

int raptr_recon_backward(raptr_recon* R, int nframes, const float* d_x, raptr_stream_t stream) {
  if (!R || !valid_frames(nframes) || nframes > R->fmax)
    return fail(RAPTR_ERR_INVALID_ARG, "bad plan or frame count");
  if (!frames_aligned(d_x, nframes))
    return fail(RAPTR_ERR_INVALID_ARG, "x must be aligned to the frame vector");
  if (nframes == 1 && R->bwd_lds == 3) return launch_bwd1_xcd(R, d_x, S(stream), nullptr);
  if (nframes == 64 && R->grp64 && R->grp_pipe == 3) return launch_bwd_xcd(R, d_x, S(stream), nullptr);
  for (int b = 0; b < R->nblk; b += R->bwd_bpl)
    RAPTR_TRY(launch_bwd_any(R, nframes, b, std::min(R->bwd_bpl, R->nblk - b), d_x, S(stream),
                             nullptr));
  return RAPTR_OK;
}

// ------------------------------------------------ frame-slot plumbing
int raptr_frame_fill(float* d_a, int64_t rows, int nframes, int slot, float value,
                     raptr_stream_t stream) {
  if (!d_a || rows < 0 || nframes < 1 || slot < 0 || slot >= nframes)
    return fail(RAPTR_ERR_INVALID_ARG, "bad frame fill arguments");
  if (rows == 0) return RAPTR_OK;
  hipLaunchKernelGGL(k_frame_fill, dim3(grid_for(rows, 256, 8192)), dim3(256), 0, S(stream),
                     d_a, rows, nframes, slot, value);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_frame_scatter(float* d_a, int64_t rows, int nframes, int slot, const int* d_idx,
                        const float* d_val, int64_t n, int* d_bad, raptr_stream_t stream) {
  if (!d_a || rows < 0 || nframes < 1 || slot < 0 || slot >= nframes || n < 0 ||
      (n > 0 && (!d_idx || !d_val || !d_bad)))
    return fail(RAPTR_ERR_INVALID_ARG, "bad frame scatter arguments");
  if (n == 0) return RAPTR_OK;
  hipLaunchKernelGGL(k_frame_scatter, dim3(grid_for(n, 256, 8192)), dim3(256), 0, S(stream),
                     d_a, rows, nframes, slot, d_idx, d_val, n, d_bad);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_frame_gather(const float* d_a, int64_t rows, int nframes, int slot, float* d_out,
                       raptr_stream_t stream) {
  if (!d_a || !d_out || rows < 0 || nframes < 1 || slot < 0 || slot >= nframes)
    return fail(RAPTR_ERR_INVALID_ARG, "bad frame gather arguments");
  if (rows == 0) return RAPTR_OK;
  hipLaunchKernelGGL(k_frame_gather, dim3(grid_for(rows, 256, 8192)), dim3(256), 0, S(stream),
                     d_a, rows, nframes, slot, d_out);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_frame_copy(float* d_a, int64_t rows, int nframes, int src_slot, int dst_slot,
                     raptr_stream_t stream) {
  if (!d_a || rows < 0 || nframes < 1 || src_slot < 0 || src_slot >= nframes || dst_slot < 0 ||
      dst_slot >= nframes)
    return fail(RAPTR_ERR_INVALID_ARG, "bad frame copy arguments");
  if (rows == 0 || src_slot == dst_slot) return RAPTR_OK;
  hipLaunchKernelGGL(k_frame_copy, dim3(grid_for(rows, 256, 8192)), dim3(256), 0, S(stream),
                     d_a, rows, nframes, src_slot, dst_slot);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

}  // extern "C"
