// recon.hpp -- the reconstruction plan (recon.hip: plan, row-blocked
// layouts, single-frame and batched kernels).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "primitives.hpp"

namespace raptr {

// --------------------------------------------------------- value types
enum { kF32 = 0, kF16 = 1, kBF16 = 2, kU16 = 3 };

__device__ __forceinline__ float h2f(uint16_t h, int vt) {
  if (vt == kF16) return __half2float(__ushort_as_half(h));
  if (vt == kU16) return float(h);
  return __uint_as_float(uint32_t(h) << 16);
}

// f in [0, 1] (value / row maximum)
__device__ __forceinline__ uint16_t f2h(float f, int vt) {
  if (vt == kF16) return __half_as_ushort(__float2half_rn(f));
  if (vt == kU16) return uint16_t(fminf(rintf(f * 65535.0f), 65535.0f));
  // BF16, round to nearest even
  uint32_t u = __float_as_uint(f);
  uint32_t r = 0x7fffu + ((u >> 16) & 1u);
  return uint16_t((u + r) >> 16);
}

template <int VT>
struct Val;
template <>
struct Val<kF32> {
  using T = float;
  __device__ static float get(const float* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
};
template <>
struct Val<kF16> {
  using T = uint16_t;
  __device__ static float get(const uint16_t* p, int64_t k) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(p + k)));
  }
};
template <>
struct Val<kBF16> {
  using T = uint16_t;
  __device__ static float get(const uint16_t* p, int64_t k) {
    return __uint_as_float(uint32_t(__builtin_nontemporal_load(p + k)) << 16);
  }
};
template <>
struct Val<kU16> {
  using T = uint16_t;
  __device__ static float get(const uint16_t* p, int64_t k) {
    return float(__builtin_nontemporal_load(p + k));
  }
};

// four consecutive matrix values (16-byte / 8-byte aligned) as floats
template <int VT>
__device__ __forceinline__ void load4(const typename Val<VT>::T* __restrict__ val, int64_t k,
                                      float& v0, float& v1, float& v2, float& v3) {
  if (VT == kF32) {
    vf4 v = __builtin_nontemporal_load(
        reinterpret_cast<const vf4*>(reinterpret_cast<const float*>(val) + k));
    v0 = v.x, v1 = v.y, v2 = v.z, v3 = v.w;
  } else {
    vu2 h = __builtin_nontemporal_load(
        reinterpret_cast<const vu2*>(reinterpret_cast<const uint16_t*>(val) + k));
    v0 = h2f(uint16_t(h.x & 0xffff), VT), v1 = h2f(uint16_t(h.x >> 16), VT);
    v2 = h2f(uint16_t(h.y & 0xffff), VT), v3 = h2f(uint16_t(h.y >> 16), VT);
  }
}
// as load4 with ordinary (cached) loads
template <int VT>
__device__ __forceinline__ void load4c(const typename Val<VT>::T* __restrict__ val, int64_t k,
                                       float& v0, float& v1, float& v2, float& v3) {
  if (VT == kF32) {
    const vf4 v = *reinterpret_cast<const vf4*>(reinterpret_cast<const float*>(val) + k);
    v0 = v.x, v1 = v.y, v2 = v.z, v3 = v.w;
  } else {
    const vu2 h = *reinterpret_cast<const vu2*>(reinterpret_cast<const uint16_t*>(val) + k);
    v0 = h2f(uint16_t(h.x & 0xffff), VT), v1 = h2f(uint16_t(h.x >> 16), VT);
    v2 = h2f(uint16_t(h.y & 0xffff), VT), v3 = h2f(uint16_t(h.y >> 16), VT);
  }
}
}  // namespace raptr

struct raptr_recon {
  int m = 0, n = 0, vt = raptr::kF32, fmax = 1;
  // frames a 64-frame grouped pass computes (16, 32 or 64: the first fa of
  // each 64-frame row); 64 except inside raptr_recon_mlem_active
  int fa = 64;
  int64_t nnz = 0, pnnz = 0;
  int blk_rows = 0, nblk = 0;
  // forward (padded CSR)
  int64_t* frp = nullptr;
  int* fcol = nullptr;
  void* fval = nullptr;
  float* rscale = nullptr;
  // backward (row-blocked transpose)
  int64_t* bptr = nullptr;  // nblk * (n+1)
  int* brow = nullptr;
  void* bval = nullptr;
  int64_t bnnz = 0;  // padded entries of the transpose
  // sensitivity and work vectors (sized for fmax frames)
  float* s = nullptr;
  float* e = nullptr;
  float* c = nullptr;
  float* xx = nullptr;
  double* partials = nullptr;
  float* scale = nullptr;
  double* norm = nullptr;
  int fwd_lanes = 64;
  int fwd_unroll = 1;  // groups of 4 entries in flight per lane / entry slot
  int bwd_unroll = 2;
  int bwd_lds = 0;     // single-frame backward: 0 block by block, 3 per-XCD row ranges
  int bwd_lpc = 16;    // lanes per column of the per-XCD backward (8 runs as 64)
  int bwd_wgs = 512;   // kept for the get/set API (unused)
  int bwd_bpl = 1;     // row blocks per backward launch (blocked layout)
  int xcd_bpl = 0;     // per-XCD 64-frame backward (kind 3): row blocks each
                       // XCD walks per launch (0: all of its blocks in one)
  int bcast64 = 0;     // 64-frame kernels with in-row broadcast of values/indices
  int layout = RAPTR_LAYOUT_BLOCKED;
  // grouped 64-frame layouts (max_frames = 64): forward groups of 4 plan
  // rows and backward groups of 4 consecutive columns per row block, each
  // an ascending union list with 4 values per union entry
  int grp64 = 0;       // use the grouped kernels for 64-frame passes
  int grp_g = 16;      // members per group (16, 32)
  int* gcols = nullptr;  // ngc * grp_g member columns of the backward groups (-1: none)
  int grp_unroll = 8;  // union entries in flight per wave (4, 8, 16)
  int grp_pipe = 1;    // kernel kind: 0 plain, 1 pipelined scalar loads, 2 matrix cores,
                       // 3 matrix cores with per-XCD row ranges in the backward
  float* c8 = nullptr;  // kind 3: 8 partial images [8][n][64]
  int grouped_only = 0;  // no padded CSR / padded transpose (64-frame grouped passes only)
  uint16_t* brow16 = nullptr;  // transpose rows as offsets in their block (blocks <= 65,536 rows)
  int bwd_r16 = 0;             // single-frame per-XCD backward reads brow16
  uint16_t* fcol16 = nullptr;  // forward columns' low 16 bits (grids of <= 4 x 65,536 voxels)
  uint64_t* fsplit = nullptr;  // per row: sorted positions where the high part steps
  int fwd_c16 = 0;             // single-frame forward reads fcol16
  int fwd_kind = 0;            // 0 rows per lane group, 2 LDS-staged x
  // LDS-staged forward: per group of lrows consecutive plan rows the
  // sorted union of its columns (lup: group offsets, lucol: columns) and
  // every padded entry's 16-bit rank in its group's union
  int64_t* lup = nullptr;
  int* lucol = nullptr;
  uint16_t* lcol = nullptr;
  int lrows = 64;              // plan rows per union group
  int lmax = 0;                // largest union
  int64_t lnu = 0;             // union entries of all groups
  // fused single-frame pass (k_fb1l + a column reduce): the groups'
  // partial column sums over their unions (cpart) and, for the per-XCD and
  // by-column reduces, the unions' positions by column (vptr: n + 1
  // offsets, vpos: positions ascending per column)
  float* cpart = nullptr;
  int64_t* vptr = nullptr;
  int* vpos = nullptr;
  int64_t* vb = nullptr;  // 7 x n: per XCD range of groups, where column j's positions start
  // the unions cut at column tiles of 2^tshift columns: segments (union
  // position ranges of one group inside one tile) sorted by (tile, group);
  // each tile's segments split into tk[t] parts of about tseg_target
  // segments, one workgroup each (twg: tile, part, first and end segment);
  // ck[k][j] takes part k's sums of column j
  int2* tseg = nullptr;
  uint16_t* tcol = nullptr;  // per union position: its column inside its tile
  int4* twg = nullptr;
  int* tk = nullptr;
  float* ck = nullptr;
  int tshift = 0, tseg_target = 0, tnwg = 0, tkmax = 0;
  int64_t nseg = 0;
  int fused = 0;      // single-frame iterations run the fused pass
  int fb_xcd = 1;     // the reduce: 0 one pass by column, 1 per-XCD group ranges
                      // (k_fb1_reduce8), 2 column tiles (k_fb1_treduce)
  int fb_waves = 4;   // waves per row group (2 or 4)
  int fb_lanes = 32;  // lanes per row (16, 32 or 64)
  int fb_steps = 2;   // register-held steps per row with 64 lanes (2, 3 or 4)
  int fb_lpc = 64;    // lanes per column of the reduce (16, 32 or 64)
  int64_t ngf = 0, ngc = 0, gfnnz = 0, gbnnz = 0;
  int64_t* gfp = nullptr;  // ngf + 1
  int* gfcol = nullptr;
  void* gfval = nullptr;   // 4 values per union entry
  int64_t* gbp = nullptr;  // nblk * ngc + 1, (block, column group) order
  int* gbrow = nullptr;
  void* gbval = nullptr;
  // a clone (raptr_recon_clone) shares its parent's read-only layouts and
  // owns only its work vectors; a parent counts its live clones
  raptr_recon* parent = nullptr;
  int nclones = 0;
  // plan row order: plan row p holds the caller's row rperm[p] (null:
  // identity); yp holds the counts in plan order
  int* rperm = nullptr;
  float* yp = nullptr;
  size_t bytes = 0;
  std::vector<void*> allocs;
  std::vector<size_t> asizes;  // bytes of allocs[i]
  // captured iterations, keyed by (frames, y, x) exactly; at most
  // kMaxGraphs, the least recently used destroyed first
  struct CachedGraph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint64_t used;
  };
  // keyed by (frames, active frames, y, x)
  std::map<std::tuple<int, int, uintptr_t, uintptr_t>, CachedGraph> graphs;
  uint64_t graph_tick = 0;
};

namespace raptr {

constexpr size_t kMaxGraphs = 16;

inline void destroy_graph(raptr_recon::CachedGraph& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
}

// every captured iteration of the plan (after a change of kernel shape);
// replays may still be queued on the caller's streams, so the device is
// drained first
inline void clear_graphs(raptr_recon* R) {
  if (R->graphs.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : R->graphs) destroy_graph(kv.second);
  R->graphs.clear();
}

template <class T>
int ralloc(raptr_recon* R, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256));
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "recon: out of device memory (%zu B)",
                count * sizeof(T));
  }
  RAPTR_HIP_TRY(e);
  R->allocs.push_back(q);
  R->asizes.push_back(count * sizeof(T));
  R->bytes += count * sizeof(T);
  *p = static_cast<T*>(q);
  return RAPTR_OK;
}

// frees a buffer of the plan before the plan ends (a format the build
// rejected); the device must be done with it
template <class T>
void rfree(raptr_recon* R, T*& p) {
  for (size_t i = 0; p && i < R->allocs.size(); ++i) {
    if (R->allocs[i] != static_cast<void*>(p)) continue;
    (void)hipFree(R->allocs[i]);
    R->bytes -= R->asizes[i];
    R->allocs.erase(R->allocs.begin() + i);
    R->asizes.erase(R->asizes.begin() + i);
    break;
  }
  p = nullptr;
}

inline int vbytes(int vt) { return vt == kF32 ? 4 : 2; }

}  // namespace raptr
