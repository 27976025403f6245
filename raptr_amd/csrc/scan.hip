// scan.hip -- device-wide exclusive scan: tile reduce, recursive scan of the
// tile sums, tile scan with offsets.  Tiles of 256 threads x 8 items, loaded
// coalesced through LDS, scanned with wave64 shuffles.
#include "primitives.hpp"

namespace raptr {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int64_t kScanTile = kScanBlock * kScanItems;

template <class TIn, class TOut>
__global__ __launch_bounds__(kScanBlock) void k_tile_reduce(const TIn* in, int64_t n,
                                                            TOut* sums) {
  __shared__ TOut lds[kScanBlock / kWave];
  int64_t base = int64_t(blockIdx.x) * kScanTile;
  TOut acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t k = base + int64_t(i) * kScanBlock + threadIdx.x;
    if (k < n) acc += TOut(in[k]);
  }
  acc = wave_reduce_sum(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    TOut t = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) t += lds[w];
    sums[blockIdx.x] = t;
  }
}

// Exclusive scan of one tile, plus an optional per-tile offset; writes the
// grand total (offset of the last tile + its sum) when total != nullptr.
template <class TIn, class TOut>
__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const TIn* in, TOut* out,
                                                          int64_t n,
                                                          const TOut* offsets,
                                                          TOut* total) {
  __shared__ TOut tile[kScanTile];
  __shared__ TOut lds[kScanBlock / kWave];
  int64_t base = int64_t(blockIdx.x) * kScanTile;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int idx = i * kScanBlock + threadIdx.x;
    int64_t k = base + idx;
    tile[idx] = k < n ? TOut(in[k]) : TOut(0);
  }
  __syncthreads();
  TOut v[kScanItems];
  TOut sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = tile[threadIdx.x * kScanItems + i];
    sum += v[i];
  }
  TOut tot;
  TOut pre = block_exclusive_scan<TOut, kScanBlock>(sum, lds, &tot);
  TOut off = offsets ? offsets[blockIdx.x] : TOut(0);
  TOut run = off + pre;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    tile[threadIdx.x * kScanItems + i] = run;
    run += v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int idx = i * kScanBlock + threadIdx.x;
    int64_t k = base + idx;
    if (k < n) out[k] = tile[idx];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tot;
}

template <class TOut>
size_t levels_bytes(int64_t n) {
  size_t b = 0;
  while (n > kScanTile) {
    int64_t t = (n + kScanTile - 1) / kScanTile;
    b += 2 * (size_t(t) * sizeof(TOut) + 256);
    n = t;
  }
  return b + 512;
}

template <class TIn, class TOut>
int scan_impl(const TIn* in, TOut* out, int64_t n, TOut* d_total, char* ws,
              size_t ws_bytes, hipStream_t st) {
  if (n <= 0) {
    if (d_total) RAPTR_HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(TOut), st));
    return RAPTR_OK;
  }
  if (n <= kScanTile) {
    hipLaunchKernelGGL((k_tile_scan<TIn, TOut>), dim3(1), dim3(kScanBlock), 0, st, in,
                       out, n, (const TOut*)nullptr, d_total);
    RAPTR_LAUNCH_CHECK();
    return RAPTR_OK;
  }
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles > INT32_MAX) return fail(RAPTR_ERR_INVALID_ARG, "scan too large");
  size_t need = 2 * (size_t(tiles) * sizeof(TOut) + 256);
  if (!ws || ws_bytes < need) return fail(RAPTR_ERR_WORKSPACE, "scan workspace too small");
  TOut* sums = reinterpret_cast<TOut*>(ws);
  char* p2 = ws + ((size_t(tiles) * sizeof(TOut) + 255) & ~size_t(255));
  TOut* offs = reinterpret_cast<TOut*>(p2);
  char* rest = p2 + ((size_t(tiles) * sizeof(TOut) + 255) & ~size_t(255));
  size_t rest_bytes = ws_bytes - size_t(rest - ws);
  hipLaunchKernelGGL((k_tile_reduce<TIn, TOut>), dim3(unsigned(tiles)), dim3(kScanBlock),
                     0, st, in, n, sums);
  RAPTR_LAUNCH_CHECK();
  RAPTR_TRY((scan_impl<TOut, TOut>(sums, offs, tiles, (TOut*)nullptr, rest, rest_bytes, st)));
  hipLaunchKernelGGL((k_tile_scan<TIn, TOut>), dim3(unsigned(tiles)), dim3(kScanBlock), 0,
                     st, in, out, n, (const TOut*)offs, d_total);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}
}  // namespace

size_t scan_workspace_bytes(int64_t n) { return levels_bytes<int64_t>(n); }

int scan_exclusive_i32_i64(const int* in, int64_t* out, int64_t n, int64_t* d_total,
                           void* ws, size_t ws_bytes, hipStream_t st) {
  return scan_impl<int, int64_t>(in, out, n, d_total, static_cast<char*>(ws), ws_bytes, st);
}
int scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* d_total,
                       void* ws, size_t ws_bytes, hipStream_t st) {
  return scan_impl<int64_t, int64_t>(in, out, n, d_total, static_cast<char*>(ws), ws_bytes,
                                     st);
}
int scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                       void* ws, size_t ws_bytes, hipStream_t st) {
  return scan_impl<uint32_t, uint32_t>(in, out, n, d_total, static_cast<char*>(ws),
                                       ws_bytes, st);
}

}  // namespace raptr
