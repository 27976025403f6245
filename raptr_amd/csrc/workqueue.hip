// workqueue.hip -- cylinder-culled (channel, voxel) workqueue on gfx950.
//
// Replaces the serial host scan getWorkqueue (getWorkqueue.hpp:39-243).  The
// (list entry x voxel) space is cut into tiles of 1024 voxels of one list
// entry; a count pass gives per-tile candidate counts (wave ballot +
// popcount), a device-wide scan turns them into output offsets, and an emit
// pass re-evaluates the same float test and writes each candidate at its
// offset plus its ballot rank, so the output keeps the reference's
// list-major, voxel-ascending order exactly.
#include "primitives.hpp"

namespace raptr {
namespace {

constexpr int kWqBlock = 256;
constexpr int kWqRounds = 4;
constexpr int kWqTile = kWqBlock * kWqRounds;  // voxels per tile

struct WqArgs {
  Geom G;
  FastDiv div_yz, div_z;
  const int* list;
  int list_first;  // cursor list entry
  int vxl_first;   // cursor voxel (applies to list_first only)
  int tiles;       // tiles per list entry
};

__device__ inline bool wq_test(const WqArgs& A, const float* p0, const float* p1, int v) {
  uint32_t ix = fdiv(uint32_t(v), A.div_yz);
  uint32_t r = uint32_t(v) - ix * A.div_yz.d;
  uint32_t iy = fdiv(r, A.div_z);
  uint32_t iz = r - iy * A.div_z.d;
  float c[3];
  voxel_centre(A.G.g, int(ix), int(iy), int(iz), c);
  return distance_point_line(p0, p1, c) < A.G.sumRadii;
}

// Block (li_rel, tile) -> candidate mask per round; shared by both passes.
__device__ inline void wq_tile(const WqArgs& A, int64_t blk, int& li, int& tile,
                               float* p0, float* p1) {
  li = int(blk / A.tiles);
  tile = int(blk - int64_t(li) * A.tiles);
  li += A.list_first;
  int c = A.list[li];
  channel_centres(A.G, (c >= 0 && c < A.G.nch) ? c : 0, p0, p1);
}

__device__ inline bool wq_valid(const WqArgs& A, int li) {
  int c = A.list[li];
  return c >= 0 && c < A.G.nch;
}

__global__ __launch_bounds__(kWqBlock) void k_wq_count(WqArgs A, int64_t blk0,
                                                       int* tile_counts) {
  __shared__ int wsum[kWqBlock / kWave];
  int64_t blk = blk0 + blockIdx.x;
  int li, tile;
  float p0[3], p1[3];
  wq_tile(A, blk, li, tile, p0, p1);
  int vstart = (li == A.list_first) ? A.vxl_first : 0;
  const bool ok = wq_valid(A, li);
  int c = 0;
#pragma unroll
  for (int r = 0; r < kWqRounds; ++r) {
    int v = tile * kWqTile + r * kWqBlock + threadIdx.x;
    bool pass = ok && v < A.G.nvox && v >= vstart && wq_test(A, p0, p1, v);
    c += __popcll(__ballot(pass));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kWqBlock / kWave; ++w) t += wsum[w];
    tile_counts[blk] = t;
  }
}

__global__ __launch_bounds__(kWqBlock) void k_wq_emit(WqArgs A, int64_t blk0,
                                                      const int64_t* tile_offsets,
                                                      int64_t limit, int* cnl_out,
                                                      int* vxl_out, int* last_pos) {
  __shared__ int wcnt[kWqRounds][kWqBlock / kWave];
  int64_t blk = blk0 + blockIdx.x;
  int li, tile;
  float p0[3], p1[3];
  wq_tile(A, blk, li, tile, p0, p1);
  int vstart = (li == A.list_first) ? A.vxl_first : 0;
  const bool ok = wq_valid(A, li);
  int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  uint64_t masks[kWqRounds];
  bool passes[kWqRounds];
#pragma unroll
  for (int r = 0; r < kWqRounds; ++r) {
    int v = tile * kWqTile + r * kWqBlock + threadIdx.x;
    passes[r] = ok && v < A.G.nvox && v >= vstart && wq_test(A, p0, p1, v);
    masks[r] = __ballot(passes[r]);
    if (lane == 0) wcnt[r][w] = __popcll(masks[r]);
  }
  __syncthreads();
  int64_t base = tile_offsets[blk];
  int cnl = A.list[li];
  int before = 0;
#pragma unroll
  for (int r = 0; r < kWqRounds; ++r) {
    int pre = before;
    for (int ww = 0; ww < w; ++ww) pre += wcnt[r][ww];
    for (int ww = 0; ww < kWqBlock / kWave; ++ww) before += wcnt[r][ww];
    if (passes[r]) {
      int64_t pos = base + pre + __popcll(masks[r] & lanemask_lt());
      if (pos < limit) {
        int v = tile * kWqTile + r * kWqBlock + threadIdx.x;
        cnl_out[pos] = cnl;
        vxl_out[pos] = v;
        if (pos == limit - 1) {
          last_pos[0] = li;
          last_pos[1] = v;
        }
      }
    }
  }
}

// Candidate count of each list entry over its full voxel range: one block
// per entry (grid-strided), all tiles of the entry in turn.
__global__ __launch_bounds__(kWqBlock) void k_wq_entry_counts(WqArgs A, int n_entries,
                                                              int64_t* out) {
  __shared__ int wsum[kWqBlock / kWave];
  for (int e = blockIdx.x; e < n_entries; e += gridDim.x) {
    float p0[3], p1[3];
    const int ce = A.list[e];
    const bool ok = ce >= 0 && ce < A.G.nch;
    channel_centres(A.G, ok ? ce : 0, p0, p1);
    int c = 0;
    for (int v = threadIdx.x; v < (ok ? A.G.nvox : 0); v += kWqBlock)
      c += __popcll(__ballot(wq_test(A, p0, p1, v)));
    if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int w = 0; w < kWqBlock / kWave; ++w) t += wsum[w];
      out[e] = t;
    }
    __syncthreads();
  }
}

int make_wq_args(const raptr_setup_t* setup, const raptr_grid_t* grid, const int* d_list,
                 int list_first, int vxl_first, WqArgs* A) {
  RAPTR_TRY(make_geom(setup, grid, &A->G));
  A->div_yz = make_fastdiv(uint32_t(grid->n[1] * grid->n[2]));
  A->div_z = make_fastdiv(uint32_t(grid->n[2]));
  A->list = d_list;
  A->list_first = list_first;
  A->vxl_first = vxl_first;
  A->tiles = (A->G.nvox + kWqTile - 1) / kWqTile;
  return RAPTR_OK;
}

// Launch a per-block kernel over nblk blocks in slices that keep the grid
// below 2^31 work-items.
template <class F>
int launch_sliced(int64_t nblk, F f) {
  const int64_t kMaxBlocks = (int64_t(1) << 31) / kWqBlock;
  for (int64_t b0 = 0; b0 < nblk; b0 += kMaxBlocks) {
    int64_t nb = nblk - b0 < kMaxBlocks ? nblk - b0 : kMaxBlocks;
    f(b0, unsigned(nb));
    RAPTR_LAUNCH_CHECK();
  }
  return RAPTR_OK;
}

}  // namespace
}  // namespace raptr

using namespace raptr;

extern "C" {

size_t raptr_workqueue_workspace_size(const raptr_grid_t* grid, int list_size) {
  if (!grid || list_size < 0) return 0;
  int64_t nvox = int64_t(grid->n[0]) * grid->n[1] * grid->n[2];
  int64_t tiles = (nvox + kWqTile - 1) / kWqTile;
  int64_t nblk = tiles * int64_t(list_size);
  Sizer z;
  z.take<int>(size_t(nblk));
  z.take<int64_t>(size_t(nblk));
  z.take<int64_t>(4);
  return z.bytes() + scan_workspace_bytes(nblk);
}

int raptr_workqueue(const raptr_setup_t* setup, const raptr_grid_t* grid, const int* d_list,
                    int list_size, int* list_id, int* vxl_id, int64_t n, int* d_cnl,
                    int* d_vxl, int64_t capacity, int64_t* n_found, void* d_ws,
                    size_t ws_bytes, raptr_stream_t stream) {
  if (!list_id || !vxl_id || !n_found) return fail(RAPTR_ERR_INVALID_ARG, "null cursor");
  *n_found = 0;
  if (list_size < 0 || (list_size > 0 && !d_list))
    return fail(RAPTR_ERR_INVALID_ARG, "bad list");
  WqArgs A;
  RAPTR_TRY(make_wq_args(setup, grid, d_list, *list_id, *vxl_id, &A));
  if (*list_id < 0 || *list_id > list_size || *vxl_id < 0 || *vxl_id >= A.G.nvox)
    return fail(RAPTR_ERR_INVALID_ARG,
                "invalid cursor (list_id %d, vxl_id %d)", *list_id, *vxl_id);
  if (n == 0 || *list_id == list_size) return RAPTR_OK;
  hipStream_t st = S(stream);
  int entries = list_size - *list_id;
  int64_t nblk = int64_t(entries) * A.tiles;
  Carver cv(d_ws, ws_bytes);
  int* tile_counts = cv.take<int>(size_t(nblk));
  int64_t* tile_offs = cv.take<int64_t>(size_t(nblk));
  int64_t* small = cv.take<int64_t>(4);
  size_t scan_bytes = scan_workspace_bytes(nblk);
  void* scan_ws = cv.take<char>(scan_bytes);
  if (!cv.ok) return fail(RAPTR_ERR_WORKSPACE, "workqueue workspace too small (%zu < %zu)",
                          ws_bytes, raptr_workqueue_workspace_size(grid, list_size));
  RAPTR_TRY(launch_sliced(nblk, [&](int64_t b0, unsigned nb) {
    hipLaunchKernelGGL(k_wq_count, dim3(nb), dim3(kWqBlock), 0, st, A, b0, tile_counts);
  }));
  RAPTR_TRY(scan_exclusive_i32_i64(tile_counts, tile_offs, nblk, &small[0], scan_ws,
                                   scan_bytes, st));
  int64_t total = 0;
  RAPTR_HIP_TRY(hipMemcpyAsync(&total, &small[0], sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RAPTR_HIP_TRY(hipStreamSynchronize(st));
  bool bounded = n >= 0;
  int64_t found = (bounded && n < total) ? n : total;
  if (found > capacity)
    return fail(RAPTR_ERR_INVALID_ARG, "output capacity %lld < %lld entries",
                (long long)capacity, (long long)found);
  if (found > 0 && (!d_cnl || !d_vxl)) return fail(RAPTR_ERR_INVALID_ARG, "null output");
  int* last = reinterpret_cast<int*>(&small[1]);
  RAPTR_TRY(launch_sliced(nblk, [&](int64_t b0, unsigned nb) {
    hipLaunchKernelGGL(k_wq_emit, dim3(nb), dim3(kWqBlock), 0, st, A, b0,
                       (const int64_t*)tile_offs, found, d_cnl, d_vxl, last);
  }));
  int cur[2] = {list_size, 0};
  if (bounded && found == n && found > 0) {
    // Cursor after the n-th entry, as getWorkqueueEntry advances it
    // (getWorkqueue.hpp:101-112).
    int lp[2];
    RAPTR_HIP_TRY(hipMemcpyAsync(lp, last, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    RAPTR_HIP_TRY(hipStreamSynchronize(st));
    if (lp[1] < A.G.nvox - 1) {
      cur[0] = lp[0];
      cur[1] = lp[1] + 1;
    } else {
      cur[0] = lp[0] + 1;
      cur[1] = 0;
    }
  } else {
    RAPTR_HIP_TRY(hipStreamSynchronize(st));
  }
  *list_id = cur[0];
  *vxl_id = cur[1];
  *n_found = found;
  return RAPTR_OK;
}

int raptr_workqueue_counts(const raptr_setup_t* setup, const raptr_grid_t* grid,
                           const int* d_list, int list_size, int64_t* d_counts,
                           raptr_stream_t stream) {
  if (list_size <= 0) return RAPTR_OK;
  if (!d_list || !d_counts) return fail(RAPTR_ERR_INVALID_ARG, "null argument");
  WqArgs A;
  RAPTR_TRY(make_wq_args(setup, grid, d_list, 0, 0, &A));
  unsigned g = unsigned(list_size < (1 << 20) ? list_size : (1 << 20));
  hipLaunchKernelGGL(k_wq_entry_counts, dim3(g), dim3(kWqBlock), 0, S(stream), A, list_size,
                     d_counts);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

}  // extern "C"
