// recon_tiled.hip -- column-tiled single-frame MLEM layout on gfx950.
//
// The voxels are cut into tiles of tile_cols consecutive columns (at most
// 36,864, so a tile of FP32 values is <= 144 KiB of LDS) and every row's
// entries are stored tile by tile: tile q holds, row after row, the
// entries of columns [q*tile_cols, (q+1)*tile_cols), each row segment
// padded to a multiple of 4 entries, with 16-bit column offsets.  Segment
// starts (tile-major, then row) are one int64 array, so the entries of one
// tile and a range of rows are contiguous.
//
// A pass launches ntile * tile_wgs workgroups of 1024 threads, one per CU
// (the LDS tile limits residency), each with one tile and a range of rows
// balanced by entry count:
//  * forward: the x tile is loaded into LDS; lane groups stream row
//    segments (4 values + 4 offsets per load) and gather from LDS; the
//    per-(tile, row) partial dot is written out.  A small kernel sums the
//    partials of each row in tile order and forms the ratio e.
//  * backward: the tile's column sums live in LDS; each row segment is
//    scaled by its row's ratio and added with LDS atomics; rows with e = 0
//    are skipped.  Each workgroup writes its partial column sums, and the
//    update kernel sums them in a fixed order, applies
//    x' = x*c/(s + (s==0)) and forms FP64 partial sums of x'.
// Neither pass gathers through L2, and the matrix is read once per pass
// (6 bytes per FP32 entry, 4 per 16-bit entry).
#include "recon.hpp"

namespace raptr {
namespace {

constexpr int kTileMaxCols = 36864;
constexpr int kTileBlock = 1024;

// per-row scale of the 16-bit types: row maximum (1 for an all-zero row)
__global__ __launch_bounds__(256) void k_row_scale(const int64_t* rp, const float* val, int m,
                                                   int vt, float* scale) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    float mx = 0.0f;
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += kWave) mx = fmaxf(mx, fabsf(val[k]));
    for (int off = kWave / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
    float sc = mx > 0.0f ? mx : 1.0f;
    if (lane == 0) scale[i] = vt == kU16 ? sc * (1.0f / 65535.0f) : sc;
  }
}

// first entry of row [b, e) with column >= c (columns ascending)
__device__ inline int64_t col_lower_bound(const int* col, int64_t b, int64_t e, int c) {
  while (b < e) {
    int64_t mid = b + ((e - b) >> 1);
    if (col[mid] < c) b = mid + 1; else e = mid;
  }
  return b;
}

// padded segment length of (tile q, row i) into cnt[q*m + i]
__global__ __launch_bounds__(256) void k_tile_count(const int64_t* rp, const int* col, int m,
                                                    int ntile, int tcols, int64_t* cnt) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    const int64_t b = rp[i], e = rp[i + 1];
    for (int q = lane; q < ntile; q += kWave) {
      int64_t lo = col_lower_bound(col, b, e, q * tcols);
      int64_t hi = col_lower_bound(col, lo, e, (q + 1) * tcols);
      cnt[int64_t(q) * m + i] = (hi - lo + 3) & ~int64_t(3);
    }
  }
}

__global__ void k_tile_last(int64_t* tptr, int64_t n, const int64_t* total) {
  tptr[n] = *total;
}

// copy each row's tile segments to their place, zero-padded
__global__ __launch_bounds__(256) void k_tile_fill(const int64_t* rp, const int* col,
                                                   const float* val, int m, int ntile,
                                                   int tcols, int vt, const float* scale,
                                                   const int64_t* tptr, uint16_t* tidx,
                                                   void* tval) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    const int64_t b = rp[i], e = rp[i + 1];
    const float sc = vt == kF32 ? 1.0f : (vt == kU16 ? scale[i] * 65535.0f : scale[i]);
    int64_t lo = col_lower_bound(col, b, e, 0);
    for (int q = 0; q < ntile; ++q) {
      const int64_t hi = col_lower_bound(col, lo, e, (q + 1) * tcols);
      const int64_t d = tptr[int64_t(q) * m + i];
      const int64_t dl = tptr[int64_t(q) * m + i + 1] - d;
      for (int64_t k = lane; k < dl; k += kWave) {
        const bool real = lo + k < hi;
        const int c = real ? col[lo + k] - q * tcols : 0;
        const float v = real ? val[lo + k] : 0.0f;
        tidx[d + k] = uint16_t(c);
        if (vt == kF32)
          static_cast<float*>(tval)[d + k] = v;
        else
          static_cast<uint16_t*>(tval)[d + k] = f2h(v / sc, vt);
      }
      lo = hi;
    }
  }
}

// row bounds of the tile's workgroups, balanced by entry count
__global__ void k_tile_rows(const int64_t* tptr, int m, int ntile, int W, int* trow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntile * (W + 1)) return;
  const int q = t / (W + 1), w = t % (W + 1);
  const int64_t* P = tptr + int64_t(q) * m;
  const int64_t base = P[0], total = P[m] - base;
  const int64_t target = base + total * w / W;
  int lo = 0, hi = m;
  while (lo < hi) {
    int mid = lo + ((hi - lo) >> 1);
    if (P[mid] < target) lo = mid + 1; else hi = mid;
  }
  trow[t] = w == W ? m : lo;
}

// x tile -> LDS (vector loads when the tile start is 16-byte aligned)
__device__ inline void load_tile(float* lds, const float* src, int cols) {
  const int t = threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    const int c4 = cols & ~3;
    for (int k = 4 * t; k < c4; k += 4 * kTileBlock) {
      vf4 v = *reinterpret_cast<const vf4*>(src + k);
      lds[k] = v.x, lds[k + 1] = v.y, lds[k + 2] = v.z, lds[k + 3] = v.w;
    }
    for (int k = c4 + t; k < cols; k += kTileBlock) lds[k] = src[k];
  } else {
    for (int k = t; k < cols; k += kTileBlock) lds[k] = src[k];
  }
}

__device__ __forceinline__ void idx4(const uint16_t* tidx, int64_t k, int& i0, int& i1, int& i2,
                                     int& i3) {
  vu2 a = __builtin_nontemporal_load(reinterpret_cast<const vu2*>(tidx + k));
  i0 = int(a.x & 0xffffu), i1 = int(a.x >> 16), i2 = int(a.y & 0xffffu), i3 = int(a.y >> 16);
}

template <int VT, int LANES, int U>
__global__ __launch_bounds__(kTileBlock) void k_tfwd(int m, int n, int tcols, int W,
                                                     const int64_t* __restrict__ tptr,
                                                     const uint16_t* __restrict__ tidx,
                                                     const typename Val<VT>::T* __restrict__ tval,
                                                     const int* __restrict__ trow,
                                                     const float* __restrict__ x,
                                                     const float* __restrict__ y,
                                                     float* __restrict__ tpart) {
  extern __shared__ float xs[];
  const int q = blockIdx.x / W, w = blockIdx.x % W;
  const int c0 = q * tcols;
  const int cols = min(tcols, n - c0);
  load_tile(xs, x + c0, cols);
  __syncthreads();
  const int r0 = trow[q * (W + 1) + w], r1 = trow[q * (W + 1) + w + 1];
  constexpr int GROUPS = kTileBlock / LANES;
  const int g = threadIdx.x % LANES, grp = threadIdx.x / LANES;
  const int64_t* P = tptr + int64_t(q) * m;
  float* out = tpart + int64_t(q) * m;
  // the next row's segment bounds and count are loaded while the current
  // row streams, so a row costs one memory round trip, not two
  int i = r0 + grp;
  int64_t b = 0, en = 0;
  float yi = 0.0f;
  if (i < r1) b = P[i], en = P[i + 1], yi = y[i];
  for (; i < r1; i += GROUPS) {
    const int inx = i + GROUPS;
    int64_t bn = 0, enn = 0;
    float yn = 0.0f;
    if (inx < r1) bn = P[inx], enn = P[inx + 1], yn = y[inx];
    float acc = 0.0f;
    if (yi != 0.0f) {  // e = 0 whatever A x is when y = 0
      int64_t k = b + 4 * g;
      for (; k + int64_t(4) * LANES * (U - 1) < en; k += int64_t(4) * LANES * U) {
        float v[U][4];
        int c[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t kk = k + int64_t(4) * LANES * u;
          load4<VT>(tval, kk, v[u][0], v[u][1], v[u][2], v[u][3]);
          idx4(tidx, kk, c[u][0], c[u][1], c[u][2], c[u][3]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          acc = fmaf(v[u][0], xs[c[u][0]], acc);
          acc = fmaf(v[u][1], xs[c[u][1]], acc);
          acc = fmaf(v[u][2], xs[c[u][2]], acc);
          acc = fmaf(v[u][3], xs[c[u][3]], acc);
        }
      }
      for (; k < en; k += int64_t(4) * LANES) {
        float v0, v1, v2, v3;
        int c0_, c1_, c2_, c3_;
        load4<VT>(tval, k, v0, v1, v2, v3);
        idx4(tidx, k, c0_, c1_, c2_, c3_);
        acc = fmaf(v0, xs[c0_], acc);
        acc = fmaf(v1, xs[c1_], acc);
        acc = fmaf(v2, xs[c2_], acc);
        acc = fmaf(v3, xs[c3_], acc);
      }
#pragma unroll
      for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    }
    if (g == 0) out[i] = acc;
    b = bn, en = enn, yi = yn;
  }
}

// e_i = y_i / (scale_i * sum_q tpart[q][i]) (0 where y or the sum is 0),
// folded with the row scale for the backward pass of 16-bit values
__global__ void k_tratio(int m, int ntile, int vt, const float* __restrict__ tpart,
                         const float* __restrict__ rscale, const float* __restrict__ y,
                         float* __restrict__ e) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += int64_t(gridDim.x) * blockDim.x) {
    float acc = 0.0f;
    for (int q = 0; q < ntile; ++q) acc += tpart[int64_t(q) * m + i];
    const float sc = vt == kF32 ? 1.0f : rscale[i];
    const float yt = sc * acc;
    const float yy = y[i];
    const float r = (yt != 0.0f && yy != 0.0f) ? yy / yt : 0.0f;
    e[i] = vt == kF32 ? r : sc * r;
  }
}

template <int VT, int LANES, int U>
__global__ __launch_bounds__(kTileBlock) void k_tbwd(int m, int n, int tcols, int W,
                                                     const int64_t* __restrict__ tptr,
                                                     const uint16_t* __restrict__ tidx,
                                                     const typename Val<VT>::T* __restrict__ tval,
                                                     const int* __restrict__ trow,
                                                     const float* __restrict__ e,
                                                     float* __restrict__ cpart) {
  extern __shared__ float cs[];
  const int q = blockIdx.x / W, w = blockIdx.x % W;
  const int c0 = q * tcols;
  const int cols = min(tcols, n - c0);
  for (int k = threadIdx.x; k < tcols; k += kTileBlock) cs[k] = 0.0f;
  __syncthreads();
  const int r0 = trow[q * (W + 1) + w], r1 = trow[q * (W + 1) + w + 1];
  constexpr int GROUPS = kTileBlock / LANES;
  const int g = threadIdx.x % LANES, grp = threadIdx.x / LANES;
  const int64_t* P = tptr + int64_t(q) * m;
  int i = r0 + grp;
  int64_t b = 0, en = 0;
  float ei = 0.0f;
  if (i < r1) b = P[i], en = P[i + 1], ei = e[i];
  for (; i < r1; i += GROUPS) {
    const int inx = i + GROUPS;
    int64_t bn = 0, enn = 0;
    float en_ = 0.0f;
    if (inx < r1) bn = P[inx], enn = P[inx + 1], en_ = e[inx];
    if (ei != 0.0f) {
      int64_t k = b + 4 * g;
      for (; k + int64_t(4) * LANES * (U - 1) < en; k += int64_t(4) * LANES * U) {
        float v[U][4];
        int c[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t kk = k + int64_t(4) * LANES * u;
          load4<VT>(tval, kk, v[u][0], v[u][1], v[u][2], v[u][3]);
          idx4(tidx, kk, c[u][0], c[u][1], c[u][2], c[u][3]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (v[u][j] != 0.0f) atomicAdd(&cs[c[u][j]], v[u][j] * ei);
      }
      for (; k < en; k += int64_t(4) * LANES) {
        float v[4];
        int c[4];
        load4<VT>(tval, k, v[0], v[1], v[2], v[3]);
        idx4(tidx, k, c[0], c[1], c[2], c[3]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (v[j] != 0.0f) atomicAdd(&cs[c[j]], v[j] * ei);
      }
    }
    b = bn, en = enn, ei = en_;
  }
  __syncthreads();
  float* out = cpart + (int64_t(q) * W + w) * tcols;
  for (int k = threadIdx.x; k < cols; k += kTileBlock) out[k] = cs[k];
}

// c_j = sum over the tile's workgroups (fixed order); x' = x*c/(s + (s==0));
// FP64 partial sum per block
__global__ __launch_bounds__(256) void k_tupdate(int n, int tcols, int W,
                                                 const float* __restrict__ cpart,
                                                 const float* __restrict__ x,
                                                 const float* __restrict__ s,
                                                 float* __restrict__ xx,
                                                 double* __restrict__ partials) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  double part = 0.0;
  if (j < n) {
    const int q = j / tcols, jl = j - q * tcols;
    const float* p = cpart + int64_t(q) * W * tcols + jl;
    float c = 0.0f;
    for (int w = 0; w < W; ++w) c += p[int64_t(w) * tcols];
    float d = s[j];
    d += float(d == 0.0f);
    const float v = x[j] * c / d;
    xx[j] = v;
    part = double(v);
  }
  __shared__ double red[256 / kWave];
  part = wave_reduce_sum(part);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 256 / kWave; ++k) t += red[k];
    partials[blockIdx.x] = t;
  }
}

template <class K>
void allow_lds(K kern) {
  static bool done = false;
  if (!done) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
      (void)hipGetLastError();
    done = true;
  }
}

}  // namespace

int tiled_create(raptr_recon* R, const int64_t* d_rp, const int* d_col, const float* d_val,
                 hipStream_t st) {
  const int m = R->m, n = R->n;
  // block_rows below the LDS limit narrows the tiles (tests use it to get
  // several tiles on small grids)
  const int cap = (R->blk_rows > 0 && R->blk_rows < kTileMaxCols) ? R->blk_rows : kTileMaxCols;
  R->ntile = (n + cap - 1) / cap;
  int tc = (n + R->ntile - 1) / R->ntile;
  R->tile_cols = (tc + 3) & ~3;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      ncu = prop.multiProcessorCount;
  }
  R->tile_wgs = std::max(1, (ncu + R->ntile - 1) / R->ntile);
  const int64_t segs = int64_t(R->ntile) * m;
  int rc;
  if (R->vt != kF32) {
    if ((rc = ralloc(R, &R->rscale, size_t(m) + 1))) return rc;
    if (m > 0) {
      hipLaunchKernelGGL(k_row_scale, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                         dim3(256), 0, st, d_rp, d_val, m, R->vt, R->rscale);
      RAPTR_LAUNCH_CHECK();
    }
  }
  if ((rc = ralloc(R, &R->tptr, size_t(segs) + 1))) return rc;
  {
    void* tmp = nullptr;
    const size_t sb = scan_workspace_bytes(std::max<int64_t>(segs, 1));
    const size_t cb = (size_t(std::max<int64_t>(segs, 1)) * 8 + 255) & ~size_t(255);
    RAPTR_HIP_TRY(hipMalloc(&tmp, cb + 256 + sb));
    int64_t* cnt = static_cast<int64_t*>(tmp);
    int64_t* tot = reinterpret_cast<int64_t*>(static_cast<char*>(tmp) + cb);
    char* sws = reinterpret_cast<char*>(tot) + 256;
    rc = RAPTR_OK;
    if (segs > 0) {
      hipLaunchKernelGGL(k_tile_count, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                         dim3(256), 0, st, d_rp, d_col, m, R->ntile, R->tile_cols, cnt);
      rc = scan_exclusive_i64(cnt, R->tptr, segs, tot, sws, sb, st);
      if (!rc)
        hipLaunchKernelGGL(k_tile_last, dim3(1), dim3(1), 0, st, R->tptr, segs,
                           (const int64_t*)tot);
    } else {
      (void)hipMemsetAsync(R->tptr, 0, 8, st);
      (void)hipMemsetAsync(tot, 0, 8, st);
    }
    int64_t h = 0;
    (void)hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st);
    hipError_t se = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    if (rc) return rc;
    RAPTR_HIP_TRY(se);
    R->tnnz = h;
  }
  if ((rc = ralloc(R, &R->tidx, size_t(R->tnnz) + 8))) return rc;
  if (R->vt == kF32) {
    float* p;
    if ((rc = ralloc(R, &p, size_t(R->tnnz) + 8))) return rc;
    R->tval = p;
  } else {
    uint16_t* p;
    if ((rc = ralloc(R, &p, size_t(R->tnnz) + 8))) return rc;
    R->tval = p;
  }
  if (m > 0 && R->tnnz > 0) {
    hipLaunchKernelGGL(k_tile_fill, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)), dim3(256),
                       0, st, d_rp, d_col, d_val, m, R->ntile, R->tile_cols, R->vt,
                       (const float*)R->rscale, (const int64_t*)R->tptr, R->tidx, R->tval);
    RAPTR_LAUNCH_CHECK();
  }
  const int nb = R->ntile * (R->tile_wgs + 1);
  if ((rc = ralloc(R, &R->trow, size_t(nb)))) return rc;
  hipLaunchKernelGGL(k_tile_rows, dim3((nb + 255) / 256), dim3(256), 0, st,
                     (const int64_t*)R->tptr, m, R->ntile, R->tile_wgs, R->trow);
  RAPTR_LAUNCH_CHECK();
  if ((rc = ralloc(R, &R->tpart, size_t(segs) + 1))) return rc;
  if ((rc = ralloc(R, &R->cpart, size_t(R->ntile) * R->tile_wgs * R->tile_cols))) return rc;
  R->pnnz = R->tnnz;
  return RAPTR_OK;
}

int tiled_forward(const raptr_recon* R, const float* x, const float* y, float* e,
                  hipStream_t st) {
  const int G = R->ntile * R->tile_wgs;
  const size_t lds = size_t(R->tile_cols) * sizeof(float);
  const int L = R->t_fwd_lanes, U = R->t_fwd_unroll;
  bool launched = false;
#define RAPTR_TF(VV, LL, UU)                                                                    \
  if (!launched && R->vt == VV && L == LL && U == UU) {                                        \
    allow_lds(k_tfwd<VV, LL, UU>);                                                              \
    hipLaunchKernelGGL((k_tfwd<VV, LL, UU>), dim3(G), dim3(kTileBlock), lds, st, R->m, R->n,    \
                       R->tile_cols, R->tile_wgs, (const int64_t*)R->tptr,                      \
                       (const uint16_t*)R->tidx, (const typename Val<VV>::T*)R->tval,           \
                       (const int*)R->trow, x, y, R->tpart);                                    \
    launched = true;                                                                            \
  }
#define RAPTR_TFV(VV)                                                                           \
  RAPTR_TF(VV, 4, 1) RAPTR_TF(VV, 4, 2) RAPTR_TF(VV, 4, 4) RAPTR_TF(VV, 8, 1) RAPTR_TF(VV, 8, 2) \
  RAPTR_TF(VV, 8, 4) RAPTR_TF(VV, 16, 1) RAPTR_TF(VV, 16, 2) RAPTR_TF(VV, 16, 4)
  RAPTR_TFV(kF32) RAPTR_TFV(kF16) RAPTR_TFV(kBF16) RAPTR_TFV(kU16)
#undef RAPTR_TFV
#undef RAPTR_TF
  if (!launched) return fail(RAPTR_ERR_INVALID_ARG, "tiled forward shape unsupported");
  RAPTR_LAUNCH_CHECK();
  if (R->m > 0) {
    hipLaunchKernelGGL(k_tratio, dim3(grid_for(R->m, 256, 8192)), dim3(256), 0, st, R->m,
                       R->ntile, R->vt, (const float*)R->tpart, (const float*)R->rscale, y, e);
    RAPTR_LAUNCH_CHECK();
  }
  return RAPTR_OK;
}

int tiled_backward(const raptr_recon* R, const float* x, const float* e, hipStream_t st,
                   int* nb) {
  const int G = R->ntile * R->tile_wgs;
  const size_t lds = size_t(R->tile_cols) * sizeof(float);
  const int L = R->t_bwd_lanes, U = R->t_bwd_unroll;
  bool launched = false;
#define RAPTR_TB(VV, LL, UU)                                                                    \
  if (!launched && R->vt == VV && L == LL && U == UU) {                                        \
    allow_lds(k_tbwd<VV, LL, UU>);                                                              \
    hipLaunchKernelGGL((k_tbwd<VV, LL, UU>), dim3(G), dim3(kTileBlock), lds, st, R->m, R->n,    \
                       R->tile_cols, R->tile_wgs, (const int64_t*)R->tptr,                      \
                       (const uint16_t*)R->tidx, (const typename Val<VV>::T*)R->tval,           \
                       (const int*)R->trow, e, R->cpart);                                       \
    launched = true;                                                                            \
  }
#define RAPTR_TBV(VV)                                                                           \
  RAPTR_TB(VV, 4, 1) RAPTR_TB(VV, 4, 2) RAPTR_TB(VV, 4, 4) RAPTR_TB(VV, 8, 1) RAPTR_TB(VV, 8, 2) \
  RAPTR_TB(VV, 8, 4) RAPTR_TB(VV, 16, 1) RAPTR_TB(VV, 16, 2) RAPTR_TB(VV, 16, 4)
  RAPTR_TBV(kF32) RAPTR_TBV(kF16) RAPTR_TBV(kBF16) RAPTR_TBV(kU16)
#undef RAPTR_TBV
#undef RAPTR_TB
  if (!launched) return fail(RAPTR_ERR_INVALID_ARG, "tiled backward shape unsupported");
  RAPTR_LAUNCH_CHECK();
  const int blocks = (R->n + 255) / 256;
  hipLaunchKernelGGL(k_tupdate, dim3(blocks), dim3(256), 0, st, R->n, R->tile_cols, R->tile_wgs,
                     (const float*)R->cpart, x, (const float*)R->s, R->xx, R->partials);
  RAPTR_LAUNCH_CHECK();
  if (nb) *nb = blocks;
  return RAPTR_OK;
}

}  // namespace raptr
