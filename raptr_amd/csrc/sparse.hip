// sparse.hip -- sparse-format kernels on gfx950: own stable LSD radix sort
// (8-bit digits; per-tile histograms, device-wide scan, stable scatter with
// wave-ballot peer ranking), COO sort, COO->CSR, CSR->ECSR, CSR transpose.
#include "primitives.hpp"

namespace raptr {
namespace {

constexpr int kRsBlock = 256;
constexpr int kRsRounds = 8;
constexpr int kRsTile = kRsBlock * kRsRounds;
constexpr int kRsBins = 256;
constexpr int kRsWaves = kRsBlock / kWave;

__global__ __launch_bounds__(kRsBlock) void k_rs_hist(const uint64_t* keys, int64_t n,
                                                      int shift, uint32_t* hist,
                                                      int ntiles) {
  __shared__ uint32_t h[kRsBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = int64_t(blockIdx.x) * kRsTile;
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    int64_t p = base + r * kRsBlock + threadIdx.x;
    if (p < n) atomicAdd(&h[(keys[p] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  hist[int64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile: rank = running count of the digit in earlier
// rounds + counts of the digit in earlier waves of this round + rank among
// equal-digit lanes of this wave (found with 8 ballots).
__global__ __launch_bounds__(kRsBlock) void k_rs_scatter(const uint64_t* kin,
                                                         const uint32_t* vin, int64_t n,
                                                         int shift, const uint32_t* offs,
                                                         int ntiles, uint64_t* kout,
                                                         uint32_t* vout) {
  __shared__ uint32_t run[kRsBins];
  __shared__ uint32_t wc[kRsWaves][kRsBins];
  const int t = threadIdx.x, w = t / kWave;
  run[t] = offs[int64_t(t) * ntiles + blockIdx.x];
  int64_t base = int64_t(blockIdx.x) * kRsTile;
  for (int r = 0; r < kRsRounds; ++r) {
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ++ww) wc[ww][t] = 0;
    __syncthreads();
    int64_t p = base + r * kRsBlock + t;
    bool valid = p < n;
    uint64_t key = valid ? kin[p] : 0;
    uint32_t val = valid ? vin[p] : 0;
    uint32_t digit = uint32_t(key >> shift) & 0xff;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (digit >> b) & 1;
      uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    uint64_t lower = m & lanemask_lt();
    if (valid && lower == 0) wc[w][digit] = uint32_t(__popcll(m));
    __syncthreads();
    if (valid) {
      uint32_t dst = run[digit] + uint32_t(__popcll(lower));
      for (int ww = 0; ww < w; ++ww) dst += wc[ww][digit];
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ++ww) add += wc[ww][t];
    run[t] += add;
    __syncthreads();
  }
}

struct SortBufs {
  uint64_t* k[2];
  uint32_t* v[2];
  uint32_t* hist;
  uint32_t* offs;
  void* scan_ws;
  size_t scan_bytes;
};

template <class C>
void carve_sort(C& c, int64_t n) {
  int64_t ntiles = (n + kRsTile - 1) / kRsTile;
  c.template take<uint64_t>(size_t(n));
  c.template take<uint64_t>(size_t(n));
  c.template take<uint32_t>(size_t(n));
  c.template take<uint32_t>(size_t(n));
  c.template take<uint32_t>(size_t(ntiles * kRsBins));
  c.template take<uint32_t>(size_t(ntiles * kRsBins));
  c.template take<char>(scan_workspace_bytes(ntiles * kRsBins));
}

bool take_sort(Carver& c, int64_t n, SortBufs* b) {
  int64_t ntiles = (n + kRsTile - 1) / kRsTile;
  b->k[0] = c.take<uint64_t>(size_t(n));
  b->k[1] = c.take<uint64_t>(size_t(n));
  b->v[0] = c.take<uint32_t>(size_t(n));
  b->v[1] = c.take<uint32_t>(size_t(n));
  b->hist = c.take<uint32_t>(size_t(ntiles * kRsBins));
  b->offs = c.take<uint32_t>(size_t(ntiles * kRsBins));
  b->scan_bytes = scan_workspace_bytes(ntiles * kRsBins);
  b->scan_ws = c.take<char>(b->scan_bytes);
  return c.ok;
}

// Sorts (k[0], v[0]) by bits [0, bits) of the key; returns the buffer index
// holding the result.
int radix_sort(SortBufs& b, int64_t n, int bits, hipStream_t st, int* result) {
  int cur = 0;
  if (n > int64_t(UINT32_MAX)) return fail(RAPTR_ERR_INVALID_ARG, "sort length exceeds 2^32");
  int64_t ntiles = (n + kRsTile - 1) / kRsTile;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(unsigned(ntiles)), dim3(kRsBlock), 0, st,
                       (const uint64_t*)b.k[cur], n, shift, b.hist, int(ntiles));
    RAPTR_LAUNCH_CHECK();
    RAPTR_TRY(scan_exclusive_u32(b.hist, b.offs, ntiles * kRsBins, nullptr, b.scan_ws,
                                 b.scan_bytes, st));
    hipLaunchKernelGGL(k_rs_scatter, dim3(unsigned(ntiles)), dim3(kRsBlock), 0, st,
                       (const uint64_t*)b.k[cur], (const uint32_t*)b.v[cur], n, shift,
                       (const uint32_t*)b.offs, int(ntiles), b.k[cur ^ 1], b.v[cur ^ 1]);
    RAPTR_LAUNCH_CHECK();
    cur ^= 1;
  }
  *result = cur;
  return RAPTR_OK;
}

int bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// ----------------------------------------------------------- COO sort
__global__ void k_minmax(const int* row, const int* col, int64_t n, int* out) {
  int rmax = 0, cmax = 0, mn = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int r = row[i], c = col[i];
    rmax = r > rmax ? r : rmax;
    cmax = c > cmax ? c : cmax;
    int m = r < c ? r : c;
    mn = m < mn ? m : mn;
  }
  atomicMax(&out[0], rmax);
  atomicMax(&out[1], cmax);
  atomicMin(&out[2], mn);
}

__global__ void k_coo_keys(const int* row, const int* col, int64_t n, int colbits,
                           const float* val, uint64_t* keys, uint32_t* idx, float* vcopy) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    keys[i] = (uint64_t(uint32_t(row[i])) << colbits) | uint64_t(uint32_t(col[i]));
    idx[i] = uint32_t(i);
    vcopy[i] = val[i];
  }
}

__global__ void k_coo_unpack(const uint64_t* keys, const uint32_t* idx, int64_t n,
                             int colbits, const float* vcopy, int* row, int* col,
                             float* val) {
  uint64_t mask = (uint64_t(1) << colbits) - 1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint64_t k = keys[i];
    row[i] = int(k >> colbits);
    col[i] = int(k & mask);
    val[i] = vcopy[idx[i]];
  }
}

// ---------------------------------------------------- coo2csr / ecsr
template <class IdxT, class PtrT>
__global__ void k_lower_bound_ptr(const IdxT* sorted, int64_t nnz, int64_t m, PtrT* ptr) {
  for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r <= m;
       r += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (int64_t(sorted[mid]) < r) lo = mid + 1; else hi = mid;
    }
    ptr[r] = PtrT(lo);
  }
}

__global__ void k_csr2ecsr(int* ecsr, const int* y, int nnzr, const int* csr, int m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnzr) ecsr[i] = csr[y[i]];
  if (i == nnzr) ecsr[nnzr] = csr[m];
}

// ----------------------- reference index arrays of a built chunk (int32)
// ecsr[i] = base + row_ptr[i] (i <= m)
__global__ void k_ecsr32(const int64_t* row_ptr, int m, int64_t base, int* ecsr) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i <= m;
       i += int64_t(gridDim.x) * blockDim.x)
    ecsr[i] = int(base + row_ptr[i]);
}

// coo[base + k] = y[i] for the entries k of row i (one wave per row)
__global__ void k_coo_rows_of(const int64_t* row_ptr, const int* y, int m, int64_t base,
                              int* coo) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    const int c = y[i];
    for (int64_t k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += kWave) coo[base + k] = c;
  }
}

// csr[c] = base + row_ptr[#listed channels < c] (c <= nch): the row pointer
// over the whole channel space that coo2csr gives for these entries
__global__ void k_csr_full(const int64_t* row_ptr, const int* y, int m, int64_t base, int nch,
                           int* csr) {
  for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c <= nch;
       c += int64_t(gridDim.x) * blockDim.x) {
    int lo = 0, hi = m;
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if (int64_t(y[mid]) < c) lo = mid + 1; else hi = mid;
    }
    csr[c] = int(base + row_ptr[lo]);
  }
}

// --------------------------------------------------------- transpose
__global__ void k_t_keys(const int* col, int64_t nnz, uint64_t* keys, uint32_t* idx) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nnz;
       i += int64_t(gridDim.x) * blockDim.x) {
    keys[i] = uint64_t(uint32_t(col[i]));
    idx[i] = uint32_t(i);
  }
}

// original row of every entry (one wave per row)
__global__ void k_expand_rows(const int64_t* row_ptr, int m, uint32_t* rowid) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw)
    for (int64_t k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += kWave) rowid[k] = uint32_t(i);
}

__global__ void k_t_gather(const uint64_t* keys, const uint32_t* idx, int64_t nnz,
                           const uint32_t* rowid, const float* val, int* t_col,
                           float* t_val, int* t_key32) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nnz;
       j += int64_t(gridDim.x) * blockDim.x) {
    uint32_t k = idx[j];
    t_col[j] = int(rowid[k]);
    t_val[j] = val[k];
    t_key32[j] = int(keys[j]);
  }
}

}  // namespace

int radix_sort_pairs(uint64_t* k0, uint64_t* k1, uint32_t* v0, uint32_t* v1, int64_t n,
                     int bits, hipStream_t st, int* result, Carver& cv) {
  SortBufs b;
  int64_t ntiles = (n + kRsTile - 1) / kRsTile;
  b.k[0] = k0;
  b.k[1] = k1;
  b.v[0] = v0;
  b.v[1] = v1;
  b.hist = cv.take<uint32_t>(size_t(ntiles * kRsBins));
  b.offs = cv.take<uint32_t>(size_t(ntiles * kRsBins));
  b.scan_bytes = scan_workspace_bytes(ntiles * kRsBins);
  b.scan_ws = cv.take<char>(b.scan_bytes);
  if (!cv.ok) return fail(RAPTR_ERR_WORKSPACE, "radix sort scratch too small");
  return radix_sort(b, n, bits, st, result);
}

}  // namespace raptr

using namespace raptr;

extern "C" {

size_t raptr_coo_sort_workspace_size(int64_t n) {
  if (n < 1) n = 1;
  Sizer z;
  z.take<int>(4);
  z.take<float>(size_t(n));
  carve_sort(z, n);
  return z.bytes();
}

int raptr_coo_sort(float* d_val, int* d_row, int* d_col, int64_t n, void* d_ws,
                   size_t ws_bytes, raptr_stream_t stream) {
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  if (n <= 1) return RAPTR_OK;  // empty or single entry: already sorted
  if (!d_val || !d_row || !d_col) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipStream_t st = S(stream);
  Carver cv(d_ws, ws_bytes);
  int* mm = cv.take<int>(4);
  float* vcopy = cv.take<float>(size_t(n));
  SortBufs b;
  if (!take_sort(cv, n, &b))
    return fail(RAPTR_ERR_WORKSPACE, "coo_sort workspace too small (%zu < %zu)", ws_bytes,
                raptr_coo_sort_workspace_size(n));
  int init[4] = {0, 0, 0, 0};
  RAPTR_HIP_TRY(hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_minmax, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, d_row, d_col,
                     n, mm);
  RAPTR_LAUNCH_CHECK();
  int h[4];
  RAPTR_HIP_TRY(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, st));
  RAPTR_HIP_TRY(hipStreamSynchronize(st));
  if (h[2] < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative row or column index");
  int colbits = bits_for(uint64_t(h[1]));
  int rowbits = bits_for(uint64_t(h[0]));
  if (colbits < 1) colbits = 1;
  hipLaunchKernelGGL(k_coo_keys, dim3(grid_for(n, 256, 1 << 16)), dim3(256), 0, st, d_row,
                     d_col, n, colbits, d_val, b.k[0], b.v[0], vcopy);
  RAPTR_LAUNCH_CHECK();
  int res = 0;
  RAPTR_TRY(radix_sort(b, n, rowbits + colbits, st, &res));
  hipLaunchKernelGGL(k_coo_unpack, dim3(grid_for(n, 256, 1 << 16)), dim3(256), 0, st,
                     (const uint64_t*)b.k[res], (const uint32_t*)b.v[res], n, colbits,
                     (const float*)vcopy, d_row, d_col, d_val);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_coo2csr(int* d_row_ptr, const int* d_row_id, int64_t nnz, int m,
                  raptr_stream_t stream) {
  if (m < 0 || nnz < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (!d_row_ptr || (nnz > 0 && !d_row_id)) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (nnz > INT32_MAX) return fail(RAPTR_ERR_INVALID_ARG, "nnz exceeds int32 row pointers");
  hipLaunchKernelGGL((k_lower_bound_ptr<int, int>), dim3(grid_for(int64_t(m) + 1, 256, 1 << 16)),
                     dim3(256), 0, S(stream), d_row_id, nnz, int64_t(m), d_row_ptr);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_csr2ecsr(int* d_ecsr, const int* d_y, int nnzr, const int* d_csr, int m,
                   raptr_stream_t stream) {
  if (nnzr < 0 || m < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (!d_ecsr || !d_csr || (nnzr > 0 && !d_y)) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_csr2ecsr, dim3(grid_for(int64_t(nnzr) + 1, 256)), dim3(256), 0,
                     S(stream), d_ecsr, d_y, nnzr, d_csr, m);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_sm_index_arrays(const int64_t* d_row_ptr, const int* d_y_row_id, int m, int64_t base,
                          int nchannels, int* d_ecsr, int* d_coo_row, int* d_csr,
                          raptr_stream_t stream) {
  if (m < 0 || base < 0 || nchannels < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (!d_row_ptr || (m > 0 && !d_y_row_id)) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  int64_t last = 0;
  RAPTR_HIP_TRY(hipMemcpyAsync(&last, d_row_ptr + m, sizeof(int64_t), hipMemcpyDeviceToHost,
                               S(stream)));
  RAPTR_HIP_TRY(hipStreamSynchronize(S(stream)));
  if (base + last > INT32_MAX)
    return fail(RAPTR_ERR_INVALID_ARG, "%lld entries exceed int32 row pointers",
                (long long)(base + last));
  if (d_ecsr) {
    hipLaunchKernelGGL(k_ecsr32, dim3(grid_for(int64_t(m) + 1, 256)), dim3(256), 0, S(stream),
                       d_row_ptr, m, base, d_ecsr);
    RAPTR_LAUNCH_CHECK();
  }
  if (d_coo_row && m > 0) {
    hipLaunchKernelGGL(k_coo_rows_of, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                       dim3(256), 0, S(stream), d_row_ptr, d_y_row_id, m, base, d_coo_row);
    RAPTR_LAUNCH_CHECK();
  }
  if (d_csr) {
    hipLaunchKernelGGL(k_csr_full, dim3(grid_for(int64_t(nchannels) + 1, 256, 1 << 16)),
                       dim3(256), 0, S(stream), d_row_ptr, d_y_row_id, m, base, nchannels,
                       d_csr);
    RAPTR_LAUNCH_CHECK();
  }
  return RAPTR_OK;
}

size_t raptr_csr_transpose_workspace_size(int64_t nnz, int m, int n) {
  (void)m;
  (void)n;
  if (nnz < 1) nnz = 1;
  Sizer z;
  z.take<uint32_t>(size_t(nnz));
  z.take<int>(size_t(nnz));
  carve_sort(z, nnz);
  return z.bytes();
}

int raptr_csr_transpose(int m, int n, int64_t nnz, const int64_t* d_row_ptr, const int* d_col,
                        const float* d_val, int64_t* d_t_row_ptr, int* d_t_col,
                        float* d_t_val, void* d_ws, size_t ws_bytes, raptr_stream_t stream) {
  if (m < 0 || n < 0 || nnz < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (!d_row_ptr || !d_t_row_ptr) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipStream_t st = S(stream);
  if (nnz == 0) {
    RAPTR_HIP_TRY(hipMemsetAsync(d_t_row_ptr, 0, sizeof(int64_t) * (size_t(n) + 1), st));
    return RAPTR_OK;
  }
  Carver cv(d_ws, ws_bytes);
  uint32_t* rowid = cv.take<uint32_t>(size_t(nnz));
  int* key32 = cv.take<int>(size_t(nnz));
  SortBufs b;
  if (!take_sort(cv, nnz, &b))
    return fail(RAPTR_ERR_WORKSPACE, "transpose workspace too small (%zu < %zu)", ws_bytes,
                raptr_csr_transpose_workspace_size(nnz, m, n));
  hipLaunchKernelGGL(k_t_keys, dim3(grid_for(nnz, 256, 1 << 16)), dim3(256), 0, st, d_col, nnz,
                     b.k[0], b.v[0]);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                     dim3(256), 0, st, d_row_ptr, m, rowid);
  RAPTR_LAUNCH_CHECK();
  int res = 0;
  RAPTR_TRY(radix_sort(b, nnz, bits_for(uint64_t(n > 0 ? n - 1 : 0)) > 0
                                   ? bits_for(uint64_t(n - 1)) : 1, st, &res));
  hipLaunchKernelGGL(k_t_gather, dim3(grid_for(nnz, 256, 1 << 16)), dim3(256), 0, st,
                     (const uint64_t*)b.k[res], (const uint32_t*)b.v[res], nnz,
                     (const uint32_t*)rowid, d_val, d_t_col, d_t_val, key32);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_lower_bound_ptr<int, int64_t>),
                     dim3(grid_for(int64_t(n) + 1, 256, 1 << 16)), dim3(256), 0, st,
                     (const int*)key32, nnz, int64_t(n), d_t_row_ptr);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

}  // extern "C"
