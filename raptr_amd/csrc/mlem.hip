// mlem.hip -- MLEM elementwise operations and the fused plain MLEM iteration
// on gfx950.
//
// Elementwise ops restate mlemOperations.hpp:44-169.  The fused iteration
// (reco.cu:230-303) runs per iteration:
//   forward   e_i  = y_i / sum_k A_ik x_k          (ratio fused into the row sum)
//   backward  xx_j = x_j * (sum_k At_jk e_k) / (s_j + (s_j == 0)),
//             FP64 partial sums of xx per block (fixed grid, fixed order)
//   norm      scale = float(1 / sum(xx)) in FP64 from the partials
//   scale     x_j = scale * xx_j
// The back-projection reads a stored transpose (gather, no atomics), so the
// result does not depend on scheduling.
#include "primitives.hpp"
#include "spmv.hpp"

namespace raptr {
namespace {

constexpr int kEwBlock = 256;
constexpr int kPartials = 2048;  // max blocks of the backward kernel

__global__ void k_divides(float* r, const float* a, const float* b, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    r[i] = a[i] / b[i];
}

__global__ void k_divides_multiplies(float* r, const float* a, const float* b, const float* c,
                                     int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float d = c[i];
    d += float(d == 0.0f);
    r[i] = a[i] * b[i] / d;
  }
}

__global__ void k_scales(float* x, float alpha, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = alpha * x[i];
}

__device__ inline double block_sum_f64(double v) {
  __shared__ double lds[kEwBlock / kWave];
  v = wave_reduce_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kEwBlock / kWave; ++w) t += lds[w];
  __syncthreads();
  return t;  // valid in thread 0
}

template <class T>
__global__ __launch_bounds__(kEwBlock) void k_partial_sums(const T* x, int64_t n,
                                                           double* partials) {
  double acc = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    acc += double(x[i]);
  double t = block_sum_f64(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Single block: total of nb partials; optional float scale = 1/total.
__global__ __launch_bounds__(kEwBlock) void k_final_sum(const double* partials, int nb,
                                                        double* total, float* scale) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kEwBlock) acc += partials[i];
  double t = block_sum_f64(acc);
  if (threadIdx.x == 0) {
    *total = t;
    if (scale) *scale = float(1.0 / t);
  }
}

// FP64 sum of every row of a CSR (one wave per row, fixed order).
__global__ void k_row_sums_f64(int64_t n, const int64_t* rp, const float* v, double* out) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t j = wid; j < n; j += nw) {
    double acc = 0.0;
    for (int64_t k = rp[j] + lane; k < rp[j + 1]; k += kWave) acc += double(v[k]);
    acc = wave_reduce_sum(acc);
    if (lane == 0) out[j] = acc;
  }
}

// s_j = float(sens_j / total)
__global__ void k_sens_norm(const double* sens, const double* total, int n, float* s) {
  double tt = *total;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s[i] = float(sens[i] / tt);
}

// ---------------------------------------------------------------------
// Fused iteration kernels.  LANES lanes of a wave64 share one row (column of
// A for the backward pass); each lane keeps UNROLL independent value/index
// loads in flight per step (the matrix stream is latency-bound at one load
// pair per lane), the matrix is read with non-temporal loads so that it
// does not evict the gathered vector from L2, and partial sums are combined
// with wave shuffles in a fixed order (deterministic results).
template <int LANES, int UNROLL>
__device__ inline float row_dot(int64_t b, int64_t en, int g, const int* __restrict__ col,
                                const float* __restrict__ val, const float* __restrict__ x) {
  float acc = 0.0f;
  int64_t k = b + g;
  constexpr int64_t kStep = int64_t(LANES) * UNROLL;
  for (; k + int64_t(UNROLL - 1) * LANES < en; k += kStep) {
    float v[UNROLL];
    int c[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      v[u] = __builtin_nontemporal_load(val + k + u * LANES);
      c[u] = __builtin_nontemporal_load(col + k + u * LANES);
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) acc = fmaf(v[u], x[c[u]], acc);
  }
  for (; k < en; k += LANES) acc = fmaf(val[k], x[col[k]], acc);
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  return acc;
}

// Block index -> logical block such that each XCD (blocks are dealt to the
// 8 XCDs round robin) works on one contiguous range of logical blocks, so
// neighbouring columns share their XCD's L2.
__device__ inline int xcd_block(int b, int G) {
  constexpr int kXcd = 8;
  int x = b % kXcd, q = G / kXcd, r = G % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

template <int LANES, int UNROLL>
__global__ __launch_bounds__(256) void k_mlem_fwd(int64_t m, const int64_t* __restrict__ rp,
                                                  const int* __restrict__ col,
                                                  const float* __restrict__ val,
                                                  const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  float* __restrict__ e) {
  const int g = threadIdx.x % LANES;
  const int64_t i = int64_t(blockIdx.x) * (256 / LANES) + threadIdx.x / LANES;
  if (i >= m) return;
  float acc = row_dot<LANES, UNROLL>(rp[i], rp[i + 1], g, col, val, x);
  if (g == 0) e[i] = acc != 0.0f ? y[i] / acc : 0.0f;
}

template <int LANES, int UNROLL>
__global__ __launch_bounds__(256) void k_mlem_bwd(int64_t n, const int64_t* __restrict__ trp,
                                                  const int* __restrict__ tcol,
                                                  const float* __restrict__ tval,
                                                  const float* __restrict__ e,
                                                  const float* __restrict__ x,
                                                  const float* __restrict__ s,
                                                  float* __restrict__ xx,
                                                  double* __restrict__ partials) {
  const int g = threadIdx.x % LANES;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t j = int64_t(lb) * (256 / LANES) + threadIdx.x / LANES;
  double part = 0.0;
  if (j < n) {
    float acc = row_dot<LANES, UNROLL>(trp[j], trp[j + 1], g, tcol, tval, e);
    if (g == 0) {
      float d = s[j];
      d += float(d == 0.0f);
      float v = x[j] * acc / d;
      xx[j] = v;
      part = double(v);
    }
  }
  double t = block_sum_f64(part);
  if (threadIdx.x == 0) partials[lb] = t;
}

__global__ void k_apply_scale(int n, const float* xx, const float* scale, float* x) {
  float sc = *scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = sc * xx[i];
}

struct MlemBufs {
  float* e;
  float* xx;
  float* s;
  double* partials;
  double* scal;  // [0] total sens, [1] norm
  float* scale;
};


// Tuning configuration code: lanes * 16 + unroll (lanes 8..64, unroll 1..8).
inline int cfg_lanes(int cfg) { return cfg >> 4; }
inline int cfg_unroll(int cfg) { return cfg & 15; }
constexpr int kCfgs[] = {8 * 16 + 4, 16 * 16 + 2, 16 * 16 + 4, 32 * 16 + 2, 32 * 16 + 4,
                         32 * 16 + 8, 64 * 16 + 1, 64 * 16 + 2, 64 * 16 + 4, 64 * 16 + 8};

int default_cfg(double mean) {
  if (mean < 24) return 8 * 16 + 4;
  if (mean < 96) return 32 * 16 + 2;
  return 64 * 16 + 4;
}

inline unsigned blocks_for(int64_t rows, int lanes) {
  return unsigned((rows + (256 / lanes) - 1) / (256 / lanes));
}

int launch_fwd(int cfg, hipStream_t st, int64_t m, const int64_t* rp, const int* col,
               const float* val, const float* x, const float* y, float* e) {
  if (m <= 0) return RAPTR_OK;
  int L = cfg_lanes(cfg), U = cfg_unroll(cfg);
  unsigned grid = blocks_for(m, L);
#define RAPTR_FWD(LL, UU)                                                                  \
  if (L == LL && U == UU) {                                                                \
    hipLaunchKernelGGL((k_mlem_fwd<LL, UU>), dim3(grid), dim3(256), 0, st, m, rp, col, val, \
                       x, y, e);                                                           \
    RAPTR_LAUNCH_CHECK();                                                                  \
    return RAPTR_OK;                                                                       \
  }
  RAPTR_FWD(8, 4) RAPTR_FWD(16, 2) RAPTR_FWD(16, 4) RAPTR_FWD(32, 2) RAPTR_FWD(32, 4)
  RAPTR_FWD(32, 8) RAPTR_FWD(64, 1) RAPTR_FWD(64, 2) RAPTR_FWD(64, 4) RAPTR_FWD(64, 8)
#undef RAPTR_FWD
  return fail(RAPTR_ERR_INVALID_ARG, "unsupported forward configuration %d", cfg);
}

int launch_bwd(int cfg, hipStream_t st, int64_t n, const int64_t* trp, const int* tcol,
               const float* tval, const float* e, const float* x, const float* s, float* xx,
               double* partials, int* nblocks) {
  int L = cfg_lanes(cfg), U = cfg_unroll(cfg);
  unsigned grid = blocks_for(n, L);
  if (nblocks) *nblocks = int(grid);
#define RAPTR_BWD(LL, UU)                                                                   \
  if (L == LL && U == UU) {                                                                 \
    hipLaunchKernelGGL((k_mlem_bwd<LL, UU>), dim3(grid), dim3(256), 0, st, n, trp, tcol,    \
                       tval, e, x, s, xx, partials);                                        \
    RAPTR_LAUNCH_CHECK();                                                                   \
    return RAPTR_OK;                                                                        \
  }
  RAPTR_BWD(8, 4) RAPTR_BWD(16, 2) RAPTR_BWD(16, 4) RAPTR_BWD(32, 2) RAPTR_BWD(32, 4)
  RAPTR_BWD(32, 8) RAPTR_BWD(64, 1) RAPTR_BWD(64, 2) RAPTR_BWD(64, 4) RAPTR_BWD(64, 8)
#undef RAPTR_BWD
  return fail(RAPTR_ERR_INVALID_ARG, "unsupported backward configuration %d", cfg);
}

// partial-sum slots of the backward kernel: one per block at the smallest
// block width (lanes = 64 -> 4 columns per block)
inline int64_t partial_slots(int n) { return int64_t(n) / 4 + 8; }

template <class C>
void carve_mlem(C& c, int m, int n) {
  c.template take<float>(size_t(m > 0 ? m : 1));
  c.template take<float>(size_t(n));
  c.template take<float>(size_t(n));
  c.template take<double>(size_t(partial_slots(n) > kPartials ? partial_slots(n) : kPartials));
  c.template take<double>(4);
  c.template take<float>(4);
}

bool known_cfg(int cfg) {
  for (int c : kCfgs)
    if (c == cfg) return true;
  return false;
}

}  // namespace
}  // namespace raptr

using namespace raptr;

extern "C" {

int raptr_divides(float* r, const float* a, const float* b, int n, raptr_stream_t stream) {
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  if (n == 0) return RAPTR_OK;
  if (!r || !a || !b) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_divides, dim3(grid_for(n, kEwBlock, 8192)), dim3(kEwBlock), 0,
                     S(stream), r, a, b, n);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_divides_multiplies(float* r, const float* a, const float* b, const float* c, int n,
                             raptr_stream_t stream) {
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  if (n == 0) return RAPTR_OK;
  if (!r || !a || !b || !c) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_divides_multiplies, dim3(grid_for(n, kEwBlock, 8192)), dim3(kEwBlock),
                     0, S(stream), r, a, b, c, n);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_scales(float* x, float alpha, int n, raptr_stream_t stream) {
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  if (n == 0) return RAPTR_OK;
  if (!x) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_scales, dim3(grid_for(n, kEwBlock, 8192)), dim3(kEwBlock), 0,
                     S(stream), x, alpha, n);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

size_t raptr_sum_workspace_size(int n) {
  (void)n;
  Sizer z;
  z.take<double>(kPartials);
  z.take<double>(2);
  return z.bytes();
}

int raptr_sum(const float* d_x, int n, float* h_sum, void* d_ws, size_t ws_bytes,
              raptr_stream_t stream) {
  if (!h_sum) return fail(RAPTR_ERR_INVALID_ARG, "null result");
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  Carver cv(d_ws, ws_bytes);
  double* partials = cv.take<double>(kPartials);
  double* tot = cv.take<double>(2);
  if (!cv.ok) return fail(RAPTR_ERR_WORKSPACE, "sum workspace too small");
  hipStream_t st = S(stream);
  unsigned nb = grid_for(n, kEwBlock, kPartials);
  hipLaunchKernelGGL(k_partial_sums<float>, dim3(nb), dim3(kEwBlock), 0, st, d_x, int64_t(n),
                     partials);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kEwBlock), 0, st, (const double*)partials,
                     int(nb), tot, (float*)nullptr);
  RAPTR_LAUNCH_CHECK();
  double h = 0.0;
  RAPTR_HIP_TRY(hipMemcpyAsync(&h, tot, sizeof(double), hipMemcpyDeviceToHost, st));
  RAPTR_HIP_TRY(hipStreamSynchronize(st));
  *h_sum = float(h);
  return RAPTR_OK;
}

int raptr_column_sums_f64(int n, const int64_t* d_t_row_ptr, const float* d_t_val,
                          double* d_out, raptr_stream_t stream) {
  if (n < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (n == 0) return RAPTR_OK;
  if (!d_t_row_ptr || !d_t_val || !d_out) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_row_sums_f64, dim3(grid_for(int64_t(n) * kWave, 256, 1 << 16)),
                     dim3(256), 0, S(stream), int64_t(n), d_t_row_ptr, d_t_val, d_out);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_mlem_forward_ratio(int m, const int64_t* d_rp, const int* d_col, const float* d_val,
                             const float* d_x, const float* d_y, float* d_e, int cfg,
                             raptr_stream_t stream) {
  if (m < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (m == 0) return RAPTR_OK;
  if (!d_rp || !d_x || !d_y || !d_e) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (cfg == 0) cfg = default_cfg(64.0);
  if (!known_cfg(cfg)) return fail(RAPTR_ERR_INVALID_ARG, "unsupported configuration %d", cfg);
  return launch_fwd(cfg, S(stream), m, d_rp, d_col, d_val, d_x, d_y, d_e);
}

int raptr_mlem_backward_update(int n, const int64_t* d_trp, const int* d_tcol,
                               const float* d_tval, const float* d_e, const float* d_x,
                               const float* d_s, float* d_xx, double* d_partials, int cfg,
                               raptr_stream_t stream) {
  if (n <= 0) return fail(RAPTR_ERR_INVALID_ARG, "bad size");
  if (!d_trp || !d_e || !d_x || !d_s || !d_xx || !d_partials)
    return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (cfg == 0) cfg = default_cfg(64.0);
  if (!known_cfg(cfg)) return fail(RAPTR_ERR_INVALID_ARG, "unsupported configuration %d", cfg);
  return launch_bwd(cfg, S(stream), n, d_trp, d_tcol, d_tval, d_e, d_x, d_s, d_xx, d_partials,
                    nullptr);
}

int raptr_mlem_partial_slots(int n) { return int(partial_slots(n)); }

size_t raptr_mlem_workspace_size(int m, int n) {
  Sizer z;
  carve_mlem(z, m, n);
  return z.bytes();
}

int raptr_mlem(int m, int n, const int64_t* d_rp, const int* d_col, const float* d_val,
               const int64_t* d_trp, const int* d_tcol, const float* d_tval, const float* d_y,
               const double* d_sens, float* d_x, int niter, int use_graph, int fwd_cfg,
               int bwd_cfg, void* d_ws, size_t ws_bytes, raptr_stream_t stream) {
  if (m < 0 || n <= 0 || niter < 0) return fail(RAPTR_ERR_INVALID_ARG, "bad sizes");
  if (!d_rp || !d_trp || !d_y || !d_sens || !d_x)
    return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  Carver cv(d_ws, ws_bytes);
  MlemBufs B;
  B.e = cv.take<float>(size_t(m > 0 ? m : 1));
  B.xx = cv.take<float>(size_t(n));
  B.s = cv.take<float>(size_t(n));
  B.partials = cv.take<double>(
      size_t(partial_slots(n) > kPartials ? partial_slots(n) : kPartials));
  B.scal = cv.take<double>(4);
  B.scale = cv.take<float>(4);
  if (!cv.ok)
    return fail(RAPTR_ERR_WORKSPACE, "mlem workspace too small (%zu < %zu)", ws_bytes,
                raptr_mlem_workspace_size(m, n));
  hipStream_t st = S(stream);
  // host-side shape facts for lane selection
  int64_t nnz = 0, tnnz = 0;
  RAPTR_HIP_TRY(hipMemcpyAsync(&nnz, d_rp + m, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RAPTR_HIP_TRY(hipMemcpyAsync(&tnnz, d_trp + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RAPTR_HIP_TRY(hipStreamSynchronize(st));
  if (nnz != tnnz) return fail(RAPTR_ERR_INVALID_ARG, "matrix and transpose differ in nnz");
  if (fwd_cfg == 0) fwd_cfg = default_cfg(m ? double(nnz) / m : 0.0);
  if (bwd_cfg == 0) bwd_cfg = default_cfg(double(nnz) / n);
  if (!known_cfg(fwd_cfg) || !known_cfg(bwd_cfg))
    return fail(RAPTR_ERR_INVALID_ARG, "unsupported configuration %d/%d", fwd_cfg, bwd_cfg);
  // normalised FP32 sensitivity
  unsigned ns = grid_for(n, kEwBlock, kPartials);
  hipLaunchKernelGGL(k_partial_sums<double>, dim3(ns), dim3(kEwBlock), 0, st, d_sens,
                     int64_t(n), B.partials);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kEwBlock), 0, st, (const double*)B.partials,
                     int(ns), &B.scal[0], (float*)nullptr);
  RAPTR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sens_norm, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, d_sens,
                     (const double*)&B.scal[0], n, B.s);
  RAPTR_LAUNCH_CHECK();

  auto iteration = [&](hipStream_t s2) -> int {
    RAPTR_TRY(launch_fwd(fwd_cfg, s2, m, d_rp, d_col, d_val, d_x, d_y, B.e));
    int gb = 0;
    RAPTR_TRY(launch_bwd(bwd_cfg, s2, n, d_trp, d_tcol, d_tval, B.e, d_x, B.s, B.xx,
                         B.partials, &gb));
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kEwBlock), 0, s2, (const double*)B.partials,
                       gb, &B.scal[1], B.scale);
    RAPTR_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_apply_scale, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s2, n,
                       (const float*)B.xx, (const float*)B.scale, d_x);
    RAPTR_LAUNCH_CHECK();
    return RAPTR_OK;
  };

  if (!use_graph || niter < 2) {
    for (int it = 0; it < niter; ++it) RAPTR_TRY(iteration(st));
    return RAPTR_OK;
  }
  // capture one iteration on a private stream ordered after the caller's work
  hipStream_t cs;
  RAPTR_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipEvent_t ev;
  RAPTR_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  int rc = RAPTR_OK;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  do {
    if (hipEventRecord(ev, st) != hipSuccess || hipStreamWaitEvent(cs, ev, 0) != hipSuccess) {
      rc = fail(RAPTR_ERR_DEVICE, "stream ordering failed");
      break;
    }
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = fail(RAPTR_ERR_DEVICE, "begin capture failed");
      break;
    }
    int irc = iteration(cs);
    hipError_t ce = hipStreamEndCapture(cs, &graph);
    if (irc != RAPTR_OK) {
      rc = irc;
      break;
    }
    if (ce != hipSuccess) {
      rc = fail(RAPTR_ERR_DEVICE, "end capture: %s", hipGetErrorString(ce));
      break;
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      rc = fail(RAPTR_ERR_DEVICE, "instantiate: %s", hipGetErrorString(ie));
      break;
    }
    for (int it = 0; it < niter; ++it) {
      hipError_t le = hipGraphLaunch(exec, cs);
      if (le != hipSuccess) {
        rc = fail(RAPTR_ERR_DEVICE, "graph launch: %s", hipGetErrorString(le));
        break;
      }
    }
    if (rc != RAPTR_OK) break;
    if (hipEventRecord(ev, cs) != hipSuccess || hipStreamWaitEvent(st, ev, 0) != hipSuccess)
      rc = fail(RAPTR_ERR_DEVICE, "stream ordering failed");
  } while (0);
  (void)hipStreamSynchronize(cs);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(cs);
  return rc;
}

int raptr_mlem_autotune(int m, int n, const int64_t* d_rp, const int* d_col,
                        const float* d_val, const int64_t* d_trp, const int* d_tcol,
                        const float* d_tval, const float* d_y, const float* d_x, int reps,
                        int* fwd_cfg, int* bwd_cfg, float* fwd_ms, float* bwd_ms, void* d_ws,
                        size_t ws_bytes, raptr_stream_t stream) {
  if (!fwd_cfg || !bwd_cfg) return fail(RAPTR_ERR_INVALID_ARG, "null result");
  if (reps <= 0) reps = 5;
  Carver cv(d_ws, ws_bytes);
  float* e = cv.take<float>(size_t(m > 0 ? m : 1));
  float* xx = cv.take<float>(size_t(n));
  float* s = cv.take<float>(size_t(n));
  double* part = cv.take<double>(
      size_t(partial_slots(n) > kPartials ? partial_slots(n) : kPartials));
  if (!cv.ok) return fail(RAPTR_ERR_WORKSPACE, "autotune workspace too small");
  hipStream_t st = S(stream);
  RAPTR_HIP_TRY(hipMemsetAsync(s, 0, size_t(n) * sizeof(float), st));
  hipEvent_t e0, e1;
  RAPTR_HIP_TRY(hipEventCreate(&e0));
  RAPTR_HIP_TRY(hipEventCreate(&e1));
  float best_f = 1e30f, best_b = 1e30f;
  int bf = 0, bb = 0, rc = RAPTR_OK;
  for (int cfg : kCfgs) {
    if (m > 0) {
      rc = launch_fwd(cfg, st, m, d_rp, d_col, d_val, d_x, d_y, e);  // warm
      if (rc) break;
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps && !rc; ++r)
        rc = launch_fwd(cfg, st, m, d_rp, d_col, d_val, d_x, d_y, e);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= float(reps);
      if (ms < best_f) best_f = ms, bf = cfg;
    }
    rc = launch_bwd(cfg, st, n, d_trp, d_tcol, d_tval, e, d_x, s, xx, part, nullptr);
    if (rc) break;
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && !rc; ++r)
      rc = launch_bwd(cfg, st, n, d_trp, d_tcol, d_tval, e, d_x, s, xx, part, nullptr);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= float(reps);
    if (ms < best_b) best_b = ms, bb = cfg;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *fwd_cfg = bf ? bf : default_cfg(64.0);
  *bwd_cfg = bb;
  if (fwd_ms) *fwd_ms = bf ? best_f : 0.0f;
  if (bwd_ms) *bwd_ms = best_b;
  return RAPTR_OK;
}

}  // extern "C"
