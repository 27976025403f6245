// spmv.hip -- CSR sparse matrix-vector products on gfx950 (no MFMA: the
// product is bandwidth-bound).  Rows are processed by groups of LANES lanes
// of a wave64 (LANES chosen from the mean row length), loads of values and
// column indices are coalesced across the group, the vector is gathered
// through the L2, partial sums are reduced with wave shuffles.
#include "primitives.hpp"
#include "spmv.hpp"

namespace raptr {
namespace {

template <class T, class PtrT, class AccT, class OutT, int LANES>
__global__ __launch_bounds__(256) void k_csr_rows(int64_t m, const PtrT* rp, const int* col,
                                                  const T* val, const T* x, T alpha, T beta,
                                                  OutT* y) {
  const int g = threadIdx.x % LANES;
  const int64_t groups = int64_t(gridDim.x) * (blockDim.x / LANES);
  for (int64_t i = int64_t(blockIdx.x) * (blockDim.x / LANES) + threadIdx.x / LANES; i < m;
       i += groups) {
    PtrT b = rp[i], e = rp[i + 1];
    AccT acc = 0;
    for (PtrT k = b + g; k < e; k += LANES) acc += AccT(val[k]) * AccT(x[col[k]]);
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    if (g == 0) {
      if (beta == T(0))
        y[i] = OutT(AccT(alpha) * acc);
      else
        y[i] = OutT(AccT(alpha) * acc + AccT(beta) * AccT(y[i]));
    }
  }
}

template <class T>
__global__ void k_scale_vec(int64_t n, T beta, T* y) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = beta == T(0) ? T(0) : beta * y[i];
}

template <class T>
__global__ void k_csr_scatter_t(int m, const int* rp, const int* col, const T* val,
                                const T* x, T alpha, T* y) {
  int64_t wid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x & (kWave - 1);
  int64_t nw = int64_t(gridDim.x) * blockDim.x / kWave;
  for (int64_t i = wid; i < m; i += nw) {
    T xi = alpha * x[i];
    for (int k = rp[i] + lane; k < rp[i + 1]; k += kWave) atomicAdd(&y[col[k]], val[k] * xi);
  }
}

template <class T, class PtrT, class AccT, class OutT>
int rows_dispatch(int64_t m, double mean, const PtrT* rp, const int* col, const T* val,
                  const T* x, T alpha, T beta, OutT* y, hipStream_t st, int lanes) {
  if (lanes <= 0) lanes = pick_lanes(mean);
  unsigned grid = grid_for(m * lanes, 256, 1 << 18);
#define RAPTR_ROWS(L)                                                                    \
  hipLaunchKernelGGL((k_csr_rows<T, PtrT, AccT, OutT, L>), dim3(grid), dim3(256), 0, st, \
                     m, rp, col, val, x, alpha, beta, y)
  switch (lanes) {
    case 4: RAPTR_ROWS(4); break;
    case 8: RAPTR_ROWS(8); break;
    case 16: RAPTR_ROWS(16); break;
    case 32: RAPTR_ROWS(32); break;
    default: RAPTR_ROWS(64); break;
  }
#undef RAPTR_ROWS
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

template <class T>
int csrmv_impl(int op, int m, int n, int64_t nnz, T alpha, const T* val, const int* rp,
               const int* col, const T* x, T beta, T* y, hipStream_t st) {
  if (m < 0 || n < 0 || nnz < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (op != RAPTR_OP_NON_TRANSPOSE && op != RAPTR_OP_TRANSPOSE)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown operation %d", op);
  if (!rp || !x || !y || (nnz > 0 && (!val || !col)))
    return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (op == RAPTR_OP_NON_TRANSPOSE) {
    if (m == 0) return RAPTR_OK;
    double mean = m ? double(nnz) / m : 0.0;
    return rows_dispatch<T, int, T, T>(m, mean, rp, col, val, x, alpha, beta, y, st, 0);
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_scale_vec<T>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st,
                       int64_t(n), beta, y);
    RAPTR_LAUNCH_CHECK();
  }
  if (m > 0 && nnz > 0) {
    hipLaunchKernelGGL(k_csr_scatter_t<T>, dim3(grid_for(int64_t(m) * kWave, 256, 1 << 16)),
                       dim3(256), 0, st, m, rp, col, val, x, alpha, y);
    RAPTR_LAUNCH_CHECK();
  }
  return RAPTR_OK;
}

}  // namespace

int pick_lanes(double mean) {
  if (mean < 6) return 4;
  if (mean < 12) return 8;
  if (mean < 24) return 16;
  if (mean < 48) return 32;
  return 64;
}

int spmv64_f32(int64_t m, const int64_t* rp, const int* col, const float* val,
               const float* x, float* y, int lanes, hipStream_t st) {
  return rows_dispatch<float, int64_t, float, float>(m, 64.0, rp, col, val, x, 1.0f, 0.0f, y,
                                                     st, lanes);
}

}  // namespace raptr

using namespace raptr;

extern "C" {

int raptr_csrmv_f32(int op, int m, int n, int64_t nnz, float alpha, const float* d_val,
                    const int* d_row_ptr, const int* d_col, const float* d_x, float beta,
                    float* d_y, raptr_stream_t stream) {
  return csrmv_impl<float>(op, m, n, nnz, alpha, d_val, d_row_ptr, d_col, d_x, beta, d_y,
                           S(stream));
}

int raptr_csrmv_f64(int op, int m, int n, int64_t nnz, double alpha, const double* d_val,
                    const int* d_row_ptr, const int* d_col, const double* d_x, double beta,
                    double* d_y, raptr_stream_t stream) {
  return csrmv_impl<double>(op, m, n, nnz, alpha, d_val, d_row_ptr, d_col, d_x, beta, d_y,
                            S(stream));
}

int raptr_spmv_csr64(int m, const int64_t* d_row_ptr, const int* d_col, const float* d_val,
                     const float* d_x, void* d_y, int acc64, raptr_stream_t stream) {
  if (m < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative size");
  if (m == 0) return RAPTR_OK;
  if (!d_row_ptr || !d_x || !d_y) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  if (acc64)
    return rows_dispatch<float, int64_t, double, double>(
        m, 64.0, d_row_ptr, d_col, d_val, d_x, 1.0f, 0.0f, static_cast<double*>(d_y),
        S(stream), 64);
  return rows_dispatch<float, int64_t, float, float>(m, 64.0, d_row_ptr, d_col, d_val, d_x,
                                                     1.0f, 0.0f, static_cast<float*>(d_y),
                                                     S(stream), 64);
}

}  // extern "C"
