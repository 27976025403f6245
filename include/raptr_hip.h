/*
 * raptr_hip.h -- C ABI of the MI355X (gfx950) raptr reconstruction hot path.
 *
 * Every entry point is re-entrant: geometry travels as a POD argument (the
 * reference keeps it in process-global __constant__ memory,
 * device_constant_memory.hpp:34-46), work is enqueued on the caller's HIP
 * stream (NULL = the null stream), device buffers belong to the caller, and
 * scratch comes from a caller-provided workspace whose size is queried first.
 * Functions return RAPTR_OK (0) or a status code; raptr_last_error() gives
 * the message of the calling thread's last failure.
 *
 * Numerics: every kernel is compiled without FP contraction and with
 * correctly rounded FP32 divide and square root, in the reference's operation
 * order, so the workqueue and the matrix values are bit-identical to the
 * host restatement of the reference on the same inputs.
 */
#ifndef RAPTR_HIP_H
#define RAPTR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
enum {
  RAPTR_OK = 0,
  RAPTR_ERR_INVALID_ARG = 1,
  RAPTR_ERR_DEVICE = 2,
  RAPTR_ERR_OUT_OF_MEMORY = 3,
  RAPTR_ERR_BUDGET = 4,
  RAPTR_ERR_WORKSPACE = 5,
  RAPTR_ERR_NOT_SUPPORTED = 6
};

/* Ray generators (RayGenerators.hpp:36-99).  RAND: Philox4x32-10 rays per
 * channel, key (seed, 0), counter (channel, ray, 0|1, 0): every voxel of a
 * channel sees the same rays (the per-channel seeding of
 * getSystemMatrixFromWorkqueue.cu:209-210).  REGU: pixel-centre rays.
 * RAND_ELEMENT: rays per (channel, voxel) element, counter (channel, ray,
 * 0|1, voxel + 1): every matrix element draws its own independent rays, the
 * semantics of the fused kernel's per-thread generators
 * (getSystemMatrixDeviceOnly.cu:396), reproducible whatever the launch
 * shape.  RAND_ELEMENT applies to matrix values; ray listings
 * (raptr_generate_rays) take RAND or REGU. */
enum { RAPTR_RAYGEN_RAND = 0, RAPTR_RAYGEN_REGU = 1, RAPTR_RAYGEN_RAND_ELEMENT = 2 };
enum { RAPTR_OP_NON_TRANSPOSE = 0, RAPTR_OP_TRANSPOSE = 1 };
enum { RAPTR_MAX_ANGLES = 360 };

typedef void* raptr_stream_t; /* hipStream_t */

/* DefaultMeasurementSetup<float> (MeasurementSetup.hpp:181-220): head x
 * positions, angle count, pixels per head (z, y), angular step [deg],
 * pixel edge lengths (x = depth, y, z). */
typedef struct raptr_setup_t {
  float pos0x, pos1x;
  int na, n0z, n0y, n1z, n1y;
  float da, segx, segy, segz;
} raptr_setup_t;

/* DefaultVoxelGrid<float> (VoxelGrid.hpp:151-176). */
typedef struct raptr_grid_t {
  float o[3];
  float d[3];
  int n[3];
} raptr_grid_t;

const char* raptr_status_string(int status);
const char* raptr_last_error(void);
int raptr_version(void);

/* ------------------------------------------------- device plumbing */
int raptr_device_count(int* count);
int raptr_set_device(int device);
int raptr_device_name(char* buf, int len);
int raptr_malloc(void** ptr, size_t bytes);
int raptr_free(void* ptr);
int raptr_memcpy_h2d(void* dst, const void* src, size_t bytes, raptr_stream_t s);
int raptr_memcpy_d2h(void* dst, const void* src, size_t bytes, raptr_stream_t s);
int raptr_memcpy_d2d(void* dst, const void* src, size_t bytes, raptr_stream_t s);
int raptr_memset(void* dst, int value, size_t bytes, raptr_stream_t s);
/* dst[i] = src[i], i < n, by a plain streaming kernel of the library (no
 * counterpart in the reference): the known-size dispatch that calibrates
 * bandwidth and the memory-traffic counters. */
int raptr_stream_copy(float* d_dst, const float* d_src, int64_t n, raptr_stream_t s);
int raptr_stream_create(raptr_stream_t* s);
int raptr_stream_destroy(raptr_stream_t s);
int raptr_stream_sync(raptr_stream_t s);
int raptr_device_sync(void);
int raptr_event_create(void** ev);
int raptr_event_destroy(void* ev);
int raptr_event_record(void* ev, raptr_stream_t s);
int raptr_event_elapsed_ms(float* ms, void* start, void* stop);
/* Work queued on s after this call waits for ev (hipStreamWaitEvent): joins
 * two streams' work, e.g. two plans run side by side. */
int raptr_stream_wait_event(raptr_stream_t s, void* ev);
int raptr_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ----------------------------------------------- geometry (host side) */
/* Per-angle sin/cos table shared by every kernel: the float argument
 * pi*ida*da/180 of MeasurementSetupTrafo2CartCoord.hpp:110-111, with the
 * correctly rounded float of its double-precision sin/cos. */
int raptr_angle_table(const raptr_setup_t* setup, float* sin_out, float* cos_out);
/* 0.5*(|voxel edges| + |pixel edges|), getWorkqueue.hpp:74-76. */
float raptr_sum_radii(const raptr_setup_t* setup, const raptr_grid_t* grid);

/* ------------------------------------------------- A6/A7: workqueue */
/* Replaces getWorkqueue (getWorkqueue.hpp:135-191 bounded, :193-243
 * unbounded).  Scans (list entry, voxel) pairs from the cursor
 * (*list_id, *vxl_id) in list-major, voxel-minor order and writes up to n
 * candidate pairs (n < 0: all) to d_cnl/d_vxl; advances the cursor exactly
 * as the reference does.  *n_found receives the count.  d_list is device
 * memory.  Workspace size from raptr_workqueue_workspace_size. */
size_t raptr_workqueue_workspace_size(const raptr_grid_t* grid, int list_size);
int raptr_workqueue(const raptr_setup_t* setup, const raptr_grid_t* grid,
                    const int* d_list, int list_size, int* list_id, int* vxl_id,
                    int64_t n, int* d_cnl, int* d_vxl, int64_t capacity,
                    int64_t* n_found, void* d_workspace, size_t workspace_bytes,
                    raptr_stream_t stream);
/* Candidate count of every list entry over the full voxel range
 * (d_counts[list_size], int64). */
int raptr_workqueue_counts(const raptr_setup_t* setup, const raptr_grid_t* grid,
                           const int* d_list, int list_size, int64_t* d_counts,
                           raptr_stream_t stream);

/* ------------------------------------- A8-A10: matrix from workqueue */
/* Replaces the getSystemMatrixFromWorkqueue kernel
 * (getSystemMatrixFromWorkqueue.cu:116-281): d_val[w] = mean over nrays rays
 * of the slab intersection length of voxel d_vxl[w] with the rays of channel
 * d_cnl[w].  Rays: Philox4x32-10, key (seed, 0), counter (channel, ray, 0|1,
 * 0), uniforms ((x>>8)+1)*2^-24, or pixel-centre rays (REGU). */
int raptr_sm_from_wq(const raptr_setup_t* setup, const raptr_grid_t* grid,
                     const int* d_cnl, const int* d_vxl, float* d_val,
                     int64_t wq_length, int nrays, uint32_t seed, int raygen,
                     raptr_stream_t stream);
/* Ray end points of (channel, ray) pairs, d_rays[6*k]: the RayGen functor
 * (RayGenerators.hpp:36-99) evaluated on the device. */
int raptr_generate_rays(const raptr_setup_t* setup, const int* d_cnl, int n_cnl,
                        int nrays, uint32_t seed, int raygen, float* d_rays,
                        raptr_stream_t stream);

/* ------------------------------------------------- A11: Siddon chords */
/* getChords (ChordsCalc_lowlevel.hpp:194-379) for nrays rays d_rays[6*r]
 * (start xyz, end xyz): per ray up to cap (chord length, voxel index ix, iy,
 * iz) in traversal order at d_chords[r*cap + k], d_voxel_ids[3*(r*cap + k)];
 * d_counts[r] = chords written (min(traversed voxels, cap)).  A ray whose
 * end points are not both outside the grid along its path, or that misses
 * it, gives 0.  cap = nx + ny + nz + 4 always suffices.  Each axis starts at
 * its first plane after the grid entry (the reference starts one plane past
 * the axis' first grid plane and yields negative chords for axes other than
 * the entry axis). */
int raptr_get_chords(const float* d_rays, int nrays, const raptr_grid_t* grid, int cap,
                     float* d_chords, int* d_voxel_ids, int* d_counts,
                     raptr_stream_t stream);

/* ---------------------- A12/A13/A23: fused matrix build, chunked CSR */
/* Builds the matrix of the listed channels (rows in list order) with exact
 * zeros dropped, columns ascending: the result of the reference's
 * systemMatrixCalculation (wrappers.hpp:185-266: getSystemMatrix, cooSort,
 * coo2csr, csr2ecsr) without its zero-valued candidates, computed in
 * channel chunks sized from the workspace.  Two calls: first with
 * d_col == NULL to get d_row_ptr (int64, list_size+1) and *nnz (count
 * mode), or directly with d_col/d_val of `capacity` entries (fill mode;
 * RAPTR_ERR_WORKSPACE if the matrix does not fit).  Values are
 * bit-identical to raptr_sm_from_wq on the same (channel, voxel) pairs. */
size_t raptr_sm_build_workspace_size(const raptr_grid_t* grid, int nrays,
                                     int64_t staging_entries);
int raptr_sm_build_csr(const raptr_setup_t* setup, const raptr_grid_t* grid,
                       const int* d_list, int list_size, int nrays,
                       uint32_t seed, int raygen, int64_t* d_row_ptr,
                       int* d_col, float* d_val, int64_t capacity,
                       int64_t* nnz, void* d_workspace, size_t workspace_bytes,
                       raptr_stream_t stream);
/* raptr_sm_build_csr with build flags.  RAPTR_SM_KEEP_ZEROS: store every
 * candidate that passes the cylinder cull, zero-valued ones included, as
 * the reference's fused kernel writes them (getSystemMatrixDeviceOnly.cu:
 * 507-509); the matrix then holds exactly the workqueue's (channel, voxel)
 * pairs (nnz = workqueue length).  flags = 0 is raptr_sm_build_csr. */
enum { RAPTR_SM_KEEP_ZEROS = 1 };
int raptr_sm_build_csr_ex(const raptr_setup_t* setup, const raptr_grid_t* grid,
                          const int* d_list, int list_size, int nrays,
                          uint32_t seed, int raygen, int flags, int64_t* d_row_ptr,
                          int* d_col, float* d_val, int64_t capacity,
                          int64_t* nnz, void* d_workspace, size_t workspace_bytes,
                          raptr_stream_t stream);
/* raptr_sm_build_csr_ex without a workspace argument and without any
 * device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The build's staging area comes from a
 * 64 MiB array of device memory that the library holds on every device it
 * runs on until the process ends; rows are built in chunks that fit it, and
 * a row with more candidates than it stages (about 7.5 million) is rejected
 * with RAPTR_ERR_WORKSPACE.  Results are bit-identical to
 * raptr_sm_build_csr_ex.  Every call on one device shares that device's
 * array: calls that overlap in time on one device (other host threads or
 * streams) race and must be serialised by the caller. */
int raptr_sm_build_csr_noalloc(const raptr_setup_t* setup, const raptr_grid_t* grid,
                               const int* d_list, int list_size, int nrays, uint32_t seed,
                               int raygen, int flags, int64_t* d_row_ptr, int* d_col,
                               float* d_val, int64_t capacity, int64_t* nnz,
                               raptr_stream_t stream);
/* Sensitivity image s_j = sum over ALL channels of the setup of A_ij, in
 * FP64 (d_sens[nvox]); the rows are generated and reduced on the fly, never
 * stored.  Replaces the sensitivity file the reference reads
 * (reco.cu:171-204).  The column sums are accumulated in 64-bit fixed
 * point (at least 2^-37 resolution at the full-size workload, the float
 * matrix values themselves exact from 2^-14 mm up) with integer atomics, so
 * the result is bit-reproducible whatever the order of the additions. */
int raptr_sensitivity_f64(const raptr_setup_t* setup, const raptr_grid_t* grid,
                          int nrays, uint32_t seed, int raygen, double* d_sens,
                          raptr_stream_t stream);

/* Matrix-free forward projection: d_out[i] = sum_j A_ij x_j in FP64 for the
 * listed channels (d_list == NULL: all channels of the setup, list_size
 * ignored), rows generated on the fly exactly as raptr_sm_build_csr does
 * (the reference's on-the-fly system matrix, reco.cu:246-259). */
/* Summed back-projection c = A^T w of the listed channels with weights
 * d_weights (the measurement), FP64, matrix-free: the quantity the
 * reference's backprojection driver accumulates chunk by chunk with
 * CSRmv(TRANSPOSE, beta = 1) (tests/backprojection.cu:144-158) before it
 * normalises by the sum.  d_out: n voxels, overwritten. */
int raptr_backproject_f64(const raptr_setup_t* setup, const raptr_grid_t* grid,
                          const int* d_list, const float* d_weights, int list_size, int nrays,
                          uint32_t seed, int raygen, double* d_out, raptr_stream_t stream);
int raptr_forward_project_f64(const raptr_setup_t* setup, const raptr_grid_t* grid,
                              const int* d_list, int list_size, int nrays,
                              uint32_t seed, int raygen, const float* d_x,
                              double* d_out, raptr_stream_t stream);

/* --------------------------------------- A14-A16: sparse formats */
/* Replaces cooSort<T> (cooSort.hpp:75-99): in-place stable sort of COO
 * entries by (row, col); values follow.  Indices must be >= 0. */
size_t raptr_coo_sort_workspace_size(int64_t n);
int raptr_coo_sort(float* d_val, int* d_row, int* d_col, int64_t n,
                   void* d_workspace, size_t workspace_bytes,
                   raptr_stream_t stream);
/* Replaces convertCoo2Csr (cusparseWrapper.hpp:53-60): zero-based row
 * pointer (m+1) of row-sorted COO row ids. */
int raptr_coo2csr(int* d_row_ptr, const int* d_row_id, int64_t nnz, int m,
                  raptr_stream_t stream);
/* Replaces convertCsr2Ecsr (convertCsr2Ecsr.hpp:51-64):
 * ecsr[i] = csr[y_row_id[i]] (i < nnzr), ecsr[nnzr] = csr[m]. */
int raptr_csr2ecsr(int* d_ecsr, const int* d_y_row_id, int nnzr,
                   const int* d_csr, int m, raptr_stream_t stream);
/* The index arrays the reference's systemMatrixCalculation leaves
 * (wrappers.hpp:185-266), in its int32 types, for a matrix of the listed
 * channels d_y_row_id[0..m) (ascending) built by raptr_sm_build_csr[_ex]
 * into arrays whose first `base` entries were already taken:
 *   d_ecsr[i] = base + d_row_ptr[i], i <= m (the ECSR row pointers,
 *     convertCsr2Ecsr.hpp:51-64);
 *   d_coo_row[base + k] = channel of entry k (the COO row ids cooSort
 *     leaves, cooSort.hpp:75-99);
 *   d_csr[c] = base + entries of listed channels < c, c <= nchannels (the
 *     CSR row pointer over the channel space, cusparseWrapper.hpp:53-60).
 * Null outputs are skipped; RAPTR_ERR_INVALID_ARG if base + nnz exceeds
 * int32. */
int raptr_sm_index_arrays(const int64_t* d_row_ptr, const int* d_y_row_id, int m, int64_t base,
                          int nchannels, int* d_ecsr, int* d_coo_row, int* d_csr,
                          raptr_stream_t stream);
/* CSR (64-bit row pointers) -> CSR of the transpose, stable: entries of
 * each transposed row keep ascending original-row order. */
size_t raptr_csr_transpose_workspace_size(int64_t nnz, int m, int n);
int raptr_csr_transpose(int m, int n, int64_t nnz, const int64_t* d_row_ptr,
                        const int* d_col, const float* d_val,
                        int64_t* d_t_row_ptr, int* d_t_col, float* d_t_val,
                        void* d_workspace, size_t workspace_bytes,
                        raptr_stream_t stream);

/* ----------------------------------------------------- A17: SpMV */
/* Replaces CSRmv<T> (csrmv.hpp:37-99, cusparse{S,D}csrmv):
 * y = alpha*op(A)*x + beta*y, A m x n zero-based CSR with int32 pointers;
 * beta == 0 does not read y.  The transposed product scatters with atomics
 * (like the legacy library routine); the reconstruction path uses a stored
 * transpose and raptr_spmv_csr64 instead. */
int raptr_csrmv_f32(int op, int m, int n, int64_t nnz, float alpha,
                    const float* d_val, const int* d_row_ptr, const int* d_col,
                    const float* d_x, float beta, float* d_y,
                    raptr_stream_t stream);
int raptr_csrmv_f64(int op, int m, int n, int64_t nnz, double alpha,
                    const double* d_val, const int* d_row_ptr, const int* d_col,
                    const double* d_x, double beta, double* d_y,
                    raptr_stream_t stream);
/* Deterministic y = A x for a CSR with 64-bit row pointers, FP32 values,
 * FP32 (acc64 = 0) or FP64 (acc64 = 1, y is double*) accumulation. */
int raptr_spmv_csr64(int m, const int64_t* d_row_ptr, const int* d_col,
                     const float* d_val, const float* d_x, void* d_y,
                     int acc64, raptr_stream_t stream);

/* ------------------------------------------- A18-A20: MLEM ops */
/* mlemOperations.hpp:44-54, :89-104, :147-153, :164-169. */
int raptr_divides(float* d_res, const float* d_dividend, const float* d_divisor,
                  int n, raptr_stream_t stream);
int raptr_divides_multiplies(float* d_res, const float* d_a, const float* d_b,
                             const float* d_c, int n, raptr_stream_t stream);
int raptr_scales(float* d_x, float alpha, int n, raptr_stream_t stream);
/* Sum (deterministic fixed-order tree with FP64 partials; the FP32 result
 * is written to *h_sum after a stream sync, like thrust::reduce returning
 * to the host).  Workspace from raptr_sum_workspace_size. */
size_t raptr_sum_workspace_size(int n);
int raptr_sum(const float* d_x, int n, float* h_sum, void* d_workspace,
              size_t workspace_bytes, raptr_stream_t stream);

/* ---------------------------------------- A21/A22: fused MLEM */
/* Column sums in FP64 (sensitivity image s_j = sum_i A_ij) of a CSR that
 * stores the transpose (one row per voxel). */
int raptr_column_sums_f64(int n, const int64_t* d_t_row_ptr,
                          const float* d_t_val, double* d_out,
                          raptr_stream_t stream);

/* Plain MLEM (reco.cu:230-303) on a device-resident matrix held both as CSR
 * (rows = measured channels) and as its transpose.  Per iteration:
 * yt = A x, e = y/yt (0 where yt == 0), c = A^T e, x = x*c/(s + (s==0)),
 * x *= 1/sum(x) with the sum in FP64.  s is the raw sensitivity in FP64,
 * Sigma-normalised inside.  d_x holds the start image and receives the
 * result.  Work buffers come from the workspace.  use_graph != 0 captures
 * the iteration into a HIP graph and replays it. */
size_t raptr_mlem_workspace_size(int m, int n);
/* The two fused kernels of one iteration, exposed for drivers and timing:
 * forward + ratio  e_i = y_i / (A x)_i (0 where (A x)_i == 0);
 * backward + update  xx_j = x_j (A^T e)_j / (s_j + (s_j==0)) with FP64
 * partial sums of xx per block in d_partials (raptr_mlem_partial_slots(n)
 * slots).  s is the Sigma-normalised FP32 sensitivity.  (0 =
 * default).  cfg: kernel configuration code (see raptr_mlem). */
int raptr_mlem_forward_ratio(int m, const int64_t* d_row_ptr, const int* d_col,
                             const float* d_val, const float* d_x, const float* d_y,
                             float* d_e, int cfg, raptr_stream_t stream);
int raptr_mlem_backward_update(int n, const int64_t* d_t_row_ptr, const int* d_t_col,
                               const float* d_t_val, const float* d_e, const float* d_x,
                               const float* d_s, float* d_xx, double* d_partials,
                               int cfg, raptr_stream_t stream);
int raptr_mlem(int m, int n, const int64_t* d_row_ptr, const int* d_col,
               const float* d_val, const int64_t* d_t_row_ptr,
               const int* d_t_col, const float* d_t_val, const float* d_y,
               const double* d_sens, float* d_x, int niter, int use_graph,
               int fwd_cfg, int bwd_cfg, void* d_workspace,
               size_t workspace_bytes, raptr_stream_t stream);
/* Kernel configurations are codes lanes*16 + unroll: lanes per row
 * (8, 16, 32, 64) and independent loads in flight per lane (1, 2, 4, 8);
 * 0 picks one from the mean row length.  The autotuner times every
 * supported code on the caller's matrix (reps launches each, HIP events)
 * and returns the fastest forward and backward codes and their times. */
int raptr_mlem_autotune(int m, int n, const int64_t* d_row_ptr, const int* d_col,
                        const float* d_val, const int64_t* d_t_row_ptr,
                        const int* d_t_col, const float* d_t_val, const float* d_y,
                        const float* d_x, int reps, int* fwd_cfg, int* bwd_cfg,
                        float* fwd_ms, float* bwd_ms, void* d_workspace,
                        size_t workspace_bytes, raptr_stream_t stream);
/* Partial-sum slots d_partials must hold for raptr_mlem_backward_update. */
int raptr_mlem_partial_slots(int n);

/* ------------------------- reconstruction plan (plain, batched, mixed) */
/* A device-resident MLEM plan built from the CSR of the measured channels
 * (64-bit row pointers, ascending columns) and the FP64 sensitivity.  The
 * plan owns two derived layouts: a row-padded CSR (16-byte aligned rows,
 * vector loads) for the forward pass and the transpose cut into row blocks
 * of block_rows channels (0: 65,536 for single-frame row-blocked plans,
 * else sized so the block's 2 MB ratio slice fits one XCD's L2) for the
 * atomic-free backward pass.  value_type: 0 = FP32,
 * 1 = FP16, 2 = BF16 (of value/row maximum), 3 = UNORM16 (fixed point
 * round(65535*value/row maximum)); 16-bit types carry a per-row FP32
 * scale and accumulate in FP32.  max_frames: batch capacity (power of two <= 64).  The
 * plan allocates its own device memory (raptr_recon_info reports it). */
typedef struct raptr_recon raptr_recon;
enum { RAPTR_VALUE_F32 = 0, RAPTR_VALUE_F16 = 1, RAPTR_VALUE_BF16 = 2, RAPTR_VALUE_U16 = 3 };
int raptr_recon_create(raptr_recon** plan, int m, int n, int64_t nnz,
                       const int64_t* d_row_ptr, const int* d_col,
                       const float* d_val, const double* d_sens, int value_type,
                       int max_frames, int block_rows, raptr_stream_t stream);
/* Layout choice: RAPTR_LAYOUT_BLOCKED, the two layouts above (any frame
 * count), is the only layout; other values are rejected. */
enum { RAPTR_LAYOUT_BLOCKED = 0 };
int raptr_recon_create_ex(raptr_recon** plan, int m, int n, int64_t nnz,
                          const int64_t* d_row_ptr, const int* d_col,
                          const float* d_val, const double* d_sens, int value_type,
                          int max_frames, int block_rows, int layout,
                          raptr_stream_t stream);
/* Plan options.  d_row_order: plan row p holds the caller's row
 * d_row_order[p] (device; a permutation of 0..m-1, checked; null: the
 * caller's order).  The interface keeps the caller's row order: the plan
 * permutes y on entry to raptr_recon_mlem / raptr_recon_forward and e on
 * the way out, and no row's sum changes.  An order that puts the rows
 * through one voxel next to each other (raptr_locality_order) lets the
 * gathers of both passes touch fewer cache lines.  grid: the voxel grid of
 * the columns (n = nx*ny*nz), so that the backward groups of the 64-frame
 * grouped layouts are bricks of voxels (null: runs of consecutive column
 * ids).  group: members per group of those layouts (16 or 32;
 * 0 = 16).
 * grouped_only (max_frames = 64): build only the grouped layouts (and the
 * row scales), not the padded CSR and padded transpose; such a plan runs
 * 64-frame passes only, and needs about 40 % less device memory (matrices
 * of many channels, e.g. serving lanes over every channel of a geometry).
 * No counterpart in the reference, which keeps the measurement list order
 * and computes per chunk (reco.cu:230-303). */
typedef struct raptr_recon_options {
  int layout;
  const int* d_row_order;
  const raptr_grid_t* grid;
  int group;
  int grouped_only;
} raptr_recon_options;
int raptr_recon_create_opts(raptr_recon** plan, int m, int n, int64_t nnz,
                            const int64_t* d_row_ptr, const int* d_col, const float* d_val,
                            const double* d_sens, int value_type, int max_frames,
                            int block_rows, const raptr_recon_options* options,
                            raptr_stream_t stream);
/* raptr_recon_create_opts with a row order only. */
int raptr_recon_create_ordered(raptr_recon** plan, int m, int n, int64_t nnz,
                               const int64_t* d_row_ptr, const int* d_col,
                               const float* d_val, const double* d_sens, int value_type,
                               int max_frames, int block_rows, int layout,
                               const int* d_row_order, raptr_stream_t stream);
/* 1 if the plan keeps its own row order. */
int raptr_recon_row_ordered(const raptr_recon* plan);
/* Row order for a two-head setup: the channels d_channels[0..m) sorted
 * (stably) by angle, then by the line's midpoint between the heads
 * (id0z + id1z, id0y + id1y) in tiles of tile x tile midpoints (the tile,
 * then the midpoint inside it; tile = 1: midpoints in (z, y) order), then
 * by its slope (id1z - id0z, id1y - id0y); d_order[p] = position in
 * d_channels.  Allocates its own scratch (about 26 bytes per channel). */
int raptr_locality_order(const raptr_setup_t* setup, const int* d_channels, int m, int tile,
                         int* d_order, raptr_stream_t stream);
/* raptr_locality_order with the key order chosen: RAPTR_ORDER_ANGLE_MAJOR
 * (the order above) or RAPTR_ORDER_Y_MAJOR: by the midpoint's y sum
 * (id0y + id1y) first, then angle, midpoint z in tiles of `tile`, slope.
 * The lines of a two-head setup rotated about y stay within a few voxels of
 * their midpoint's y, so a range of consecutive y-major rows crosses only a
 * band of the grid: the rows one XCD processes at a time touch a fraction
 * of the image (batched forward) and the voxels that gather one row's
 * ratios share that band (batched backward).  RAPTR_ORDER_SLOPE_TILES: the
 * angle-major order with the slopes (id1z - id0z, id1y - id0y) in 4 x 4
 * tiles, so 16 consecutive rows through one midpoint have nearly the same
 * direction and cover nearly the same voxels.  RAPTR_ORDER_ANGLE_SLOPE_TILES:
 * the slope-tiled order over angles in groups of 4 (ida / 4 first, ida % 4
 * last): one channel at neighbouring angles is nearly the same line inside
 * the grid, so a run of 16 rows is 4 slopes x 4 angles of one midpoint (on
 * the full-size geometry 7.5 instead of 5.9 of 16 rows per column of a
 * 16-row union). */
enum {
  RAPTR_ORDER_ANGLE_MAJOR = 0,
  RAPTR_ORDER_Y_MAJOR = 1,
  RAPTR_ORDER_SLOPE_TILES = 2,
  RAPTR_ORDER_ANGLE_SLOPE_TILES = 3
};
int raptr_locality_order_ex(const raptr_setup_t* setup, const int* d_channels, int m, int tile,
                            int kind, int* d_order, raptr_stream_t stream);
int raptr_recon_layout(const raptr_recon* plan);
/* Fails (RAPTR_ERR_INVALID_ARG) while clones of the plan live. */
int raptr_recon_destroy(raptr_recon* plan);
/* A second plan over the same matrix (no counterpart in the reference): it
 * shares the plan's read-only layouts (matrix formats, unions, row order,
 * sensitivity) and owns only its work vectors (ratios, partial sums,
 * images, normalisation; 1.3 GB for a 64-frame plan at full size, against
 * 63 GB of layouts), with the plan's current kernel shapes, which it can
 * change on its own.  Clones run on their own streams at the same time as
 * the plan and each other: two 64-frame batches side by side overlap their
 * gather-bound passes (full size: 80 against 73 frames/s).  A clone of a
 * clone shares the first plan's layouts.  The plan must outlive its
 * clones (raptr_recon_destroy refuses it while they live); clone and
 * destroy calls on one plan must not run concurrently. */
int raptr_recon_clone(raptr_recon* plan, raptr_recon** out, raptr_stream_t stream);
/* Live clones of a plan (-1 for a null plan). */
int raptr_recon_clone_count(const raptr_recon* plan);
int raptr_recon_info(const raptr_recon* plan, int64_t* device_bytes,
                     int64_t* padded_nnz, int* row_blocks, int* fwd_lanes,
                     int* fwd_unroll, int* bwd_unroll);
/* Kernel shape: lanes per forward row (16, 32, 64; single frame), and the
 * groups of 4 entries in flight per lane or entry slot in the forward and
 * the backward (1, 2, 4).  The transposed segments of the blocked layout
 * are padded to multiples of 4 entries (value 0) for 16-byte loads. */
int raptr_recon_set_config(raptr_recon* plan, int fwd_lanes, int fwd_unroll,
                           int bwd_unroll);
/* Single-frame backward pass.  staged = 3 (per-XCD row ranges): each XCD
 * takes one eighth of the row blocks (workgroup i runs on XCD i % 8), so
 * its ratio slices stay in its own L2; per-XCD partial column sums (8 x n
 * floats) are added in a fixed order; lanes_per_column 16, 32 or 64 (8
 * runs as 64); the default of single-frame row-blocked plans with >= 8 row
 * blocks (32 lanes, unroll 2; with fewer blocks some XCDs would idle).
 * staged = 0: the block-by-block pass.  workgroups is kept for the
 * signature and must be positive; no kernel reads it. */
int raptr_recon_set_backward(raptr_recon* plan, int staged, int lanes_per_column,
                             int workgroups);
int raptr_recon_get_backward(const raptr_recon* plan, int* staged, int* lanes_per_column,
                             int* workgroups);
/* 16-bit indices of single-frame blocked plans (6 bytes per FP32 entry
 * instead of 8): with block_rows <= 65536 the transpose's rows as offsets in
 * their row block (read by the per-XCD backward, staged = 3); with <= 262,144
 * voxels (and rows of <= 65,535 entries) the forward's columns as their low
 * 16 bits plus, per row, the sorted positions where the high part steps.
 * on (the default when present) switches both.  get: bit 0 = rows on, bit
 * 1 = columns on; -1: none present. */
int raptr_recon_set_compact_rows(raptr_recon* plan, int on);
int raptr_recon_get_compact_rows(const raptr_recon* plan);
/* Kernel of the single-frame forward pass (both give the same sums, bit
 * for bit, for the same lanes/unroll):
 *   0  one matrix entry per lane and gather (32-bit or, when on, 16-bit
 *      columns);
 *   2  x staged in LDS: each workgroup copies x over the sorted column
 *      union of its 64 consecutive plan rows into LDS and gathers from
 *      there through 16-bit ranks.  Built for single-frame blocked plans
 *      of <= 258,048 voxels whose unions all hold <= 16,384 voxels (rows
 *      in an order that keeps neighbouring lines together, e.g.
 *      raptr_locality_order); the default there.
 * available: bit k set if kind k can run on this plan; union_entries: the
 * column unions' total size (0: none).  Null outputs are skipped. */
int raptr_recon_set_forward_kind(raptr_recon* plan, int kind);
int raptr_recon_get_forward_kind(const raptr_recon* plan, int* kind, int* available,
                                 int64_t* union_entries);
/* Fused single-frame iteration.  Replaces the pair of products
 * cusparseScsrmv(NON_TRANSPOSE) / cusparseScsrmv(TRANSPOSE) of one MLEM
 * iteration (reco.cu:257-270, csrmv.hpp:72-83) by one pass over the
 * matrix: per group of consecutive plan rows (raptr_recon_set_union_rows),
 * x over the group's column union in LDS, each row's ratio
 * e_i = y_i / (A x)_i, and A_ij e_i added into column sums over the union
 * in LDS, written as the group's partial sums; a second kernel adds each
 * column's partials in a fixed order (raptr_recon_set_fused_reduce) and
 * applies the update.  Bit-reproducible run to run (fixed orders, no
 * atomics).  Built with the LDS-staged forward's unions where no row
 * repeats a column; adds 8 bytes per union entry, 8 per segment of the
 * column-tile reduce and 8 per voxel to the plan.
 *   on: 0 the separate forward and backward passes; 1 fused, one column-sum
 *       region per wave (8 x (rows + 1) + 4 x (largest union + waves x
 *       (largest union + 1)) bytes of LDS must fit 160 KiB); 2 fused, one
 *       region per workgroup that the waves add to in turn (8 x (rows + 1)
 *       + 4 x (2 x largest union + 1) bytes; the default);
 *   waves: waves per row group (2 or 4; 8 with on = 2; default 4 where five
 *          such workgroups fit a CU's 160 KiB of LDS, else 8);
 *   lanes_per_row: lanes per matrix row (16, 32 or 64);
 *   lanes_per_column: lanes per column of the reduce, or per segment slot
 *       of the column-tile reduce (16, 32 or 64).
 * get: available = 1 if the plan has the fused pass; null outputs are
 * skipped. */
int raptr_recon_set_fused(raptr_recon* plan, int on, int waves, int lanes_per_row,
                          int lanes_per_column);
int raptr_recon_get_fused(const raptr_recon* plan, int* available, int* on, int* waves,
                          int* lanes_per_row, int* lanes_per_column);
/* Column reduce of the fused iteration (the transposed product's sums,
 * reco.cu:268-270), by mode:
 *   0: all of a column's group partials in one pass, by column;
 *   1: the groups split into 8 ranges, one per XCD (workgroup i runs on
 *      XCD i % 8), so each XCD's L2 holds an eighth of the partial sums;
 *      the 8 range sums of a column added in order;
 *   2 (the default where the plan's segment table gives at least 512
 *      workgroups; else 1): column tiles:
 *      per part of a tile's groups, their partials read in the order the
 *      products kernel wrote them and added into per-slot column sums in
 *      LDS, then in slot and part order (raptr_recon_set_fused_tiles).
 * All are fixed-order; their sums differ in rounding.  get: the mode (-1
 * for a null plan). */
int raptr_recon_set_fused_reduce(raptr_recon* plan, int mode);
int raptr_recon_get_fused_reduce(const raptr_recon* plan);
/* Plan rows per column-union group (the LDS-staged forward and the fused
 * iteration): rows = 0 lets the plan choose (among 256, 128, 64, 32 and 16
 * rows giving at least 1,024 groups, else 16, the size whose fused
 * iteration a bytes-and-occupancy model of the measured shapes puts
 * fastest: 256 at full size, 32 at 96^3), else a power of two in 16..512.  Rebuilds the unions, the
 * fused pass's index and the column-tile table, drops the plan's captured
 * iterations and restores the fused defaults.  Refused on a clone or a plan
 * with clones, which share them, and on plans other than single-frame
 * blocked ones.  get: rows, the largest union and the union entries of all
 * groups (zeros: no unions).  No counterpart in the reference. */
int raptr_recon_set_union_rows(raptr_recon* plan, int rows, raptr_stream_t stream);
int raptr_recon_get_unions(const raptr_recon* plan, int* rows, int* largest, int64_t* entries);
/* Shape of the column-tile reduce: columns per tile (a power of two,
 * 128-2048) and segments per workgroup (64-1,048,576; a tile's segments,
 * a group's union positions inside the tile, 8 bytes each, are cut into
 * parts of about that many, one workgroup and one partial image row each).
 * The plan builds 1,024 columns and 512 segments.  Rebuilds the table and
 * drops the plan's captured iterations.  Refused on a clone or a plan with
 * clones, which share the table.  get: the shape (0: no table), the
 * segments, workgroups and the largest number of parts of a tile; null
 * outputs are skipped.  No counterpart in the reference. */
int raptr_recon_set_fused_tiles(raptr_recon* plan, int columns, int segments_per_workgroup,
                                raptr_stream_t stream);
int raptr_recon_get_fused_tiles(const raptr_recon* plan, int* columns, int* segments_per_workgroup,
                                int64_t* segments, int* workgroups, int* max_parts);
/* Register-held steps of 4 x 64 entries per row of the fused products
 * kernel with 64 lanes per row (2, 3 or 4; rows past them continue from
 * memory).  The plan's default: (mean padded row + 64) / 256, rounded up
 * and clamped to 2-4.  With 16 or 32 lanes per row the steps are fixed (6
 * and 3).  get: the setting (-1 for a null plan). */
int raptr_recon_set_fused_steps(raptr_recon* plan, int steps);
int raptr_recon_get_fused_steps(const raptr_recon* plan);
/* The fused iteration's kernels alone (timing, tests), as the plan is set:
 * part 0 the products and the column reduce with the update into the
 * plan's internal image buffer (d_y in the caller's row order, as
 * raptr_recon_forward); 1 the products alone (k_fb1l: ratios and the
 * groups' partial sums; d_y in the plan's row order); 2 the column reduce
 * and update alone (from the last products; d_y unused).  No counterpart
 * in the reference, whose iteration is the two csrmv calls. */
int raptr_recon_fused_pass(raptr_recon* plan, const float* d_x, const float* d_y, int part,
                           raptr_stream_t stream);
/* Row blocks per backward launch of the blocked layout (1..64): a wave
 * walks its column through that many consecutive blocks, keeping the sum in
 * registers (longer waves, fewer launches). */
int raptr_recon_set_blocks_per_launch(raptr_recon* plan, int blocks);
int raptr_recon_get_blocks_per_launch(const raptr_recon* plan);
/* 64-frame batches: kernels that load the values and indices of 64
 * entries with one instruction each and broadcast them across the 16 lanes
 * of an entry slot (on != 0), or the general kernels (0). */
int raptr_recon_set_broadcast(raptr_recon* plan, int on);
int raptr_recon_get_broadcast(const raptr_recon* plan);
/* 64-frame plans (max_frames = 64) also hold grouped layouts: groups of G
 * consecutive plan rows (forward) and of G voxels (a brick of the options'
 * grid, else consecutive columns) per row block (backward), each stored as
 * one ascending union list with G values per union entry (0 where a member
 * has no entry).  One gathered 256-byte frame vector then serves G rows or
 * G voxels.  on: use them for 64-frame passes (the default when present);
 * unroll: union entries per step of a wave; kind 2 = matrix cores
 * (G = 16 only: v_mfma_f32_16x16x4f32 over 4 union entries x 16 frames,
 * keys and values from coalesced vector loads; the backward adds each
 * launch's row blocks into one shared image sum), kind 3 (the default) =
 * kind 2 with a backward in which each XCD takes one eighth of the row
 * blocks and writes its own partial image (8 x n x 64 floats, allocated
 * with the plan); a fixed-order reduce adds them.  Shapes per G: 16: unroll
 * 8, 16 or 32, kind 2 or 3; 32: unroll 8 or 16, kind 3
 * (v_mfma_f32_32x32x2f32, 2 entries x 32 frames, whole 128-byte lines per
 * gathered frame block).  Voxel bricks: 4x1x4 (G = 16), 4x2x4 (G = 32).
 * Kinds 0 and 1 (vector FMAs with the keys and values in scalar
 * registers, groups of 4 and 8) were removed: the scalar-cache stream
 * capped them near 0.8 TB/s.  unroll applies to the forward (backward
 * steps are 8 entries).  get: group = G (0: no grouped layouts), fwd/
 * bwd_entries = union entries of each layout (lists padded to multiples of
 * 32 / 8). */
int raptr_recon_set_grouped(raptr_recon* plan, int on, int unroll, int kind);
/* Kind 3's backward in several launches: each XCD walks `blocks` of its
 * row blocks per launch (0: all in one launch), adding to its partial
 * image between launches.  The bricks resident on an XCD then stay within
 * a few blocks of each other, so the ratio rows they gather fit its L2;
 * each extra launch re-reads and re-writes the partial images.  (The
 * single-frame per-XCD backward always runs in one launch: windows there
 * were 2-40 % slower at full size.) */
int raptr_recon_set_xcd_blocks(raptr_recon* plan, int blocks);
int raptr_recon_get_xcd_blocks(const raptr_recon* plan);
int raptr_recon_get_grouped(const raptr_recon* plan, int* group, int* on, int* unroll,
                            int* kind, int64_t* fwd_entries, int64_t* bwd_entries);
/* Entries of the padded row-blocked transpose. */
int64_t raptr_recon_transpose_nnz(const raptr_recon* plan);
/* niter MLEM iterations (reco.cu:230-303) of nframes frames at once
 * (power of two <= max_frames).  d_y: counts [m][nframes]; d_x: images
 * [n][nframes], start images in, results out.  Per frame and iteration:
 * yt = A x, e = y/yt (0 where y or yt is 0), c = A^T e,
 * x = x*c/(s + (s==0)), x *= 1/sum(x) (FP64 sum).  With use_graph and a
 * non-null stream the iteration is captured once into a HIP graph (cached
 * per frame count and buffers, raptr_recon_graph_count) and replayed. */
int raptr_recon_mlem(raptr_recon* plan, int nframes, const float* d_y, float* d_x,
                     int niter, int use_graph, raptr_stream_t stream);
/* raptr_recon_mlem over the first `active` frames (16 or 32) of a 64-frame
 * pass: d_y and d_x keep their [rows][64] layout, frames active .. 63 of
 * d_x are left as they are, and frames 0 .. active - 1 get exactly the
 * images raptr_recon_mlem gives them (the same arithmetic per frame).
 * The matrix-core kernels then gather and multiply active/64 of the frame
 * vectors.  Needs the grouped layouts of 16 on, kind 3 (the default of a
 * 64-frame plan with a grid).  active = nframes runs raptr_recon_mlem.
 * Graphs are cached per (frames, active, buffers).  No counterpart in the
 * reference; for serving lanes whose busy slots fit the low frames. */
int raptr_recon_mlem_active(raptr_recon* plan, int nframes, int active, const float* d_y,
                            float* d_x, int niter, int use_graph, raptr_stream_t stream);
/* Captured iterations the plan holds (raptr_recon_mlem with use_graph):
 * one per distinct (nframes, active frames, d_y, d_x), at most 16, the least recently
 * used destroyed first; all are destroyed when a kernel shape changes.
 * -1 for a null plan. */
int raptr_recon_graph_count(const raptr_recon* plan);
/* The two passes alone (timing, tests): forward + ratio into d_e
 * ([m][nframes]); backward over all row blocks with the update into the
 * plan's internal image buffer. */
int raptr_recon_forward(raptr_recon* plan, int nframes, const float* d_x,
                        const float* d_y, float* d_e, raptr_stream_t stream);
int raptr_recon_backward(raptr_recon* plan, int nframes, const float* d_x,
                         raptr_stream_t stream);

/* ----------------------- frame slots of [rows][nframes] arrays (serving) */
/* No counterpart in the reference (one frame per run there): a request
 * occupies one frame slot of a batched plan's y ([m][F]) and x ([n][F]).
 * fill: column `slot` = value.  scatter: column `slot`[d_idx[i]] = d_val[i]
 * (rows outside [0, rows) are skipped and set *d_bad = 1).  gather:
 * d_out[i] = column `slot`[i].  copy: column dst_slot = column src_slot
 * (a request moved to a lower slot, so a lane's busy slots fit the first
 * 16 or 32 frames of raptr_recon_mlem_active). */
int raptr_frame_fill(float* d_a, int64_t rows, int nframes, int slot, float value,
                     raptr_stream_t stream);
int raptr_frame_scatter(float* d_a, int64_t rows, int nframes, int slot, const int* d_idx,
                        const float* d_val, int64_t n, int* d_bad, raptr_stream_t stream);
int raptr_frame_gather(const float* d_a, int64_t rows, int nframes, int slot, float* d_out,
                       raptr_stream_t stream);
int raptr_frame_copy(float* d_a, int64_t rows, int nframes, int src_slot, int dst_slot,
                     raptr_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
