/*
 * raptr_oracle.h -- CPU restatement of the raptr reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * gfx950 HIP path.  Only the test-suite, the smoke check and the benchmark's
 * CPU-baseline leg may load it; no product code calls it, and the product
 * path fails loudly (instead of falling back here) when its HIP library is
 * missing.
 *
 * Every function restates the reference's behaviour from its source (cited as
 * file:line relative to the raptr source tree) in plain C, written from
 * scratch.  Floating-point expressions keep the reference's operation order
 * and are compiled with -ffp-contract=off so that FP32 results are
 * bit-reproducible against the GPU kernels, which use the same order.
 *
 * Pinning (how each part is checked independently of its own formula):
 *   geometry / index maps   closed-form round trips, trafo closed forms
 *   philox4x32_10           Random123 known-answer vectors
 *   workqueue               reference getWorkqueue compiled from its own
 *                           sources (oracle/_ref), FP64 brute force, prefix
 *                           (resume) invariant
 *   intersection_length     closed forms (axis-aligned, body diagonal, miss),
 *                           brute-force numeric integration
 *   sm_from_wq              FP64 restatement on identical rays, row-sum
 *                           identity (sum over voxels = in-grid ray length)
 *   get_chords (Siddon)     each chord equals intersection_length of its voxel,
 *                           chords sum to the in-grid segment length
 *   coo_sort/coo2csr/ecsr   reference known answers (cooSortUT.cu,
 *                           test_cusparseWrapper.cu, test_convertCsr2Ecsr.cu),
 *                           std-sort brute force
 *   csrmv                   dense brute force
 *   mlem ops                test_mlemOperations.cu known answers
 *   mlem loop               fixed point for consistent data, sum(x)=1,
 *                           monotone Poisson log-likelihood
 */
#ifndef RAPTR_ORACLE_H
#define RAPTR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DefaultMeasurementSetup, MeasurementSetup.hpp:181-220 (11 scalars). */
typedef struct ro_setup {
  float pos0x, pos1x;
  int na, n0z, n0y, n1z, n1y;
  float da, segx, segy, segz;
} ro_setup;

/* DefaultVoxelGrid, VoxelGrid.hpp:151-176 (origin, pitch, count). */
typedef struct ro_grid {
  float o[3];
  float d[3];
  int n[3];
} ro_grid;

enum { RO_RAYGEN_RAND = 0, RO_RAYGEN_REGU = 1, RO_RAYGEN_RAND_ELEMENT = 2 };
enum { RO_SM_KEEP_ZEROS = 1 };

/* ---- A2/A3: index maps ---------------------------------------------- */
int64_t ro_nchannels(const ro_setup* s);
void ro_cnl_split(const ro_setup* s, int cnl, int* ida, int* id0z, int* id0y,
                  int* id1z, int* id1y);
int ro_cnl_lin(const ro_setup* s, int id0z, int id0y, int id1z, int id1y,
               int ida);
void ro_vxl_split(const ro_grid* g, int vxl, int* ix, int* iy, int* iz);
int ro_vxl_lin(const ro_grid* g, int ix, int iy, int iz);

/* ---- A4: per-angle sin/cos table and pixel->world trafo ------------- */
void ro_angle_table(const ro_setup* s, float* sin_out, float* cos_out);
void ro_trafo(const ro_setup* s, int det, float sn, float cs, int idz, int idy,
              const float box[3], float out[3]);

/* ---- A5: point-line distance ---------------------------------------- */
float ro_distance(const float a[3], const float b[3], const float p[3]);

/* ---- A6/A7: workqueue (resumable, optionally bounded by n) ----------- */
/* Returns number of entries written.  n < 0 means unbounded.  cnl_out and
 * vxl_out must have room for the result (n, or the full count). */
int64_t ro_workqueue(const ro_setup* s, const ro_grid* g, const int* list,
                     int list_size, int* list_id, int* vxl_id, int64_t n,
                     int* cnl_out, int* vxl_out);
/* Per list entry candidate counts (full voxel range), OpenMP parallel. */
void ro_workqueue_counts(const ro_setup* s, const ro_grid* g, const int* list,
                         int list_size, int64_t* counts);

/* ---- A8: Philox4x32-10 and ray generators --------------------------- */
void ro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
float ro_u01(uint32_t x);
void ro_set_threads(int n);
int ro_get_threads(void);
void ro_ray(const ro_setup* s, const float* sn, const float* cs, int raygen,
            uint32_t seed, int cnl, int ray, float out[6]);
/* ray of one (channel, element) pair: element is the voxel for
 * RO_RAYGEN_RAND_ELEMENT, ignored otherwise */
void ro_ray_element(const ro_setup* s, const float* sn, const float* cs, int raygen,
                    uint32_t seed, int cnl, int element, int ray, float out[6]);

/* ---- A9: slab intersection length ------------------------------------ */
float ro_intersection_length(const float vxl[6], const float ray[6]);
double ro_intersection_length_f64(const double vxl[6], const double ray[6]);
void ro_voxel_box(const ro_grid* g, int vxl, float box[6]);

/* ---- A10: system matrix from workqueue -------------------------------- */
void ro_sm_from_wq(const ro_setup* s, const ro_grid* g, const int* cnl,
                   const int* vxl, float* val, int64_t wq_len, int nrays,
                   uint32_t seed, int raygen);
void ro_sm_from_wq_f64(const ro_setup* s, const ro_grid* g, const int* cnl,
                       const int* vxl, double* val, int64_t wq_len, int nrays,
                       uint32_t seed, int raygen);

/* ---- A11: Siddon chords (documented correction of the initial plane) -- */
int ro_get_chords(float* chords, int* voxel_ids, int n_chords_cap,
                  const float ray[6], const float grid_o[3],
                  const float grid_d[3], const int grid_n[3]);

/* ---- A12/A13: matrix of measured rows, zeros dropped, CSR ------------- */
/* Two-call protocol: with row_ptr only (col/val NULL) fills row_ptr; then
 * with col/val sized row_ptr[list_size]. */
int64_t ro_build_csr(const ro_setup* s, const ro_grid* g, const int* list,
                     int list_size, int nrays, uint32_t seed, int raygen,
                     int64_t* row_ptr, int* col, float* val, double* val64);
int64_t ro_build_csr_ex(const ro_setup* s, const ro_grid* g, const int* list,
                        int list_size, int nrays, uint32_t seed, int raygen, int flags,
                        int64_t* row_ptr, int* col, float* val, double* val64);

/* ---- A14-A16: COO sort, COO->CSR, CSR->ECSR --------------------------- */
void ro_coo_sort(float* val, int* row, int* col, int64_t n);
void ro_coo2csr(int* row_ptr, const int* row_id, int64_t nnz, int m);
void ro_csr2ecsr(int* ecsr, const int* y_row_id, int nnzr, const int* csr,
                 int m);

/* ---- A17: csrmv (0 = non-transpose, 1 = transpose) ------------------- */
void ro_csrmv_f32(int trans, int m, int n, float alpha, const float* val,
                  const int* row_ptr, const int* col, const float* x,
                  float beta, float* y);
void ro_csrmv_f64(int trans, int m, int n, double alpha, const double* val,
                  const int64_t* row_ptr, const int* col, const double* x,
                  double beta, double* y);

/* ---- A18-A20: MLEM elementwise ops ------------------------------------ */
void ro_divides(float* res, const float* dvd, const float* dvs, int n);
void ro_divides_multiplies(float* res, const float* a, const float* b,
                           const float* c, int n);
void ro_scales(float* x, float alpha, int n);
float ro_sum(const float* x, int n);

/* ---- A21/A22: MLEM ---------------------------------------------------- */
/* FP64 recurrence of reco.cu:230-303 on a CSR of measured rows (64-bit
 * row pointers).  x in/out (N), sens (N, already Sigma-normalised or not:
 * normalised here), y (m). */
void ro_mlem_f64(int m, int nvox, const int64_t* row_ptr, const int* col,
                 const double* val, const double* y, const double* sens,
                 double* x, int niter);
/* the same recurrence in FP32 (the reference's precision; CPU baseline) */
void ro_mlem_f32(int m, int nvox, const int64_t* rp, const int* col,
                 const float* val, const float* y, const double* sens,
                 float* x, int niter);
void ro_sensitivity_f64(int m, int nvox, const int64_t* row_ptr,
                        const int* col, const double* val, double* sens);

#ifdef __cplusplus
}
#endif
#endif
