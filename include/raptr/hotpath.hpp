// raptr/hotpath.hpp -- host mirror of the reference's hot-path call
// signatures, served by the gfx950 kernels behind include/raptr_hip.h:
//
//   getWorkqueue               getWorkqueue.hpp:135-243
//   getSystemMatrixFromWorkqueue  getSystemMatrixFromWorkqueue.cu:116-122
//                              (a host launcher here: device pointers in,
//                              geometry passed explicitly, no __constant__)
//   getChords<T>               ChordsCalc_lowlevel.hpp:194-204 (a batched
//                              device call here: nRays rays, nChords per ray)
//   cooSort<T>                 cooSort.hpp:75-99
//   customizeMatDescr / convertCoo2Csr   cusparseWrapper.hpp:36-60
//   convertCsr2Ecsr            convertCsr2Ecsr.hpp:51-64
//   CSRmv<T>                   csrmv.hpp:37-99
//   divides / dividesMultiplies / scales / sum   mlemOperations.hpp:44-169
//   systemMatrixCalculation    wrappers.hpp:185-266 (the reference's
//                              template and parameter list, geometry from
//                              grid_const / setup_const / nrays_const)
//   RandRayGen / ReguRayGen    RayGenerators.hpp:36-99
//   grid_const / setup_const / nrays_const   device_constant_memory.hpp:34-46
//                              (host variables here)
//   nInChunk / chunkPtr / nChunks   wrappers.hpp:429-464
//
// Error behaviour follows the reference: device errors throw
// raptr::DeviceError (its cudaException, CUDA_HandleError.hpp:39-53);
// sparse-library style calls return a status; getWorkqueue reports misuse on
// stderr and exits (getWorkqueue.hpp:109-111, 156-159).  Pointers named
// *_devi / documented as device are device memory.  Helpers that the
// reference gets from thrust (temporary storage) allocate and free their
// own scratch, as thrust does.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "raptr/geometry.hpp"
#include "raptr_hip.h"

namespace raptr {

class DeviceError : public std::runtime_error {
 public:
  DeviceError(int status, const std::string& where)
      : std::runtime_error(where + ": " + raptr_status_string(status) + ": " +
                           raptr_last_error()),
        status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* where) {
  if (status != RAPTR_OK) throw DeviceError(status, where);
}

// Owning device buffer (host helper; the C ABI itself never allocates).
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() {}
  explicit DeviceBuffer(size_t n) : n_(n) {
    void* p = nullptr;
    check(raptr_malloc(&p, n * sizeof(T) ? n * sizeof(T) : 1), "raptr_malloc");
    p_ = static_cast<T*>(p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
  ~DeviceBuffer() {
    if (p_) raptr_free(p_);
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n) { check(raptr_memcpy_h2d(p_, h, n * sizeof(T), nullptr), "h2d"); }
  void download(T* h, size_t n) const {
    check(raptr_memcpy_d2h(h, p_, n * sizeof(T), nullptr), "d2h");
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// The precompiled kernels implement the Default* plugin family.
template <typename MS, typename VG>
void require_default_geometry() {
  static_assert(std::is_same<MS, DefaultMeasurementSetup<float> >::value,
                "the gfx950 kernels implement DefaultMeasurementSetup<float>");
  static_assert(std::is_same<VG, DefaultVoxelGrid<float> >::value,
                "the gfx950 kernels implement DefaultVoxelGrid<float>");
}

}  // namespace raptr

// ----------------------------------------------- device constant memory
// device_constant_memory.hpp:34-46 keeps the voxel grid, the measurement
// setup and the ray count in __constant__ memory; the drivers set them with
// cudaMemcpyToSymbol (reco.cu:85-93) and getSystemMatrix and the ray
// generators read them.  Here they are host variables with the same names,
// read by the calls that take no geometry argument (systemMatrixCalculation
// and the RayGen functors); a driver sets them by assignment or with
// raptr::setConstant.  The geometry is copied into every launch's
// arguments, so no device state exists and calls may run concurrently on
// different streams as long as these variables are not changed meanwhile.
inline DefaultVoxelGrid<float> grid_const;
inline DefaultMeasurementSetup<float> setup_const;
inline int nrays_const = 1;

namespace raptr {
// Generator key of systemMatrixCalculation's rays (the reference compiles
// RANDOM_SEED = 1251 into its kernels, getSystemMatrixDeviceOnly.cu:43; the
// workloads of this project use 1).
inline uint32_t random_seed = 1;

template <typename T>
void setConstant(T& symbol, const T& value) {
  symbol = value;
}

namespace detail {
// Philox4x32-10 (Salmon et al., SC'11), the counter-based generator of every
// ray of this project; u01 maps a 32-bit draw to (0, 1].
inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t* out) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
inline float u01(uint32_t x) { return float((x >> 8) + 1u) * (1.0f / 16777216.0f); }
}  // namespace detail
}  // namespace raptr

// ------------------------------------------------------- ray generators
// RandRayGen (RayGenerators.hpp:36-67): rays with end points uniform in the
// two pixel volumes.  The reference draws six cuRAND XORWOW uniforms from
// the stream curand_init(seed, setNr, 0); here the k-th call of a generator
// constructed with (seed, setNr) draws them from Philox4x32-10 with key
// (seed, 0) and counters (setNr, k, 0|1, element + 1), element = -1 unless
// set with setElement.  With setNr = channel id the k-th ray is the k-th
// RAPTR_RAYGEN_RAND ray of that channel (raptr_generate_rays,
// raptr_sm_from_wq); with an element set, the RAPTR_RAYGEN_RAND_ELEMENT ray
// of that (channel, voxel) pair.  The trafos read setup_const, as the
// reference's do.
template <typename T, typename MSTrafo0_inplace, typename MSTrafo1_inplace>
struct RandRayGen {
  RandRayGen(int const seed, int const setNr)
      : key_(uint32_t(seed)), set_(uint32_t(setNr)), word3_(0u), k_(0u) {}

  void setElement(int const element) { word3_ = uint32_t(element) + 1u; }

  void operator()(T* const ray, int const& id0z, int const& id0y, int const& id1z,
                  int const& id1y, int const& ida, MSTrafo0_inplace const& trafo0,
                  MSTrafo1_inplace const& trafo1) {
    uint32_t r0[4], r1[4];
    raptr::detail::philox4x32_10(set_, k_, 0u, word3_, key_, 0u, r0);
    raptr::detail::philox4x32_10(set_, k_, 1u, word3_, key_, 0u, r1);
    ++k_;
    ray[0] = T(raptr::detail::u01(r0[0]));
    ray[1] = T(raptr::detail::u01(r0[1]));
    ray[2] = T(raptr::detail::u01(r0[2]));
    ray[3] = T(raptr::detail::u01(r0[3]));
    ray[4] = T(raptr::detail::u01(r1[0]));
    ray[5] = T(raptr::detail::u01(r1[1]));
    trafo0(&ray[0], id0z, id0y, ida, &setup_const);
    trafo1(&ray[3], id1z, id1y, ida, &setup_const);
  }

 private:
  uint32_t key_, set_, word3_, k_;
};

// ReguRayGen (RayGenerators.hpp:73-99): pixel-centre rays (RAPTR_RAYGEN_REGU).
template <typename T, typename MSTrafo0_inplace, typename MSTrafo1_inplace>
struct ReguRayGen {
  ReguRayGen(int const seed, int const setNr) {
    (void)seed;
    (void)setNr;
  }

  void operator()(T* const ray, int const& id0z, int const& id0y, int const& id1z,
                  int const& id1y, int const& ida, MSTrafo0_inplace const& trafo0,
                  MSTrafo1_inplace const& trafo1) const {
    for (int i = 0; i < 6; ++i) ray[i] = T(0.5);
    trafo0(&ray[0], id0z, id0y, ida, &setup_const);
    trafo1(&ray[3], id1z, id1y, ida, &setup_const);
  }
};

namespace raptr {
// The device generator a RayGen functor type selects in
// systemMatrixCalculation: RandRayGen -> per-element streams (the reference's
// getSystemMatrix gives each thread its own generator,
// getSystemMatrixDeviceOnly.cu:396, so every element draws its own rays),
// ReguRayGen -> pixel centres.
template <typename RayGen>
struct raygen_kind;
template <typename T, typename A, typename B>
struct raygen_kind<RandRayGen<T, A, B> > {
  static constexpr int value = RAPTR_RAYGEN_RAND_ELEMENT;
};
template <typename T, typename A, typename B>
struct raygen_kind<ReguRayGen<T, A, B> > {
  static constexpr int value = RAPTR_RAYGEN_REGU;
};
}  // namespace raptr

// ------------------------------------------------------------ workqueue
// Resumable scan of (list entry, voxel) pairs from the cursor (listId,
// vxlId); appends up to n (or all) candidate pairs and advances the cursor.
template <typename T, typename ConcreteList, typename ConcreteVG, typename ConcreteVGidx,
          typename ConcreteVGidy, typename ConcreteVGidz, typename ConcreteMS,
          typename ConcreteMSid0z, typename ConcreteMSid0y, typename ConcreteMSid1z,
          typename ConcreteMSid1y, typename ConcreteMSida, typename Trafo0, typename Trafo1>
int getWorkqueue(std::vector<int>& returnCnlId, std::vector<int>& returnVxlId, int& listId,
                 int& vxlId, ConcreteList const* const list, ConcreteVG const* const grid,
                 ConcreteMS const* const setup, int const n = -1) {
  raptr::require_default_geometry<ConcreteMS, ConcreteVG>();
  if (!returnCnlId.empty() || !returnVxlId.empty()) {
    std::cerr << "raptr: getWorkqueue needs empty output vectors" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  const int size = list->size();
  raptr_setup_t s = setup->to_c();
  raptr_grid_t g = grid->to_c();
  if (vxlId < 0 || vxlId >= grid->nVoxels()) {
    std::cerr << "raptr: getWorkqueue cursor voxel " << vxlId << " out of range" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  if (listId >= size) return 0;
  raptr::DeviceBuffer<int> d_list(size_t(size) + 1);
  d_list.upload(list->cnlData(), size_t(size));
  std::vector<int64_t> counts(size_t(size - listId));
  raptr::DeviceBuffer<int64_t> d_counts(counts.size());
  raptr::check(raptr_workqueue_counts(&s, &g, d_list.get() + listId, size - listId,
                                      d_counts.get(), nullptr),
               "raptr_workqueue_counts");
  d_counts.download(counts.data(), counts.size());
  int64_t cap = 0;
  for (int64_t c : counts) cap += c;
  if (n >= 0 && cap > n) cap = n;
  raptr::DeviceBuffer<int> d_cnl(size_t(cap) + 1), d_vxl(size_t(cap) + 1);
  size_t wsb = raptr_workqueue_workspace_size(&g, size);
  raptr::DeviceBuffer<char> ws(wsb);
  int64_t found = 0;
  raptr::check(raptr_workqueue(&s, &g, d_list.get(), size, &listId, &vxlId, n, d_cnl.get(),
                               d_vxl.get(), cap, &found, ws.get(), wsb, nullptr),
               "raptr_workqueue");
  returnCnlId.resize(size_t(found));
  returnVxlId.resize(size_t(found));
  if (found) {
    d_cnl.download(returnCnlId.data(), size_t(found));
    d_vxl.download(returnVxlId.data(), size_t(found));
  }
  return int(found);
}

// ---------------------------------------------- matrix from workqueue
// Device pointers cnl, vxl (wqLength entries) in, val out.
template <typename T, typename ConcreteVG, typename ConcreteVGIdx, typename ConcreteVGIdy,
          typename ConcreteVGIdz, typename ConcreteMS, typename ConcreteMSid0z,
          typename ConcreteMSid0y, typename ConcreteMSid1z, typename ConcreteMSid1y,
          typename ConcreteMSida, typename Trafo0, typename Trafo1>
void getSystemMatrixFromWorkqueue(int const* const globalCnl, int const* const globalVxl,
                                  T* const globalVal, int const wqLength, int const nrays,
                                  ConcreteVG const& grid, ConcreteMS const& setup,
                                  uint32_t seed = 1, int raygen = RAPTR_RAYGEN_RAND,
                                  raptr_stream_t stream = nullptr) {
  static_assert(std::is_same<T, float>::value, "matrix values are float");
  raptr::require_default_geometry<ConcreteMS, ConcreteVG>();
  raptr_setup_t s = setup.to_c();
  raptr_grid_t g = grid.to_c();
  raptr::check(raptr_sm_from_wq(&s, &g, globalCnl, globalVxl, globalVal, wqLength, nrays,
                                seed, raygen, stream),
               "raptr_sm_from_wq");
}

// Siddon chords of nRays rays (device pointers): per ray up to nChords
// (length, ix, iy, iz); returns nothing, counts per ray in nFound_devi.
// The reference's form is one ray per call with host or device pointers
// (ChordsCalc_lowlevel.hpp:194-204); a device batch replaces its loop.
template <typename T>
void getChords(T* const chords_devi, int* const voxelIds_devi, int const nChords,
               T const* const rays_devi, int const nRays, T const* const gridO,
               T const* const gridD, int const* const gridN, int* const nFound_devi,
               raptr_stream_t stream = nullptr) {
  static_assert(std::is_same<T, float>::value, "rays are float");
  raptr_grid_t g = {{gridO[0], gridO[1], gridO[2]},
                    {gridD[0], gridD[1], gridD[2]},
                    {gridN[0], gridN[1], gridN[2]}};
  raptr::check(raptr_get_chords(rays_devi, nRays, &g, nChords, chords_devi, voxelIds_devi,
                                nFound_devi, stream),
               "raptr_get_chords");
}

// ------------------------------------------------------- sparse formats
template <typename T>
void cooSort(T* const cooVal, int* const cooRowId, int* const cooColId, int const N) {
  static_assert(std::is_same<T, float>::value, "COO values are float");
  if (N <= 0 || !cooVal) return;
  size_t wsb = raptr_coo_sort_workspace_size(N);
  raptr::DeviceBuffer<char> ws(wsb);
  raptr::check(raptr_coo_sort(cooVal, cooRowId, cooColId, N, ws.get(), wsb, nullptr),
               "raptr_coo_sort");
}

// Sparse "library" handle and descriptor of this project (no CUDA types).
struct raptr_sparse_context {
  raptr_stream_t stream = nullptr;
};
typedef raptr_sparse_context* raptr_sparse_handle_t;
struct raptr_mat_descr {
  int general = 1;
  int index_base = 0;
};
typedef raptr_mat_descr* raptr_mat_descr_t;
enum raptr_sparse_status_t { RAPTR_SPARSE_STATUS_SUCCESS = 0, RAPTR_SPARSE_STATUS_ERROR = 1 };
enum raptr_operation_t {
  RAPTR_SPARSE_OPERATION_NON_TRANSPOSE = RAPTR_OP_NON_TRANSPOSE,
  RAPTR_SPARSE_OPERATION_TRANSPOSE = RAPTR_OP_TRANSPOSE
};

inline raptr_sparse_status_t raptr_sparse_create(raptr_sparse_handle_t* h) {
  *h = new raptr_sparse_context();
  return RAPTR_SPARSE_STATUS_SUCCESS;
}
inline void raptr_sparse_destroy(raptr_sparse_handle_t h) { delete h; }
inline raptr_sparse_status_t raptr_sparse_create_descr(raptr_mat_descr_t* d) {
  *d = new raptr_mat_descr();
  return RAPTR_SPARSE_STATUS_SUCCESS;
}
inline void raptr_sparse_destroy_descr(raptr_mat_descr_t d) { delete d; }

// general matrix, zero-based indices
inline raptr_sparse_status_t customizeMatDescr(raptr_mat_descr_t& descr,
                                               raptr_sparse_handle_t const& handle) {
  (void)handle;
  descr->general = 1;
  descr->index_base = 0;
  return RAPTR_SPARSE_STATUS_SUCCESS;
}

inline raptr_sparse_status_t convertCoo2Csr(int* const csrRowPtr, int const* const cooRowId,
                                            raptr_sparse_handle_t const& handle,
                                            int const nnz, int const m) {
  int st = raptr_coo2csr(csrRowPtr, cooRowId, nnz, m, handle ? handle->stream : nullptr);
  return st == RAPTR_OK ? RAPTR_SPARSE_STATUS_SUCCESS : RAPTR_SPARSE_STATUS_ERROR;
}

inline void convertCsr2Ecsr(int* const ecsrRowPtr, int* const yRowId, int const nnzr,
                            int* const csrRowPtr, int const m) {
  raptr::check(raptr_csr2ecsr(ecsrRowPtr, yRowId, nnzr, csrRowPtr, m, nullptr),
               "raptr_csr2ecsr");
}

// y = alpha*op(A)*x + beta*y; alpha and beta are host pointers
template <typename T>
struct CSRmv;

template <>
struct CSRmv<float> {
  raptr_sparse_status_t operator()(raptr_sparse_handle_t handle, raptr_operation_t transA,
                                   int m, int n, int nnz, float* const alpha,
                                   raptr_mat_descr_t descrA, float* const csrValA,
                                   int* const csrRowPtrA, int* const csrColIdA, float* const x,
                                   float* const beta, float* const y) {
    (void)descrA;
    int st = raptr_csrmv_f32(int(transA), m, n, nnz, *alpha, csrValA, csrRowPtrA, csrColIdA, x,
                             *beta, y, handle ? handle->stream : nullptr);
    return st == RAPTR_OK ? RAPTR_SPARSE_STATUS_SUCCESS : RAPTR_SPARSE_STATUS_ERROR;
  }
};

template <>
struct CSRmv<double> {
  raptr_sparse_status_t operator()(raptr_sparse_handle_t handle, raptr_operation_t transA,
                                   int m, int n, int nnz, double* const alpha,
                                   raptr_mat_descr_t descrA, double* const csrValA,
                                   int* const csrRowPtrA, int* const csrColIdA, double* const x,
                                   double* const beta, double* const y) {
    (void)descrA;
    int st = raptr_csrmv_f64(int(transA), m, n, nnz, *alpha, csrValA, csrRowPtrA, csrColIdA, x,
                             *beta, y, handle ? handle->stream : nullptr);
    return st == RAPTR_OK ? RAPTR_SPARSE_STATUS_SUCCESS : RAPTR_SPARSE_STATUS_ERROR;
  }
};

// ----------------------------------------------------------- MLEM ops
template <typename T>
void divides(T* const& result, T* const& dividend, T* const& divisor, int const n) {
  static_assert(std::is_same<T, float>::value, "float vectors");
  raptr::check(raptr_divides(result, dividend, divisor, n, nullptr), "raptr_divides");
}

template <typename T>
void dividesMultiplies(T* const& result, T* const& a, T* const& b, T* const& c, int const n) {
  static_assert(std::is_same<T, float>::value, "float vectors");
  raptr::check(raptr_divides_multiplies(result, a, b, c, n, nullptr),
               "raptr_divides_multiplies");
}

template <typename T>
void scales(T* const& x, T const alpha, int const n) {
  static_assert(std::is_same<T, float>::value, "float vectors");
  raptr::check(raptr_scales(x, alpha, n, nullptr), "raptr_scales");
}

template <typename T>
T sum(T* const& x, int const n) {
  static_assert(std::is_same<T, float>::value, "float vectors");
  size_t wsb = raptr_sum_workspace_size(n);
  raptr::DeviceBuffer<char> ws(wsb);
  float out = 0.0f;
  raptr::check(raptr_sum(x, n, &out, ws.get(), wsb, nullptr), "raptr_sum");
  return out;
}

// ------------------------------------------------------- chunk math
template <typename ListSizeType, typename ChunkGridSizeType>
ListSizeType nInChunk(ChunkGridSizeType idChunk, ListSizeType const nAll,
                      ListSizeType const nFullChunk) {
  ChunkGridSizeType nFull = ChunkGridSizeType(nAll / nFullChunk);
  ChunkGridSizeType nTotal = ChunkGridSizeType((nAll + nFullChunk - 1) / nFullChunk);
  if (idChunk < nFull) return nFullChunk;
  if (idChunk < nTotal) return nAll % nFullChunk;
  return 0;
}

template <typename ListSizeType, typename ChunkGridSizeType>
ListSizeType chunkPtr(ChunkGridSizeType const chunkId, ListSizeType const nFullChunk) {
  return ListSizeType(chunkId) * nFullChunk;
}

template <typename ChunkGridSizeType, typename MemArrSizeType>
ChunkGridSizeType nChunks(MemArrSizeType const maxNnz, MemArrSizeType const maxNChunk) {
  return ChunkGridSizeType((maxNnz + maxNChunk - 1) / maxNChunk);
}

// ------------------------------------------- system matrix of a chunk
// The reference's call (wrappers.hpp:185-266, called at reco.cu:246-254,
// backprojection.cu:145-152, pureSMCalculation.cu:149-155): the matrix of
// the *nY_host channels yRowId_devi[0..nY) (ascending channel ids) with the
// geometry of grid_const / setup_const / nrays_const and the generator of
// the RayGen type (raygen_kind), every culled candidate stored (zero-valued
// ones too, getSystemMatrixDeviceOnly.cu:507-509), so the count is the
// workqueue's.  Entries go to aVxlId_devi / aVal_devi after the first
// *nnz_devi entries (device in/out accumulator: it receives its initial
// value plus the count, wrappers.hpp:172-174), rows in list order, columns
// ascending (the cooSort order).  Also written, in the reference's int32
// types: aEcsrCnlPtr_devi[0..nY] (ECSR row pointers), aCnlId_devi (COO row
// ids of the new entries) and aCsrCnlPtr_devi[0..NCHANNELS] (row pointer
// over the channel space).  The arrays must hold the entries, as in the
// reference (LIMNNZ sizing, reco.cu:47-51).  Work goes on the handle's
// stream and is complete on return.  Scratch (the 64-bit row pointers and
// the build's staging) is allocated and freed here, as the reference
// allocates its nY_devi temporary.
template <typename T, typename ListSizeType, typename GridSizeType, typename MemArrSizeType,
          typename RayGen>
void systemMatrixCalculation(int* const aEcsrCnlPtr_devi, GridSizeType* const aVxlId_devi,
                             T* const aVal_devi, MemArrSizeType* const nnz_devi,
                             int* const aCnlId_devi, int* const aCsrCnlPtr_devi,
                             int* const yRowId_devi, ListSizeType const* const& nY_host,
                             raptr_sparse_handle_t const& handle) {
  static_assert(std::is_same<T, float>::value, "matrix values are float");
  static_assert(sizeof(GridSizeType) == sizeof(int) && std::is_integral<GridSizeType>::value,
                "voxel ids are 32-bit integers");
  static_assert(std::is_integral<MemArrSizeType>::value, "nnz counter is an integer");
  const raptr_stream_t st = handle ? handle->stream : nullptr;
  const int nY = int(*nY_host);
  raptr_setup_t s = setup_const.to_c();
  raptr_grid_t g = grid_const.to_c();
  MemArrSizeType nnz0 = 0;
  raptr::check(raptr_memcpy_d2h(&nnz0, nnz_devi, sizeof(MemArrSizeType), st), "d2h nnz");
  raptr::check(raptr_stream_sync(st), "stream sync");
  raptr::DeviceBuffer<int64_t> rp(size_t(nY) + 1);
  const size_t wsb = raptr_sm_build_workspace_size(&g, nrays_const, int64_t(1) << 22);
  raptr::DeviceBuffer<char> ws(wsb);
  const int kind = raptr::raygen_kind<RayGen>::value;
  int64_t count = 0;
  raptr::check(raptr_sm_build_csr_ex(&s, &g, yRowId_devi, nY, nrays_const, raptr::random_seed,
                                     kind, RAPTR_SM_KEEP_ZEROS, rp.get(), nullptr, nullptr, 0,
                                     &count, ws.get(), wsb, st),
               "raptr_sm_build_csr_ex (count)");
  int64_t filled = 0;
  raptr::check(raptr_sm_build_csr_ex(&s, &g, yRowId_devi, nY, nrays_const, raptr::random_seed,
                                     kind, RAPTR_SM_KEEP_ZEROS, rp.get(),
                                     reinterpret_cast<int*>(aVxlId_devi) + nnz0, aVal_devi + nnz0,
                                     count, &filled, ws.get(), wsb, st),
               "raptr_sm_build_csr_ex (fill)");
  raptr::check(raptr_sm_index_arrays(rp.get(), yRowId_devi, nY, int64_t(nnz0),
                                     setup_const.nChannels(), aEcsrCnlPtr_devi, aCnlId_devi,
                                     aCsrCnlPtr_devi, st),
               "raptr_sm_index_arrays");
  const MemArrSizeType nnz1 = nnz0 + MemArrSizeType(filled);
  raptr::check(raptr_memcpy_h2d(nnz_devi, &nnz1, sizeof(MemArrSizeType), st), "h2d nnz");
  raptr::check(raptr_stream_sync(st), "stream sync");
}

// The same matrix with explicit geometry and exact zeros dropped (the
// compact form the reconstruction plans are built from), int32 ECSR row
// pointers, *nnz_host = count.
inline void systemMatrixCalculation(int* aEcsrCnlPtr_devi, int* aVxlId_devi, float* aVal_devi,
                                    int64_t capacity, int64_t* nnz_host,
                                    const int* yRowId_devi, int nY,
                                    const DefaultMeasurementSetup<float>& setup,
                                    const DefaultVoxelGrid<float>& grid, int nrays,
                                    uint32_t seed = 1, raptr_stream_t stream = nullptr) {
  raptr_setup_t s = setup.to_c();
  raptr_grid_t g = grid.to_c();
  raptr::DeviceBuffer<int64_t> rp(size_t(nY) + 1);
  size_t wsb = raptr_sm_build_workspace_size(&g, nrays, 1 << 22);
  raptr::DeviceBuffer<char> ws(wsb);
  raptr::check(raptr_sm_build_csr(&s, &g, yRowId_devi, nY, nrays, seed, RAPTR_RAYGEN_RAND,
                                  rp.get(), aVxlId_devi, aVal_devi, capacity, nnz_host,
                                  ws.get(), wsb, stream),
               "raptr_sm_build_csr");
  if (*nnz_host > INT32_MAX) throw std::runtime_error("ECSR chunk exceeds int32 pointers");
  std::vector<int64_t> h(size_t(nY) + 1);
  rp.download(h.data(), h.size());
  std::vector<int> e(h.begin(), h.end());
  raptr::check(raptr_memcpy_h2d(aEcsrCnlPtr_devi, e.data(), e.size() * sizeof(int), stream),
               "h2d");
}
