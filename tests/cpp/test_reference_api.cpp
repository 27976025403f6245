// test_reference_api.cpp -- the reference's own unit tests, restated against
// the C++ mirror (include/raptr/*.hpp) of its interfaces.  With --cpu-only
// it runs the host-side geometry checks only (no device calls).
//
//   cooSortUT.cu:68-307              COO sort: dense shuffled, empty, sparse
//   test_convertCsr2Ecsr.cu:27-46    known answer {0, 3, 3, 5}
//   test_cusparseWrapper.cu:46-75    known answer {0, 1, 4, 6, 9}
//   test_mlemOperations.cu:59-114    divides, dividesMultiplies, sum, scales
//   test_getWorkqueue.cu:124-219     bounded queue = prefix of unbounded
//   test_MeasurementSetupLinIndex.cu:57-66, test_VoxelGridLinIndex.cu:57-64
//   MeasurementSetupTrafo2CartCoordUT.cu:91-137, 309-320
//   RayGenerators.hpp:36-99          RandRayGen / ReguRayGen against the
//                                    device's rays of the same streams
//   reco.cu:238-259                  systemMatrixCalculation called with
//                                    the reference's template list and
//                                    arguments, then CSRmv on its ECSR
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "raptr/hotpath.hpp"

using raptr::check;

static int g_fail = 0;
#define EXPECT(cond)                                                             \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++g_fail;                                                                  \
    }                                                                            \
  } while (0)

typedef float val_t;
typedef DefaultMeasurementSetup<val_t> MS;
typedef DefaultVoxelGrid<val_t> VG;
typedef DefaultMeasurementList<val_t> ML;
typedef DefaultVoxelGridIdx<VG> Idx;
typedef DefaultVoxelGridIdy<VG> Idy;
typedef DefaultVoxelGridIdz<VG> Idz;
typedef DefaultMeasurementSetupId0z<MS> Id0z;
typedef DefaultMeasurementSetupId0y<MS> Id0y;
typedef DefaultMeasurementSetupId1z<MS> Id1z;
typedef DefaultMeasurementSetupId1y<MS> Id1y;
typedef DefaultMeasurementSetupIda<MS> Ida;
typedef DefaultMeasurementSetupTrafo2CartCoordFirstPixel<val_t, MS> Trafo0;
typedef DefaultMeasurementSetupTrafo2CartCoordSecndPixel<val_t, MS> Trafo1;
typedef DefaultMeasurementSetupTrafo2CartCoordFirstPixel_inplace<val_t, MS> Trafo0_inplace;
typedef DefaultMeasurementSetupTrafo2CartCoordSecndPixel_inplace<val_t, MS> Trafo1_inplace;
typedef uint32_t ListSizeType;                   // typedefs_array_sizes.hpp:44
typedef unsigned long long int MemArrSizeType;   // typedefs_array_sizes.hpp:36

// ---------------------------------------------------------------- host
static void test_index_round_trips() {
  MS ms(-230, 230, 5, 8, 8, 8, 8, 36, 20, 4, 4);
  DefaultMeasurementSetupLinId<MS> lin;
  for (int c = 0; c < ms.nChannels(); c += 7) {
    int a = Ida()(c, &ms), z0 = Id0z()(c, &ms), y0 = Id0y()(c, &ms), z1 = Id1z()(c, &ms),
        y1 = Id1y()(c, &ms);
    EXPECT(lin(z0, y0, z1, y1, a, &ms) == c);
  }
  VG vg(0, 0, 0, 1, 1, 1, 5, 7, 3);
  DefaultVoxelGridLinId<VG> vlin;
  for (int v = 0; v < vg.nVoxels(); ++v)
    EXPECT(vlin(Idx()(v, &vg), Idy()(v, &vg), Idz()(v, &vg), &vg) == v);
}

static void test_trafo_closed_forms() {
  MS ms(-35.0, 35.0, 6, 4, 4, 4, 4, 60.0, 10.0, 10.0, 10.0);
  const double box[][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {.1, .1, .1}};
  for (int det = 0; det < 2; ++det)
    for (int iz = 0; iz < 4; ++iz)
      for (int iy = 0; iy < 4; ++iy)
        for (auto& b : box)
          for (int a = 0; a < 6; ++a) {
            val_t in[3] = {val_t(b[0]), val_t(b[1]), val_t(b[2])}, out[3];
            if (det == 0)
              Trafo0()(out, in, iz, iy, a, &ms);
            else
              Trafo1()(out, in, iz, iy, a, &ms);
            double x = (det ? 35.0 : -35.0) + (-0.5 + b[0]) * 10.0;
            double y = ((-0.5 * 4 + iy) + b[1]) * 10.0;
            double z = ((-0.5 * 4 + iz) + b[2]) * 10.0;
            double th = a * 60.0 * M_PI / 180.0;
            double wx = std::sin(th) * z + std::cos(th) * x;
            double wz = std::cos(th) * z - std::sin(th) * x;
            EXPECT(std::fabs(out[0] - wx) < 5e-5 && std::fabs(out[1] - y) < 5e-5 &&
                   std::fabs(out[2] - wz) < 5e-5);
          }
}

static void test_chunk_math() {
  EXPECT((nInChunk<int, int>(0, 10, 4) == 4));
  EXPECT((nInChunk<int, int>(2, 10, 4) == 2));
  EXPECT((nInChunk<int, int>(3, 10, 4) == 0));
  EXPECT((chunkPtr<int, int>(2, 4) == 8));
  EXPECT((nChunks<int, long>(10, 4) == 3));
}

// -------------------------------------------------------------- device
static void test_coo_sort() {
  {  // dense 2x2, shuffled
    val_t v[] = {4, 1, 3, 2};
    int r[] = {1, 0, 1, 0}, c[] = {1, 0, 0, 1};
    raptr::DeviceBuffer<val_t> dv(4);
    raptr::DeviceBuffer<int> dr(4), dc(4);
    dv.upload(v, 4), dr.upload(r, 4), dc.upload(c, 4);
    cooSort<val_t>(dv.get(), dr.get(), dc.get(), 4);
    dv.download(v, 4), dr.download(r, 4), dc.download(c, 4);
    EXPECT(r[0] == 0 && c[0] == 0 && v[0] == 1);
    EXPECT(r[1] == 0 && c[1] == 1 && v[1] == 2);
    EXPECT(r[2] == 1 && c[2] == 0 && v[2] == 3);
    EXPECT(r[3] == 1 && c[3] == 1 && v[3] == 4);
  }
  cooSort<val_t>(nullptr, nullptr, nullptr, 0);  // empty matrix: no-op
  {  // sparse 60x50 at density 0.7, shuffled
    std::mt19937 rng(5);
    std::vector<int> r, c;
    std::vector<val_t> v;
    for (int i = 0; i < 60; ++i)
      for (int j = 0; j < 50; ++j)
        if (std::uniform_real_distribution<double>(0, 1)(rng) < 0.7) {
          r.push_back(i), c.push_back(j), v.push_back(val_t(i * 1000 + j));
        }
    std::vector<size_t> p(r.size());
    std::iota(p.begin(), p.end(), 0);
    std::shuffle(p.begin(), p.end(), rng);
    std::vector<int> rs(p.size()), cs(p.size());
    std::vector<val_t> vs(p.size());
    for (size_t k = 0; k < p.size(); ++k) rs[k] = r[p[k]], cs[k] = c[p[k]], vs[k] = v[p[k]];
    size_t n = rs.size();
    raptr::DeviceBuffer<val_t> dv(n);
    raptr::DeviceBuffer<int> dr(n), dc(n);
    dv.upload(vs.data(), n), dr.upload(rs.data(), n), dc.upload(cs.data(), n);
    cooSort<val_t>(dv.get(), dr.get(), dc.get(), int(n));
    dv.download(vs.data(), n), dr.download(rs.data(), n), dc.download(cs.data(), n);
    for (size_t k = 0; k < n; ++k) {
      EXPECT(vs[k] == val_t(rs[k] * 1000 + cs[k]));
      if (k) EXPECT(rs[k - 1] < rs[k] || (rs[k - 1] == rs[k] && cs[k - 1] < cs[k]));
    }
  }
}

static void test_coo2csr_and_ecsr() {
  int cooRow[] = {0, 1, 1, 1, 2, 2, 3, 3, 3};
  raptr::DeviceBuffer<int> dr(9), dptr(5);
  dr.upload(cooRow, 9);
  raptr_sparse_handle_t h;
  raptr_mat_descr_t d;
  raptr_sparse_create(&h);
  raptr_sparse_create_descr(&d);
  EXPECT(customizeMatDescr(d, h) == RAPTR_SPARSE_STATUS_SUCCESS);
  EXPECT(convertCoo2Csr(dptr.get(), dr.get(), h, 9, 4) == RAPTR_SPARSE_STATUS_SUCCESS);
  int ptr[5];
  dptr.download(ptr, 5);
  const int want[] = {0, 1, 4, 6, 9};
  EXPECT(std::equal(ptr, ptr + 5, want));
  // CSRmv: the 4x5 matrix of test_cusparseWrapper.cu:40-46
  val_t vals[] = {1, 5, 4, 6, 3, 2, 8, 7, 9};
  int cols[] = {1, 0, 2, 4, 1, 3, 0, 3, 4};
  val_t x[] = {1, 2, 3, 4, 5}, y[4] = {0, 0, 0, 0};
  raptr::DeviceBuffer<val_t> dv(9), dx(5), dy(4);
  raptr::DeviceBuffer<int> dc(9);
  dv.upload(vals, 9), dc.upload(cols, 9), dx.upload(x, 5), dy.upload(y, 4);
  val_t one = 1, zero = 0;
  EXPECT((CSRmv<val_t>()(h, RAPTR_SPARSE_OPERATION_NON_TRANSPOSE, 4, 5, 9, &one, d, dv.get(),
                         dptr.get(), dc.get(), dx.get(), &zero, dy.get()) ==
          RAPTR_SPARSE_STATUS_SUCCESS));
  dy.download(y, 4);
  EXPECT(y[0] == 2 && y[1] == 5 + 12 + 30 && y[2] == 6 + 8 && y[3] == 8 + 28 + 45);
  val_t e[] = {1, 1, 1, 1}, c[5] = {0, 0, 0, 0, 0};
  raptr::DeviceBuffer<val_t> de(4), dcv(5);
  de.upload(e, 4), dcv.upload(c, 5);
  EXPECT((CSRmv<val_t>()(h, RAPTR_SPARSE_OPERATION_TRANSPOSE, 4, 5, 9, &one, d, dv.get(),
                         dptr.get(), dc.get(), de.get(), &one, dcv.get()) ==
          RAPTR_SPARSE_STATUS_SUCCESS));
  dcv.download(c, 5);
  EXPECT(c[0] == 13 && c[1] == 4 && c[2] == 4 && c[3] == 9 && c[4] == 15);
  raptr_sparse_destroy_descr(d);
  raptr_sparse_destroy(h);
  // convertCsr2Ecsr known answer
  int yRid[] = {2, 4, 6}, aRptr[] = {0, 0, 0, 3, 3, 3, 3, 5};
  raptr::DeviceBuffer<int> dy2(3), dr2(8), de2(4);
  dy2.upload(yRid, 3), dr2.upload(aRptr, 8);
  convertCsr2Ecsr(de2.get(), dy2.get(), 3, dr2.get(), 7);
  int ecsr[4];
  de2.download(ecsr, 4);
  const int wantE[] = {0, 3, 3, 5};
  EXPECT(std::equal(ecsr, ecsr + 4, wantE));
}

static void test_mlem_operations() {
  const int N = 10;
  val_t A[N], B[N], C[N], D[N];
  for (int i = 0; i < N; ++i) A[i] = val_t((i + 1) * (i + 1)), C[i] = val_t(i + 1), B[i] = 10;
  raptr::DeviceBuffer<val_t> dA(N), dB(N), dC(N), dD(N);
  dA.upload(A, N), dB.upload(B, N), dC.upload(C, N);
  val_t *a = dA.get(), *b = dB.get(), *c = dC.get(), *d = dD.get();
  divides<val_t>(d, a, c, N);
  dD.download(D, N);
  for (int i = 0; i < N; ++i) EXPECT(D[i] == val_t(i + 1));
  dividesMultiplies<val_t>(d, a, b, c, N);
  dD.download(D, N);
  for (int i = 0; i < N; ++i) EXPECT(D[i] == val_t(10 * (i + 1)));
  val_t norm = sum<val_t>(d, N);
  EXPECT(norm == 550);
  scales<val_t>(d, val_t(1. / norm), N);
  dD.download(D, N);
  for (int i = 0; i < N; ++i) EXPECT(std::fabs(D[i] - (i + 1) / 55.0) < 1e-6);
}

static void test_workqueue_prefix() {
  MS ms(-230, 230, 5, 8, 8, 8, 8, 36, 20, 4, 4);
  VG vg(-40, -40, -40, 2.5, 2.5, 2.5, 32, 32, 32);
  ML ml(40);
  for (int i = 0; i < 40; ++i) ml.set(i, 123 * i, 1);
  std::vector<int> fc, fv;
  int li = 0, vi = 0;
  int nall = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida, Trafo0,
                          Trafo1>(fc, fv, li, vi, &ml, &vg, &ms);
  EXPECT(nall == int(fc.size()) && li == 40 && vi == 0 && nall > 0);
  for (int nb : {1, 999, 40000}) {
    std::vector<int> pc, pv;
    int l2 = 0, v2 = 0;
    int found = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida,
                             Trafo0, Trafo1>(pc, pv, l2, v2, &ml, &vg, &ms, nb);
    EXPECT(found == std::min(nb, nall));
    for (int k = 0; k < found; ++k) EXPECT(pc[k] == fc[k] && pv[k] == fv[k]);
    std::vector<int> rc, rv;
    getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida, Trafo0,
                 Trafo1>(rc, rv, l2, v2, &ml, &vg, &ms);
    EXPECT(int(pc.size() + rc.size()) == nall);
    for (size_t k = 0; k < rc.size(); ++k) EXPECT(rc[k] == fc[found + k]);
  }
  // matrix from the workqueue: finite, non-negative, some non-zero
  raptr::DeviceBuffer<int> dc(fc.size()), dv(fv.size());
  raptr::DeviceBuffer<val_t> dval(fc.size());
  dc.upload(fc.data(), fc.size()), dv.upload(fv.data(), fv.size());
  getSystemMatrixFromWorkqueue<val_t, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida, Trafo0,
                               Trafo1>(dc.get(), dv.get(), dval.get(), int(fc.size()), 10, vg,
                                       ms);
  std::vector<val_t> val(fc.size());
  dval.download(val.data(), val.size());
  int nz = 0;
  for (val_t a : val) {
    EXPECT(std::isfinite(a) && a >= 0);
    nz += a > 0;
  }
  EXPECT(nz > 0);
}

// RayGenerators.hpp:42-63 usage: a generator per (seed, setNr), one call per
// ray; with setNr = channel its rays are the device's RAND rays of that
// channel, bit for bit; ReguRayGen gives the pixel centres.
static void test_ray_generators() {
  MS ms(-230, 230, 5, 8, 8, 8, 8, 36, 20, 4, 4);
  raptr::setConstant(setup_const, ms);
  const int nrays = 10;
  const std::vector<int> chans = {0, 1, 777, 12345, ms.nChannels() - 1};
  raptr::DeviceBuffer<int> dc(chans.size());
  dc.upload(chans.data(), chans.size());
  raptr::DeviceBuffer<val_t> dr(chans.size() * nrays * 6);
  raptr_setup_t s = ms.to_c();
  std::vector<val_t> dev(chans.size() * nrays * 6);
  Trafo0_inplace trafo0;
  Trafo1_inplace trafo1;
  for (int kind : {RAPTR_RAYGEN_RAND, RAPTR_RAYGEN_REGU}) {
    raptr::check(raptr_generate_rays(&s, dc.get(), int(chans.size()), nrays, 1, kind, dr.get(),
                                     nullptr),
                 "raptr_generate_rays");
    dr.download(dev.data(), dev.size());
    for (size_t i = 0; i < chans.size(); ++i) {
      const int c = chans[i];
      const int ida = Ida()(c, &ms), id0z = Id0z()(c, &ms), id0y = Id0y()(c, &ms),
                id1z = Id1z()(c, &ms), id1y = Id1y()(c, &ms);
      RandRayGen<val_t, Trafo0_inplace, Trafo1_inplace> rand(1, c);
      ReguRayGen<val_t, Trafo0_inplace, Trafo1_inplace> regu(1, c);
      for (int r = 0; r < nrays; ++r) {
        val_t ray[6];
        if (kind == RAPTR_RAYGEN_RAND)
          rand(ray, id0z, id0y, id1z, id1y, ida, trafo0, trafo1);
        else
          regu(ray, id0z, id0y, id1z, id1y, ida, trafo0, trafo1);
        for (int q = 0; q < 6; ++q) EXPECT(ray[q] == dev[(i * nrays + r) * 6 + q]);
      }
    }
  }
  // an element stream differs from the channel stream
  RandRayGen<val_t, Trafo0_inplace, Trafo1_inplace> a(1, 777), b(1, 777);
  b.setElement(42);
  val_t ra[6], rb[6];
  a(ra, 0, 1, 2, 3, 4, trafo0, trafo1);
  b(rb, 0, 1, 2, 3, 4, trafo0, trafo1);
  EXPECT(std::memcmp(ra, rb, sizeof(ra)) != 0);
}

// reco.cu:238-259: chunks of the measured channel list through
// systemMatrixCalculation<val_t, ListSizeType, int, MemArrSizeType,
// RandRayGen<...> > with the constants set as in reco.cu:84-93, then the
// forward CSRmv on the ECSR.
static void test_system_matrix_calculation() {
  MS setup(-230, 230, 5, 8, 8, 8, 8, 36, 20, 4, 4);
  VG grid(-40, -40, -40, 2.5, 2.5, 2.5, 32, 32, 32);
  const int nrays = 10;
  raptr::setConstant(nrays_const, nrays);
  raptr::setConstant(setup_const, setup);
  raptr::setConstant(grid_const, grid);
  const int VGRIDSIZE = grid.nVoxels(), NCHANNELS = setup.nChannels();
  // measured channels (ascending), chunks of LIMM rows
  std::vector<int> yRowId_host;
  for (int c = 3; c < NCHANNELS; c += 97) yRowId_host.push_back(c);
  const ListSizeType effM = ListSizeType(yRowId_host.size()), LIMM = 64;
  const MemArrSizeType LIMNNZ = MemArrSizeType(LIMM) * MemArrSizeType(VGRIDSIZE);
  raptr::DeviceBuffer<int> yRowId(effM);
  yRowId.upload(yRowId_host.data(), effM);
  raptr::DeviceBuffer<int> ecsr(LIMM + 1), vxl(2 * LIMNNZ), cnlId(2 * LIMNNZ), csr(NCHANNELS + 1);
  raptr::DeviceBuffer<val_t> val(2 * LIMNNZ), x(VGRIDSIZE), yTilde(LIMM);
  raptr::DeviceBuffer<MemArrSizeType> nnz(1);
  int* aEcsrCnlPtr_devi = ecsr.get();
  int* aVxlId_devi = vxl.get();
  int* aCnlId_devi = cnlId.get();
  int* aCsrCnlPtr_devi = csr.get();
  int* yRowId_devi = yRowId.get();
  val_t* aVal_devi = val.get();
  MemArrSizeType* nnz_devi = nnz.get();
  raptr_sparse_handle_t handle;
  raptr_sparse_create(&handle);
  raptr_mat_descr_t A;
  raptr_sparse_create_descr(&A);
  customizeMatDescr(A, handle);
  std::vector<val_t> x_host(VGRIDSIZE);
  for (int j = 0; j < VGRIDSIZE; ++j) x_host[j] = val_t(1 + (j % 7));
  x.upload(x_host.data(), VGRIDSIZE);
  val_t* x_devi = x.get();
  val_t* yTildeVal_devi = yTilde.get();
  val_t one = 1, zero = 0;
  raptr_setup_t s = setup.to_c();
  raptr_grid_t g = grid.to_c();
  const unsigned NChunks = nChunks<unsigned, MemArrSizeType>(MemArrSizeType(effM), LIMM);
  for (unsigned chunkId = 0; chunkId < NChunks; ++chunkId) {
    ListSizeType m = nInChunk(chunkId, effM, LIMM);
    ListSizeType ptr = chunkPtr(chunkId, LIMM);
    MemArrSizeType nnz_host[1] = {0};
    check(raptr_memcpy_h2d(nnz_devi, nnz_host, sizeof(MemArrSizeType), nullptr), "h2d");
    systemMatrixCalculation<val_t, ListSizeType, int, MemArrSizeType,
                            RandRayGen<val_t, Trafo0_inplace, Trafo1_inplace> >(
        aEcsrCnlPtr_devi, aVxlId_devi, aVal_devi, nnz_devi, aCnlId_devi, aCsrCnlPtr_devi,
        &(yRowId_devi[ptr]), &m, handle);
    check(raptr_memcpy_d2h(nnz_host, nnz_devi, sizeof(MemArrSizeType), nullptr), "d2h");
    // the reference's nnz: the workqueue length of the chunk's channels
    std::vector<int64_t> cnt(m);
    raptr::DeviceBuffer<int64_t> dcnt(m);
    check(raptr_workqueue_counts(&s, &g, yRowId_devi + ptr, int(m), dcnt.get(), nullptr),
          "raptr_workqueue_counts");
    dcnt.download(cnt.data(), m);
    const int64_t wq = std::accumulate(cnt.begin(), cnt.end(), int64_t(0));
    EXPECT(int64_t(nnz_host[0]) == wq);
    const int64_t nnz_c = int64_t(nnz_host[0]);
    std::vector<int> e(m + 1), col(nnz_c), rows(nnz_c), c(NCHANNELS + 1);
    std::vector<val_t> v(nnz_c);
    ecsr.download(e.data(), m + 1);
    vxl.download(col.data(), nnz_c);
    cnlId.download(rows.data(), nnz_c);
    csr.download(c.data(), NCHANNELS + 1);
    val.download(v.data(), nnz_c);
    EXPECT(e[0] == 0 && e[m] == nnz_c);
    for (ListSizeType i = 0; i < m; ++i) {
      EXPECT(e[i + 1] - e[i] == cnt[i]);
      const int ch = yRowId_host[ptr + i];
      EXPECT(c[ch + 1] - c[ch] == cnt[i] && c[ch] == e[i]);
      for (int k = e[i]; k < e[i + 1]; ++k) {
        EXPECT(rows[k] == ch);
        if (k > e[i]) EXPECT(col[k] > col[k - 1]);
      }
    }
    EXPECT(c[0] == 0 && c[NCHANNELS] == nnz_c);
    // values: the two-stage kernel on the same pairs with per-element rays
    raptr::DeviceBuffer<val_t> v2(nnz_c);
    check(raptr_sm_from_wq(&s, &g, aCnlId_devi, aVxlId_devi, v2.get(), nnz_c, nrays,
                           raptr::random_seed, RAPTR_RAYGEN_RAND_ELEMENT, nullptr),
          "raptr_sm_from_wq");
    std::vector<val_t> w(nnz_c);
    v2.download(w.data(), nnz_c);
    EXPECT(std::memcmp(w.data(), v.data(), sizeof(val_t) * nnz_c) == 0);
    // forward projection of the chunk (reco.cu:257-259)
    EXPECT(CSRmv<val_t>()(handle, RAPTR_SPARSE_OPERATION_NON_TRANSPOSE, int(m), VGRIDSIZE,
                          int(nnz_c), &one, A, aVal_devi, aEcsrCnlPtr_devi, aVxlId_devi, x_devi,
                          &zero, yTildeVal_devi) == RAPTR_SPARSE_STATUS_SUCCESS);
    std::vector<val_t> yt(m);
    yTilde.download(yt.data(), m);
    for (ListSizeType i = 0; i < m; ++i) {
      double want = 0.0;
      for (int k = e[i]; k < e[i + 1]; ++k) want += double(v[k]) * x_host[col[k]];
      EXPECT(std::fabs(yt[i] - want) <= 1e-5 * std::fabs(want) + 1e-6);
    }
  }
  // nnz_devi accumulates: a second call appends after the first's entries
  {
    ListSizeType m = 5;
    MemArrSizeType nnz_host[1] = {0};
    check(raptr_memcpy_h2d(nnz_devi, nnz_host, sizeof(MemArrSizeType), nullptr), "h2d");
    systemMatrixCalculation<val_t, ListSizeType, int, MemArrSizeType,
                            ReguRayGen<val_t, Trafo0_inplace, Trafo1_inplace> >(
        aEcsrCnlPtr_devi, aVxlId_devi, aVal_devi, nnz_devi, aCnlId_devi, aCsrCnlPtr_devi,
        yRowId_devi, &m, handle);
    MemArrSizeType first = 0, second = 0;
    check(raptr_memcpy_d2h(&first, nnz_devi, sizeof(first), nullptr), "d2h");
    systemMatrixCalculation<val_t, ListSizeType, int, MemArrSizeType,
                            ReguRayGen<val_t, Trafo0_inplace, Trafo1_inplace> >(
        aEcsrCnlPtr_devi, aVxlId_devi, aVal_devi, nnz_devi, aCnlId_devi, aCsrCnlPtr_devi,
        yRowId_devi + 5, &m, handle);
    check(raptr_memcpy_d2h(&second, nnz_devi, sizeof(second), nullptr), "d2h");
    std::vector<int> e(m + 1);
    ecsr.download(e.data(), m + 1);
    EXPECT(first > 0 && second > first && MemArrSizeType(e[0]) == first &&
           MemArrSizeType(e[m]) == second);
  }
  raptr_sparse_destroy_descr(A);
  raptr_sparse_destroy(handle);
}

int main(int argc, char** argv) {
  bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu-only") == 0;
  test_index_round_trips();
  test_trafo_closed_forms();
  test_chunk_math();
  if (!cpu_only) {
    try {
      test_coo_sort();
      test_coo2csr_and_ecsr();
      test_mlem_operations();
      test_workqueue_prefix();
      test_ray_generators();
      test_system_matrix_calculation();
    } catch (const std::exception& e) {
      std::fprintf(stderr, "exception: %s\n", e.what());
      return 2;
    }
  }
  std::printf("%s: %d failure(s)\n", cpu_only ? "cpu" : "all", g_fail);
  return g_fail ? 1 : 0;
}
