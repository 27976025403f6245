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
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "raptr/hotpath.hpp"

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
    } catch (const std::exception& e) {
      std::fprintf(stderr, "exception: %s\n", e.what());
      return 2;
    }
  }
  std::printf("%s: %d failure(s)\n", cpu_only ? "cpu" : "all", g_fail);
  return g_fail ? 1 : 0;
}
