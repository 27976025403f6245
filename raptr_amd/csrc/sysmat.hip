// sysmat.hip -- Monte Carlo ray-sampled system matrix on gfx950.
//
// Two entry points follow the reference:
//  * raptr_sm_from_wq: the two-stage path (getSystemMatrixFromWorkqueue.cu:
//    116-281), one thread per workqueue entry, rays of the entry's channel
//    regenerated from the counter-based Philox stream.
//  * raptr_sm_build_csr: the production path (what getSystemMatrix +
//    cooSort + coo2csr + csr2ecsr produce, getSystemMatrixDeviceOnly.cu:
//    357-562, wrappers.hpp:185-266), one workgroup per channel:
//      1. the channel's rays are generated once into LDS, with per-ray plane
//         tables alpha[dim][k] = (o + k*d - r0)/(r1 - r0) for every grid plane
//         (the voxel-box divisions of the slab method depend only on the
//         plane index, so one division per plane and ray replaces six per
//         voxel and ray, with identical float results);
//      2. the cylinder cull runs over the grid in rounds of 1024 voxels and
//         compacts candidates into LDS with wave ballots (ascending voxel
//         order preserved);
//      3. candidates are sampled densely (all lanes busy), exact zeros are
//         dropped with a second ballot compaction into an LDS row buffer;
//      4. row buffers are flushed as pieces into a staging area and packed
//         into a CSR with 64-bit row pointers per channel chunk.
#include <cmath>
#include <mutex>

#include "primitives.hpp"

namespace raptr {

// Scratch of the allocation-free build (raptr_sm_build_csr_noalloc): the
// staging area, counters, piece table and scan space of the fused build in
// device memory that exists from module load to process end, one copy per
// device.  `scratch` is a host-side stand-in in host code: the code below
// takes the array's device address once per device (hipGetSymbolAddress,
// with that device current) and uses only that address.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// the current device's address of `scratch` (looked up on its first use)
int scratch_address(void** out) {
  int dev = 0;
  RAPTR_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(RAPTR_ERR_NOT_SUPPORTED, "device %d beyond the %d scratch slots", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(RAPTR_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RAPTR_OK;
}

constexpr int kSmBlock = 256;
constexpr int kCandCap = 3072;  // >= 2048 + one cull round (1024)
constexpr int kCandFlush = 2048;
constexpr int kRowBuf = 3072;   // >= kCandCap: one batch always fits
constexpr int kCullRound = 1024;

// ------------------------------------------------- two-stage kernels
__global__ __launch_bounds__(256) void k_sm_from_wq(Geom G, const int* cnl, const int* vxl,
                                                    float* val, int64_t n, int nrays,
                                                    uint32_t seed, int raygen) {
  for (int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; w < n;
       w += int64_t(gridDim.x) * blockDim.x) {
    int c = cnl[w], vv = vxl[w];
    if (c < 0 || c >= G.nch || vv < 0 || vv >= G.nvox) {
      val[w] = 0.0f;  // out-of-range pair: no geometry, no value
      continue;
    }
    int ix, iy, iz;
    vxl_split(G.g, vv, ix, iy, iz);
    float box[6] = {plane(G.g, 0, ix),     plane(G.g, 1, iy),     plane(G.g, 2, iz),
                    plane(G.g, 0, ix + 1), plane(G.g, 1, iy + 1), plane(G.g, 2, iz + 1)};
    float a = 0.0f;
    for (int r = 0; r < nrays; ++r) {
      float ray[6];
      make_ray(G, raygen, seed, c, r, ray, vv);
      a += slab_length(box, ray);
    }
    a /= float(nrays);
    val[w] = a;
  }
}

__global__ void k_generate_rays(Geom G, const int* cnl, int n_cnl, int nrays, uint32_t seed,
                                int raygen, float* rays) {
  int64_t total = int64_t(n_cnl) * nrays;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += int64_t(gridDim.x) * blockDim.x) {
    int64_t ci = k / nrays;
    int r = int(k - ci * nrays);
    int c = cnl[ci];
    if (c < 0 || c >= G.nch) {
      for (int q = 0; q < 6; ++q) rays[6 * k + q] = 0.0f;
      continue;
    }
    make_ray(G, raygen, seed, c, r, rays + 6 * k);
  }
}

// ----------------------------------------------------- fused build
struct Piece {
  int row;       // row within chunk
  int within;    // offset of the piece inside its row
  int64_t src;   // staging offset
  int len;
  int pad;
};

struct BuildArgs {
  Geom G;
  FastDiv div_yz, div_z;
  const int* list;  // nullptr: channel = list index (all channels)
  int first;        // first list entry of this launch
  int rows;         // list entries in this launch
  int nrays;
  uint32_t seed;
  int raygen;
  int keep_zeros;   // store zero-valued candidates too (RAPTR_SM_KEEP_ZEROS)
  int pl;           // planes per ray (nx+1 + ny+1 + nz+1)
  // outputs
  int* row_nnz;     // [rows]
  int* st_col;      // staging
  float* st_val;
  int64_t st_cap;
  unsigned long long* st_count;
  Piece* pieces;
  int piece_cap;
  unsigned* piece_count;
  unsigned* flags;  // bit0 staging overflow, bit1 piece overflow
  double* colsum;   // sensitivity / back-projection mode
  // sensitivity: exact fixed-point column sums (value * 2^fx_bits rounded to
  // an integer, added with 64-bit integer atomics, so the result does not
  // depend on the order in which the atomics land); null: FP64 atomics
  unsigned long long* colsum_fx;
  double fx_scale;
  const float* wgt; // back-projection mode: weight per list entry (nullptr: 1)
  const float* xin; // forward-projection mode: image
  double* fwd;      // forward-projection mode: per-row FP64 projection
};

enum { kModeCount = 0, kModeFill = 1, kModeColsum = 2, kModeFwd = 3 };

__device__ inline float sample_voxel(const float* rinfo, const float* alpha, int nrays,
                                     int pl, int ix, int iy, int iz, int off1, int off2) {
  float a = 0.0f;
  for (int r = 0; r < nrays; ++r) {
    const float* ri = rinfo + 8 * r;
    const float* al = alpha + size_t(r) * pl;
    int bits = __float_as_int(ri[7]);
    float amin = 0.0f, amax = 0.0f;
    bool good = false;
    if (bits & 1) {
      float lo = al[ix], hi = al[ix + 1];
      if (hi < lo) {
        float t = lo;
        lo = hi;
        hi = t;
      }
      amin = lo;
      amax = hi;
      good = true;
    }
    if (bits & 2) {
      float lo = al[off1 + iy], hi = al[off1 + iy + 1];
      if (hi < lo) {
        float t = lo;
        lo = hi;
        hi = t;
      }
      amin = good ? (amin < lo ? lo : amin) : lo;
      amax = good ? (hi < amax ? hi : amax) : hi;
      good = true;
    }
    if (bits & 4) {
      float lo = al[off2 + iz], hi = al[off2 + iz + 1];
      if (hi < lo) {
        float t = lo;
        lo = hi;
        hi = t;
      }
      amin = good ? (amin < lo ? lo : amin) : lo;
      amax = good ? (hi < amax ? hi : amax) : hi;
    }
    float len = (amin < amax) ? (amax - amin) * ri[6] : 0.0f;
    a += len;
  }
  return a / float(nrays);
}

// RAPTR_RAYGEN_RAND_ELEMENT: the element's own nrays rays, slab method on
// the voxel box (the arithmetic of k_sm_from_wq, so both paths agree bit
// for bit).
__device__ inline float sample_element(const Geom& G, uint32_t seed, int nrays, int cnl,
                                       int v, int ix, int iy, int iz) {
  float box[6] = {plane(G.g, 0, ix),     plane(G.g, 1, iy),     plane(G.g, 2, iz),
                  plane(G.g, 0, ix + 1), plane(G.g, 1, iy + 1), plane(G.g, 2, iz + 1)};
  float a = 0.0f;
  for (int r = 0; r < nrays; ++r) {
    float ray[6];
    make_ray(G, RAPTR_RAYGEN_RAND_ELEMENT, seed, cnl, r, ray, v);
    a += slab_length(box, ray);
  }
  return a / float(nrays);
}

template <int MODE>
__global__ __launch_bounds__(kSmBlock) void k_sm_build(BuildArgs A) {
  extern __shared__ float smem[];
  __shared__ int wcnt[kSmBlock / kWave];
  __shared__ int s_scan[kSmBlock / kWave];
  __shared__ int s_lo[kSmBlock], s_pre[kSmBlock];
  __shared__ long long s_off;
  const int nr = A.nrays, pl = A.pl;
  float* rinfo = smem;                                   // nr * 8
  float* alpha = rinfo + 8 * nr;                         // nr * pl
  int* cand = reinterpret_cast<int*>(alpha + size_t(nr) * pl);
  float* cval = reinterpret_cast<float*>(cand + kCandCap);
  int* rb_col = reinterpret_cast<int*>(cval + kCandCap);
  float* rb_val = reinterpret_cast<float*>(rb_col + kRowBuf);

  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const int nx = A.G.g.n[0], ny = A.G.g.n[1];
  const int off1 = nx + 1, off2 = nx + 1 + ny + 1;
  const int nvox = A.G.nvox;

  for (int rl = blockIdx.x; rl < A.rows; rl += gridDim.x) {
    const int li = A.first + rl;
    int cnl = A.list ? A.list[li] : li;
    // an out-of-range channel id yields an empty row (uniform branch)
    const bool bad = cnl < 0 || cnl >= A.G.nch;
    if (bad) cnl = 0;
    // 1. rays and plane tables
    for (int r = t; r < nr; r += kSmBlock) {
      float ray[6];
      make_ray(A.G, A.raygen, A.seed, cnl, r, ray);
      float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
      int bits = (ray[0] != ray[3] ? 1 : 0) | (ray[1] != ray[4] ? 2 : 0) |
                 (ray[2] != ray[5] ? 4 : 0);
      float* ri = rinfo + 8 * r;
      ri[0] = ray[0];
      ri[1] = ray[1];
      ri[2] = ray[2];
      ri[3] = dx;
      ri[4] = dy;
      ri[5] = dz;
      ri[6] = sqrtf(dx * dx + dy * dy + dz * dz);
      ri[7] = __int_as_float(bits);
    }
    __syncthreads();
    for (int k = t; k < nr * pl; k += kSmBlock) {
      int r = k / pl, p = k - r * pl;
      int dim = p < off1 ? 0 : (p < off2 ? 1 : 2);
      int idx = p - (dim == 0 ? 0 : (dim == 1 ? off1 : off2));
      const float* ri = rinfo + 8 * r;
      int bits = __float_as_int(ri[7]);
      float v = 0.0f;
      if (bits & (1 << dim)) v = (plane(A.G.g, dim, idx) - ri[dim]) / ri[3 + dim];
      alpha[k] = v;
    }
    float p0[3], p1[3];
    channel_centres(A.G, cnl, p0, p1);
    __syncthreads();

    int cand_n = 0, rb_n = 0, flushed = 0;
    double proj = 0.0;
    // sample, then keep (or count / accumulate) the candidates in cand[]
    auto process = [&]() {
      // 3. sample candidates densely
      for (int c = t; c < cand_n; c += kSmBlock) {
        int v = cand[c];
        uint32_t qx = fdiv(uint32_t(v), A.div_yz);
        uint32_t rem = uint32_t(v) - qx * A.div_yz.d;
        uint32_t qy = fdiv(rem, A.div_z);
        const int iz = int(rem - qy * A.div_z.d);
        cval[c] = A.raygen == RAPTR_RAYGEN_RAND_ELEMENT
                      ? sample_element(A.G, A.seed, nr, cnl, v, int(qx), int(qy), iz)
                      : sample_voxel(rinfo, alpha, nr, pl, int(qx), int(qy), iz, off1, off2);
      }
      __syncthreads();
      if (MODE == kModeColsum) {
        const double wt = A.wgt ? double(A.wgt[li]) : 1.0;
        for (int c = t; c < cand_n; c += kSmBlock) {
          float a = cval[c];
          if (A.colsum_fx) {
            if (a != 0.0f)
              atomicAdd(&A.colsum_fx[cand[c]],
                        (unsigned long long)llrint(double(a) * A.fx_scale));
          } else if (a != 0.0f && wt != 0.0) {
            atomicAdd(&A.colsum[cand[c]], double(a) * wt);
          }
        }
        __syncthreads();
        cand_n = 0;
        return;
      }
      if (MODE == kModeFwd) {
        for (int c = t; c < cand_n; c += kSmBlock)
          proj += double(cval[c]) * double(A.xin[cand[c]]);
        __syncthreads();
        cand_n = 0;
        return;
      }
      // flush the row buffer first if the candidates could overflow it
      if (MODE == kModeFill && rb_n + cand_n > kRowBuf) {
        if (t == 0) {
          unsigned long long off = atomicAdd(A.st_count, (unsigned long long)rb_n);
          long long o = (long long)off;
          if (o + rb_n > A.st_cap) {
            atomicOr(A.flags, 1u);
            o = -1;
          } else {
            unsigned pc = atomicAdd(A.piece_count, 1u);
            if (int(pc) >= A.piece_cap) {
              atomicOr(A.flags, 2u);
              o = -1;
            } else {
              A.pieces[pc] = Piece{rl, flushed, o, rb_n, 0};
            }
          }
          s_off = o;
        }
        __syncthreads();
        long long o = s_off;
        if (o >= 0)
          for (int k = t; k < rb_n; k += kSmBlock) {
            A.st_col[o + k] = rb_col[k];
            A.st_val[o + k] = rb_val[k];
          }
        flushed += rb_n;
        rb_n = 0;
        __syncthreads();
      }
      // compact non-zeros in candidate order
      for (int c0 = 0; c0 < cand_n; c0 += kSmBlock) {
        int c = c0 + t;
        bool nz = c < cand_n && (A.keep_zeros || cval[c] != 0.0f);
        uint64_t m = __ballot(nz);
        if (lane == 0) wcnt[w] = __popcll(m);
        __syncthreads();
        int pre = 0, tot = 0;
        for (int ww = 0; ww < kSmBlock / kWave; ++ww) {
          int cc = wcnt[ww];
          pre += ww < w ? cc : 0;
          tot += cc;
        }
        if (MODE == kModeFill && nz) {
          int pos = rb_n + pre + __popcll(m & lanemask_lt());
          rb_col[pos] = cand[c];
          rb_val[pos] = cval[c];
        }
        rb_n += tot;
        __syncthreads();
      }
      if (MODE == kModeCount) {
        flushed += rb_n;
        rb_n = 0;
      }
      cand_n = 0;
    };
    // 2. cull.  For a voxel column (ix, iy) the centres closer to the line
    // than the radius form one z-interval (a convex set meets a line in an
    // interval); it is found in FP64, widened by a relative radius slack and
    // one voxel at each end, and every voxel in it gets the exact float test
    // of the reference.  Columns run in ascending (ix, iy) and z ascends
    // within a column, so the candidates and their order are those of a
    // scan over all voxels.
    const int nz = A.G.g.n[2];
    const int ncols = nx * ny;
    double u[3], len2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      u[d] = double(p1[d]) - double(p0[d]);
      len2 += u[d] * u[d];
    }
    const double il = len2 > 0.0 ? 1.0 / sqrt(len2) : 0.0;
    for (int d = 0; d < 3; ++d) u[d] *= il;
    const double rr = double(A.G.sumRadii) * (1.0 + 1e-4) + 1e-4;
    const double qa = u[0] * u[0] + u[1] * u[1];
    for (int cb = 0; cb < (bad ? 0 : ncols); cb += kSmBlock) {
      const int cc = cb + t;
      int lo = 0, cnt = 0;
      if (cc < ncols) {
        const int ix = cc / ny, iy = cc - (cc / ny) * ny;
        const double ax = double(A.G.g.o[0]) + (ix + 0.5) * double(A.G.g.d[0]) - double(p0[0]);
        const double ay = double(A.G.g.o[1]) + (iy + 0.5) * double(A.G.g.d[1]) - double(p0[1]);
        const double sp = ax * u[0] + ay * u[1];
        const double qc = ax * ax + ay * ay - sp * sp - rr * rr;
        int hi = -1;
        if (qa < 1e-12) {
          if (qc < 0.0) lo = 0, hi = nz - 1;
        } else {
          const double qb = -2.0 * sp * u[2];
          const double disc = qb * qb - 4.0 * qa * qc;
          if (disc >= 0.0) {
            const double sq = sqrt(disc);
            const double z1 = double(p0[2]) + (-qb - sq) / (2.0 * qa);
            const double z2 = double(p0[2]) + (-qb + sq) / (2.0 * qa);
            const double f1 = (z1 - double(A.G.g.o[2])) / double(A.G.g.d[2]) - 0.5;
            const double f2 = (z2 - double(A.G.g.o[2])) / double(A.G.g.d[2]) - 0.5;
            lo = int(fmax(floor(f1) - 1.0, 0.0));
            hi = int(fmin(ceil(f2) + 1.0, double(nz - 1)));
          }
        }
        cnt = hi >= lo ? hi - lo + 1 : 0;
      }
      int tot;
      const int pre = block_exclusive_scan<int, kSmBlock>(cnt, s_scan, &tot);
      s_lo[t] = lo;
      s_pre[t] = pre;
      __syncthreads();
      for (int k0 = 0; k0 < tot; k0 += kSmBlock) {
        const int k = k0 + t;
        bool pass = false;
        int v = 0;
        if (k < tot) {
          int jl = 0, jh = kSmBlock - 1;  // last column with s_pre <= k
          while (jl < jh) {
            const int mid = (jl + jh + 1) >> 1;
            if (s_pre[mid] <= k) jl = mid; else jh = mid - 1;
          }
          const int col = cb + jl;
          const int ix = col / ny, iy = col - (col / ny) * ny;
          const int iz = s_lo[jl] + (k - s_pre[jl]);
          v = col * nz + iz;
          float c[3];
          voxel_centre(A.G.g, ix, iy, iz, c);
          pass = distance_point_line(p0, p1, c) < A.G.sumRadii;
        }
        uint64_t m = __ballot(pass);
        if (lane == 0) wcnt[w] = __popcll(m);
        __syncthreads();
        int wpre = 0, wtot = 0;
        for (int ww = 0; ww < kSmBlock / kWave; ++ww) {
          const int c = wcnt[ww];
          wpre += ww < w ? c : 0;
          wtot += c;
        }
        if (pass) cand[cand_n + wpre + __popcll(m & lanemask_lt())] = v;
        cand_n += wtot;
        __syncthreads();
        if (cand_n >= kCandFlush) process();
      }
      __syncthreads();
    }
    if (cand_n > 0) process();
    if (MODE == kModeFill && rb_n > 0) {
      if (t == 0) {
        unsigned long long off = atomicAdd(A.st_count, (unsigned long long)rb_n);
        long long o = (long long)off;
        if (o + rb_n > A.st_cap) {
          atomicOr(A.flags, 1u);
          o = -1;
        } else {
          unsigned pc = atomicAdd(A.piece_count, 1u);
          if (int(pc) >= A.piece_cap) {
            atomicOr(A.flags, 2u);
            o = -1;
          } else {
            A.pieces[pc] = Piece{rl, flushed, o, rb_n, 0};
          }
        }
        s_off = o;
      }
      __syncthreads();
      long long o = s_off;
      if (o >= 0)
        for (int k = t; k < rb_n; k += kSmBlock) {
          A.st_col[o + k] = rb_col[k];
          A.st_val[o + k] = rb_val[k];
        }
      flushed += rb_n;
      rb_n = 0;
    }
    if (MODE == kModeFwd) {
      // fixed-order block reduction of the per-thread FP64 partials
      proj = wave_reduce_sum(proj);
      __shared__ double wred[kSmBlock / kWave];
      if (lane == 0) wred[w] = proj;
      __syncthreads();
      if (t == 0) {
        double tot = 0.0;
        for (int ww = 0; ww < kSmBlock / kWave; ++ww) tot += wred[ww];
        A.fwd[li] = tot;
      }
    }
    if ((MODE == kModeCount || MODE == kModeFill) && t == 0) A.row_nnz[rl] = flushed;
    __syncthreads();
  }
}

// fixed-point column sums -> FP64, in place
__global__ void k_fx_to_f64(unsigned long long* v, int n, double inv_scale) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x) {
    const unsigned long long q = v[j];
    reinterpret_cast<double*>(v)[j] = double(q) * inv_scale;
  }
}

// row_ptr[first + i] = base + excl[i]
__global__ void k_rowptr_fill(const int64_t* excl, int rows, int64_t base, int64_t* row_ptr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows) row_ptr[i] = base + excl[i];
}

// one block per piece: copy staging -> final CSR position
__global__ __launch_bounds__(256) void k_pieces_copy(const Piece* pieces, int npieces,
                                                     const int64_t* row_ptr_chunk,
                                                     const int* st_col, const float* st_val,
                                                     int* col, float* val) {
  for (int p = blockIdx.x; p < npieces; p += gridDim.x) {
    Piece pc = pieces[p];
    int64_t dst = row_ptr_chunk[pc.row] + pc.within;
    for (int k = threadIdx.x; k < pc.len; k += 256) {
      col[dst + k] = st_col[pc.src + k];
      val[dst + k] = st_val[pc.src + k];
    }
  }
}

size_t build_lds_bytes(int nrays, int pl) {
  return size_t(nrays) * (8 + size_t(pl)) * sizeof(float) +
         size_t(kCandCap) * (sizeof(int) + sizeof(float)) +
         size_t(kRowBuf) * (sizeof(int) + sizeof(float));
}

int make_build_args(const raptr_setup_t* setup, const raptr_grid_t* grid, int nrays,
                    uint32_t seed, int raygen, BuildArgs* A, size_t* lds) {
  std::memset(A, 0, sizeof(BuildArgs));
  RAPTR_TRY(make_geom(setup, grid, &A->G));
  if (nrays <= 0) return fail(RAPTR_ERR_INVALID_ARG, "nrays must be positive");
  if (raygen != RAPTR_RAYGEN_RAND && raygen != RAPTR_RAYGEN_REGU &&
      raygen != RAPTR_RAYGEN_RAND_ELEMENT)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown ray generator %d", raygen);
  A->div_yz = make_fastdiv(uint32_t(grid->n[1] * grid->n[2]));
  A->div_z = make_fastdiv(uint32_t(grid->n[2]));
  A->nrays = nrays;
  A->seed = seed;
  A->raygen = raygen;
  A->pl = grid->n[0] + grid->n[1] + grid->n[2] + 3;
  *lds = build_lds_bytes(nrays, A->pl);
  if (*lds > 160 * 1024)
    return fail(RAPTR_ERR_NOT_SUPPORTED,
                "nrays=%d x planes=%d needs %zu B of LDS (> 160 KiB)", nrays, A->pl, *lds);
  return RAPTR_OK;
}

template <int MODE>
int launch_build(const BuildArgs& A, size_t lds, hipStream_t st) {
  static bool configured = false;
  if (!configured) {
    // Raise the dynamic-LDS limit once; not fatal where the default suffices.
    if (hipFuncSetAttribute((const void*)k_sm_build<MODE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024) != hipSuccess)
      (void)hipGetLastError();
    configured = true;
  }
  unsigned g = unsigned(A.rows);
  hipLaunchKernelGGL(k_sm_build<MODE>, dim3(g), dim3(kSmBlock), lds, st, A);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}


// Workspace layout of the fused build for a staging area of S entries:
// counters, staging (col, val), per-row nnz and scan, piece table, scan ws.
struct BuildLayout {
  int64_t staging, max_rows, piece_cap;
};

inline BuildLayout build_layout(int64_t S) {
  BuildLayout L;
  L.staging = S;
  L.max_rows = S / 16 + 1;
  if (L.max_rows > (1 << 20)) L.max_rows = 1 << 20;
  L.piece_cap = 2 * L.max_rows + 1024;
  return L;
}

template <class C>
void carve_build(C& c, const BuildLayout& L) {
  c.template take<unsigned long long>(4);
  c.template take<int>(size_t(L.staging));
  c.template take<float>(size_t(L.staging));
  c.template take<int>(size_t(L.max_rows));
  c.template take<int64_t>(size_t(L.max_rows));
  c.template take<char>(size_t(L.piece_cap) * 24);
  c.template take<char>(scan_workspace_bytes(L.max_rows));
}

inline size_t build_bytes(int64_t S) {
  Sizer z;
  carve_build(z, build_layout(S));
  return z.bytes();
}

}  // namespace
}  // namespace raptr

using namespace raptr;

extern "C" {

int raptr_sm_from_wq(const raptr_setup_t* setup, const raptr_grid_t* grid, const int* d_cnl,
                     const int* d_vxl, float* d_val, int64_t wq_length, int nrays,
                     uint32_t seed, int raygen, raptr_stream_t stream) {
  Geom G;
  RAPTR_TRY(make_geom(setup, grid, &G));
  if (nrays <= 0) return fail(RAPTR_ERR_INVALID_ARG, "nrays must be positive");
  if (wq_length < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative length");
  if (raygen != RAPTR_RAYGEN_RAND && raygen != RAPTR_RAYGEN_REGU &&
      raygen != RAPTR_RAYGEN_RAND_ELEMENT)
    return fail(RAPTR_ERR_INVALID_ARG, "unknown ray generator %d", raygen);
  if (wq_length == 0) return RAPTR_OK;
  if (!d_cnl || !d_vxl || !d_val) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_sm_from_wq, dim3(grid_for(wq_length, 256, 1 << 16)), dim3(256), 0,
                     S(stream), G, d_cnl, d_vxl, d_val, wq_length, nrays, seed, raygen);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_generate_rays(const raptr_setup_t* setup, const int* d_cnl, int n_cnl, int nrays,
                        uint32_t seed, int raygen, float* d_rays, raptr_stream_t stream) {
  raptr_grid_t g = {{0, 0, 0}, {1, 1, 1}, {1, 1, 1}};
  Geom G;
  RAPTR_TRY(make_geom(setup, &g, &G));
  if (raygen != RAPTR_RAYGEN_RAND && raygen != RAPTR_RAYGEN_REGU)
    return fail(RAPTR_ERR_INVALID_ARG, "ray listings take RAND or REGU (got %d)", raygen);
  if (n_cnl <= 0 || nrays <= 0) return RAPTR_OK;
  hipLaunchKernelGGL(k_generate_rays, dim3(grid_for(int64_t(n_cnl) * nrays, 256, 1 << 16)),
                     dim3(256), 0, S(stream), G, d_cnl, n_cnl, nrays, seed, raygen, d_rays);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

size_t raptr_sm_build_workspace_size(const raptr_grid_t* grid, int nrays,
                                     int64_t staging_entries) {
  (void)grid;
  (void)nrays;
  if (staging_entries < kRowBuf) staging_entries = kRowBuf;
  return build_bytes(staging_entries);
}

int raptr_sm_build_csr(const raptr_setup_t* setup, const raptr_grid_t* grid,
                       const int* d_list, int list_size, int nrays, uint32_t seed,
                       int raygen, int64_t* d_row_ptr, int* d_col, float* d_val,
                       int64_t capacity, int64_t* nnz, void* d_ws, size_t ws_bytes,
                       raptr_stream_t stream) {
  return raptr_sm_build_csr_ex(setup, grid, d_list, list_size, nrays, seed, raygen, 0,
                               d_row_ptr, d_col, d_val, capacity, nnz, d_ws, ws_bytes,
                               stream);
}

int raptr_sm_build_csr_ex(const raptr_setup_t* setup, const raptr_grid_t* grid,
                          const int* d_list, int list_size, int nrays, uint32_t seed,
                          int raygen, int flags, int64_t* d_row_ptr, int* d_col, float* d_val,
                          int64_t capacity, int64_t* nnz, void* d_ws, size_t ws_bytes,
                          raptr_stream_t stream) {
  if (!nnz || !d_row_ptr) return fail(RAPTR_ERR_INVALID_ARG, "null output");
  *nnz = 0;
  if (list_size < 0 || (list_size > 0 && !d_list))
    return fail(RAPTR_ERR_INVALID_ARG, "bad list");
  if (flags & ~int(RAPTR_SM_KEEP_ZEROS))
    return fail(RAPTR_ERR_INVALID_ARG, "unknown build flags 0x%x", flags);
  BuildArgs A;
  size_t lds = 0;
  RAPTR_TRY(make_build_args(setup, grid, nrays, seed, raygen, &A, &lds));
  A.keep_zeros = (flags & RAPTR_SM_KEEP_ZEROS) ? 1 : 0;
  hipStream_t st = S(stream);
  const bool fill = d_col != nullptr;
  if (fill && !d_val) return fail(RAPTR_ERR_INVALID_ARG, "null values");
  // largest staging area whose layout fits the workspace
  int64_t lo = 0, hi = int64_t(ws_bytes / 8) + 1;
  while (lo + 1 < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (build_bytes(mid) <= ws_bytes) lo = mid; else hi = mid;
  }
  if (lo < kRowBuf || !d_ws)
    return fail(RAPTR_ERR_WORKSPACE, "sm_build workspace too small (%zu B)", ws_bytes);
  BuildLayout L = build_layout(lo);
  Carver cv(d_ws, ws_bytes);
  unsigned long long* counters = cv.take<unsigned long long>(4);
  int* st_col = cv.take<int>(size_t(L.staging));
  float* st_val = cv.take<float>(size_t(L.staging));
  int* row_nnz = cv.take<int>(size_t(L.max_rows));
  int64_t* excl = cv.take<int64_t>(size_t(L.max_rows));
  Piece* pieces = reinterpret_cast<Piece*>(cv.take<char>(size_t(L.piece_cap) * 24));
  size_t sb = scan_workspace_bytes(L.max_rows);
  void* scan_ws = cv.take<char>(sb);
  if (!cv.ok) return fail(RAPTR_ERR_WORKSPACE, "sm_build workspace carve failed");
  const int64_t staging = L.staging, max_rows = L.max_rows;
  A.list = d_list;
  A.row_nnz = row_nnz;
  A.st_col = st_col;
  A.st_val = st_val;
  A.st_cap = staging;
  A.st_count = &counters[0];
  A.piece_count = reinterpret_cast<unsigned*>(&counters[1]);
  A.flags = reinterpret_cast<unsigned*>(&counters[2]);
  A.pieces = pieces;

  int64_t total = 0;
  int first = 0;
  int rows = int(max_rows < 4096 ? max_rows : 4096);
  if (rows > list_size) rows = list_size;
  while (first < list_size) {
    if (rows > list_size - first) rows = list_size - first;
    A.first = first;
    A.rows = rows;
    A.piece_cap = int(L.piece_cap);
    RAPTR_HIP_TRY(hipMemsetAsync(counters, 0, 4 * sizeof(unsigned long long), st));
    if (fill)
      RAPTR_TRY(launch_build<kModeFill>(A, lds, st));
    else
      RAPTR_TRY(launch_build<kModeCount>(A, lds, st));
    unsigned long long hc[4];
    RAPTR_HIP_TRY(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    RAPTR_HIP_TRY(hipStreamSynchronize(st));
    unsigned flags = unsigned(hc[2] & 0xffffffffu);
    if (flags) {
      if (rows == 1)
        return fail(RAPTR_ERR_WORKSPACE, "row %d does not fit the staging area", first);
      rows = rows / 2;
      continue;
    }
    int64_t chunk_nnz = 0;
    RAPTR_TRY(scan_exclusive_i32_i64(row_nnz, excl, rows, &reinterpret_cast<int64_t*>(counters)[3],
                                     scan_ws, sb, st));
    hipLaunchKernelGGL(k_rowptr_fill, dim3(grid_for(rows, 256)), dim3(256), 0, st,
                       (const int64_t*)excl, rows, total, d_row_ptr + first);
    RAPTR_LAUNCH_CHECK();
    RAPTR_HIP_TRY(hipMemcpyAsync(&chunk_nnz, &counters[3], sizeof(int64_t),
                                 hipMemcpyDeviceToHost, st));
    RAPTR_HIP_TRY(hipStreamSynchronize(st));
    if (fill) {
      if (total + chunk_nnz > capacity)
        return fail(RAPTR_ERR_WORKSPACE, "matrix exceeds capacity %lld after row %d",
                    (long long)capacity, first + rows);
      unsigned np = unsigned(hc[1] & 0xffffffffu);
      if (np)
        hipLaunchKernelGGL(k_pieces_copy, dim3(np < 65536 ? np : 65536), dim3(256), 0, st,
                           (const Piece*)pieces, int(np), (const int64_t*)(d_row_ptr + first),
                           (const int*)st_col, (const float*)st_val, d_col, d_val);
      RAPTR_LAUNCH_CHECK();
    }
    total += chunk_nnz;
    first += rows;
    // next chunk: fill ~60% of the staging at the observed density
    double avg = rows ? double(chunk_nnz) / rows : 0.0;
    int64_t want = avg > 0 ? int64_t(0.6 * double(staging) / (avg * 2.0)) : max_rows;
    if (want > max_rows) want = max_rows;
    if (want < 1) want = 1;
    rows = int(want);
  }
  int64_t tot = total;
  RAPTR_HIP_TRY(hipMemcpyAsync(d_row_ptr + list_size, &tot, sizeof(int64_t),
                               hipMemcpyHostToDevice, st));
  RAPTR_HIP_TRY(hipStreamSynchronize(st));
  *nnz = total;
  return RAPTR_OK;
}

int raptr_sm_build_csr_noalloc(const raptr_setup_t* setup, const raptr_grid_t* grid,
                               const int* d_list, int list_size, int nrays, uint32_t seed,
                               int raygen, int flags, int64_t* d_row_ptr, int* d_col,
                               float* d_val, int64_t capacity, int64_t* nnz,
                               raptr_stream_t stream) {
  void* ws = nullptr;
  RAPTR_TRY(scratch_address(&ws));
  const int st = raptr_sm_build_csr_ex(setup, grid, d_list, list_size, nrays, seed, raygen,
                                       flags, d_row_ptr, d_col, d_val, capacity, nnz, ws,
                                       kNoAllocScratchBytes, stream);
  if (st == RAPTR_ERR_WORKSPACE && std::strstr(raptr_last_error(), "staging")) {
    char why[256];
    std::snprintf(why, sizeof(why), "%s", raptr_last_error());  // fail() rewrites it
    return fail(RAPTR_ERR_WORKSPACE,
                "raptr_sm_build_csr_noalloc: a row has more candidates than the %zu-byte "
                "scratch stages (%s)",
                kNoAllocScratchBytes, why);
  }
  return st;
}

int raptr_sensitivity_f64(const raptr_setup_t* setup, const raptr_grid_t* grid, int nrays,
                          uint32_t seed, int raygen, double* d_sens, raptr_stream_t stream) {
  BuildArgs A;
  size_t lds = 0;
  RAPTR_TRY(make_build_args(setup, grid, nrays, seed, raygen, &A, &lds));
  if (!d_sens) return fail(RAPTR_ERR_INVALID_ARG, "null output");
  hipStream_t st = S(stream);
  RAPTR_HIP_TRY(hipMemsetAsync(d_sens, 0, size_t(A.G.nvox) * sizeof(double), st));
  int64_t nch = int64_t(setup->na) * setup->n0z * setup->n0y * setup->n1z * setup->n1y;
  // Fixed point with `bits` fraction bits: no column sum exceeds (channels)
  // x (voxel diagonal, the longest mean chord), so 2^bits times that bound
  // stays below 2^62.  A float value of at least 2^(23 - bits) is a whole
  // multiple of 2^-bits (exact); smaller ones are rounded by at most
  // 2^-(bits + 1).  The integer sums are the same whatever the order of the
  // additions, so the sensitivity is reproducible bit for bit.
  const double diag = std::sqrt(double(grid->d[0]) * grid->d[0] +
                                double(grid->d[1]) * grid->d[1] +
                                double(grid->d[2]) * grid->d[2]);
  const double bound = double(nch > 0 ? nch : 1) * (diag > 0.0 ? diag : 1.0);
  int bits = int(std::floor(62.0 - std::log2(bound)));
  if (bits > 52) bits = 52;
  if (bits < 8)
    return fail(RAPTR_ERR_NOT_SUPPORTED, "sensitivity bound %.3g too large for fixed point",
                bound);
  A.list = nullptr;
  A.wgt = nullptr;
  A.colsum = d_sens;
  A.colsum_fx = reinterpret_cast<unsigned long long*>(d_sens);
  A.fx_scale = std::ldexp(1.0, bits);
  const int64_t kRowsPerLaunch = 1 << 20;
  for (int64_t f = 0; f < nch; f += kRowsPerLaunch) {
    A.first = int(f);
    A.rows = int(nch - f < kRowsPerLaunch ? nch - f : kRowsPerLaunch);
    RAPTR_TRY(launch_build<kModeColsum>(A, lds, st));
  }
  hipLaunchKernelGGL(k_fx_to_f64, dim3(grid_for(A.G.nvox, 256)), dim3(256), 0, st,
                     A.colsum_fx, A.G.nvox, std::ldexp(1.0, -bits));
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_backproject_f64(const raptr_setup_t* setup, const raptr_grid_t* grid,
                          const int* d_list, const float* d_weights, int list_size, int nrays,
                          uint32_t seed, int raygen, double* d_out, raptr_stream_t stream) {
  BuildArgs A;
  size_t lds = 0;
  RAPTR_TRY(make_build_args(setup, grid, nrays, seed, raygen, &A, &lds));
  if (!d_out || list_size < 0 || (list_size > 0 && (!d_list || !d_weights)))
    return fail(RAPTR_ERR_INVALID_ARG, "bad back-projection arguments");
  hipStream_t st = S(stream);
  RAPTR_HIP_TRY(hipMemsetAsync(d_out, 0, size_t(A.G.nvox) * sizeof(double), st));
  A.list = d_list;
  A.wgt = d_weights;
  A.colsum = d_out;
  const int64_t kRowsPerLaunch = 1 << 20;
  for (int64_t f = 0; f < list_size; f += kRowsPerLaunch) {
    A.first = int(f);
    A.rows = int(list_size - f < kRowsPerLaunch ? list_size - f : kRowsPerLaunch);
    RAPTR_TRY(launch_build<kModeColsum>(A, lds, st));
  }
  return RAPTR_OK;
}

int raptr_forward_project_f64(const raptr_setup_t* setup, const raptr_grid_t* grid,
                              const int* d_list, int list_size, int nrays, uint32_t seed,
                              int raygen, const float* d_x, double* d_out,
                              raptr_stream_t stream) {
  BuildArgs A;
  size_t lds = 0;
  RAPTR_TRY(make_build_args(setup, grid, nrays, seed, raygen, &A, &lds));
  if (!d_x || !d_out) return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  int64_t nch = int64_t(setup->na) * setup->n0z * setup->n0y * setup->n1z * setup->n1y;
  int64_t rows = d_list ? int64_t(list_size) : nch;
  if (rows < 0) return fail(RAPTR_ERR_INVALID_ARG, "negative list size");
  A.list = d_list;
  A.xin = d_x;
  A.fwd = d_out;
  const int64_t kRowsPerLaunch = 1 << 20;
  for (int64_t f = 0; f < rows; f += kRowsPerLaunch) {
    A.first = int(f);
    A.rows = int(rows - f < kRowsPerLaunch ? rows - f : kRowsPerLaunch);
    RAPTR_TRY(launch_build<kModeFwd>(A, lds, S(stream)));
  }
  return RAPTR_OK;
}

}  // extern "C"
