// runtime.hip -- status/error plumbing, device memory and stream helpers,
// host-side geometry tables.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "common.hpp"

namespace raptr {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int make_geom(const raptr_setup_t* setup, const raptr_grid_t* grid, Geom* G) {
  if (!setup || !grid || !G) return fail(RAPTR_ERR_INVALID_ARG, "null geometry");
  const raptr_setup_t& s = *setup;
  const raptr_grid_t& g = *grid;
  if (s.na <= 0 || s.na > RAPTR_MAX_ANGLES)
    return fail(RAPTR_ERR_INVALID_ARG, "na=%d outside [1, %d]", s.na, RAPTR_MAX_ANGLES);
  if (s.n0z <= 0 || s.n0y <= 0 || s.n1z <= 0 || s.n1y <= 0)
    return fail(RAPTR_ERR_INVALID_ARG, "pixel counts must be positive");
  int64_t nch = int64_t(s.na) * s.n0z * s.n0y * s.n1z * s.n1y;
  if (nch > INT32_MAX) return fail(RAPTR_ERR_INVALID_ARG, "channel count exceeds int32");
  for (int k = 0; k < 3; ++k)
    if (g.n[k] <= 0 || !(g.d[k] > 0.0f))
      return fail(RAPTR_ERR_INVALID_ARG, "grid dimension %d invalid", k);
  int64_t nv = int64_t(g.n[0]) * g.n[1] * g.n[2];
  if (nv > INT32_MAX / 2) return fail(RAPTR_ERR_INVALID_ARG, "voxel count too large");
  std::memset(G, 0, sizeof(Geom));
  G->s = s;
  G->g = g;
  G->nvox = int(nv);
  G->nch = int(nch);
  G->sumRadii = raptr_sum_radii(setup, grid);
  raptr_angle_table(setup, G->sn, G->cs);
  return RAPTR_OK;
}

}  // namespace raptr

using namespace raptr;

namespace {
// plain streaming copy (16-byte accesses when aligned): the known-size
// dispatch for bandwidth and traffic-counter calibration
__global__ void k_stream_copy(float* __restrict__ dst, const float* __restrict__ src, int64_t n,
                              int vec) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t t0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (vec) {
    typedef float vf4_t __attribute__((ext_vector_type(4)));
    const int64_t n4 = n / 4;
    for (int64_t i = t0; i < n4; i += stride)
      reinterpret_cast<vf4_t*>(dst)[i] = reinterpret_cast<const vf4_t*>(src)[i];
    for (int64_t i = 4 * n4 + t0; i < n; i += stride) dst[i] = src[i];
  } else {
    for (int64_t i = t0; i < n; i += stride) dst[i] = src[i];
  }
}
}  // namespace

extern "C" {

const char* raptr_status_string(int st) {
  switch (st) {
    case RAPTR_OK: return "ok";
    case RAPTR_ERR_INVALID_ARG: return "invalid argument";
    case RAPTR_ERR_DEVICE: return "device error";
    case RAPTR_ERR_OUT_OF_MEMORY: return "out of memory";
    case RAPTR_ERR_BUDGET: return "HBM budget exceeded";
    case RAPTR_ERR_WORKSPACE: return "workspace too small";
    case RAPTR_ERR_NOT_SUPPORTED: return "not supported";
    default: return "unknown status";
  }
}

const char* raptr_last_error(void) { return g_err; }
int raptr_version(void) { return 100; }

int raptr_device_count(int* count) {
  RAPTR_HIP_TRY(hipGetDeviceCount(count));
  return RAPTR_OK;
}
int raptr_set_device(int device) {
  RAPTR_HIP_TRY(hipSetDevice(device));
  return RAPTR_OK;
}
int raptr_device_name(char* buf, int len) {
  int dev = 0;
  RAPTR_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  RAPTR_HIP_TRY(hipGetDeviceProperties(&p, dev));
  snprintf(buf, size_t(len), "%s (%s, %d CUs)", p.name, p.gcnArchName,
           p.multiProcessorCount);
  return RAPTR_OK;
}
int raptr_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(RAPTR_ERR_INVALID_ARG, "null ptr");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    return fail(RAPTR_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): out of memory", bytes);
  }
  RAPTR_HIP_TRY(e);
  return RAPTR_OK;
}
int raptr_free(void* ptr) {
  if (ptr) RAPTR_HIP_TRY(hipFree(ptr));
  return RAPTR_OK;
}
int raptr_memcpy_h2d(void* dst, const void* src, size_t bytes, raptr_stream_t s) {
  if (!bytes) return RAPTR_OK;
  RAPTR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
  RAPTR_HIP_TRY(hipStreamSynchronize(S(s)));
  return RAPTR_OK;
}
int raptr_memcpy_d2h(void* dst, const void* src, size_t bytes, raptr_stream_t s) {
  if (!bytes) return RAPTR_OK;
  RAPTR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
  RAPTR_HIP_TRY(hipStreamSynchronize(S(s)));
  return RAPTR_OK;
}
int raptr_memcpy_d2d(void* dst, const void* src, size_t bytes, raptr_stream_t s) {
  if (!bytes) return RAPTR_OK;
  RAPTR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return RAPTR_OK;
}
int raptr_stream_copy(float* dst, const float* src, int64_t n, raptr_stream_t s) {
  if (n < 0 || (n > 0 && (!dst || !src))) return fail(RAPTR_ERR_INVALID_ARG, "bad copy");
  if (n == 0) return RAPTR_OK;
  const bool vec = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
  hipLaunchKernelGGL(k_stream_copy, dim3(grid_for(n / 4 + 1, 256, 1 << 16)), dim3(256), 0, S(s),
                     dst, src, n, vec ? 1 : 0);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}

int raptr_memset(void* dst, int value, size_t bytes, raptr_stream_t s) {
  if (!bytes) return RAPTR_OK;
  RAPTR_HIP_TRY(hipMemsetAsync(dst, value, bytes, S(s)));
  return RAPTR_OK;
}
int raptr_stream_create(raptr_stream_t* s) {
  hipStream_t h;
  RAPTR_HIP_TRY(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = h;
  return RAPTR_OK;
}
int raptr_stream_destroy(raptr_stream_t s) {
  if (s) RAPTR_HIP_TRY(hipStreamDestroy(S(s)));
  return RAPTR_OK;
}
int raptr_stream_sync(raptr_stream_t s) {
  RAPTR_HIP_TRY(hipStreamSynchronize(S(s)));
  return RAPTR_OK;
}
int raptr_device_sync(void) {
  RAPTR_HIP_TRY(hipDeviceSynchronize());
  return RAPTR_OK;
}
int raptr_event_create(void** ev) {
  hipEvent_t e;
  RAPTR_HIP_TRY(hipEventCreate(&e));
  *ev = e;
  return RAPTR_OK;
}
int raptr_event_destroy(void* ev) {
  if (ev) RAPTR_HIP_TRY(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
  return RAPTR_OK;
}
int raptr_event_record(void* ev, raptr_stream_t s) {
  RAPTR_HIP_TRY(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), S(s)));
  return RAPTR_OK;
}
int raptr_stream_wait_event(raptr_stream_t s, void* ev) {
  if (!ev) return fail(RAPTR_ERR_INVALID_ARG, "null event");
  RAPTR_HIP_TRY(hipStreamWaitEvent(S(s), reinterpret_cast<hipEvent_t>(ev), 0));
  return RAPTR_OK;
}
int raptr_event_elapsed_ms(float* ms, void* a, void* b) {
  RAPTR_HIP_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(b)));
  RAPTR_HIP_TRY(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(a),
                                    reinterpret_cast<hipEvent_t>(b)));
  return RAPTR_OK;
}
int raptr_mem_info(size_t* free_bytes, size_t* total_bytes) {
  RAPTR_HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return RAPTR_OK;
}

int raptr_angle_table(const raptr_setup_t* setup, float* sn, float* cs) {
  if (!setup || !sn || !cs) return fail(RAPTR_ERR_INVALID_ARG, "null argument");
  const float pi_f = 3.14159265358979323846f;
  for (int a = 0; a < setup->na; ++a) {
    float arg = pi_f * float(a);
    arg = arg * setup->da;
    arg = arg / 180.0f;
    // the host C library's sinf/cosf of the float argument: the values the
    // reference's host code uses (glibc's are not correctly rounded at every
    // angle, so a double-precision evaluation would differ by an ulp there)
    sn[a] = ::sinf(arg);
    cs[a] = ::cosf(arg);
  }
  return RAPTR_OK;
}

float raptr_sum_radii(const raptr_setup_t* s, const raptr_grid_t* g) {
  float v = absolute3(g->d[0], g->d[1], g->d[2]);
  float p = absolute3(s->segx, s->segy, s->segz);
  return 0.5f * (v + p);
}

}  // extern "C"
