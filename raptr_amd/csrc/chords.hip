// chords.hip -- Siddon ray traversal (getChords) on gfx950.
//
// One thread per ray walks the grid planes in parameter order and writes
// (chord length, 3-D voxel index) per traversed voxel: the behaviour of
// ChordsCalc_lowlevel.hpp:194-379 with its helpers :395-676, with each axis
// started at its first plane strictly after the grid entry (the reference
// starts at the plane after the axis' first grid plane, which lies before
// the entry for all but the entry axis), the output bounded by the caller's
// capacity, the last chord clipped at the exit and the stepping bounded by
// the grid's plane count.  Float operations follow the host restatement in
// order, so results are bit-identical to it.
#include "primitives.hpp"

namespace raptr {
namespace {

__global__ __launch_bounds__(256) void k_get_chords(const float* __restrict__ rays, int nrays,
                                                    raptr_grid_t g, int cap,
                                                    float* __restrict__ chords,
                                                    int* __restrict__ ids,
                                                    int* __restrict__ counts) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  float ray[6];
  for (int k = 0; k < 6; ++k) ray[k] = rays[6 * int64_t(r) + k];
  const float* go = g.o;
  const float* gd = g.d;
  const int* gn = g.n;
  float* ch = chords + int64_t(r) * cap;
  int* id3 = ids + int64_t(r) * cap * 3;
  int crosses[3];
  int good_par = 1;
  for (int d = 0; d < 3; ++d) {
    crosses[d] = (ray[d + 3] - ray[d] != 0.0f);
    good_par &= crosses[d] || (ray[d] >= go[d] && ray[d] < go[d] + gd[d] * float(gn[d]));
  }
  for (int d = 0; d < 3; ++d) crosses[d] &= good_par;
  float amin_d[3], amax_d[3], up[3];
  int idup[3];
  for (int d = 0; d < 3; ++d) {
    float div = ray[d + 3] - ray[d];
    div += float(div == 0.0f);
    float a0 = (go[d] + 0.0f * gd[d] - ray[d]) / div;
    float a1 = (go[d] + float(gn[d]) * gd[d] - ray[d]) / div;
    amin_d[d] = a0 < a1 ? a0 : a1;
    amax_d[d] = a0 < a1 ? a1 : a0;
    float dd = sqrtf((ray[d + 3] - ray[d]) * (ray[d + 3] - ray[d]));
    dd += float(dd == 0.0f);
    up[d] = gd[d] / dd;
    idup[d] = (ray[d] < ray[d + 3]) ? 1 : -1;
  }
  int good = 0;
  float amin = 0.0f, amax = 0.0f;
  for (int d = 0; d < 3; ++d) {
    if (!crosses[d]) continue;
    amin = good ? (amin < amin_d[d] ? amin_d[d] : amin) : amin_d[d];
    amax = good ? (amax_d[d] < amax ? amax_d[d] : amax) : amax_d[d];
    good = 1;
  }
  if (!good || !(amin >= 0.0f && amin <= 1.0f) || !(amax >= 0.0f && amax <= 1.0f) ||
      amin > amax) {
    counts[r] = 0;
    return;
  }
  float next[3];
  for (int d = 0; d < 3; ++d) {
    next[d] = amin_d[d] + up[d];
    if (crosses[d])
      for (int k = 0; next[d] <= amin && k < gn[d] + 4; ++k) next[d] += up[d];
  }
  float anext = 2.0f;
  for (int d = 0; d < 3; ++d)
    if (crosses[d] && next[d] < anext) anext = next[d];
  const float amid = 0.5f * (amin + (anext < amax ? anext : amax));
  int id[3];
  for (int d = 0; d < 3; ++d)
    id[d] = int(floorf((ray[d] + amid * (ray[d + 3] - ray[d]) - go[d]) / gd[d]));
  const float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
  const float length = sqrtf(dx * dx + dy * dy + dz * dz);
  float acur = amin;
  int n = 0;
  const int max_steps = gn[0] + gn[1] + gn[2] + 8;
  for (int step = 0; acur < amax && step < max_steps; ++step) {
    anext = 2.0f;
    for (int d = 0; d < 3; ++d)
      if (crosses[d] && next[d] < anext) anext = next[d];
    if (anext > amax) anext = amax;
    const bool inside = id[0] >= 0 && id[0] < gn[0] && id[1] >= 0 && id[1] < gn[1] &&
                        id[2] >= 0 && id[2] < gn[2];
    if (inside && n < cap) {
      ch[n] = (anext - acur) * length;
      id3[3 * n + 0] = id[0];
      id3[3 * n + 1] = id[1];
      id3[3 * n + 2] = id[2];
    }
    if (inside) ++n;
    acur = anext;
    for (int d = 0; d < 3; ++d) {
      if (crosses[d] && next[d] == anext) {
        next[d] += up[d];
        id[d] += idup[d];
      }
    }
  }
  counts[r] = n < cap ? n : cap;
}

}  // namespace
}  // namespace raptr

using namespace raptr;

extern "C" int raptr_get_chords(const float* d_rays, int nrays, const raptr_grid_t* grid,
                                int cap, float* d_chords, int* d_voxel_ids, int* d_counts,
                                raptr_stream_t stream) {
  if (!grid || nrays < 0 || cap < 0) return fail(RAPTR_ERR_INVALID_ARG, "bad arguments");
  for (int d = 0; d < 3; ++d)
    if (grid->n[d] <= 0 || !(grid->d[d] > 0.0f))
      return fail(RAPTR_ERR_INVALID_ARG, "grid needs positive counts and pitches");
  if (nrays == 0) return RAPTR_OK;
  if (!d_rays || !d_counts || (cap > 0 && (!d_chords || !d_voxel_ids)))
    return fail(RAPTR_ERR_INVALID_ARG, "null buffer");
  hipLaunchKernelGGL(k_get_chords, dim3((nrays + 255) / 256), dim3(256), 0, S(stream), d_rays,
                     nrays, *grid, cap, d_chords, d_voxel_ids, d_counts);
  RAPTR_LAUNCH_CHECK();
  return RAPTR_OK;
}
