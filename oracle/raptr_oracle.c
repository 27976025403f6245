/*
 * raptr_oracle.c -- CPU restatement of the raptr reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see raptr_oracle.h): loaded by the test-suite,
 * the smoke check and the benchmark's CPU-baseline leg; never by the product.
 *
 * Build: gcc -O2 -std=c11 -fPIC -shared -fopenmp -ffp-contract=off
 *        -fno-fast-math (oracle/Makefile).  No FMA contraction anywhere: the
 *        FP32 results must match the GPU kernels bit for bit.
 */
#include "raptr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes for the pin tests (tests/test_oracle_pins.py): a build
 * with -DRO_MUTANT=k carries mistake k; the default build has none. */
#ifndef RO_MUTANT
#define RO_MUTANT 0
#endif

#define RO_PI_F 3.14159265358979323846f

/* ======================================================================
 * A2: channel index maps.  MeasurementSetupLinIndex.hpp:78-86 (linId),
 * :117-124 (ida), :156-165 (id0z), :196-206 (id0y), :237-248 (id1z),
 * :278-320 (id1y).  Order of volatility: (ida, id0z, id0y, id1z, id1y).
 * ==================================================================== */
int64_t ro_nchannels(const ro_setup* s) {
  return (int64_t)s->na * s->n0z * s->n0y * s->n1z * s->n1y;
}

void ro_cnl_split(const ro_setup* s, int cnl, int* ida, int* id0z, int* id0y,
                  int* id1z, int* id1y) {
  int p4 = s->n1y * s->n1z * s->n0y * s->n0z;
  int p3 = s->n1y * s->n1z * s->n0y;
  int p2 = s->n1y * s->n1z;
  int p1 = s->n1y;
  int t = cnl;
  *ida = t / p4;
  t %= p4;
  *id0z = t / p3;
  t %= p3;
  *id0y = t / p2;
  t %= p2;
  *id1z = t / p1;
  *id1y = t % p1;
  if (RO_MUTANT == 2) { /* transposed index */
    int q = *id1z;
    *id1z = *id1y;
    *id1y = q;
  }
}

int ro_cnl_lin(const ro_setup* s, int id0z, int id0y, int id1z, int id1y,
               int ida) {
  return ida * (s->n1y * s->n1z * s->n0y * s->n0z) +
         id0z * (s->n1y * s->n1z * s->n0y) + id0y * (s->n1y * s->n1z) +
         id1z * (s->n1y) + id1y;
}

/* A3: voxel index maps, VoxelGridLinIndex.hpp:73-80 (lin = idx*ny*nz +
 * idy*nz + idz, z fastest), :114-120, :154-162, :196-205. */
void ro_vxl_split(const ro_grid* g, int vxl, int* ix, int* iy, int* iz) {
  int pyz = g->n[1] * g->n[2];
  *ix = vxl / pyz;
  int t = vxl % pyz;
  *iy = t / g->n[2];
  *iz = t % g->n[2];
}

int ro_vxl_lin(const ro_grid* g, int ix, int iy, int iz) {
  return ix * (g->n[2] * g->n[1]) + iy * g->n[2] + iz;
}

/* ======================================================================
 * A4: angle table.  The reference evaluates sinf/cosf of
 * float(pi)*ida*da/180 inside every trafo call
 * (MeasurementSetupTrafo2CartCoord.hpp:110-111).  The argument keeps that
 * float evaluation order and the function value is the C library's
 * sinf/cosf of that float argument, i.e. the reference's host evaluation
 * (glibc's sinf/cosf are not correctly rounded everywhere: at 8 of the 180
 * one-degree angles they differ from the rounded double-precision value by
 * one ulp).  Tabulated per angle on the host and shared by the oracle and
 * every GPU kernel.
 * ==================================================================== */
void ro_angle_table(const ro_setup* s, float* sn, float* cs) {
  for (int a = 0; a < s->na; ++a) {
    float arg = RO_PI_F * (float)a;
    arg = arg * s->da;
    arg = arg / 180.0f;
    sn[a] = sinf(arg);
    cs[a] = cosf(arg);
  }
}

/* Pixel unit-cube -> world coordinates.
 * MeasurementSetupTrafo2CartCoord.hpp:94-148 (det 0) and :305-359 (det 1):
 * pixel centre y/z = (id - 0.5*(n-1))*seg, x = pos; affine 3x4 with a
 * rotation about y by the channel angle; products summed left to right. */
void ro_trafo(const ro_setup* s, int det, float sn, float cs, int idz, int idy,
              const float box[3], float out[3]) {
  float e0 = s->segx, e1 = s->segy, e2 = s->segz;
  float p0, p1, p2;
  if (det == 0) {
    p0 = s->pos0x;
    p1 = ((float)idy - 0.5f * (float)(s->n0y - 1)) * e1;
    p2 = ((float)idz - 0.5f * (float)(s->n0z - 1)) * e2;
  } else {
    p0 = s->pos1x;
    p1 = ((float)idy - 0.5f * (float)(s->n1y - 1)) * e1;
    p2 = ((float)idz - 0.5f * (float)(s->n1z - 1)) * e2;
  }
  float t[12];
  t[0] = cs * e0;
  t[1] = 0.0f;
  t[2] = sn * e2;
  t[3] = cs * (p0 - 0.5f * e0) + sn * (p2 - 0.5f * e2);
  t[4] = 0.0f;
  t[5] = e1;
  t[6] = 0.0f;
  t[7] = p1 - 0.5f * e1;
  t[8] = RO_MUTANT == 1 ? sn * e0 : -sn * e0; /* 1: wrong sign */
  t[9] = 0.0f;
  t[10] = cs * e2;
  t[11] = -sn * (p0 - 0.5f * e0) + cs * (p2 - 0.5f * e2);
  float b0 = box[0], b1 = box[1], b2 = box[2];
  out[0] = t[0] * b0 + t[1] * b1 + t[2] * b2 + t[3] * 1.0f;
  out[1] = t[4] * b0 + t[5] * b1 + t[6] * b2 + t[7] * 1.0f;
  out[2] = t[8] * b0 + t[9] * b1 + t[10] * b2 + t[11] * 1.0f;
}

/* ======================================================================
 * A5: distance of p from the line through a and b,
 * distancePointLine.h:38-43 (absolute), :65-72 (cross), :87-101 (distance).
 * ==================================================================== */
static float ro_absolute(const float v[3]) {
  return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

float ro_distance(const float a[3], const float b[3], const float p[3]) {
  float ap[3], ab[3], c[3];
  for (int d = 0; d < 3; ++d) {
    ap[d] = p[d] - a[d];
    ab[d] = b[d] - a[d];
  }
  for (int d = 0; d < 3; ++d)
    c[d] = ab[(d + 1) % 3] * ap[(d + 2) % 3] - ab[(d + 2) % 3] * ap[(d + 1) % 3];
  return ro_absolute(c) / ro_absolute(RO_MUTANT == 9 ? ap : ab); /* 9: wrong operand */
}

/* ======================================================================
 * A6/A7: workqueue.  getWorkqueue.hpp:39-133 (entry scan), :135-191
 * (bounded), :193-243 (unbounded).  Candidate iff the distance of the voxel
 * centre from the line through the two pixel centres is strictly below
 * 0.5*(|voxel edges| + |pixel edges|).
 * ==================================================================== */
typedef struct {
  float pix0[3], pix1[3];
} ro_cnl_geom;

static float ro_sum_radii(const ro_setup* s, const ro_grid* g) {
  float ve[3] = {g->d[0], g->d[1], g->d[2]};
  float pe[3] = {s->segx, s->segy, s->segz};
  return 0.5f * (ro_absolute(ve) + (RO_MUTANT == 3 ? 0.0f : ro_absolute(pe))); /* 3: dropped term */
}

static void ro_channel_centres(const ro_setup* s, const float* sn,
                               const float* cs, int cnl, ro_cnl_geom* cg) {
  int ida, id0z, id0y, id1z, id1y;
  ro_cnl_split(s, cnl, &ida, &id0z, &id0y, &id1z, &id1y);
  const float c[3] = {0.5f, 0.5f, 0.5f};
  ro_trafo(s, 0, sn[ida], cs[ida], id0z, id0y, c, cg->pix0);
  ro_trafo(s, 1, sn[ida], cs[ida], id1z, id1y, c, cg->pix1);
}

static inline void ro_voxel_centre(const ro_grid* g, int vxl, float vc[3]) {
  int ix, iy, iz;
  ro_vxl_split(g, vxl, &ix, &iy, &iz);
  vc[0] = g->o[0] + ((float)ix + 0.5f) * g->d[0];
  vc[1] = g->o[1] + ((float)iy + 0.5f) * g->d[1];
  vc[2] = g->o[2] + ((float)iz + 0.5f) * g->d[2];
}

int64_t ro_workqueue(const ro_setup* s, const ro_grid* g, const int* list,
                     int list_size, int* list_id, int* vxl_id, int64_t n,
                     int* cnl_out, int* vxl_out) {
  int nvox = g->n[0] * g->n[1] * g->n[2];
  float* sn = (float*)malloc(sizeof(float) * s->na);
  float* cs = (float*)malloc(sizeof(float) * s->na);
  ro_angle_table(s, sn, cs);
  float rsum = ro_sum_radii(s, g);
  int64_t found = 0;
  int li = *list_id, vi = *vxl_id;
  int cur = -1;
  ro_cnl_geom cg;
  while (li < list_size && (n < 0 || found < n)) {
    int cnl = list[li];
    if (li != cur) {
      ro_channel_centres(s, sn, cs, cnl, &cg);
      cur = li;
    }
    float vc[3];
    ro_voxel_centre(g, vi, vc);
    int pass = ro_distance(cg.pix0, cg.pix1, vc) < rsum;
    if (pass) {
      cnl_out[found] = cnl;
      vxl_out[found] = vi;
      ++found;
    }
    if (vi < nvox - 1) {
      ++vi;
    } else {
      vi = 0;
      ++li;
    }
  }
  *list_id = li;
  *vxl_id = vi;
  free(sn);
  free(cs);
  return found;
}

void ro_workqueue_counts(const ro_setup* s, const ro_grid* g, const int* list,
                         int list_size, int64_t* counts) {
  int nvox = g->n[0] * g->n[1] * g->n[2];
  float* sn = (float*)malloc(sizeof(float) * s->na);
  float* cs = (float*)malloc(sizeof(float) * s->na);
  ro_angle_table(s, sn, cs);
  float rsum = ro_sum_radii(s, g);
#pragma omp parallel for schedule(dynamic, 4)
  for (int li = 0; li < list_size; ++li) {
    ro_cnl_geom cg;
    ro_channel_centres(s, sn, cs, list[li], &cg);
    int64_t c = 0;
    for (int v = 0; v < nvox; ++v) {
      float vc[3];
      ro_voxel_centre(g, v, vc);
      c += ro_distance(cg.pix0, cg.pix1, vc) < rsum;
    }
    counts[li] = c;
  }
  free(sn);
  free(cs);
}

/* ======================================================================
 * A8: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with
 * M0=0xD2511F53, M1=0xCD9E8D57, key schedule bumped by the Weyl constants
 * 0x9E3779B9, 0xBB67AE85.  Replaces the reference's cuRAND XORWOW stream
 * (RayGenerators.hpp:44,57; getSystemMatrixFromWorkqueue.cu:209-224).
 * ==================================================================== */
void ro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < (RO_MUTANT == 7 ? 9 : 10); ++r) { /* 7: a round short */
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
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

/* uniform in (0, 1]: the top 24 bits plus one, times 2^-24 (exact in float). */
/* Thread count of the OpenMP regions (CPU-baseline timing). */
void ro_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int ro_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

float ro_u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

/* RandRayGen (RayGenerators.hpp:36-67): six uniforms, the first three placed
 * in pixel 0's volume by trafo0, the last three in pixel 1's by trafo1.
 * Counter = (channel, ray, block, 0), key = (seed, 0); block 0 gives u0..u3,
 * block 1 gives u4, u5.  ReguRayGen (RayGenerators.hpp:73-99): all six = 0.5. */
void ro_ray_element(const ro_setup* s, const float* sn, const float* cs, int raygen,
                    uint32_t seed, int cnl, int element, int ray, float out[6]) {
  float u[6];
  if (raygen == RO_RAYGEN_REGU) {
    for (int i = 0; i < 6; ++i) u[i] = 0.5f;
  } else {
    /* per-element streams (RAND_ELEMENT): counter word 3 = element + 1, the
     * per-thread generators of getSystemMatrixDeviceOnly.cu:396 restated as
     * one independent stream per (channel, voxel) */
    uint32_t w3 = raygen == RO_RAYGEN_RAND_ELEMENT ? (uint32_t)element + 1u : 0u;
    uint32_t key[2] = {seed, 0u};
    uint32_t c0[4] = {(uint32_t)cnl, (uint32_t)ray, 0u, w3};
    uint32_t c1[4] = {(uint32_t)cnl, (uint32_t)ray, 1u, w3};
    uint32_t r0[4], r1[4];
    ro_philox4x32_10(c0, key, r0);
    ro_philox4x32_10(c1, key, r1);
    u[0] = ro_u01(r0[0]);
    u[1] = ro_u01(r0[1]);
    u[2] = ro_u01(r0[2]);
    u[3] = ro_u01(r0[3]);
    u[4] = ro_u01(r1[0]);
    u[5] = ro_u01(r1[1]);
  }
  int ida, id0z, id0y, id1z, id1y;
  ro_cnl_split(s, cnl, &ida, &id0z, &id0y, &id1z, &id1y);
  ro_trafo(s, 0, sn[ida], cs[ida], id0z, id0y, &u[0], &out[0]);
  ro_trafo(s, 1, sn[ida], cs[ida], id1z, id1y, &u[3], &out[3]);
}

void ro_ray(const ro_setup* s, const float* sn, const float* cs, int raygen,
            uint32_t seed, int cnl, int ray, float out[6]) {
  ro_ray_element(s, sn, cs, raygen, seed, cnl, -1, ray, out);
}

/* ======================================================================
 * A9: slab-method intersection length of the infinite line through the
 * ray's end points with an axis-aligned box,
 * getSystemMatrixFromWorkqueue.cu:44-89.  Axes with equal end-point
 * coordinates are excluded from both the entry max and the exit min (they
 * are not range-checked, as in the reference).  Masked max/min follow
 * ChordsCalc_lowlevel.hpp:49-175.
 * ==================================================================== */
float ro_intersection_length(const float v[6], const float r[6]) {
  int sx[3];
  for (int d = 0; d < 3; ++d) sx[d] = r[d] != r[3 + d];
  if (!(sx[0] || sx[1] || sx[2])) return 0.0f;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    if (sx[d]) {
      lo[d] = (v[d] - r[d]) / (r[3 + d] - r[d]);
      hi[d] = (v[3 + d] - r[d]) / (r[3 + d] - r[d]);
      if (hi[d] < lo[d]) {
        float t = lo[d];
        lo[d] = hi[d];
        hi[d] = t;
      }
    }
  }
  int good = 0;
  float amin = 0.0f, amax = 0.0f;
  for (int d = 0; d < 3; ++d) {
    if (!sx[d]) continue;
    amin = good ? (amin < lo[d] ? lo[d] : amin) : lo[d];
    amax = good ? (hi[d] < amax ? hi[d] : amax) : hi[d];
    good = 1;
  }
  if (!(amin < amax)) return 0.0f;
  float dx = r[3] - r[0], dy = r[4] - r[1], dz = r[5] - r[2];
  if (RO_MUTANT == 4) return amax - amin; /* dropped factor |ray| */
  return (amax - amin) * sqrtf(dx * dx + dy * dy + dz * dz);
}

double ro_intersection_length_f64(const double v[6], const double r[6]) {
  int sx[3];
  for (int d = 0; d < 3; ++d) sx[d] = r[d] != r[3 + d];
  if (!(sx[0] || sx[1] || sx[2])) return 0.0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    if (sx[d]) {
      lo[d] = (v[d] - r[d]) / (r[3 + d] - r[d]);
      hi[d] = (v[3 + d] - r[d]) / (r[3 + d] - r[d]);
      if (hi[d] < lo[d]) {
        double t = lo[d];
        lo[d] = hi[d];
        hi[d] = t;
      }
    }
  }
  int good = 0;
  double amin = 0.0, amax = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!sx[d]) continue;
    amin = good ? (amin < lo[d] ? lo[d] : amin) : lo[d];
    amax = good ? (hi[d] < amax ? hi[d] : amax) : hi[d];
    good = 1;
  }
  if (!(amin < amax)) return 0.0;
  double dx = r[3] - r[0], dy = r[4] - r[1], dz = r[5] - r[2];
  return (amax - amin) * sqrt(dx * dx + dy * dy + dz * dz);
}

/* Voxel box corners, getSystemMatrixFromWorkqueue.cu:181-186:
 * lo = o + i*d, hi = o + (i+1)*d, evaluated in float. */
void ro_voxel_box(const ro_grid* g, int vxl, float b[6]) {
  int id[3];
  ro_vxl_split(g, vxl, &id[0], &id[1], &id[2]);
  for (int d = 0; d < 3; ++d) {
    b[d] = g->o[d] + (float)id[d] * g->d[d];
    b[3 + d] = g->o[d] + (float)(id[d] + 1) * g->d[d];
  }
}

/* ======================================================================
 * A10: matrix element = (sum over rays of A9) / nrays, rays of the entry's
 * channel, summed in ray order (getSystemMatrixFromWorkqueue.cu:209-264).
 * All voxels of one channel see the same rays (the stream is keyed by the
 * channel, :209-210).
 * ==================================================================== */
void ro_sm_from_wq(const ro_setup* s, const ro_grid* g, const int* cnl,
                   const int* vxl, float* val, int64_t wq_len, int nrays,
                   uint32_t seed, int raygen) {
  float* sn = (float*)malloc(sizeof(float) * s->na);
  float* cs = (float*)malloc(sizeof(float) * s->na);
  ro_angle_table(s, sn, cs);
#pragma omp parallel
  {
    float* rays = (float*)malloc(sizeof(float) * 6 * (nrays > 0 ? nrays : 1));
    int cur = -1;
#pragma omp for schedule(static)
    for (int64_t w = 0; w < wq_len; ++w) {
      if (raygen == RO_RAYGEN_RAND_ELEMENT) {
        for (int r = 0; r < nrays; ++r)
          ro_ray_element(s, sn, cs, raygen, seed, cnl[w], vxl[w], r, &rays[6 * r]);
        cur = -1;
      } else if (cnl[w] != cur) {
        cur = cnl[w];
        for (int r = 0; r < nrays; ++r)
          ro_ray(s, sn, cs, raygen, seed, cur, r, &rays[6 * r]);
      }
      float box[6];
      ro_voxel_box(g, vxl[w], box);
      float a = 0.0f;
      for (int r = 0; r < nrays; ++r) a += ro_intersection_length(box, &rays[6 * r]);
      a /= (float)nrays;
      val[w] = a;
    }
    free(rays);
  }
  free(sn);
  free(cs);
}

void ro_sm_from_wq_f64(const ro_setup* s, const ro_grid* g, const int* cnl,
                       const int* vxl, double* val, int64_t wq_len, int nrays,
                       uint32_t seed, int raygen) {
  float* sn = (float*)malloc(sizeof(float) * s->na);
  float* cs = (float*)malloc(sizeof(float) * s->na);
  ro_angle_table(s, sn, cs);
#pragma omp parallel
  {
    double* rays = (double*)malloc(sizeof(double) * 6 * (nrays > 0 ? nrays : 1));
    int cur = -1;
#pragma omp for schedule(static)
    for (int64_t w = 0; w < wq_len; ++w) {
      if (raygen == RO_RAYGEN_RAND_ELEMENT || cnl[w] != cur) {
        cur = raygen == RO_RAYGEN_RAND_ELEMENT ? -1 : cnl[w];
        for (int r = 0; r < nrays; ++r) {
          float rf[6];
          ro_ray_element(s, sn, cs, raygen, seed, cnl[w], vxl[w], r, rf);
          for (int k = 0; k < 6; ++k) rays[6 * r + k] = (double)rf[k];
        }
      }
      float bf[6];
      double box[6];
      ro_voxel_box(g, vxl[w], bf);
      for (int k = 0; k < 6; ++k) box[k] = (double)bf[k];
      double a = 0.0;
      for (int r = 0; r < nrays; ++r)
        a += ro_intersection_length_f64(box, &rays[6 * r]);
      val[w] = a / (double)nrays;
    }
    free(rays);
  }
  free(sn);
  free(cs);
}

/* ======================================================================
 * A11: Siddon traversal, ChordsCalc_lowlevel.hpp:194-379 with helpers
 * :395-676.  Restated with one documented correction: the reference starts
 * each axis at the plane after that axis' first grid plane
 * (aDimnext = aDimmin + aDimup, :264), which lies before the grid entry for
 * every axis other than the entry axis and yields negative chords; here each
 * axis starts at its first plane strictly after the entry parameter.  The
 * output is bounded by n_chords_cap (the reference ignores nChords) and the
 * last chord is clipped at the exit parameter; the plane stepping is bounded
 * by the grid's plane count.  Returns the chord count (at most cap).
 * ==================================================================== */
int ro_get_chords(float* chords, int* voxel_ids, int cap, const float ray[6],
                  const float go[3], const float gd[3], const int gn[3]) {
  int crosses[3];
  int good_par = 1;
  for (int d = 0; d < 3; ++d) {
    crosses[d] = (ray[d + 3] - ray[d] != 0.0f);
    good_par &= crosses[d] || (ray[d] >= go[d] && ray[d] < go[d] + gd[d] * (float)gn[d]);
  }
  for (int d = 0; d < 3; ++d) crosses[d] &= good_par;
  float amin_d[3], amax_d[3], up[3];
  int idup[3];
  for (int d = 0; d < 3; ++d) {
    float div = ray[d + 3] - ray[d];
    div += (float)(div == 0.0f);
    float a0 = (go[d] + 0.0f * gd[d] - ray[d]) / div;
    float a1 = (go[d] + (float)gn[d] * gd[d] - ray[d]) / div;
    amin_d[d] = a0 < a1 ? a0 : a1;
    amax_d[d] = a0 < a1 ? a1 : a0;
    float dd = sqrtf((ray[d + 3] - ray[d]) * (ray[d + 3] - ray[d]));
    dd += (float)(dd == 0.0f);
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
      amin > amax)
    return 0;
  float next[3];
  for (int d = 0; d < 3; ++d) {
    next[d] = amin_d[d] + up[d];
    if (crosses[d])
      for (int k = 0; next[d] <= amin && k < gn[d] + 4; ++k) next[d] += up[d];
  }
  float anext = 2.0f;
  for (int d = 0; d < 3; ++d)
    if (crosses[d] && next[d] < anext) anext = next[d];
  float amid = 0.5f * (amin + (anext < amax ? anext : amax));
  int id[3];
  for (int d = 0; d < 3; ++d)
    id[d] = (int)floorf((ray[d] + amid * (ray[d + 3] - ray[d]) - go[d]) / gd[d]);
  float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
  float length = sqrtf(dx * dx + dy * dy + dz * dz);
  float acur = amin;
  int n = 0;
  /* every step crosses at least one grid plane: bounded by the plane count */
  const int max_steps = gn[0] + gn[1] + gn[2] + 8;
  for (int step = 0; acur < amax && step < max_steps; ++step) {
    anext = 2.0f;
    for (int d = 0; d < 3; ++d)
      if (crosses[d] && next[d] < anext) anext = next[d];
    if (anext > amax) anext = amax;
    int inside = id[0] >= 0 && id[0] < gn[0] && id[1] >= 0 && id[1] < gn[1] &&
                 id[2] >= 0 && id[2] < gn[2];
    if (inside && n < cap) {
      chords[n] = (anext - acur) * length;
      voxel_ids[3 * n + 0] = id[0];
      voxel_ids[3 * n + 1] = id[1];
      voxel_ids[3 * n + 2] = id[2];
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
  return n < cap ? n : cap;
}

/* ======================================================================
 * A12/A13: matrix of the measured rows with exact zeros dropped, in CSR
 * (rows in list order, columns ascending).  Values per entry are A10's; the
 * candidate set per row is A6's.  This is what the reference's
 * systemMatrixCalculation (wrappers.hpp:185-266) produces after cooSort,
 * coo2csr and csr2ecsr, minus the zero-valued candidates it stores.
 * ro_build_csr_ex with RO_SM_KEEP_ZEROS keeps them too
 * (getSystemMatrixDeviceOnly.cu:507-509): the matrix then holds exactly the
 * workqueue's pairs.  RO_RAYGEN_RAND_ELEMENT samples each element with its
 * own rays.
 * ==================================================================== */
int64_t ro_build_csr(const ro_setup* s, const ro_grid* g, const int* list,
                     int list_size, int nrays, uint32_t seed, int raygen,
                     int64_t* row_ptr, int* col, float* val, double* val64) {
  return ro_build_csr_ex(s, g, list, list_size, nrays, seed, raygen, 0, row_ptr, col, val,
                         val64);
}

int64_t ro_build_csr_ex(const ro_setup* s, const ro_grid* g, const int* list,
                        int list_size, int nrays, uint32_t seed, int raygen, int flags,
                        int64_t* row_ptr, int* col, float* val, double* val64) {
  const int keep_zeros = (flags & RO_SM_KEEP_ZEROS) != 0;
  const int elem = raygen == RO_RAYGEN_RAND_ELEMENT;
  int nvox = g->n[0] * g->n[1] * g->n[2];
  float* sn = (float*)malloc(sizeof(float) * s->na);
  float* cs = (float*)malloc(sizeof(float) * s->na);
  ro_angle_table(s, sn, cs);
  float rsum = ro_sum_radii(s, g);
  int fill = (col != NULL);
  int64_t* counts = NULL;
  if (!fill) counts = (int64_t*)calloc((size_t)list_size + 1, sizeof(int64_t));
#pragma omp parallel
  {
    float* rays = (float*)malloc(sizeof(float) * 6 * (nrays > 0 ? nrays : 1));
    double* rays64 = (double*)malloc(sizeof(double) * 6 * (nrays > 0 ? nrays : 1));
#pragma omp for schedule(dynamic, 4)
    for (int li = 0; li < list_size; ++li) {
      int cnl = list[li];
      ro_cnl_geom cg;
      ro_channel_centres(s, sn, cs, cnl, &cg);
      for (int r = 0; r < nrays; ++r) {
        ro_ray(s, sn, cs, raygen, seed, cnl, r, &rays[6 * r]);
        for (int k = 0; k < 6; ++k) rays64[6 * r + k] = (double)rays[6 * r + k];
      }
      int64_t pos = fill ? row_ptr[li] : 0;
      int64_t c = 0;
      for (int v = 0; v < nvox; ++v) {
        float vc[3];
        ro_voxel_centre(g, v, vc);
        if (!(ro_distance(cg.pix0, cg.pix1, vc) < rsum)) continue;
        float box[6];
        ro_voxel_box(g, v, box);
        if (elem)
          for (int r = 0; r < nrays; ++r) {
            ro_ray_element(s, sn, cs, raygen, seed, cnl, v, r, &rays[6 * r]);
            for (int k = 0; k < 6; ++k) rays64[6 * r + k] = (double)rays[6 * r + k];
          }
        float a = 0.0f;
        for (int r = 0; r < nrays; ++r) a += ro_intersection_length(box, &rays[6 * r]);
        a /= (float)nrays;
        if (a == 0.0f && !keep_zeros) continue;
        if (fill) {
          col[pos + c] = v;
          val[pos + c] = a;
          if (val64) {
            double b[6], a64 = 0.0;
            for (int k = 0; k < 6; ++k) b[k] = (double)box[k];
            for (int r = 0; r < nrays; ++r)
              a64 += ro_intersection_length_f64(b, &rays64[6 * r]);
            val64[pos + c] = a64 / (double)nrays;
          }
        }
        ++c;
      }
      if (!fill) counts[li] = c;
    }
    free(rays);
    free(rays64);
  }
  int64_t total = 0;
  if (!fill) {
    for (int li = 0; li < list_size; ++li) {
      row_ptr[li] = total;
      total += counts[li];
    }
    row_ptr[list_size] = total;
    free(counts);
  } else {
    total = row_ptr[list_size];
  }
  free(sn);
  free(cs);
  return total;
}

/* ======================================================================
 * A14: in-place sort of COO entries by (row, col), values follow
 * (cooSort.hpp:47-65 comparator, :75-99 sort_by_key).  Stable, so equal
 * (row, col) keys keep their input order.
 * ==================================================================== */
typedef struct {
  int row, col;
  int64_t pos;
  float val;
} ro_coo_item;

static int ro_coo_cmp(const void* pa, const void* pb) {
  const ro_coo_item* a = (const ro_coo_item*)pa;
  const ro_coo_item* b = (const ro_coo_item*)pb;
  if (RO_MUTANT == 8 && a->col != b->col) return a->col < b->col ? -1 : 1; /* keys swapped */
  if (a->row != b->row) return a->row < b->row ? -1 : 1;
  if (a->col != b->col) return a->col < b->col ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

void ro_coo_sort(float* val, int* row, int* col, int64_t n) {
  if (n <= 0 || !val || !row || !col) return;
  ro_coo_item* it = (ro_coo_item*)malloc(sizeof(ro_coo_item) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    it[i].row = row[i];
    it[i].col = col[i];
    it[i].pos = i;
    it[i].val = val[i];
  }
  qsort(it, (size_t)n, sizeof(ro_coo_item), ro_coo_cmp);
  for (int64_t i = 0; i < n; ++i) {
    row[i] = it[i].row;
    col[i] = it[i].col;
    val[i] = it[i].val;
  }
  free(it);
}

/* A15: row pointer of sorted COO rows (cusparseXcoo2csr semantics as called
 * in cusparseWrapper.hpp:53-60, zero-based): row_ptr[r] = #entries with
 * row < r. */
void ro_coo2csr(int* row_ptr, const int* row_id, int64_t nnz, int m) {
  for (int r = 0; r <= m; ++r) row_ptr[r] = 0;
  for (int64_t i = 0; i < nnz; ++i) row_ptr[row_id[i] + 1]++;
  for (int r = 0; r < m; ++r) row_ptr[r + 1] += row_ptr[r];
}

/* A16: effective row pointer, convertCsr2Ecsr.hpp:57-63. */
void ro_csr2ecsr(int* ecsr, const int* y, int nnzr, const int* csr, int m) {
  for (int i = 0; i < nnzr; ++i) ecsr[i] = csr[RO_MUTANT == 5 ? i : y[i]]; /* 5: wrong index */
  ecsr[nnzr] = csr[m];
}

/* ======================================================================
 * A17: y = alpha*op(A)*x + beta*y for a zero-based CSR matrix (csrmv.hpp:
 * 37-99).  beta == 0 overwrites y without reading it.
 * ==================================================================== */
void ro_csrmv_f32(int trans, int m, int n, float alpha, const float* val,
                  const int* rp, const int* col, const float* x, float beta,
                  float* y) {
  if (!trans) {
    for (int i = 0; i < m; ++i) {
      float acc = 0.0f;
      for (int k = rp[i]; k < rp[i + 1]; ++k) acc += val[k] * x[col[k]];
      y[i] = beta == 0.0f ? alpha * acc : alpha * acc + beta * y[i];
    }
  } else {
    for (int j = 0; j < n; ++j) y[j] = beta == 0.0f ? 0.0f : beta * y[j];
    for (int i = 0; i < m; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k) y[col[k]] += alpha * val[k] * x[i];
  }
}

void ro_csrmv_f64(int trans, int m, int n, double alpha, const double* val,
                  const int64_t* rp, const int* col, const double* x,
                  double beta, double* y) {
  if (!trans) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) acc += val[k] * x[col[k]];
      y[i] = beta == 0.0 ? alpha * acc : alpha * acc + beta * y[i];
    }
  } else {
    for (int j = 0; j < n; ++j) y[j] = beta == 0.0 ? 0.0 : beta * y[j];
    for (int i = 0; i < m; ++i)
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) y[col[k]] += alpha * val[k] * x[i];
  }
}

/* ======================================================================
 * A18-A20: mlemOperations.hpp:44-54 (divides, no zero guard), :61-104
 * (dividesMultiplies: a*b/(c + (c==0))), :111-153 (scales: alpha*x),
 * :163-169 (sum).
 * ==================================================================== */
void ro_divides(float* res, const float* dvd, const float* dvs, int n) {
  for (int i = 0; i < n; ++i) res[i] = dvd[i] / dvs[i];
}

void ro_divides_multiplies(float* res, const float* a, const float* b,
                           const float* c, int n) {
  for (int i = 0; i < n; ++i) {
    float d = c[i];
    d += (float)(d == 0.0f);
    res[i] = a[i] * b[i] / d;
  }
}

void ro_scales(float* x, float alpha, int n) {
  for (int i = 0; i < n; ++i) x[i] = alpha * x[i];
}

float ro_sum(const float* x, int n) {
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s += x[i];
  return s;
}

/* ======================================================================
 * A21/A22: MLEM recurrence of reco.cu:230-303 in FP64, on one CSR of all
 * measured rows (the reference's chunks partition the rows; their
 * correction images add up to the same c).  Sensitivity is normalised to
 * sum 1 (reco.cu:203-204).  Per iteration:
 *   yt = A x;  e = y / yt;  c = A^T e;  x = x*c/(s + (s==0));  x /= sum(x).
 * A row with yt == 0 contributes e = 0 (the reference divides unguarded;
 * such a row has no non-zero entries to back-project through).
 * ==================================================================== */
void ro_sensitivity_f64(int m, int nvox, const int64_t* rp, const int* col,
                        const double* val, double* sens) {
  for (int j = 0; j < nvox; ++j) sens[j] = 0.0;
  for (int i = 0; i < m; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) sens[col[k]] += val[k];
}

void ro_mlem_f64(int m, int nvox, const int64_t* rp, const int* col,
                 const double* val, const double* y, const double* sens,
                 double* x, int niter) {
  double* s = (double*)malloc(sizeof(double) * nvox);
  double* c = (double*)malloc(sizeof(double) * nvox);
  double* e = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  /* per-thread partial back-projections, reduced in thread order */
  double* part = (double*)malloc(sizeof(double) * (size_t)nvox * (size_t)nt);
  double ss = 0.0;
  for (int j = 0; j < nvox; ++j) ss += sens[j];
  for (int j = 0; j < nvox; ++j) s[j] = sens[j] / ss;
  int used = 1; /* team size of the back-projection region */
  for (int it = 0; it < niter; ++it) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) acc += val[k] * x[col[k]];
      e[i] = acc != 0.0 ? y[i] / acc : 0.0;
    }
#pragma omp parallel num_threads(nt)
    {
      int t = 0, T = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      T = omp_get_num_threads();
#endif
      if (t == 0) used = T;
      double* p = part + (size_t)nvox * (size_t)t;
      for (int j = 0; j < nvox; ++j) p[j] = 0.0;
      int64_t lo = (int64_t)m * t / T, hi = (int64_t)m * (t + 1) / T;
      for (int64_t i = lo; i < hi; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) p[col[k]] += val[k] * e[i];
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nvox; ++j) {
      double acc = 0.0;
      for (int t = 0; t < used; ++t) acc += part[(size_t)nvox * (size_t)t + j];
      c[j] = acc;
    }
    double norm = 0.0;
    for (int j = 0; j < nvox; ++j) {
      double d = s[j];
      d += (double)(d == 0.0);
      x[j] = x[j] * c[j] / d;
      norm += x[j];
    }
    if (RO_MUTANT != 6) /* 6: normalisation dropped */
      for (int j = 0; j < nvox; ++j) x[j] = x[j] * (1.0 / norm);
  }
  free(part);
  free(s);
  free(c);
  free(e);
}

/* ======================================================================
 * A21 in the reference's own precision: the MLEM recurrence of
 * reco.cu:230-303 with FP32 matrix values, vectors and reductions, the way
 * the reference runs it (cusparseScsrmv, thrust transform/reduce on float):
 * s <- s/sum(s) (reco.cu:203-204), then per iteration yt = A x,
 * e = y/yt (0 where yt == 0, the guard of the GPU path), c = A^T e,
 * x <- x*c/(s + (s==0)), x <- x*(1/sum(x)) (mlemOperations.hpp:44-169).
 * The CPU baseline of the benchmark; the FP64 recurrence above is the
 * accuracy reference.
 * ==================================================================== */
void ro_mlem_f32(int m, int nvox, const int64_t* rp, const int* col,
                 const float* val, const float* y, const double* sens,
                 float* x, int niter) {
  float* s = (float*)malloc(sizeof(float) * nvox);
  float* c = (float*)malloc(sizeof(float) * nvox);
  float* e = (float*)malloc(sizeof(float) * (m > 0 ? m : 1));
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  float* part = (float*)malloc(sizeof(float) * (size_t)nvox * (size_t)nt);
  float ss = 0.0f;
  for (int j = 0; j < nvox; ++j) ss += (float)sens[j];
  for (int j = 0; j < nvox; ++j) s[j] = (float)sens[j] / ss;
  int used = 1;
  for (int it = 0; it < niter; ++it) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
      float acc = 0.0f;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) acc += val[k] * x[col[k]];
      e[i] = acc != 0.0f ? y[i] / acc : 0.0f;
    }
#pragma omp parallel num_threads(nt)
    {
      int t = 0, T = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      T = omp_get_num_threads();
#endif
      if (t == 0) used = T;
      float* p = part + (size_t)nvox * (size_t)t;
      for (int j = 0; j < nvox; ++j) p[j] = 0.0f;
      int64_t lo = (int64_t)m * t / T, hi = (int64_t)m * (t + 1) / T;
      for (int64_t i = lo; i < hi; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) p[col[k]] += val[k] * e[i];
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nvox; ++j) {
      float acc = 0.0f;
      for (int t = 0; t < used; ++t) acc += part[(size_t)nvox * (size_t)t + j];
      c[j] = acc;
    }
    float norm = 0.0f;
    for (int j = 0; j < nvox; ++j) {
      float d = s[j];
      d += (float)(d == 0.0f);
      x[j] = x[j] * c[j] / d;
      norm += x[j];
    }
    const float inv = 1.0f / norm;
    for (int j = 0; j < nvox; ++j) x[j] = x[j] * inv;
  }
  free(part);
  free(s);
  free(c);
  free(e);
}
