// Internal data layouts shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/twos.h"

namespace twos {

enum { kBall = TWOS_BALL, kBox = TWOS_BOX, kMesh = TWOS_MESH };
enum { kSquare = TWOS_SQUARE, kPaired = TWOS_PAIRED, kCross = TWOS_CROSS, kCoupled = TWOS_COUPLED };
enum { kRngPhilox = TWOS_RNG_PHILOX, kRngMtSite = TWOS_RNG_MT_SITE, kRngMtPaired = TWOS_RNG_MT_PAIRED,
       kRngMtKey = TWOS_RNG_MT_KEY, kRngMtFd = TWOS_RNG_MT_FD };
// Per-path extras a walk kernel is built with: none, the first-step gradient
// samples, the coupled finite-cavity pair of paths, or none and no volume
// source on either side (the source code is left out of the kernel).
enum { kChanNone = 0, kChanGrad = 1, kChanCoupled = 2, kChanBare = 3 };
// Kernel specialisations by geometry: any container (runtime switch), mesh
// container only (no analytic holes), analytic container only.
enum { kGeomGeneric = 0, kGeomMesh = 1, kGeomAnalytic = 2 };

// One side's data as the kernels read it: the C-ABI struct plus binary32
// copies of the payoff and affine-source coefficients (each rounded to
// nearest once on the host, which is the conversion a binary32 walk applied
// to the binary64 value at every use), so a binary32 step loads floats and
// converts nothing.
struct DevSide : twos_side {
  float poly_f[10];
  float cls_const_f[8];
  float src_f[7];
  float pad_f;
  DevSide() = default;
  DevSide(const twos_side &s) : twos_side(s) {
    for (int k = 0; k < 10; ++k) poly_f[k] = float(s.poly[k]);
    for (int k = 0; k < 8; ++k) cls_const_f[k] = float(s.cls_const[k]);
    for (int k = 0; k < 7; ++k) src_f[k] = float(s.src[k]);
    pad_f = 0.f;
  }
};

// One resident geometry as the kernels see it.  Analytic rows are kept in
// binary32 (float4) and binary64 (4 doubles, cavities padded) copies.
struct DevGeom {
  int32_t container, n_spheres, n_cylinders, n_cavities;
  int32_t mesh_radius_f32, n_tris;
  int32_t grid_corners;  // grid mode: 0 cell centres, 1 cell corners, 2 binary16 corner records per cell
  int32_t pad;
  double ball[4];
  double box_lo[3];
  double box_hi[3];
  double cavity_radius;
  const float4 *spheres_f, *cylinders_f, *cavities_f;
  const double *spheres_d, *cylinders_d, *cavities_d;
  const float4 *nodes;
  const float4 *nodes4;
  const float *tris_f;
  const double *tris_d;
  const int32_t *tri_ids;
  const float *grid;     // mesh distance at cell centres or corners (conservative-radius grid), or null
  float grid_lo[3];
  float grid_h;
  int32_t grid_n[3];
  float grid_inv_h;
  // Record grids: per cell, the 4-wide node to start an exact query at (the
  // deepest whose leaf range holds every triangle within grid_start_reach of
  // the cell centre), valid while the query's best distance so far is at most
  // sqrt(grid_start_r2); null without them.
  const int32_t *grid_start;
  float grid_start_r2;
  int32_t pad2;
};

// Per-site-slot state resident on the device (capacity = server max_sites).
struct SlotArrays {
  double *pos;        // [C][3]
  int32_t *geom;      // index into the geometry table
  int32_t *problem;   // index into the side table: sides 2*problem, 2*problem+1
  int32_t *combine;   // the request's combine mode (the reduction's estimate)
  uint2 *keys;        // [C][2] Philox keys of sides 0 and 1
  int64_t *site;      // site index used by the mt19937_64 key layouts
  uint64_t *seed;     // request seed (mt19937_64 layouts)
  int32_t *repeat;    // repeat index (MT_SITE)
  int32_t *walks_max;
  int32_t *hint0;     // sorted index of the closest triangle at the site (mesh), or -1
  float *grid_kappa;  // > 0: the request's distance-grid threshold in cell sizes; 0: exact radius
  int32_t *max_steps;
  double *eps;
  double *ci_rel;
  double *ci_z;
  double *acc;        // [C][8] binary64 power sums
  double *acc_grad;   // [C][28] gradient-channel sums (A, B, C, second moments)
  int64_t *acc_walks;
  int64_t *acc_steps;
  int64_t *acc_caps;
  twos_site_result *result;  // [C]
};

// One round's work list, filled by the host before the walk launch.
struct RoundArrays {
  int32_t *act_slot;     // [n_active]
  int64_t *act_prefix;   // [n_active + 1], walk units
  int32_t *walk_base;    // [n_active]: first walk index of this round
  int32_t *n_active;     // device scalar
  int64_t *total_jobs;   // device scalar
  unsigned long long *counter;  // persistent ticket counter
  twos_site_result *compact;    // [n_active] results gathered for the host
  int32_t *decision;            // [n_active] 0 done, 1 more walks, 2 budget exhausted
};

struct WalkParams {
  int32_t combine;
  int32_t rng;
  int32_t jobs_per_walk;  // 2 for CROSS, 1 otherwise
  int32_t pad0;
  int64_t job_begin;      // simple kernel: first job of this launch
  int64_t job_end;
  double *val;            // [total walks][2] binary64 samples (a, b)
  int32_t *steps;         // [total walks][2] step counts, bit 30 = capped
  double *trace;          // optional rows (x, y, z, r) for the trace entry point
  int32_t trace_rows;
  int32_t pad;
  unsigned long long *counters;  // nodes, tris, grid lookups (when counting)
  double *grad;                  // [total walks][2][3] gradient samples, or null
  int32_t refill;                // persistent kernels: idle lanes refill once this many are idle
  int32_t pad2;
};

// Certified corner grid of a mesh: at corner (i, j, k) = lo + (i, j, k) h
// (binary64), the certified distance to the union of the triangles' bounding
// boxes rounded down to binary32.  tau: certified walks take the grid's bound
// as their distance when it is at least tau (a launch parameter).
struct CertGrid {
  const float *d = nullptr;
  double lo[3] = {0.0, 0.0, 0.0};
  double h = 0.0, inv_h = 0.0;
  int32_t n[3] = {0, 0, 0};  // corners per axis
  int32_t pad = 0;
  double tau = 0.0;
};

// One launch of the differential kernels: jobs = sites of the chunk x walks.
struct DiffParams {
  const double *sites;   // [n][3] all sites of the request
  double *val;           // [jobs][4] per-job normal and V / ell
  int32_t *rec;          // [jobs][6] per-job cavity, flags and counters
  double *out;           // [n][n_cavities][8] folded derivatives
  int64_t *counts;       // [n][9] per-site counters
  int64_t jobs;
  int64_t walks;
  int64_t site_base;     // first site of the chunk
  int64_t site_offset;   // added to the site index in the key
  uint64_t seed;
  int32_t repeat;
  int32_t max_steps;
  double eps;
  double ell;            // offset length
  // Philox form (diff_persistent) only:
  unsigned long long *counter;  // job tickets, zero before the launch
  uint64_t pseed;        // splitmix64(seed + 0x6a09e667f3bcc909) ^ splitmix64(repeat + 0xbb67ae8584caa73b)
  float grid_kappa;      // > 0: mesh steps take the grid bound when it clears max(kappa h, 2 eps)
  int32_t safety_batch;  // lanes waiting for an offset-safety test before the wave runs the tests
};

}  // namespace twos
