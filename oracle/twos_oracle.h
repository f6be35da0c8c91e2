/*
 * CPU oracle for the Topological Walk on Spheres (TWoS) estimator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this code;
 * it is loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, and there only as the checker.  Every walk runs in IEEE binary64
 * with no floating-point contraction (the recipe in oracle/Makefile passes
 * -ffp-contract=off), as the reference build does (no -march, no FMA:
 * benchmarks/compiled_bunny/CMakeLists.txt:13).
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   splitmix64 / mt19937_64 / uniform draws / random_unit_vector / walk:
 *       replay of benchmarks/compiled_bunny/output/walk-traces.csv (positions and
 *       radii) and field-*.csv (integer step counts, u, score, score_se).
 *   closest_point_triangle and the BVH: brute force over all triangles, and the
 *       traced radii above.
 *   Philox4x32-10: published known-answer vectors (Salmon et al., SC'11).
 *   cross-index combine and jackknife: brute force O(N^2) / O(N) on small N and
 *       the reference field fixtures' score / score_se columns.
 *   harmonic Dirichlet payoff: unit-ball closed forms (u = g for harmonic g).
 *   ray-parity inside test: analytic ball / star-shaped meshes.
 *   og_certify / og_cert_distance: the retained --mode validate samples of
 *       the minimum-benefit pilot (every value, step count and flag).
 *   og_differential: the retained wih_study --diff 1 requests on the branch
 *       mesh (request totals of steps, distance queries and parity rays
 *       exactly; derivative columns, most sites bitwise).
 */
#ifndef TWOS_ORACLE_H
#define TWOS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry container kinds. */
enum { OG_BALL = 0, OG_BOX = 1, OG_MESH = 2 };
/* Combine modes. */
enum { OG_SQUARE = 0, OG_PAIRED = 1, OG_CROSS = 2, OG_COUPLED = 3 };
/* Stream modes. */
enum { OG_RNG_PHILOX = 0, OG_RNG_MT_SITE = 1, OG_RNG_MT_PAIRED = 2, OG_RNG_MT_KEY = 3, OG_RNG_MT_FD = 4,
       OG_RNG_MT_DIFF = 5 };

typedef struct og_mesh og_mesh;

typedef struct {
  int32_t container;
  int32_t n_spheres;   /* analytic sphere holes, rows of (cx, cy, cz, r); class 6 */
  int32_t n_cylinders; /* axis-aligned infinite cylinder holes, rows of (axis, ca, cb, r); class 6 */
  int32_t n_cavities;  /* reference-style cavities of one common radius; class 7 */
  int32_t mesh_radius_f32; /* 1: round the mesh distance to binary32, as the retained
                              reference outputs do (Embree's float query radius) */
  int32_t pad;
  double ball[4];      /* cx, cy, cz, R */
  double box_lo[3];
  double box_hi[3];
  double cavity_radius;
  const double *spheres;
  const double *cylinders;
  const double *cavities;
  const og_mesh *mesh;
} og_geom;

/* Volume-source term types for source_kind 2, rows {type, cx, cy, cz, p0, p1, p2, w}
 * with t = max(0, 1 - |y - c|^2 / (p0 p0)):
 *   OG_TERM_BUMP     w * (t t)                       (benchmarks/submission_study/study.cpp:29-36)
 *   OG_TERM_UNIT     (t t) / p1, p1 the bump's mass   (study.cpp:44-51)
 *   OG_TERM_GAUSS    w * exp(-((x-cx)/p0)^2 - ((y-cy)/p1)^2 - ((z-cz)/p2)^2)
 *                                                     (benchmarks/compiled_bunny/src/wih_twos.cpp:553-566) */
enum { OG_TERM_BUMP = 0, OG_TERM_UNIT = 1, OG_TERM_GAUSS = 2 };
#define OG_MAX_TERMS 8

/* Boundary data and volume source of one side of the problem. */
typedef struct {
  double poly[10];      /* c0 + cx x + cy y + cz z + cxx x^2 + cyy y^2 + czz z^2 + cxy xy + cxz xz + cyz yz */
  double cls_const[8];  /* per boundary class constant */
  uint32_t cls_poly_mask; /* bit c set: class c adds poly(y) */
  int32_t source_kind;  /* 0 none; 1 affine f = ((s0 + sx (x+ox)) + sy (y+oy)) + sz (z+oz) at the
                           ball centre; 2 f = s0 + sum of terms at a Green-distributed point;
                           3 the exact Green-ball integral of s0 and each BUMP / UNIT row */
  double src[7];        /* s0, sx, sy, sz, ox, oy, oz */
  int32_t n_terms;
  int32_t pad;
  double terms[OG_MAX_TERMS][8];
} og_side;

typedef struct {
  int32_t combine;
  int32_t rng;
  uint64_t seed;
  int32_t repeat;
  int32_t max_steps;
  double eps;
  int32_t walks;       /* fixed budget, or the first rung of the adaptive ladder */
  int32_t walks_max;   /* > walks enables doubling until ci_z*se <= ci_rel*|score| */
  double ci_rel;
  double ci_z;
  int32_t threads;
  int32_t gradient;    /* 1: first-step gradient channel (affine or no source) */
} og_params;

typedef struct {
  double u, v, score, score_se, td, same;
  int64_t walks, steps, caps;
  int32_t status; /* 0 finished (fixed budget or CI met), 1 budget exhausted before the CI target */
  int32_t pad;
  double grad_u[3], grad_v[3]; /* gradient channel: mean samples of each side (NaN without it) */
  double grad_dot, grad_dot_se; /* cross-index grad u . grad v and its delta-method SE */
} og_site;

uint64_t og_splitmix64(uint64_t x);
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void og_philox_key(uint64_t seed, int64_t site, int32_t side, uint32_t key[2]);
void og_mt19937_64(uint64_t seed, uint64_t *out, int32_t n);
void og_mt_uniforms(uint64_t seed, double *out, int32_t n);

og_mesh *og_mesh_create(const double *vertices, int64_t nv, const uint32_t *faces, int64_t nf);
void og_mesh_destroy(og_mesh *mesh);
double og_mesh_closest(const og_mesh *mesh, const double p[3], double out_point[3], int64_t *out_tri,
                       int32_t brute_force, int64_t *nodes_visited, int64_t *tris_tested);
void og_closest_point_triangle(const double p[3], const double a[3], const double b[3],
                               const double c[3], double out[3]);
int32_t og_mesh_inside(const og_mesh *mesh, const double p[3]);
int32_t og_mesh_inside_many(const og_mesh *mesh, const double *p, int64_t n, int32_t *out, int32_t threads);

double og_closest_boundary(const og_geom *g, const double p[3], double out_point[3], int32_t *out_cls);

int32_t og_estimate(const og_geom *g, const og_side sides[2], const og_params *prm,
                    const double *sites, int64_t n, og_site *out,
                    double *dump_a, double *dump_b, int32_t *dump_steps, int64_t dump_stride);

int32_t og_trace(const og_geom *g, const og_side *side, const og_params *prm,
                 const double start[3], uint64_t key, int64_t site, int32_t walk,
                 double *rows, int32_t max_rows);

void og_combine(const double *a, const double *b, int64_t n, int32_t square, double out[6]);
void og_combine_powersums(const double sums[8], int64_t n, double out[6]);
/* Differential WoS cavity-shape derivatives (evaluate_differential_site,
 * wih_twos.cpp:725-928) on a mesh with cavities, the reference's streams
 * (rng = OG_RNG_MT_DIFF, repeat as given) or Philox streams (rng =
 * OG_RNG_PHILOX: sample n's outer path draws direction `step` from counter
 * (step, n, 0, 0) under philox_key(dseed, site, 0), its nested path is walk n
 * of philox_key(dseed, site, 1), dseed = splitmix64(seed + 0x6a09e667f3bcc909)
 * ^ splitmix64(repeat + 0xbb67ae8584caa73b)): out[n][n_cavities][8] = dc x y z,
 * their SEs, dR, its SE; counts[n][9] = outer steps, nested steps, cavity
 * terminations, near-seam, segment-unsafe, unsafe, limits, distance queries,
 * parity rays. */
int32_t og_differential(const og_geom *g, const og_side *side, const og_params *prm, const double *sites, int64_t n,
                        double ell, double *out, int64_t *counts);
/* Certified minimum-benefit sampling (run_minimum_benefit / cert_walk,
 * benchmarks/compiled_bunny/src/minimum_benefit.hpp:122-188) on a mesh with
 * cavities: per site and sample j the clipped value, its steps and flags
 * (1 capped, 2 clipped, 4 containment failure), [n][walks] each. */
typedef struct {
  uint64_t seed;
  int32_t walks;
  int32_t max_steps;
  double eps;
  double radius_cap;   /* --validation-radius-cap */
  double value_cap;    /* --validation-value-cap */
  double slope;        /* affine source slope (source 0) */
  int32_t source;      /* 0 affine 1 + slope (y + 0.5), 1 housing chips */
  int32_t threads;
  int64_t site_offset; /* global index of sites[0] in the keys */
  int32_t rng;         /* OG_RNG_MT_KEY: the reference's keys (minimum_benefit.hpp:168);
                          OG_RNG_PHILOX: key philox_key(splitmix64(seed + 0xa1224), site, 0),
                          counter (step, sample, draw, 0) */
  int32_t pad;
} og_cert_params;
int32_t og_certify(const og_geom *g, const og_cert_params *prm, const double *sites, int64_t n, double *value,
                   int32_t *steps, int32_t *flags);
/* The certified lower bound of the domain distance (cert_domain_distance,
 * minimum_benefit.hpp:83-88). */
int32_t og_cert_distance(const og_geom *g, const double *points, int64_t n, double *out);
/* Exact integral of the Green function of B(x, radius) against the bump
 * amplitude (1 - |y - c|^2 / support^2)^2 (source_kind 3; study.cpp:56-76). */
double og_bump_green_ball(const double x[3], double radius, const double c[3], double support, double amplitude);
/* Gradient channel of n walks from per-walk samples ga, gb ([n][3] each):
 * out = {grad_u[3], grad_v[3], grad_dot, grad_dot_se}. */
void og_combine_grad(const double *ga, const double *gb, int64_t n, double out[8]);

#ifdef __cplusplus
}
#endif

#endif
