/*
 * twos.h -- C-ABI of the MI355X Topological Walk on Spheres (TWoS) estimator.
 *
 * One call estimates, for every candidate site z, the Dirichlet-cavity
 * topological derivative D_T(z) = -(u(z) - h) p(z) by Walk on Spheres, with the
 * leave-diagonal-out cross-index combine and a jackknife standard error.  It is
 * the in-process replacement of the reference's compiled backend, which is a
 * process-level plugin (flags + ASCII PLY/CSV in, CSV/JSON out):
 *
 *   twos_estimate     replaces  `wih_twos --mode twos` (benchmarks/compiled_bunny/
 *                               src/wih_twos.cpp:1799-1864: evaluate_site :674-717
 *                               over an OpenMP site loop :1828-1836) and the paired
 *                               estimator of `wih_study` (benchmarks/submission_study/
 *                               study.cpp:80-120, :334-344)
 *   twos_mesh_create  replaces  EmbreeMesh(path) (wih_twos.cpp:298-325); the caller
 *                               parses the PLY (load_ascii_ply :437-493)
 *   twos_geometry_create / twos_geometry_union_sphere
 *                     replace   --cavities/--cavity-radius (wih_twos.cpp:585-626) and
 *                               add the analytic sphere/box/cylinder domains
 *   twos_inside       replaces  EmbreeMesh::inside (wih_twos.cpp:361-394)
 *   twos_closest      replaces  EmbreeMesh::closest_point / domain_distance
 *                               (wih_twos.cpp:339-359, :585-593)
 *   twos_trace        replaces  write_trace (wih_twos.cpp:1725-1744)
 *   twos_server_*     replace   the `--serve` loop (study.cpp:358-381), one
 *                               request at a time there, continuous batching here
 *
 * Conventions: plain pointers to caller-owned host memory; results are written
 * before return.  Status codes mirror the reference's exits: usage errors
 * (exit 2) -> TWOS_ERR_INVALID_ARGUMENT; "one or more paths hit --max-steps"
 * (:1846-1848) -> TWOS_ERR_STEP_CAP with results still written; a failed
 * closest-point query or ray parity overflow (:355-357, :390-392) ->
 * TWOS_ERR_GEOMETRY.  Nothing throws across the ABI; twos_last_error() gives
 * the message of the calling thread's last failure.
 */
#ifndef TWOS_H
#define TWOS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWOS_OK 0
#define TWOS_ERR_INVALID_ARGUMENT 2
#define TWOS_ERR_STEP_CAP 3
#define TWOS_ERR_GEOMETRY 4
#define TWOS_ERR_ADMISSION 5
#define TWOS_ERR_DEVICE 6
#define TWOS_ERR_NOT_FOUND 7

/* Domain containers. */
#define TWOS_BALL 0
#define TWOS_BOX 1
#define TWOS_MESH 2

/* Combine: SQUARE is evaluate_site's compliance score u^2 (a = b = one path,
 * wih_twos.cpp:687-715); PAIRED accumulates both sides on one path
 * (study.cpp:91-117); CROSS runs independent primal and adjoint streams. */
#define TWOS_SQUARE 0
#define TWOS_PAIRED 1
#define TWOS_CROSS 2
/* COUPLED is the coupled finite-cavity difference of study.cpp --fd
 * (study.cpp:317-332): each walk draws a start point x uniformly in the ball
 * of sides[1]'s single TWOS_TERM_UNIT row (the objective's density q, a
 * unit-mass bump), weight w = (1 - rho^2)^2 35/8 (q relative to the uniform
 * density, rho = |x - c| / R), and runs two paths of side 0 from x on the
 * same stream: a in the geometry, b in the geometry minus a ball of the
 * geometry's cavity_radius r at the site.  The sample is
 * d = w (a - b) / (4 pi r): score = mean d (the change of J = integral q u per
 * unit capacity when the cavity is cut), score_se = its standard error,
 * u = v = score, same = mean d^2.  Both paths count in steps and caps. */
#define TWOS_COUPLED 3

/* Streams: PHILOX keyed by (seed, site, side) with counter (step, walk);
 * MT_SITE / MT_PAIRED / MT_KEY / MT_FD replay the reference's mt19937_64
 * layouts (wih_twos.cpp:681-685, study.cpp:89, wih_twos.cpp:1737,
 * study.cpp:322-323), binary64 only; MT_FD goes with COUPLED. */
#define TWOS_RNG_PHILOX 0
#define TWOS_RNG_MT_SITE 1
#define TWOS_RNG_MT_PAIRED 2
#define TWOS_RNG_MT_KEY 3
#define TWOS_RNG_MT_FD 4

/* Schedules: PLAIN launches one kernel per site and round, one thread per
 * walk; PERSISTENT runs every site of every admitted request in one
 * persistent wavefront launch per round with walk regeneration. */
#define TWOS_PLAIN 0
#define TWOS_PERSISTENT 1

typedef struct twos_ctx twos_ctx;
typedef int32_t twos_status;

/* Source term types (source_kind 2), rows {type, cx, cy, cz, p0, p1, p2, w},
 * t = max(0, 1 - |y - c|^2 / (p0 p0)):
 *   TWOS_TERM_BUMP   w * (t t)             heat bump (benchmarks/submission_study/study.cpp:29-36)
 *   TWOS_TERM_UNIT   (t t) / p1            unit-mass sensor bump, p1 = 32 pi p0^3 / 105 (study.cpp:44-51)
 *   TWOS_TERM_GAUSS  w * exp(-((x-cx)/p0)^2 - ((y-cy)/p1)^2 - ((z-cz)/p2)^2)
 *                                          housing chips (benchmarks/compiled_bunny/src/wih_twos.cpp:553-566) */
#define TWOS_TERM_BUMP 0
#define TWOS_TERM_UNIT 1
#define TWOS_TERM_GAUSS 2
#define TWOS_MAX_TERMS 8

/* Boundary data and source of one side: the Dirichlet payoff at the closest
 * boundary point y of class c (0-5 container faces; ball and mesh use 0,
 * box faces are -x,+x,-y,+y,-z,+z; 6 analytic holes; 7 cavities) is
 *   cls_const[c] + (bit c of cls_poly_mask) * poly(y),
 * poly(y) = c0 + cx x + cy y + cz z + cxx x^2 + cyy y^2 + czz z^2 + cxy xy + cxz xz + cyz yz;
 * source_kind 1 adds (r^2/6) f(x) per step with the affine
 * f = ((s0 + sx (x+ox)) + sy (y+oy)) + sz (z+oz) at the ball centre (the
 * reference's 1 + slope (y + 0.5) is s0 = 1, sy = slope, oy = 0.5;
 * wih_twos.cpp:546-552); source_kind 2 adds (r^2/6) f(a) with
 * f = s0 + sum of the n_terms rows, at a Green-distributed point a of the
 * ball (a = x + r t omega, t ~ Beta(2,2); wih_twos.cpp:648-653,
 * study.cpp:106).  Both sides of a PAIRED path share one point a.
 * source_kind 3 (the reference's analytic-source variant, study.cpp:56-76,
 * :98-103) adds per step, instead of a sampled point, the exact integral of
 * the ball's Green function against s0 and each BUMP or UNIT row (amplitude w,
 * or 1/p1 for UNIT): for a bump of radius R and amplitude A centred at
 * distance d from the ball centre, Phi(d) minus the mean of Phi over the
 * ball's sphere, Phi the bump's Newtonian potential (binary64 in every
 * precision; no random draw). */
typedef struct {
  double poly[10];
  double cls_const[8];
  uint32_t cls_poly_mask;
  int32_t source_kind;
  double src[7];
  int32_t n_terms;
  int32_t pad;
  double terms[TWOS_MAX_TERMS][8];
} twos_side;

typedef struct {
  int32_t container;
  int32_t n_spheres;       /* holes: rows (cx, cy, cz, r) */
  int32_t n_cylinders;     /* holes: rows (axis 0|1|2, ca, cb, r), infinite, axis-aligned */
  int32_t n_cavities;      /* cavities: rows (cx, cy, cz) of radius cavity_radius */
  int32_t mesh;            /* mesh id for TWOS_MESH */
  int32_t mesh_radius_f32; /* 1: binary32-rounded mesh distance (the reference's recorded outputs) */
  double ball[4];
  double box_lo[3];
  double box_hi[3];
  double cavity_radius;
  const double *spheres;
  const double *cylinders;
  const double *cavities;
} twos_geometry_desc;

typedef struct {
  int32_t combine;
  int32_t rng;
  int32_t precision;    /* 32 or 64: arithmetic of walk positions and distances */
  int32_t schedule;
  uint64_t seed;
  int32_t repeat;       /* the reference's repeat index in MT_SITE keys */
  int32_t max_steps;
  double eps;
  int32_t walks;        /* fixed budget per side, or the first rung */
  int32_t walks_max;    /* > walks: double until ci_z * se <= ci_rel * |score| */
  double ci_rel;
  double ci_z;
  int32_t use_graph;    /* capture each round's launches in a hipGraph */
  int32_t block;        /* threads per block of the walk kernel (multiple of 64,
                           at most 256; 0 = 256) */
  int32_t waves_per_simd; /* persistent binary32 occupancy variant: 5, 6 or 8
                           waves per SIMD, 0 = the compiler's allocation */
  int32_t site_offset;  /* global index of sites[0] in stream keys */
  int32_t flags;        /* TWOS_FLAG_* */
  float grid_kappa;     /* distance grid: a step uses the grid bound when it is at
                           least grid_kappa cell sizes (0 = 4) */
  int32_t lds_stack;    /* binary32 persistent kernel: LDS traversal-stack entries
                           per lane, 16, 24 or 32 (0 = 24) */
  int32_t step_batch;   /* binary32 persistent kernel: 0, or 24, 32, 40 -- a lane
                           whose query has ended takes its walk step once that
                           many lanes of its wave are ready (1: at least as many
                           as are still traversing, and traversing lanes take
                           their next node visit once they outnumber the ready
                           ones or none is ready), or none traverses;
                           lds_stack 24 and waves_per_simd 0 or 5 only.  No
                           result depends on it. */
  int32_t refill_batch; /* persistent kernels: the idle lanes of a wave take new
                           paths once at least this many are idle (or none is
                           active), 1..64; 0 = 8.  No result depends on it. */
} twos_params;

/* Count BVH nodes fetched and triangles tested (slower; for roofline figures). */
#define TWOS_FLAG_COUNT_TRAVERSAL 1
/* Mesh steps deep inside the domain take their radius from a conservative
 * distance grid (a lower bound of the distance, so still a valid Walk on
 * Spheres step) instead of an exact closest-point query; steps near the
 * boundary stay exact.  Unbiased, but paths differ from the exact walk. */
#define TWOS_FLAG_DISTANCE_GRID 2
/* Gradient channel: each path's first step also gives a gradient sample
 * (3 / r0) (V - s0) omega0 + (r0^2 / 10) grad f, with r0, omega0 the first
 * sphere's radius and direction, V the path value, s0 its first source term
 * and grad f the affine source's gradient (the ball solution of -Lap w = f,
 * w = 0 on the sphere, has gradient r0^2/10 grad f at the centre).  Per site:
 * the means grad_u, grad_v and the cross-index product
 * grad_dot = (A.B - C)/(N(N-1)) (A, B the sums of the two sides' samples, C the
 * sum of their per-walk dot products), with a first-order (delta-method) SE.
 * No counterpart in the reference (an extension of the estimator); not
 * available with sampled (source_kind 2) sources. */
#define TWOS_FLAG_GRADIENT 4

typedef struct {
  double u;         /* mean of side-0 samples */
  double v;         /* mean of side-1 samples (= u for SQUARE) */
  double score;     /* (Sa Sb - Sab) / (N (N-1)): the reference's score column */
  double score_se;  /* leave-one-out jackknife SE of score; NaN if N < 4 */
  double td;        /* -score */
  double same;      /* Sab / N: same-index product (study.cpp same_index) */
  int64_t walks;    /* N per side */
  int64_t steps;    /* sphere steps over all of this site's paths */
  int64_t caps;     /* paths that hit max_steps */
  int32_t status;   /* 0 done, 1 budget exhausted before the CI target */
  int32_t pad;
  double grad_u[3]; /* TWOS_FLAG_GRADIENT: mean gradient samples of side 0 ... */
  double grad_v[3]; /* ... and side 1 (NaN without the flag) */
  double grad_dot;  /* cross-index estimate of grad u . grad v */
  double grad_dot_se;
} twos_site_result;

typedef struct {
  double seconds;           /* wall time of the estimate, host clock */
  double walk_kernel_ms;    /* summed duration of the walk kernel launches (HIP events) */
  double reduce_kernel_ms;
  int64_t walk_launches;
  int64_t steps;            /* all sphere steps of all paths */
  int64_t paths;
  int64_t nodes_visited;    /* BVH nodes fetched (when counting is on) */
  int64_t tris_tested;
  int32_t rounds;
  int32_t block;
  int32_t waves_per_simd;
  int32_t grid;
  int64_t grid_lookups;     /* distance-grid lookups (when counting is on) */
  int64_t grid_hits;        /* lookups whose bound was the step's radius, no traversal (counting) */
  int64_t top_nodes_visited; /* visits to the top three levels of the 4-wide BVH (counting) */
} twos_stats;

const char *twos_version(void);
const char *twos_last_error(void);

twos_status twos_create(int32_t device, uint64_t hbm_budget_bytes, twos_ctx **out);
void twos_destroy(twos_ctx *ctx);

/* A mesh whose vertex and face arrays equal a resident one bit for bit
 * (content hash, then a full comparison) returns that mesh's id, its LBVH and
 * distance grid shared, with *build_ms = 0; every create is matched by one
 * twos_mesh_release, and the last release frees it. */
twos_status twos_mesh_create(twos_ctx *ctx, const double *vertices, int64_t nv, const uint32_t *faces,
                             int64_t nf, int32_t *mesh_id, double *build_ms);
twos_status twos_mesh_info(twos_ctx *ctx, int32_t mesh_id, int32_t *depth, int64_t *nodes, int64_t *bytes);
/* Rebuilds the mesh's conservative distance grid with `cells` cells along the
 * longest axis of its bounding box (8..1024; meshes start at 128 cells, values
 * at cell centres) and updates every geometry on the mesh.  corners 1 samples
 * the distance at the cell corners and bounds a step by the best of a cell's
 * eight corners; corners 2 stores those eight values per cell rounded down to
 * binary16 in one 16-byte record (one load per lookup, four times the
 * memory).  Asking for the grid the mesh already has returns at once with
 * *build_ms = 0.  No counterpart in the reference (Embree has no distance
 * grid); it is a tuning control for TWOS_FLAG_DISTANCE_GRID. */
twos_status twos_mesh_set_distance_grid(twos_ctx *ctx, int32_t mesh_id, int32_t cells, int32_t corners,
                                        double *build_ms);
twos_status twos_mesh_release(twos_ctx *ctx, int32_t mesh_id);
/* Device memory the context accounts for: resident geometry (LBVHs, distance
 * grids with their start nodes, geometry rows, the geometry table), engines
 * (walk buffers, round blocks and the slots of admitted requests), the HBM
 * budget given to twos_create (0 there: a quarter of the free memory) and
 * the bytes one admitted site costs.  No counterpart in the reference (the
 * reference reports mesh_build_seconds and peak RSS, wih_twos.cpp:1746-1793). */
twos_status twos_memory_info(twos_ctx *ctx, int64_t *geometry_bytes, int64_t *engine_bytes, int64_t *budget_bytes,
                             int64_t *slot_bytes);

twos_status twos_geometry_create(twos_ctx *ctx, const twos_geometry_desc *desc, int32_t *geom_id);
twos_status twos_geometry_union_sphere(twos_ctx *ctx, int32_t base, const double center[3], double radius,
                                       int32_t *geom_id);
twos_status twos_geometry_release(twos_ctx *ctx, int32_t geom_id);

twos_status twos_estimate(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                          const double *sites, int64_t n, twos_site_result *out, twos_stats *stats);

/* As twos_estimate with the sites and results in device memory (for example
 * torch tensors' data_ptr(), ordered on their stream): d_sites holds n x 3
 * binary64 positions, d_out receives n twos_site_result records, both
 * ordered on `stream` (a hipStream_t; null for the null stream): the estimate
 * starts after the work already queued on it, and work queued on it after
 * the call sees the results.  Positions are not checked on the host: a
 * non-finite one ends its paths capped (TWOS_ERR_STEP_CAP).  Replaces the same
 * call as twos_estimate (wih_twos.cpp:1799-1864, the site loop :1828-1836), with
 * the caller-owned device buffers of the intended compiled-kernel boundary
 * (benchmarks/internal_cooling/README.md:75-76). */
twos_status twos_estimate_device(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                                 const double *d_sites, int64_t n, twos_site_result *d_out, twos_stats *stats,
                                 void *stream);

/* As twos_estimate_device, allocating nothing: no hipMalloc, hipFree or other
 * allocation of device or pinned memory in the call.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Every
 * device buffer of the estimate (site slots, round blocks, per-walk samples)
 * lives in one static 32-MiB device array per GPU, laid out when the context
 * is created; sites run in chunks of 4096, and walks or walks_max per side
 * must fit one round of that array (TWOS_ERR_ADMISSION with the limit
 * otherwise); use_graph and TWOS_FLAG_GRADIENT are rejected; stats carry no
 * kernel times.  Every context on a GPU shares that GPU's array: calls that
 * overlap on one GPU (from two contexts or threads) race; calls on one
 * context are serialised. */
twos_status twos_estimate_device_static(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2],
                                        const twos_params *params, const double *d_sites, int64_t n,
                                        twos_site_result *d_out, twos_stats *stats, void *stream);

/* As twos_estimate, also writing each path's samples a, b (binary64) and step
 * count (bit 30 set when capped) as [n][stride] arrays, walks >= stride dropped. */
twos_status twos_estimate_dump(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2],
                               const twos_params *params, const double *sites, int64_t n,
                               twos_site_result *out, twos_stats *stats, double *dump_a, double *dump_b,
                               int32_t *dump_steps, int64_t stride);

twos_status twos_inside(twos_ctx *ctx, int32_t mesh_id, const double *points, int64_t n, int32_t *out);
twos_status twos_closest(twos_ctx *ctx, int32_t geom_id, int32_t precision, const double *points, int64_t n,
                         double *distance, double *closest, int32_t *cls);
/* The conservative lower bound of the mesh distance that the walk takes from
 * the geometry's distance grid at each point (-inf without a grid).  No
 * counterpart in the reference: a diagnostic of TWOS_FLAG_DISTANCE_GRID. */
twos_status twos_distance_bound(twos_ctx *ctx, int32_t geom_id, int32_t precision, const double *points, int64_t n,
                                double *bound);
twos_status twos_trace(twos_ctx *ctx, int32_t geom_id, const twos_side *side, const twos_params *params,
                       const double start[3], int64_t site, int32_t walk, double *rows, int32_t max_rows,
                       int32_t *n_rows);

/* Differential Walk on Spheres: the derivative of u at each site with respect
 * to each cavity's centre and radius (replaces evaluate_differential_site,
 * benchmarks/compiled_bunny/src/wih_twos.cpp:725-928, over the sites as
 * wih_study --diff 1 runs it, benchmarks/submission_study/study.cpp:303-316).
 * geom_id: a mesh container, minus cavities only (no spheres or cylinders;
 * without cavities every sample is zero and only the counters move).  Per
 * site, params->walks samples: an outer path walks to the boundary without a
 * source; one that ends on cavity k at the point h with outward normal n
 * (into the cavity) is offset to off = h - n ell and, unless that offset is
 * unsafe (near the seam of two overlapping cavities, its segment crossing
 * another cavity, or off outside the domain / within eps of its boundary),
 * runs one nested path of `side` from off with value V; the sample adds
 * n V / ell to cavity k's centre derivative and -V / ell to its radius
 * derivative (other cavities get zero).  Streams: params->rng
 * TWOS_RNG_PHILOX is the production form -- Philox streams keyed by
 * (params->seed, params->repeat, params->site_offset + site) with the sample
 * and step in the counter, one persistent launch, params->precision 32 or 64,
 * and with TWOS_FLAG_DISTANCE_GRID mesh steps take the mesh's grid bound when
 * it clears max(grid_kappa cells, 2 eps) (cavity distances stay exact); any
 * other rng replays the reference's mt19937_64 keys (params->seed,
 * params->repeat, params->site_offset + site, sample) in binary64 (precision
 * must be 64).  out: [n][n_cavities][8] = the
 * centre derivative x y z, their standard errors, the radius derivative and
 * its standard error (means over all walks, unsafe and non-cavity samples
 * counting as zero).  counts (may be null): [n][9] = outer steps, nested
 * steps, cavity terminations, near-seam, segment-unsafe, unsafe, max_steps
 * caps, distance queries, ray-parity queries. */
twos_status twos_differential(twos_ctx *ctx, int32_t geom_id, const twos_side *side, const twos_params *params,
                              const double *sites, int64_t n, double ell, double *out, int64_t *counts);

/* Certified minimum-benefit sampling (replaces cert_walk / the sample loop of
 * run_minimum_benefit, benchmarks/compiled_bunny/src/minimum_benefit.hpp:
 * 122-188, wih_twos --mode validate).  Fail-closed walks, binary64 on the
 * reference's mt19937_64 keys: the distance is the smallest rounded-down
 * distance to a triangle's bounding box or a cavity, each step's radius
 * min(radius_cap, 0.95 d rounded down), and each step adds a rounded-down
 * lower bound of (r^2 / 6) f at the Green point; a sample stops at eps, at
 * value_cap (clipped, value = value_cap), at max_steps (capped) or when a
 * step's points cannot be shown inside the ball (containment failure: the
 * record must not be used).  Sources: TWOS_CERT_AFFINE 1 + slope (y + 0.5)
 * with slope >= 0, TWOS_CERT_HOUSING the built-in housing chips. */
#define TWOS_CERT_AFFINE 0
#define TWOS_CERT_HOUSING 1
#define TWOS_CERT_CAPPED 1
#define TWOS_CERT_CLIPPED 2
#define TWOS_CERT_CONTAINMENT 4

typedef struct {
  uint64_t seed;
  int32_t walks;         /* samples per site in this call */
  int32_t max_steps;
  double eps;
  double radius_cap;     /* --validation-radius-cap */
  double value_cap;      /* --validation-value-cap */
  double slope;          /* TWOS_CERT_AFFINE slope (--source-y-slope) */
  int32_t source;        /* TWOS_CERT_AFFINE or TWOS_CERT_HOUSING */
  int32_t philox;        /* 0: the reference's mt19937_64 keys (one thread per
                            sample); 1: Philox streams, the production form
                            (one persistent launch; sample j of site s draws
                            step k from counter (k, j, draw, 0) under the key
                            of (splitmix64(seed + 0xa1224), s); sample_begin +
                            walks at most 2^32) */
  int64_t site_offset;   /* global index of sites[0] in the keys */
  int64_t sample_begin;  /* index of the first sample of this call */
  int32_t grid_cells;    /* philox only: 0, or 8..512 -- a step whose certified
                            grid bound (the mesh's corner grid with this many
                            cells along its longest axis; built once per mesh
                            and resolution) is at least grid_kappa cells takes
                            it as its mesh distance instead of the exact
                            box-union query.  No counterpart in the reference. */
  float grid_kappa;      /* 0 = 2 cells; the threshold is at least 2 eps */
} twos_cert_params;

/* geom_id: a mesh container minus cavities only.  value, steps, flags:
 * [n][walks] for samples sample_begin .. sample_begin + walks - 1 of each
 * site (flags TWOS_CERT_*). */
twos_status twos_certify(twos_ctx *ctx, int32_t geom_id, const twos_cert_params *params, const double *sites,
                         int64_t n, double *value, int32_t *steps, int32_t *flags);
/* The certified lower bound of the domain distance at each point
 * (cert_domain_distance, minimum_benefit.hpp:83-88; the summary's
 * center_clearance_lower). */
twos_status twos_cert_distance(twos_ctx *ctx, int32_t geom_id, const double *points, int64_t n, double *out);
/* The certified grid's lower bound of the mesh's box-union distance at each
 * point (the grid with `cells` cells along the mesh's longest axis, built if
 * the mesh does not hold it yet): max over the point's cell corners c of
 * down(v(c) - up(|p - c|)).  No counterpart in the reference: a diagnostic of
 * twos_cert_params.grid_cells. */
twos_status twos_cert_grid_lower(twos_ctx *ctx, int32_t mesh_id, int32_t cells, const double *points, int64_t n,
                                 double *out);

/* Continuous-batching server: requests are admitted between rounds as device
 * memory allows, all admitted sites share each persistent launch, and each
 * result depends only on (seed, site, side, walk), alone or in the mix.
 * Admission by memory: a queued request is admitted when its slots are free
 * and its slot bytes fit the context's HBM budget beside everything resident
 * on the context (meshes with their distance grids, geometries, every
 * engine's buffers, the requests already admitted).  A geometry cannot be
 * released while a queued or active request uses it. */
typedef struct twos_server twos_server;
twos_status twos_server_create(twos_ctx *ctx, int64_t max_sites, int64_t max_walks_per_round,
                               twos_server **out);
void twos_server_destroy(twos_server *srv);
twos_status twos_server_submit(twos_server *srv, int32_t geom_id, const twos_side sides[2],
                               const twos_params *params, const double *sites, int64_t n,
                               twos_site_result *out, int64_t *request_id);
/* Runs one round over every admitted request; *n_done receives how many
 * requests completed in this round, their ids written to done_ids. */
twos_status twos_server_step(twos_server *srv, int64_t *done_ids, int32_t max_done, int32_t *n_done);
/* As twos_server_step, also writing each completed request's status to
 * done_status: TWOS_OK; TWOS_ERR_STEP_CAP when a path of one of its sites hit
 * max_steps (results written; the reference's serve loop answers DONE 1 and
 * exits, study.cpp:373-374, here the other requests go on); or the error that
 * kept it from running (TWOS_ERR_ADMISSION: its slots do not fit the HBM
 * budget beside the resident geometry even with nothing else admitted).  A
 * request that ended either way has released its slots. */
twos_status twos_server_step_status(twos_server *srv, int64_t *done_ids, int32_t *done_status, int32_t max_done,
                                    int32_t *n_done);
twos_status twos_server_pending(twos_server *srv, int64_t *queued, int64_t *active);

#ifdef __cplusplus
}
#endif

#endif
