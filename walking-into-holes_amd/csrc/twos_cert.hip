// Certified minimum-benefit sampling (run_minimum_benefit / cert_walk,
// benchmarks/compiled_bunny/src/minimum_benefit.hpp:122-188): fail-closed
// walks whose geometric and source quantities are rounded toward the safe
// side, binary64 on the reference's mt19937_64 streams.
//
// Distance.  The certified distance is the smallest rounded-down distance
// from the point to a triangle's bounding box (CertDistanceTree,
// minimum_benefit.hpp:39-81), then the cavities' rounded-down distances.  The
// per-box bound is monotone under box containment and every rounding in it
// is monotone, so it is found here over the mesh's 4-wide BVH (whose
// outward-rounded binary32 boxes contain the binary64 triangle boxes),
// pruning a node whose bound is no better than the best so far: the result
// is the minimum over all triangles, whatever the tree.  The square root is
// monotone too, so the search compares rounded-down squared sums and takes
// one root at the end.
//
// Work model.  The reference's streams: one thread runs one (site, sample)
// walk (cert_walk).  Philox streams (the production form): a persistent
// kernel (cert_persistent).  Each walk writes its value, step count and flags
// (1 capped, 2 clipped, 4 containment failure).

#include <algorithm>

#include "twos_device.h"
#include "twos_launch.h"

namespace twos {

namespace {

// nextafter(x, -inf) and nextafter(x, +inf) on the binary64 pattern.
__device__ __forceinline__ double dn(double x) {
  if (x != x || x == -__builtin_huge_val()) return x;
  if (x == 0.0) return -0x1p-1074;
  const long long b = __double_as_longlong(x);
  return __longlong_as_double(x > 0.0 ? b - 1 : b + 1);
}
__device__ __forceinline__ double upw(double x) {
  if (x != x || x == __builtin_huge_val()) return x;
  if (x == 0.0) return 0x1p-1074;
  const long long b = __double_as_longlong(x);
  return __longlong_as_double(x > 0.0 ? b + 1 : b - 1);
}
__device__ __forceinline__ double pdn(double x) { return fmax(0.0, dn(x)); }

// The rounded-down squared sum of box_distance_lower (minimum_benefit.hpp:18-25)
// before its root.
__device__ __forceinline__ double box_sum_lower(double lx, double ly, double lz, double hx, double hy, double hz,
                                                V3<double> p) {
  const double gx = fmax(0.0, fmax(dn(lx - p.x), dn(p.x - hx)));
  const double gy = fmax(0.0, fmax(dn(ly - p.y), dn(p.y - hy)));
  const double gz = fmax(0.0, fmax(dn(lz - p.z), dn(p.z - hz)));
  double sum = pdn(0.0 + pdn(gx * gx));
  sum = pdn(sum + pdn(gy * gy));
  return pdn(sum + pdn(gz * gz));
}

__device__ __forceinline__ double tri_sum_lower(const MeshView &m, int32_t k, V3<double> p) {
  V3<double> a, b, c;
  load_tri<double>(m, k, a, b, c);
  return box_sum_lower(fmin(fmin(a.x, b.x), c.x), fmin(fmin(a.y, b.y), c.y), fmin(fmin(a.z, b.z), c.z),
                       fmax(fmax(a.x, b.x), c.x), fmax(fmax(a.y, b.y), c.y), fmax(fmax(a.z, b.z), c.z), p);
}

// CertDistanceTree::distance: the smallest rounded-down triangle-box distance.
// Warm start: the search starts from `hint`'s value (the previous step's
// nearest triangle box; a walk moves by less than its distance, so this bound
// is tight) and leaves the nearest box's triangle there.  The minimum over all
// triangles does not depend on where the search starts.
__device__ double cert_mesh_distance(const MeshView &m, V3<double> p, int32_t &hint) {
  double best = __builtin_huge_val();
  if (m.n_tris == 1) return pdn(sqrt(tri_sum_lower(m, 0, p)));
  int32_t best_k = -1;
  if (hint >= 0 && hint < m.n_tris) {
    best = tri_sum_lower(m, hint, p);
    best_k = hint;
  }
  int32_t stk_node[kStackDepth];
  double stk_d[kStackDepth];
  int sp = 0;
  int32_t node = 0;
  int32_t visits = m.n_tris + 8;  // 4-wide nodes < triangles; guards a corrupt tree
  while (--visits >= 0) {
    const float4 *q = m.nodes4 + size_t(kNode4Quads) * node;
    const float4 lx = q[0], ly = q[1], lz = q[2], hx = q[3], hy = q[4], hz = q[5];
    const int4 ch = *reinterpret_cast<const int4 *>(q + 6);
    double d[4] = {box_sum_lower(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, p),
                   box_sum_lower(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, p),
                   box_sum_lower(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, p),
                   box_sum_lower(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, p)};
    int32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (c[i] < 0 && c[i] != kEmptyChild && d[i] < best) {
        const double v = tri_sum_lower(m, ~c[i], p);
        if (v < best) {
          best = v;
          best_k = ~c[i];
        }
      }
    double key[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) key[i] = (c[i] >= 0 && d[i] < best) ? d[i] : __builtin_huge_val();
    auto cswap = [&](int i, int j) {
      if (key[j] < key[i]) {
        const double tk = key[i]; key[i] = key[j]; key[j] = tk;
        const int32_t tc = c[i]; c[i] = c[j]; c[j] = tc;
      }
    };
    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
    const double inf = __builtin_huge_val();
    if (key[0] < inf) {
#pragma unroll
      for (int i = 3; i >= 1; --i)
        if (key[i] < inf && sp < kStackDepth) {
          stk_node[sp] = c[i];
          stk_d[sp] = key[i];
          ++sp;
        }
      node = c[0];
      continue;
    }
    bool found = false;
    while (sp > 0) {
      --sp;
      if (stk_d[sp] < best) {
        node = stk_node[sp];
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  hint = best_k;
  return pdn(sqrt(best));
}

// point_distance_bound (minimum_benefit.hpp:26-35).
__device__ __forceinline__ double point_bound(V3<double> a, V3<double> b, bool upper) {
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double d = fabs(comp(a, k) - comp(b, k));
    d = upper ? upw(d) : pdn(d);
    const double q = upper ? upw(d * d) : pdn(d * d);
    sum = upper ? upw(sum + q) : pdn(sum + q);
  }
  return upper ? upw(sqrt(sum)) : pdn(sqrt(sum));
}

// An upper bound of |a - b| from round-to-nearest arithmetic (the production
// form; the replay form keeps point_bound's per-operation upward rounding).
// With u = 2^-53, the three differences, three squares, two sums of
// non-negative terms and the square root (within one ulp) leave the computed
// s >= |a - b| (1 - 5u); the product s (1 + 2^-49) rounded to nearest is at
// least s (1 + 16u)(1 - u) > |a - b|.  The absolute 2^-500 covers squares
// below the normal range (a difference below 2^-511).  Eight operations
// instead of about sixty.
__device__ __forceinline__ double dist_upper_fast(V3<double> a, V3<double> b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const double sum = dx * dx + dy * dy + dz * dz;
  return sqrt(sum) * (1.0 + 0x1p-49) + 0x1p-500;
}

// Certified grid.  The exact box-union distance B is 1-Lipschitz, and a
// corner value v(c) = down(D(c)) <= B(c), so for any corner c,
// B(p) >= v(c) - |p - c| >= down(v(c) - U) for any upper bound U of |p - c|
// (dist_upper_fast).  The bound takes the best
// of the eight corners of p's cell (any cell gives a valid bound, so the index
// only has to be in range).  Corners are recomputed from the same binary64
// expression in the build and the lookup (this file has no contraction), so
// a lookup measures from the exact point the value belongs to.
__device__ __forceinline__ V3<double> cert_corner(const CertGrid &cg, int i, int j, int k) {
  return {cg.lo[0] + double(i) * cg.h, cg.lo[1] + double(j) * cg.h, cg.lo[2] + double(k) * cg.h};
}

__device__ __forceinline__ int cert_cell(double x, double lo, double inv_h, int corners) {
  return int(fmin(fmax(floor((x - lo) * inv_h), 0.0), double(corners - 2)));
}

__device__ double cert_grid_lower(const CertGrid &cg, V3<double> p) {
  const int i = cert_cell(p.x, cg.lo[0], cg.inv_h, cg.n[0]);
  const int j = cert_cell(p.y, cg.lo[1], cg.inv_h, cg.n[1]);
  const int k = cert_cell(p.z, cg.lo[2], cg.inv_h, cg.n[2]);
  double best = -__builtin_huge_val();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double v = double(cg.d[(size_t(i + a) * cg.n[1] + size_t(j + b)) * cg.n[2] + size_t(k + c)]);
        best = fmax(best, dn(v - dist_upper_fast(p, cert_corner(cg, i + a, j + b, k + c))));
      }
  return best;
}

// cert_domain_distance (:83-88).  mesh_lb >= 0: a certified lower bound of the
// mesh's box-union distance to take instead of the exact query.
__device__ double cert_domain_distance(const DevGeom &g, V3<double> p, int32_t &hint, double mesh_lb = -1.0) {
  const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
  double d = mesh_lb >= 0.0 ? mesh_lb : cert_mesh_distance(m, p, hint);
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const Vec4<double> s = load4<double>(g, 2, i);
    d = fmin(d, dn(point_bound(p, V3<double>{s.x, s.y, s.z}, false) - g.cavity_radius));
  }
  return fmax(0.0, d);
}

// (1 - t/1024)^1024 <= exp(-t), every operation rounded down (:92-96).
__device__ __forceinline__ double exp_negative_lower(double t_upper) {
  double a = fmax(0.0, dn(1.0 - upw(t_upper / 1024.0)));
  for (int i = 0; i < 10; ++i) a = pdn(a * a);
  return a;
}

// cert_source_lower (:97-116).
__device__ double cert_source_lower(V3<double> p, int32_t source, double slope) {
  if (source == TWOS_CERT_AFFINE) return fmax(0.0, dn(1.0 + dn(slope * dn(p.y + 0.5))));
  const double centers[3][3] = {{-.48, -.18, -.08}, {.43, .19, -.07}, {.02, -.31, -.02}};
  const double widths[3][3] = {{.24, .22, .16}, {.21, .20, .15}, {.30, .16, .19}};
  const double amplitude[3] = {5.2, 4.6, 2.0};
  double value = .08;
  for (int j = 0; j < 3; ++j) {
    double exponent = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double absdiff = upw(fabs(comp(p, a) - centers[j][a]));
      const double q = upw(absdiff / widths[j][a]);
      exponent = upw(exponent + upw(q * q));
    }
    value = dn(value + pdn(amplitude[j] * exp_negative_lower(exponent)));
  }
  return fmax(0.0, value);
}

struct CertLaunch {
  const double *sites;
  double *value;
  int32_t *steps;
  int32_t *flags;
  int64_t jobs;
  int64_t count;        // samples per site in this launch
  int64_t sample_begin;
  int64_t site_offset;
  uint64_t seed;
  int32_t max_steps;
  int32_t source;
  double eps, radius_cap, value_cap, slope;
  unsigned long long *counter;  // cert_persistent: job tickets, zero before the launch
};

// One step of a certified walk at `point` with certified distance d > eps:
// the step radius s, the source's Green point y (the housing chips only) and
// the next point from the step's draws; returns false when a point cannot be
// shown inside the ball (containment failure), else adds the step's
// rounded-down source term to value (cert_walk, :126-141).
// kFast (the production form): the containment check bounds |next - point|
// and |y - point| with dist_upper_fast; the replay form with point_bound, as
// the reference does.
template <bool kFast>
__device__ __forceinline__ bool cert_step(const CertLaunch &p, V3<double> &point, double d, V3<double> green_dir,
                                          double green_t, V3<double> dir, double &value) {
  const double s = fmin(p.radius_cap, pdn(0.95 * d));
  V3<double> y = point;
  if (p.source == TWOS_CERT_HOUSING) y = point + green_dir * (s * green_t);
  const V3<double> next = point + dir * s;
  const double un = kFast ? dist_upper_fast(point, next) : point_bound(point, next, true);
  const double uy = kFast ? dist_upper_fast(point, y) : point_bound(point, y, true);
  if (!(un < d) || !(uy < d)) return false;
  const double weight = pdn(pdn(s * s) / 6.0);
  const double increment = pdn(weight * cert_source_lower(y, p.source, p.slope));
  value = pdn(value + increment);
  point = next;
  return true;
}

// cert_walk (:122-143), sample j = sample_begin + job % count of site
// job / count, key (:168).
__global__ __launch_bounds__(64) void cert_walk(const DevGeom *geom, CertLaunch p) {
  const int64_t job = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (job >= p.jobs) return;
  const DevGeom &g = *geom;
  const int64_t site = job / p.count;
  const int64_t j = p.sample_begin + (job - site * p.count);
  const uint64_t key = splitmix64(p.seed) ^ splitmix64(uint64_t(p.site_offset + site) + 0xa1224ULL) ^
                       splitmix64(uint64_t(j) + 0x6242cULL);
  Mt64 mt;
  mt.seed(splitmix64(key));
  V3<double> point{p.sites[3 * site], p.sites[3 * site + 1], p.sites[3 * site + 2]};
  double value = 0.0;
  int32_t steps = 0, flags = 1, hint = -1;
  for (int32_t k = 0; k < p.max_steps; ++k) {
    const double d = cert_domain_distance(g, point, hint);
    if (d <= p.eps) {
      flags = 0;
      break;
    }
    // random_unit_vector(rng) * (s * beta_22(rng)): the Beta draw first.
    double t = 0.0;
    V3<double> gdir{0.0, 0.0, 0.0};
    if (p.source == TWOS_CERT_HOUSING) {
      t = mt.beta22();
      gdir = mt.direction();
    }
    if (!cert_step<false>(p, point, d, gdir, t, mt.direction(), value)) {
      flags = 4;
      break;
    }
    ++steps;
    if (value >= p.value_cap) {
      value = p.value_cap;
      flags = 2;
      break;
    }
  }
  p.value[job] = value;
  p.steps[job] = steps;
  p.flags[job] = flags;
}

// Production form on Philox streams: sample j of site s draws the move's
// direction of step k from counter (k, j, 0, 0) under the key of
// (splitmix64(seed + 0xa1224), s, 0), the Green point's direction from
// (k, j, 1, 0) and its Beta(2,2) (the median of three 24-bit uniforms) from
// (k, j, 2, 0).  Each wave takes jobs in chunks of kCertChunk tickets and each
// loop iteration advances every lane by one step of its walk, so lanes never
// wait for the wave's longest sample; every walk ends within max_steps steps.
constexpr int kCertBlock = 256;
// Tickets per chunk: a launch holds about 1 M jobs, so small chunks shorten
// the tail of unevenly filled waves (measured: 16, 32, 64 and 128 gave 30.99,
// 31.39, 32.54 and 34.68 ms per step of the certify workload).
constexpr int kCertChunk = 16;

__device__ __forceinline__ double cert_beta22(uint2 key, uint32_t j, uint32_t k) {
  const uint4 w = philox4x32_10(make_uint4(k, j, 2u, 0u), key);
  const double a = double(w.x >> 8) * 0x1.0p-24, b = double(w.y >> 8) * 0x1.0p-24, c = double(w.z >> 8) * 0x1.0p-24;
  return fmax(fmin(a, b), fmin(fmax(a, b), c));
}

// With a certified grid (grid.d), a step whose grid bound is at least
// grid.tau (> eps, so such a step never ends a walk) takes it as the mesh's
// distance; the cavities stay exact, and every walk ends on an exact query.
__global__ __launch_bounds__(kCertBlock) void cert_persistent(const DevGeom *geom, CertLaunch p, uint64_t cseed,
                                                              CertGrid grid) {
  const DevGeom &g = *geom;
  const int lane = int(threadIdx.x & 63);
  int64_t chunk_cur = 0, chunk_end = 0;  // wave-uniform
  bool drained = false;                  // wave-uniform
  bool active = false;
  int64_t job = 0;
  uint32_t j = 0;
  uint2 key = make_uint2(0u, 0u);
  V3<double> point{0.0, 0.0, 0.0};
  double value = 0.0;
  int32_t steps = 0, hint = -1;
  while (true) {
    uint64_t need = __ballot(!active);
    while (need != 0 && !drained) {
      if (chunk_cur >= chunk_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(p.counter, (unsigned long long)kCertChunk);
        base = __shfl(base, 0);
        if (int64_t(base) >= p.jobs) {
          drained = true;
          break;
        }
        chunk_cur = int64_t(base);
        chunk_end = min(int64_t(base) + kCertChunk, p.jobs);
      }
      const int take = int(min<int64_t>(chunk_end - chunk_cur, __popcll(need)));
      const int rank = __popcll(need & ((1ull << lane) - 1ull));
      if (!active && rank < take) {
        job = chunk_cur + rank;
        const int64_t site = job / p.count;
        j = uint32_t(p.sample_begin + (job - site * p.count));
        const uint64_t k = splitmix64(splitmix64(cseed) + (uint64_t(p.site_offset + site) << 1));
        key = make_uint2(uint32_t(k), uint32_t(k >> 32));
        point = V3<double>{p.sites[3 * site], p.sites[3 * site + 1], p.sites[3 * site + 2]};
        value = 0.0;
        steps = 0;
        hint = -1;
        active = true;
      }
      chunk_cur += take;
      need = __ballot(!active);
    }
    if (!__any(active)) break;
    if (!active) continue;
    int32_t flags = -1;  // -1: the walk goes on
    if (steps >= p.max_steps) {
      flags = 1;
    } else {
      const double lb = grid.d ? cert_grid_lower(grid, point) : -1.0;
      const double d = cert_domain_distance(g, point, hint, lb >= grid.tau ? lb : -1.0);
      if (d <= p.eps) {
        flags = 0;
      } else {
        double t = 0.0;
        V3<double> gdir{0.0, 0.0, 0.0};
        if (p.source == TWOS_CERT_HOUSING) {
          t = cert_beta22(key, j, uint32_t(steps));
          gdir = philox_direction(philox4x32_10(make_uint4(uint32_t(steps), j, 1u, 0u), key), 0.0);
        }
        const V3<double> dir = philox_direction(philox4x32_10(make_uint4(uint32_t(steps), j, 0u, 0u), key), 0.0);
        if (!cert_step<true>(p, point, d, gdir, t, dir, value)) {
          flags = 4;
        } else {
          ++steps;
          if (value >= p.value_cap) {
            value = p.value_cap;
            flags = 2;
          }
        }
      }
    }
    if (flags >= 0) {
      p.value[job] = value;
      p.steps[job] = steps;
      p.flags[job] = flags;
      active = false;
    }
  }
}

// A thread takes kCertGridRun consecutive corners along the first index,
// each query warm-started at the previous corner's nearest box (the minimum
// does not depend on the start); the last index is fastest across threads, so
// neighbouring lanes take neighbouring corners and similar paths through the
// tree.
constexpr int kCertGridRun = 8;

__global__ void cert_grid_build(MeshView m, CertGrid cg, float *out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n12 = int64_t(cg.n[1]) * cg.n[2];
  const int runs = (cg.n[0] + kCertGridRun - 1) / kCertGridRun;
  if (t >= int64_t(runs) * n12) return;
  const int r = int(t / n12), j = int((t - int64_t(r) * n12) / cg.n[2]), k = int(t % cg.n[2]);
  int32_t hint = -1;
  for (int i = r * kCertGridRun; i < min(cg.n[0], (r + 1) * kCertGridRun); ++i) {
    const double d = cert_mesh_distance(m, cert_corner(cg, i, j, k), hint);
    // Nearest binary32, then one ulp down when that rounded up (d >= 0).
    float f = float(d);
    if (double(f) > d) f = __uint_as_float(__float_as_uint(f) - 1u);
    out[int64_t(i) * n12 + int64_t(j) * cg.n[2] + k] = f;
  }
}

__global__ void cert_grid_lower_kernel(CertGrid cg, const double *pts, int64_t n, double *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = cert_grid_lower(cg, V3<double>{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
}

__global__ void cert_distance_kernel(const DevGeom *geom, const double *pts, int64_t n, double *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t hint = -1;
  out[i] = cert_domain_distance(*geom, V3<double>{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, hint);
}

inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

hipError_t launch_certify(const DevGeom *d_geom, const twos_cert_params &prm, const double *d_sites, int64_t n_sites,
                          int64_t sample_begin, int64_t count, double *d_value, int32_t *d_steps, int32_t *d_flags,
                          unsigned long long *counter, int cus, const CertGrid &grid, hipStream_t stream) {
  CertLaunch p{};
  p.sites = d_sites;
  p.value = d_value;
  p.steps = d_steps;
  p.flags = d_flags;
  p.jobs = n_sites * count;
  p.count = count;
  p.sample_begin = sample_begin;
  p.site_offset = prm.site_offset;
  p.seed = prm.seed;
  p.max_steps = prm.max_steps;
  p.source = prm.source;
  p.eps = prm.eps;
  p.radius_cap = prm.radius_cap;
  p.value_cap = prm.value_cap;
  p.slope = prm.slope;
  if (p.jobs <= 0) return hipSuccess;
  if (prm.philox) {
    p.counter = counter;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cert_persistent, kCertBlock, 0);
    if (e != hipSuccess) return e;
    const int64_t chunks = (p.jobs + 63) / 64;
    const int64_t blocks = std::min<int64_t>(int64_t(std::max(1, per_cu)) * std::max(1, cus),
                                             (chunks + kCertBlock / 64 - 1) / (kCertBlock / 64));
    e = hipMemsetAsync(counter, 0, sizeof(unsigned long long), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cert_persistent, dim3(unsigned(blocks)), dim3(kCertBlock), 0, stream, d_geom, p,
                       splitmix64(prm.seed + 0xa1224ULL), grid);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(cert_walk, dim3(blocks_for(p.jobs, 64)), dim3(64), 0, stream, d_geom, p);
  return hipGetLastError();
}

hipError_t build_cert_grid(const MeshBuffers &mb, const CertGrid &grid, hipStream_t stream) {
  const MeshView m{mb.nodes, mb.nodes4, mb.tris_f, mb.tris_d, mb.tri_ids, mb.n_tris};
  const int64_t n = int64_t((grid.n[0] + kCertGridRun - 1) / kCertGridRun) * grid.n[1] * grid.n[2];
  hipLaunchKernelGGL(cert_grid_build, dim3(blocks_for(n, 64)), dim3(64), 0, stream, m, grid,
                     const_cast<float *>(grid.d));
  return hipGetLastError();
}

hipError_t launch_cert_grid_lower(const CertGrid &grid, const double *d_pts, int64_t n, double *d_out,
                                  hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(cert_grid_lower_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, stream, grid, d_pts, n, d_out);
  return hipGetLastError();
}

hipError_t launch_cert_distance(const DevGeom *d_geom, const double *d_pts, int64_t n, double *d_out,
                                hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(cert_distance_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, stream, d_geom, d_pts, n, d_out);
  return hipGetLastError();
}

}  // namespace twos
