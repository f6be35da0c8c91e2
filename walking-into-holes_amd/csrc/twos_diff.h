// Differential Walk on Spheres: cavity-shape derivatives of the estimator's
// value (evaluate_differential_site, benchmarks/compiled_bunny/src/
// wih_twos.cpp:725-928, as wih_study --diff 1 runs it,
// benchmarks/submission_study/study.cpp:303-316).  Device code shared by the
// binary64 file (twos_diff.hip: the reference's mt19937_64 streams, Philox
// binary64, the fold) and the binary32 file (twos_diff_f32.hip).
//
// A job is one (site, sample) pair:
//   1. the outer path walks to the boundary with no source (walk_to_boundary,
//      wih_twos.cpp:725-740), keeping the last query's closest point, its
//      outward normal and which cavity (if any) it lies on;
//   2. a path that ended on a cavity is offset by ell into the domain,
//      off = hit - n ell, and classified (classify_offset_safety,
//      wih_twos.cpp:764-833): near the seam circle of an overlapping cavity,
//      a segment hit - off that passes through another cavity, or an offset
//      point outside the domain make the sample unsafe;
//   3. a safe sample runs one nested path from off with the side's boundary
//      data and source (walk(), wih_twos.cpp:635-664); its value V gives the
//      centre derivative sample n V / ell and the radius sample -V / ell.
// Each job writes its sample and counters to a per-job record; diff_fold then
// folds each (site, cavity) pair's samples in sample order, so a result
// depends only on (seed, repeat, site, sample), never on which lane ran it.
#pragma once

#include "twos_device.h"
#include "twos_ray.h"

namespace twos {

// Job record flags.
enum : int32_t {
  kDiffCavity = 1,    // the outer path ended on a cavity
  kDiffSeam = 2,      // ... near the seam of two overlapping cavities
  kDiffSegment = 4,   // ... with an unsafe offset segment or offset point
  kDiffUnsafe = 8,    // the sample was dropped (seam or segment)
  kDiffLimit = 16,    // a path reached max_steps
};

// domain_distance (wih_twos.cpp:585-593): one mesh query, then the cavities.
__device__ __forceinline__ double diff_domain_distance(const DevGeom &g, V3<double> p, int32_t &hint) {
  const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
  V3<double> q;
  double best = mesh_distance<double>(g, mesh_closest<double, false>(m, p, q, hint, nullptr));
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const Vec4<double> s = load4<double>(g, 2, i);
    const V3<double> d = p - V3<double>{s.x, s.y, s.z};
    best = fmin(best, sqrt(dot(d, d)) - g.cavity_radius);
  }
  return best;
}

// classify_offset_safety (wih_twos.cpp:764-833) with the enclosure fallback
// the study executable compiles in (study.cpp:8), binary64 in both
// precisions: returns kDiffSeam and/or kDiffSegment.
__device__ __forceinline__ int32_t offset_safety(const DevGeom &g, int32_t hit, V3<double> bp, V3<double> off,
                                                 double ell, double eps, int32_t &queries, int32_t &rays,
                                                 int32_t &hint) {
  bool near_seam = false, segment_unsafe = false;
  const double r = g.cavity_radius;
  const V3<double> seg = off - bp;
  const double seg2 = dot(seg, seg);
  const double tol = 1.0e-10 * fmax(1.0, fabs(r));
  const double coll2 = (r + tol) * (r + tol);
  const Vec4<double> h4 = load4<double>(g, 2, hit);
  const V3<double> hc{h4.x, h4.y, h4.z};
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    if (i == hit) continue;
    const Vec4<double> c4 = load4<double>(g, 2, i);
    const V3<double> ci{c4.x, c4.y, c4.z};
    const V3<double> delta = ci - hc;
    const double cd = sqrt(dot(delta, delta));
    if (cd > tol && cd < 2.0 * r) {
      // Distance from the hit point to the seam circle of two equal spheres.
      const V3<double> axis = delta * (1.0 / cd);
      const V3<double> seam_c = (hc + ci) * 0.5;
      const double seam_r = sqrt(fmax(0.0, r * r - 0.25 * (cd * cd)));
      const V3<double> rel = bp - seam_c;
      const double ax = dot(rel, axis);
      const V3<double> rad = rel - axis * ax;
      const double nr = sqrt(dot(rad, rad)) - seam_r;
      near_seam = near_seam || sqrt(ax * ax + nr * nr) <= ell;
    }
    if (seg2 > 0.0) {
      const double t = fmin(fmax(dot(ci - bp, seg) / seg2, 0.0), 1.0);
      const V3<double> cl = bp + seg * t;
      if (dot(cl - ci, cl - ci) <= coll2) segment_unsafe = true;
    }
  }
  const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
  bool inside = inside_parity(m, off, &rays) == 1;
  if (!inside && inside_parity(m, hc, &rays) == 1) {
    V3<double> q;
    ++queries;
    const double dc = mesh_distance<double>(g, mesh_closest<double, false>(m, hc, q, hint, nullptr));
    const V3<double> oc = off - hc;
    if (dc > r + ell + eps && sqrt(dot(oc, oc)) <= r + ell + eps * 0.5) inside = true;
  }
  if (!inside) {
    segment_unsafe = true;
  } else {
    ++queries;
    if (!(diff_domain_distance(g, off, hint) > eps)) segment_unsafe = true;
  }
  return (near_seam ? kDiffSeam : 0) | (segment_unsafe ? kDiffSegment : 0);
}

// Writes job `job`'s record: val[4 job + 0..2] = the outer hit's normal,
// val[4 job + 3] = V / ell; rec[6 job + 0..5] = cavity (-1: no sample),
// flags, outer steps, nested steps, distance queries, ray queries.
template <typename R>
__device__ __forceinline__ void diff_record(const DiffParams &p, int64_t job, V3<R> normal, double scaled,
                                            int32_t cavity, int32_t flags, int32_t outer, int32_t nested,
                                            int32_t queries, int32_t rays) {
  double *v = p.val + 4 * job;
  v[0] = double(normal.x);
  v[1] = double(normal.y);
  v[2] = double(normal.z);
  v[3] = scaled;
  int32_t *r = p.rec + 6 * job;
  r[0] = cavity;
  r[1] = flags;
  r[2] = outer;
  r[3] = nested;
  r[4] = queries;
  r[5] = rays;
}

// ---------------------------------------------------------------------------
// Production form: Philox streams, binary32 or binary64 walks, optional
// distance grid.  Streams: sample n's outer path draws the direction of step
// k from counter (k, n, 0, 0) under the key of (dseed, site, 0); its nested
// path is walk n of the key of (dseed, site, 1) (directions (k, n, 0, 0), the
// Green point (k, n, 1..2, 0)), with dseed = splitmix64(seed +
// 0x6a09e667f3bcc909) ^ splitmix64(repeat + 0xbb67ae8584caa73b) computed on
// the host (DiffParams::pseed).

__device__ __forceinline__ uint2 diff_philox_key(uint64_t dseed, int64_t site, int32_t side) {
  const uint64_t k = splitmix64(splitmix64(dseed) + ((uint64_t(site) << 1) | uint64_t(side & 1)));
  return make_uint2(uint32_t(k), uint32_t(k >> 32));
}

// Cavities after the mesh: each wins only when strictly closer, as
// closest_boundary orders them (wih_twos.cpp:602-626).  cavity: the winning
// cavity (-1: the mesh), rad its centre-to-p vector and len that vector's
// length.
template <typename R>
__device__ __forceinline__ R diff_cavities(const DevGeom &g, V3<R> p, R best, int32_t &cavity, V3<R> &rad, R &len) {
  cavity = -1;
  const R cr = R(g.cavity_radius);
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const Vec4<R> s = load4<R>(g, 2, i);
    const V3<R> d = p - V3<R>{s.x, s.y, s.z};
    const R l = wsqrt(dot(d, d));
    const R dist = l - cr;
    if (dist < best && l > R(0)) {
      best = dist;
      cavity = i;
      rad = d;
      len = l;
    }
  }
  return best;
}

// The exact mesh distance at p (point: the closest point), bounded by the
// grid's upper bound and started at the cell's start node when the grid is
// on (tau finite).
template <typename R>
__device__ __forceinline__ R diff_mesh_exact(const DevGeom &g, const MeshView &m, V3<R> p, R tau, int32_t &hint,
                                             V3<R> &point) {
  R ub = Inf<R>::v();
  if (tau < Inf<R>::v()) (void)grid_lower<R>(g, p, &ub);
  const R init2 = ub < Inf<R>::v() ? ub * ub : ub;
  return mesh_distance<R>(g, mesh_closest<R, false>(m, p, point, hint, nullptr, init2, grid_start_node<R>(g, p, init2)));
}

// Persistent kernel: each wave takes jobs in chunks of kDiffChunk consecutive tickets
// (one atomic per chunk; consecutive lanes start at the same site).  A lane
// is ready to step, waiting for an exact closest-point query, or waiting for
// the offset-safety test of an outer path that ended on a cavity; each loop
// iteration the wave runs one of three code blocks for the lanes in that
// state: the safety tests once at least p.safety_batch lanes wait for one
// (or nothing else is left), the exact queries once at least as many lanes
// wait for one as are ready to step, else the step (whose distance is the
// grid's lower bound when that clears tau; a lane whose bound does not waits
// for an exact query).  So the traversal and the binary64 safety code run for
// many lanes at once instead of stalling the wave for one lane at a time.
// Which code block runs when changes no path: each path's distances and draws
// are the same.  A wave leaves once the tickets are exhausted and none of its
// lanes holds a job; every job ends within 2 max_steps steps.
constexpr int kDiffBlock = 256;
// Tickets per chunk: a launch holds about 1 M jobs, so small chunks shorten
// the tail of unevenly filled waves (measured: 16, 32, 64 and 128 gave 12.53,
// 13.69, 15.70 and 21.17 ms per step of the Differential WoS workload).
constexpr int kDiffChunk = 16;

template <typename R>
__global__ __launch_bounds__(kDiffBlock) void diff_persistent(const DevGeom *geom, const DevSide *side, DiffParams p) {
  const DevGeom &g = *geom;
  const DevSide &s = *side;
  const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
  const int lane = int(threadIdx.x & 63);
  const R eps = R(p.eps);
  const R tau = (p.grid_kappa > 0.f && g.grid) ? fmax(R(p.grid_kappa) * R(g.grid_h), R(2) * eps) : Inf<R>::v();
  const R cr = R(g.cavity_radius);
  const bool sampled = s.source_kind == 2;
  int64_t chunk_cur = 0, chunk_end = 0;  // wave-uniform
  bool drained = false;                  // wave-uniform
  bool active = false;
  // Lane state.  pend: 0 ready to step, 1 waiting for an exact query, 2
  // waiting for the safety test of its outer hit (hit point in x, normal in
  // normal, cavity in cavity).
  V3<R> x{R(0), R(0), R(0)}, normal{R(0), R(0), R(0)};
  R value = R(0);
  int64_t job = 0;
  int64_t site = 0;
  uint2 key = make_uint2(0u, 0u);
  int32_t sample = 0, steps = 0, hint = -1, phase = 0, cavity = -1, pend = 0;
  int32_t flags = 0, outer = 0, nested = 0, queries = 0, rays = 0;
  while (true) {
    uint64_t need = __ballot(!active);
    while (need != 0 && !drained) {
      if (chunk_cur >= chunk_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(p.counter, (unsigned long long)kDiffChunk);
        base = __shfl(base, 0);
        if (int64_t(base) >= p.jobs) {
          drained = true;
          break;
        }
        chunk_cur = int64_t(base);
        chunk_end = min(int64_t(base) + kDiffChunk, p.jobs);
      }
      const int take = int(min<int64_t>(chunk_end - chunk_cur, __popcll(need)));
      const int rank = __popcll(need & ((1ull << lane) - 1ull));
      if (!active && rank < take) {
        job = chunk_cur + rank;
        const int64_t local = job / p.walks;
        sample = int32_t(job - local * p.walks);
        site = p.site_base + local;
        x = V3<R>{R(p.sites[3 * site]), R(p.sites[3 * site + 1]), R(p.sites[3 * site + 2])};
        key = diff_philox_key(p.pseed, p.site_offset + site, 0);
        value = R(0);
        steps = 0;
        hint = -1;
        phase = 0;
        cavity = -1;
        pend = 0;
        flags = outer = nested = queries = rays = 0;
        active = true;
      }
      chunk_cur += take;
      need = __ballot(!active);
    }
    if (!__any(active)) break;
    // Wave-uniform choice of the code block.
    const int nr = __popcll(__ballot(active && pend == 0));
    const int nq = __popcll(__ballot(active && pend == 1));
    const int ns = __popcll(__ballot(active && pend == 2));
    const int mode = (ns > 0 && (ns >= p.safety_batch || nr + nq == 0)) ? 2 : ((nq > 0 && nq >= nr) ? 1 : 0);
    if (!active) continue;
    bool done = false;
    if (mode == 2) {
      if (pend == 2) {
        // The outer path ended on `cavity` at x with outward normal `normal`
        // (boundary_hit's operation order, wih_twos.cpp:602-626): classify
        // the offset point in binary64.
        const V3<double> bpd{double(x.x), double(x.y), double(x.z)};
        const V3<double> nd{double(normal.x), double(normal.y), double(normal.z)};
        const V3<double> off = bpd - nd * p.ell;
        const int32_t safety = offset_safety(g, cavity, bpd, off, p.ell, p.eps, queries, rays, hint);
        flags |= safety;
        pend = 0;
        if (safety) {
          flags |= kDiffUnsafe;
          done = true;
        } else {
          // The nested path from the offset point.
          phase = 1;
          x = V3<R>{R(off.x), R(off.y), R(off.z)};
          key = diff_philox_key(p.pseed, p.site_offset + site, 1);
          steps = 0;
        }
      }
    } else if (mode == 1 ? pend == 1 : pend == 0) {
      if (mode == 0 && steps >= p.max_steps) {
        // The cap check precedes the query (walk_to_boundary, walk()).
        flags |= kDiffLimit;
        done = true;
        if (phase == 0) cavity = -1;
      } else {
        V3<R> bp = x, rad{R(0), R(0), R(0)};
        R len = R(0);
        int32_t cav = -1;
        R r;
        bool have = true;
        if (mode == 1 || !(tau < Inf<R>::v())) {
          // Without the grid every step is exact, so it runs in the step block.
          r = diff_mesh_exact<R>(g, m, x, tau, hint, bp);
          pend = 0;
        } else {
          const R lb = tau < Inf<R>::v() ? grid_lower<R>(g, x) : R(-1);
          r = lb;
          have = lb >= tau;  // a grid step never ends a path (tau > eps)
          if (!have) pend = 1;
        }
        if (have) {
          r = diff_cavities<R>(g, x, r, cav, rad, len);
          ++queries;
          if (phase == 0) {
            if (!(r > eps)) {
              // The outer path ended: the hit's point and outward normal.
              if (cav >= 0) {
                const V3<R> dir = rad * (R(1) / len);
                const Vec4<R> c4 = load4<R>(g, 2, cav);
                x = V3<R>{c4.x, c4.y, c4.z} + dir * cr;
                normal = dir * R(-1);
                flags |= kDiffCavity;
                cavity = cav;
                pend = 2;
              } else {
                normal = r > R(0) ? (bp - x) * (R(1) / r) : V3<R>{R(0), R(0), R(0)};
                done = true;
              }
            } else {
              x = x + philox_direction(philox4x32_10(make_uint4(uint32_t(steps), uint32_t(sample), 0u, 0u), key), R(0)) * r;
              ++steps;
              ++outer;
            }
          } else if (!isfinite(r)) {
            // A failed query ends the nested path as capped (wih_twos.cpp:355-357).
            flags |= kDiffLimit;
            done = true;
          } else if (!(r > eps)) {
            int cls = 0;
            if (cav >= 0) {
              const Vec4<R> c4 = load4<R>(g, 2, cav);
              bp = V3<R>{c4.x, c4.y, c4.z} + rad * (cr / len);
              cls = 7;
            }
            value += side_payoff<R>(s, bp, cls);
            done = true;
          } else {
            if (s.source_kind) value += side_green_term<R>(s, x, sampled ? x + philox_green<R>(key, sample, steps, r) : x, r, false);
            x = x + philox_direction(philox4x32_10(make_uint4(uint32_t(steps), uint32_t(sample), 0u, 0u), key), R(0)) * r;
            ++steps;
            ++nested;
          }
        }
      }
    }
    if (done) {
      const bool has = phase == 1;
      diff_record<R>(p, job, normal, has ? double(value) / p.ell : 0.0, has ? cavity : -1, flags, outer, nested,
                     queries, rays);
      active = false;
    }
  }
}

}  // namespace twos

namespace twos {

// One persistent launch over the tickets of p (p.counter zeroed first): as
// many blocks as fit on the device at once, and no more than the job chunks.
template <typename R>
hipError_t launch_diff_persistent(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int cus,
                                  hipStream_t stream) {
  if (p.jobs <= 0) return hipSuccess;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, diff_persistent<R>, kDiffBlock, 0);
  if (e != hipSuccess) return e;
  const int64_t chunks = (p.jobs + 63) / 64;
  const int64_t blocks = std::min<int64_t>(int64_t(std::max(1, per_cu)) * std::max(1, cus),
                                           (chunks + kDiffBlock / 64 - 1) / (kDiffBlock / 64));
  hipLaunchKernelGGL(diff_persistent<R>, dim3(unsigned(blocks)), dim3(kDiffBlock), 0, stream, d_geom, d_side, p);
  return hipGetLastError();
}

}  // namespace twos
