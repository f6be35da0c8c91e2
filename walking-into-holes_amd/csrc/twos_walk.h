// Walk on Spheres kernels (templates; instantiated per precision in
// twos_walk_f32.hip and twos_walk_f64.hip).
//
// Work model.  A round hands the kernels a list of active site slots with a
// walk count each (RoundArrays).  Walk t of the round belongs to slot
// act_slot[k] where act_prefix[k] <= t < act_prefix[k+1]; its global walk
// index is walk_base[k] + (t - act_prefix[k]).  A job is one path: walk t
// side s (CROSS runs two independent paths per walk, SQUARE and PAIRED one).
// Every job writes its binary64 sample(s) to val[2t + 0/1] and its step count
// to steps[2t + s]; the reduction then folds each site's samples in walk
// order, so a result depends only on (seed, site, side, walk), never on which
// lane, wave or launch ran the path.
#pragma once

#include "twos_device.h"

namespace twos {

struct Tables {
  const DevGeom *geoms;
  const DevSide *sides;
  SlotArrays slots;
  RoundArrays round;
};

__device__ __forceinline__ int32_t find_site(const int64_t *prefix, int32_t n, int64_t t, int32_t lo = 0) {
  // Largest k with prefix[k] <= t (prefix[0] = 0, prefix[n] = total), given
  // prefix[lo] <= t: most calls land in site lo itself.
  if (prefix[lo + 1] > t) return lo;
  int32_t hi = n;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename R>
struct Job {
  V3<R> x;
  R acc0, acc1;
  int32_t steps;
  int32_t side;
  int32_t walk;
  int32_t slot;
  int32_t max_steps;
  int32_t n_cav;  // the geometry's cavity count (read once per path, not per step)
  R tau;         // grid-radius threshold (infinite when the grid is off)
  int32_t hint;  // sorted index of the previous closest triangle, -1 before the first query
  R eps;
  int64_t t;
  uint2 key;
  const DevGeom *g;
  const DevSide *s0;
  const DevSide *s1;
  MeshView mv;  // the geometry's mesh buffers, kept per lane (unused fields fold away)
  // (kPipe kernels) the direction of the path's next step, drawn while the
  // distance-grid sample of its current position is in flight.
  V3<R> dn;
  // Gradient channel (kernels built with kChanGrad only): the first sphere's
  // 3 omega0 / r0, its radius (0 until the first step) and the path values
  // after the first step's source term.
  V3<R> gw;
  R gr, gs0, gs1;
  // Coupled finite-cavity pair (kChanCoupled only): the walk's start point,
  // the candidate cavity (centre = the site, radius), the start point's
  // weight, the first path's value and step word, which path runs (0: the
  // geometry, 1: minus the cavity) and the start point's closest triangle.
  V3<R> x0, z;
  R rc, wgt, a0;
  int32_t st0, phase, hint_start;
};

// site_lo: a site whose first walk is at or before this job's walk (0, or
// the site of the wave's ticket chunk start).
template <typename R>
__device__ __forceinline__ void job_init(Job<R> &j, const Tables &tb, const WalkParams &wp, int64_t u,
                                         int32_t site_lo = 0) {
  const int32_t n_active = *tb.round.n_active;
  const int64_t t = wp.jobs_per_walk == 2 ? (u >> 1) : u;  // 1 or 2 jobs per walk
  j.side = int32_t(u - t * wp.jobs_per_walk);
  const int32_t k = find_site(tb.round.act_prefix, n_active, t, site_lo);
  const int32_t slot = tb.round.act_slot[k];
  j.slot = slot;
  j.t = t;
  j.walk = tb.round.walk_base[k] + int32_t(t - tb.round.act_prefix[k]);
  const double *p = tb.slots.pos + 3 * size_t(slot);
  j.x = {R(p[0]), R(p[1]), R(p[2])};
  j.acc0 = R(0);
  j.acc1 = R(0);
  j.steps = 0;
  j.gr = R(0);
  j.hint = tb.slots.hint0[slot];
  j.max_steps = tb.slots.max_steps[slot];
  j.eps = R(tb.slots.eps[slot]);
  j.key = tb.slots.keys[2 * size_t(slot) + (wp.combine == kCross ? j.side : 0)];
  j.g = tb.geoms + tb.slots.geom[slot];
  j.n_cav = j.g->n_cavities;
  j.mv = MeshView{j.g->nodes, j.g->nodes4, j.g->tris_f, j.g->tris_d, j.g->tri_ids, j.g->n_tris};
  // A grid step (radius = grid lower bound >= tau) must never end the path,
  // so tau stays above 2 eps whatever the threshold.
  const float kappa = tb.slots.grid_kappa[slot];
  j.tau = (kappa > 0.f && j.g->grid) ? fmax(R(kappa) * R(j.g->grid_h), R(2) * j.eps) : Inf<R>::v();
  const int32_t prob = tb.slots.problem[slot];
  j.s0 = tb.sides + 2 * prob + (wp.combine == kCross ? j.side : 0);
  j.s1 = tb.sides + 2 * prob + 1;
}

// Gradient sample of one side: (3/r0)(V - s0) omega0 + (r0^2/10) grad f for an
// affine source; zero for a path that ended at its first query.
template <typename R>
__device__ __forceinline__ V3<R> grad_sample(const Job<R> &j, R value, R first, const DevSide &s) {
  if (!(j.gr > R(0))) return {R(0), R(0), R(0)};
  V3<R> g = j.gw * (value - first);
  if (s.source_kind == 1) {
    const R c = j.gr * j.gr / R(10);
    g = g + V3<R>{R(s.src[1]) * c, R(s.src[2]) * c, R(s.src[3]) * c};
  }
  return g;
}

__device__ __forceinline__ void put3(double *o, V3<float> v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
__device__ __forceinline__ void put3(double *o, V3<double> v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

// kGrad: the kernel is a gradient-channel variant, the only kind the
// launchers run with wp.grad set; other variants skip the gradient store at
// compile time, so its zero vector does not hold registers in their loops.
template <typename R, bool kGrad>
__device__ __forceinline__ void job_store(const Job<R> &j, const WalkParams &wp, bool capped) {
  const int32_t st = j.steps | (capped ? (1 << 30) : 0);
  if (wp.combine == kCross) {
    wp.val[2 * j.t + j.side] = double(j.acc0);
    wp.steps[2 * j.t + j.side] = st;
  } else {
    wp.val[2 * j.t + 0] = double(j.acc0);
    wp.val[2 * j.t + 1] = double(wp.combine == kPaired ? j.acc1 : j.acc0);
    wp.steps[2 * j.t + 0] = st;
    wp.steps[2 * j.t + 1] = 0;
  }
  if (kGrad && wp.grad) {
    double *o = wp.grad + 6 * j.t;
    if (wp.combine == kCross) {
      put3(o + 3 * j.side, grad_sample(j, j.acc0, j.gs0, *j.s0));
    } else {
      const V3<R> a = grad_sample(j, j.acc0, j.gs0, *j.s0);
      put3(o, a);
      put3(o + 3, wp.combine == kPaired ? grad_sample(j, j.acc1, j.gs1, *j.s1) : a);
    }
  }
}

// First step of a path taken in direction d with radius r: keep what the
// gradient sample needs (after the step's source term was added).
template <typename R, int kChan>
__device__ __forceinline__ void job_first_step(Job<R> &j, V3<R> d, R r) {
  if constexpr (kChan == kChanGrad) {
    if (j.steps == 0) {
      j.gw = d * (R(3) / r);
      j.gr = r;
      j.gs0 = j.acc0;
      j.gs1 = j.acc1;
    }
  }
}

// One step of a path.  Returns 0 while the path continues, 1 when it ended in
// the eps-shell (payoff added), 2 when it was capped at max_steps.  The cap
// check precedes the query, so a capped path makes max_steps queries and a
// terminated one steps + 1, as walk() does (wih_twos.cpp:644-663).
// The Green-ball source point of a sampled source (source_kind 2); both
// sides of a paired path share it.
template <typename R>
__device__ __forceinline__ bool job_sampled(const Job<R> &j, const WalkParams &wp) {
  return j.s0->source_kind == 2 || (wp.combine == kPaired && j.s1->source_kind == 2);
}

template <typename R>
__device__ __forceinline__ void job_source(Job<R> &j, const WalkParams &wp, R r, V3<R> a) {
  const bool paired_form = wp.rng == kRngMtPaired;
  if (j.s0->source_kind) j.acc0 += side_green_term<R>(*j.s0, j.x, a, r, paired_form);
  if (wp.combine == kPaired && j.s1->source_kind) j.acc1 += side_green_term<R>(*j.s1, j.x, a, r, paired_form);
}

// green(r) returns the Green-ball offset of this step from the path's stream.
template <typename R, int GK, bool kCount, class Green>
__device__ __forceinline__ int job_step(Job<R> &j, const WalkParams &wp, TraversalCount *cnt, R *radius_out,
                                        Green &&green) {
  if (j.steps >= j.max_steps) return 2;
  V3<R> bp;
  int cls;
  const R r = closest_boundary<R, GK, kCount>(*j.g, j.x, bp, cls, j.hint, cnt, j.tau);
  if (radius_out) *radius_out = r;
  // A non-finite distance (a path that left a closed mesh and overflowed)
  // ends the path as capped; the reference treats it as a failed query
  // (wih_twos.cpp:355-357).
  if (!isfinite(r)) return 2;
  const DevSide &s0 = *j.s0;
  if (!(r > j.eps)) {
    j.acc0 += side_payoff<R>(s0, bp, cls);
    if (wp.combine == kPaired) j.acc1 += side_payoff<R>(*j.s1, bp, cls);
    return 1;
  }
  job_source(j, wp, r, job_sampled(j, wp) ? j.x + green(r) : j.x);
  return 0;
}

template <typename R>
__device__ __forceinline__ V3<R> philox_dir(const Job<R> &j) {
  const uint4 w = philox4x32_10(make_uint4(uint32_t(j.steps), uint32_t(j.walk), 0u, 0u), j.key);
  return philox_direction(w, R(0));
}

// ---------------------------------------------------------------------------
// Coupled finite-cavity difference (study.cpp:317-332).  A walk starts at a
// point of the probe ball (sides[1]'s TERM_UNIT row: centre, radius R),
// x = c + omega (R rho) with rho the cube root of a uniform, weight
// (1 - rho^2)^2 35/8; path 0 runs in the geometry, then path 1 restarts from
// the same point on the same stream with the candidate cavity (the site, the
// geometry's cavity radius) appended after the geometry's own cavities.

template <typename R>
__device__ __forceinline__ void coupled_start(Job<R> &j, R rho, V3<R> omega) {
  const double *t = j.s1->terms[0];
  j.z = j.x;
  j.rc = R(j.g->cavity_radius);
  const R q = R(1) - rho * rho;
  j.wgt = q * q * R(35) / R(8);
  j.x = V3<R>{R(t[1]), R(t[2]), R(t[3])} + omega * (R(t[4]) * rho);
  j.x0 = j.x;
  j.a0 = R(0);
  j.st0 = 0;
  j.phase = 0;
  j.hint_start = j.hint;
}

// Philox start point: the radial uniform from words 0-1 of counter block
// (0xFFFFFFFF, walk, 0, 0), the direction from block (0xFFFFFFFF, walk, 1, 0)
// (no step of a path reaches counter 0xFFFFFFFF).
__device__ __forceinline__ double probe_uniform(uint4 b, double) {
  return double(((uint64_t(b.y) << 32) | b.x) >> 11) * 0x1.0p-53;
}
__device__ __forceinline__ float probe_uniform(uint4 b, float) { return float(b.y >> 8) * 0x1.0p-24f; }

template <typename R>
__device__ __forceinline__ void coupled_start_philox(Job<R> &j) {
  const uint4 b = philox4x32_10(make_uint4(0xFFFFFFFFu, uint32_t(j.walk), 0u, 0u), j.key);
  const R rho = cbrt(probe_uniform(b, R(0)));
  const V3<R> omega = philox_direction(philox4x32_10(make_uint4(0xFFFFFFFFu, uint32_t(j.walk), 1u, 0u), j.key), R(0));
  coupled_start(j, rho, omega);
}

// The candidate cavity in path 1, after the geometry's own boundary (it wins
// only when strictly closer, as the last cavity of `added`).
template <typename R>
__device__ __forceinline__ R coupled_cavity(const Job<R> &j, R best, V3<R> &point, int &cls) {
  if (j.phase == 0) return best;
  const V3<R> d = j.x - j.z;
  const R len = wsqrt(dot(d, d));
  const R dist = len - j.rc;
  if (dist < best && len > R(0)) {
    best = dist;
    point = j.z + d * (j.rc / len);
    cls = 7;
  }
  return best;
}

// Path 0 ended: keep its value and step word, restart from the start point.
template <typename R>
__device__ __forceinline__ void coupled_second_path(Job<R> &j, bool capped) {
  j.a0 = j.acc0;
  j.st0 = j.steps | (capped ? (1 << 30) : 0);
  j.x = j.x0;
  j.acc0 = R(0);
  j.steps = 0;
  j.phase = 1;
  j.hint = j.hint_start;
}

// Both paths ended: d = w (a - b) / (4 pi r) in binary64, as study.cpp:327
// evaluates it, in both sample slots; both step words.
template <typename R>
__device__ __forceinline__ void coupled_store(const Job<R> &j, const WalkParams &wp, bool capped) {
  const double d = double(j.wgt) * (double(j.a0) - double(j.acc0)) / (4 * 3.141592653589793 * j.g->cavity_radius);
  wp.val[2 * j.t + 0] = d;
  wp.val[2 * j.t + 1] = d;
  wp.steps[2 * j.t + 0] = j.st0;
  wp.steps[2 * j.t + 1] = j.steps | (capped ? (1 << 30) : 0);
}

// job_step of a coupled pair: the same step with the candidate cavity in
// path 1 (sampled sources of side 0 only).
template <typename R, int GK, class Green>
__device__ __forceinline__ int coupled_step(Job<R> &j, const WalkParams &wp, R *radius_out, Green &&green) {
  if (j.steps >= j.max_steps) return 2;
  V3<R> bp;
  int cls;
  R r = closest_boundary<R, GK, false>(*j.g, j.x, bp, cls, j.hint, nullptr, j.tau);
  r = coupled_cavity(j, r, bp, cls);
  if (radius_out) *radius_out = r;
  if (!isfinite(r)) return 2;
  if (!(r > j.eps)) {
    j.acc0 += side_payoff<R>(*j.s0, bp, cls);
    return 1;
  }
  if (j.s0->source_kind) j.acc0 += side_green_term<R>(*j.s0, j.x, j.s0->source_kind == 2 ? j.x + green(r) : j.x, r, false);
  return 0;
}

// ---------------------------------------------------------------------------
// Simple kernel: one thread per job, the whole path in one go.  Used by the
// plain schedule (one launch per site and round), by the mt19937_64 replay
// streams and by the trace entry point.
template <typename R, int GK, bool kMt, int kChan = kChanNone>
__global__ __launch_bounds__(256) void walk_simple(Tables tb, WalkParams wp) {
  const int64_t u = wp.job_begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= wp.job_end) return;
  Job<R> j;
  job_init(j, tb, wp, u);
  Mt64 mt;
  if constexpr (kChan == kChanCoupled) {
    if constexpr (kMt) {
      // Start point from mt19937_64(splitmix64(key)): the cube root of a
      // uniform, then random_unit_vector (study.cpp:323-324).
      mt.seed(splitmix64(mt_fd_key(tb.slots.seed[j.slot], tb.slots.site[j.slot], j.walk)));
      const double rho = cbrt(mt.uniform());
      const V3<double> om = mt.direction();
      coupled_start(j, R(rho), V3<R>{R(om.x), R(om.y), R(om.z)});
    } else {
      coupled_start_philox(j);
    }
  }
  if constexpr (kMt) {
    const int32_t slot = j.slot;
    mt.seed(mt_walk_seed(wp.rng, tb.slots.seed[slot], tb.slots.repeat[slot], tb.slots.site[slot], j.walk));
  }
  int row = 0;
  while (true) {
    R r = R(0);
    auto green = [&](R rr) -> V3<R> {
      if constexpr (kMt) {
        const V3<double> o = mt.green(double(rr));
        return {R(o.x), R(o.y), R(o.z)};
      } else {
        return philox_green<R>(j.key, j.walk, j.steps, rr);
      }
    };
    const int state = kChan == kChanCoupled ? coupled_step<R, GK>(j, wp, &r, green)
                                            : job_step<R, GK, false>(j, wp, nullptr, &r, green);
    if constexpr (kChan == kChanCoupled) {
      if (state != 0) {
        if (j.phase == 0) {
          coupled_second_path(j, state == 2);
          if constexpr (kMt) {
            const int32_t slot = j.slot;
            mt.seed(mt_walk_seed(wp.rng, tb.slots.seed[slot], tb.slots.repeat[slot], tb.slots.site[slot], j.walk));
          }
          continue;
        }
        coupled_store(j, wp, state == 2);
        break;
      }
    }
    if (wp.trace && row < wp.trace_rows && state != 2) {
      double *o = wp.trace + 4 * row++;
      o[0] = double(j.x.x); o[1] = double(j.x.y); o[2] = double(j.x.z); o[3] = double(r);
    }
    if (state == 1) { job_store<R, kChan == kChanGrad>(j, wp, false); break; }
    if (state == 2) {
      if (wp.trace && row < wp.trace_rows) {
        V3<R> bp;
        int cls;
        const R rr = closest_boundary<R, GK, false>(*j.g, j.x, bp, cls, j.hint, nullptr, j.tau);
        double *o = wp.trace + 4 * row++;
        o[0] = double(j.x.x); o[1] = double(j.x.y); o[2] = double(j.x.z); o[3] = double(rr);
      }
      job_store<R, kChan == kChanGrad>(j, wp, true);
      break;
    }
    V3<R> d;
    if constexpr (kMt) {
      const V3<double> dd = mt.direction();
      d = {R(dd.x), R(dd.y), R(dd.z)};
    } else {
      d = philox_dir(j);
    }
    job_first_step<R, kChan>(j, d, r);
    j.x = j.x + d * r;
    ++j.steps;
  }
}

// ---------------------------------------------------------------------------
// Persistent wavefront kernel with walk regeneration (Philox streams).
//
// Each loop iteration advances every lane by one unit of work: one 4-wide
// node of its current closest-point query, or -- when that query has just
// ended -- the walk step it decides (payoff and retirement inside the
// eps-shell, otherwise the Green-ball source, a Philox direction and the
// move) and the start of the next query.  Lanes therefore never wait for the
// slowest traversal of their wave.  A lane without a path takes the next job
// ticket: each wave keeps a chunk of kChunk consecutive tickets taken from the
// round's counter with one atomic and hands them out in lane order, so
// consecutive lanes start at the same site.  A wave leaves the loop when the
// counter is exhausted and none of its lanes holds a path; every wave gets
// there because every path ends within max_steps steps.
// Tickets per chunk: two waves' worth, so a chunk serves several refills of
// one site with one atomic (measured on C3: 32, 64, 128, 256 and 512 gave
// 20.19, 19.53, 19.36, 19.50 and 20.02 ms per launch).
constexpr int kChunk = 128;
// Launches with fewer jobs than a quarter of the resident lanes (small serving
// rounds) take chunks of kSmallChunk, so the jobs spread over twice as many
// waves (p50 serving latency 4.6 against 5.3 ms with 128).
constexpr int kSmallChunk = 64;
constexpr int kWalkBlock = 256;  // largest block of the persistent kernel
// Idle lanes of a wave take new paths once at least wp.refill of them are idle
// (or none is active): the ticket and path set-up code then runs for several
// lanes at once instead of in nearly every iteration for one or two.  Paths
// are keyed by job, so which lane runs which path changes nothing.
constexpr int kRefillDefault = 8;
constexpr int kLdsDefault = 24;  // default LDS stack entries per lane

// Dynamic LDS bytes of a persistent launch: the lanes' traversal stacks.
constexpr size_t persistent_lds_bytes(int lds, int block) { return size_t(lds) * size_t(block) * 4u; }

// Applies a finished boundary query (distance r, closest point bp of class
// cls) to the path: 0 moved on, 1 ended in the eps-shell, 2 capped.  With
// kDefer a path that ends in the shell keeps its closest point in j.x and the
// class in j.hint (neither is needed any more) and job_retire adds the payoff
// later, so the payoff code runs in the wave's batched refill instead of in
// the walk step.
template <typename R, int kChan = kChanNone, bool kDefer = false, bool kPipe = false>
__device__ __forceinline__ int job_advance(Job<R> &j, const WalkParams &wp, R r, V3<R> bp, int cls) {
  if (!isfinite(r)) return 2;  // as in job_step: a failed query ends the path capped
  const DevSide &s0 = *j.s0;
  if (!(r > j.eps)) {
    if constexpr (kDefer) {
      j.x = bp;
      j.hint = cls;
    } else {
      j.acc0 += side_payoff<R>(s0, bp, cls);
      if (wp.combine == kPaired) j.acc1 += side_payoff<R>(*j.s1, bp, cls);
    }
    return 1;
  }
  if constexpr (kChan != kChanBare)
    job_source(j, wp, r, job_sampled(j, wp) ? j.x + philox_green<R>(j.key, j.walk, j.steps, r) : j.x);
  V3<R> d;
  if constexpr (kPipe) d = j.dn;  // drawn when this step's query started
  else d = philox_dir(j);
  job_first_step<R, kChan>(j, d, r);
  j.x = j.x + d * r;
  ++j.steps;
  return j.steps >= j.max_steps ? 2 : 0;
}

// A path ended with job_advance<..., kDefer = true> state `end` (1 in the
// eps-shell, 2 capped): its payoff, then its samples and step word.
template <typename R, bool kGrad>
__device__ __forceinline__ void job_retire(Job<R> &j, const WalkParams &wp, int end) {
  if (end == 1) {
    j.acc0 += side_payoff<R>(*j.s0, j.x, j.hint);
    if (wp.combine == kPaired) j.acc1 += side_payoff<R>(*j.s1, j.x, j.hint);
  }
  job_store<R, kGrad>(j, wp, end == 2);
}

template <typename R, int GK>
__device__ __forceinline__ bool is_mesh(const Job<R> &j) {
  return GK == kGeomMesh || (GK == kGeomGeneric && j.g->container == kMesh);
}

// Starts the query at the path's position; true while a traversal is pending.
// kPipe: also draws the direction of the path's next step (j.dn), between
// issuing the distance-grid load and using it, so the draw overlaps the load.
template <typename R, int GK, bool kCount, bool kPipe = false>
__device__ __forceinline__ bool query_start(Job<R> &j, MeshQuery<R> &q, TraversalCount *cnt) {
  if (GK == kGeomAnalytic || !is_mesh<R, GK>(j)) {
    if constexpr (kPipe) j.dn = philox_dir(j);
    return false;
  }
  if (!(isfinite(j.x.x) && isfinite(j.x.y) && isfinite(j.x.z))) {
    q.grid = false;
    q.best = Inf<R>::v();  // no query: the path ends capped
    return false;
  }
  R init2 = Inf<R>::v();
  int32_t start = 0;
  if (j.tau < Inf<R>::v()) {
    R ub;
    R lb;
    if constexpr (kPipe) lb = grid_lower<R>(*j.g, j.x, &ub, [&j]() { j.dn = philox_dir(j); });
    else lb = grid_lower<R>(*j.g, j.x, &ub);
    if (kCount) cnt->grid += 1;
    if (lb >= j.tau) {
      if (kCount) cnt->hits += 1;
      q.grid = true;
      q.best = lb;
      return false;
    }
    // The exact query searches below the grid's upper bound and may start
    // below the root.
    init2 = ub < Inf<R>::v() ? ub * ub : ub;
    start = grid_start_node<R>(*j.g, j.x, init2);
  } else if constexpr (kPipe) {
    j.dn = philox_dir(j);
  }
  return !mq_begin<R, kCount>(j.mv, j.x, q, j.hint, cnt, init2, start);
}

// kWaves is the occupancy the register allocator is asked to allow
// (amdgpu_waves_per_eu); 1 leaves the choice to the compiler.  kLds is the
// per-lane LDS stack depth; the launch passes persistent_lds_bytes(kLds,
// blockDim.x) of dynamic LDS.  kBatch > 0 batches the walk steps: a lane whose
// query has ended waits (one node visit of the others at a time) until at
// least kBatch lanes of its wave are ready (kBatch = 1: at least as many as are
// still traversing) or none is traversing, so the wave runs the step code less
// often.  kBatch = 1 batches the node visits the same way: traversing lanes
// wait while ready lanes step, until they outnumber the ready lanes (or none
// is ready), so most iterations run either the step code or the visit code
// with the larger group, not both with a few lanes each.  Which lane takes
// which step, and every path, is unchanged.
template <typename R, int GK, bool kCount, int kWaves, int kLds, int kChan = kChanNone, int kBatch = 0>
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void walk_persistent(
    Tables tb, WalkParams wp) {
  extern __shared__ uint32_t lds_stack[];
  const uint32_t stride = blockDim.x;
  uint32_t *stack = lds_stack + threadIdx.x;
  // binary32 walks without the gradient or coupled channels draw each step's
  // direction in the shadow of the previous query's grid load.
  constexpr bool kPipe = sizeof(R) == 4 && (kChan == kChanNone || kChan == kChanBare);
  uint32_t overflow[kStackDepth - kLds];
  const int lane = int(threadIdx.x & 63);
  const int64_t total = *tb.round.total_jobs;
  const int chunk = 4 * total < int64_t(gridDim.x) * blockDim.x ? kSmallChunk : kChunk;  // wave-uniform
  Job<R> j;
  MeshQuery<R> q;
  // fresh: a new path whose first query has not started; it starts with the
  // next walk steps, so the query-start code runs in one place.
  bool active = false, in_query = false, fresh = false;
  // A path that ended waits, idle, for the wave's next refill to add its
  // payoff and store its samples (coupled pairs store at once).
  constexpr bool kDefer = kChan != kChanCoupled;
  int end = 0;  // 1 or 2: an ended path's samples are not stored yet
  int64_t chunk_cur = 0, chunk_end = 0;  // wave-uniform
  int32_t chunk_site = 0;                // wave-uniform: the site of the chunk's first ticket
  bool drained = false;                  // wave-uniform
  TraversalCount cnt{0, 0, 0, 0, 0};
  while (true) {
    uint64_t need = __ballot(!active);
    if (__popcll(need) < wp.refill && need != ~0ull) need = 0;  // wave-uniform
    if (kDefer && need != 0 && end != 0) {
      job_retire<R, kChan == kChanGrad>(j, wp, end);
      end = 0;
    }
    while (need != 0 && !drained) {
      if (chunk_cur >= chunk_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(tb.round.counter, (unsigned long long)chunk);
        base = __shfl(base, 0);
        if (int64_t(base) >= total) {
          drained = true;
          break;
        }
        chunk_cur = int64_t(base);
        chunk_end = min(int64_t(base) + chunk, total);
        chunk_site = find_site(tb.round.act_prefix, *tb.round.n_active,
                               wp.jobs_per_walk == 2 ? (chunk_cur >> 1) : chunk_cur);
      }
      const int64_t avail = chunk_end - chunk_cur;
      const int take = int(min<int64_t>(avail, __popcll(need)));
      const int rank = __popcll(need & ((1ull << lane) - 1ull));
      if (!active && rank < take) {
        job_init(j, tb, wp, chunk_cur + rank, chunk_site);
        if constexpr (kChan == kChanCoupled) coupled_start_philox(j);
        active = true;
        fresh = true;
        in_query = false;
      }
      chunk_cur += take;
      need = __ballot(!active);
    }
    if (!__any(active)) break;
    bool visit = true;  // wave-uniform
    if constexpr (kBatch == 1) {
      const int busy = __popcll(__ballot(active && in_query));
      const int ready = __popcll(__ballot(active && !in_query));
      visit = ready == 0 || busy > ready;
    }
    if (active) {
      if (in_query && visit) in_query = !mq_visit<R, kCount, kLds>(j.mv, j.x, q, stack, stride, overflow, &cnt);
      bool go = !in_query;
      if constexpr (kBatch > 0) {
        // Wave-uniform: both masks are ballots over the active lanes.
        const uint64_t busy = __ballot(in_query);
        const int ready = __popcll(__ballot(!in_query));
        go = go && (busy == 0 || ready >= (kBatch == 1 ? __popcll(busy) : kBatch));
      }
      if (go) {
        bool start = fresh;
        if (!fresh) {
          V3<R> bp;
          int cls;
          R r;
          if (GK != kGeomAnalytic && is_mesh<R, GK>(j)) {
            cls = 0;
            if (q.grid) {
              r = q.best;
              bp = j.x;
            } else {
              r = mesh_distance<R>(*j.g, q.best);
              bp = q.pt;
              j.hint = q.best_k;
            }
          } else {
            r = container_analytic<R>(*j.g, j.x, bp, cls);
          }
          r = holes_and_cavities<R, GK>(*j.g, j.x, r, bp, cls, j.n_cav);
          if constexpr (kChan == kChanCoupled) r = coupled_cavity(j, r, bp, cls);
          const int st = job_advance<R, kChan, kDefer, kPipe>(j, wp, r, bp, cls);
          if (st == 0) {
            start = true;
          } else if (kChan == kChanCoupled && j.phase == 0) {
            coupled_second_path(j, st == 2);
            start = true;
          } else {
            if constexpr (kChan == kChanCoupled) coupled_store(j, wp, st == 2);
            else end = st;
            active = false;
          }
        }
        fresh = false;
        if (start) in_query = query_start<R, GK, kCount, kPipe>(j, q, &cnt);
      }
    }
  }
  if (kCount) {
    atomicAdd(wp.counters + 0, (unsigned long long)cnt.nodes);
    atomicAdd(wp.counters + 1, (unsigned long long)cnt.tris);
    atomicAdd(wp.counters + 2, (unsigned long long)cnt.grid);
    atomicAdd(wp.counters + 3, (unsigned long long)cnt.hits);
    atomicAdd(wp.counters + 4, (unsigned long long)cnt.top);
  }
}

// Closest triangle at each new site: every path of a site starts with the
// same query, so it is answered once at admission and seeds each path's hint.
// A site where the distance grid's bound already clears the step threshold
// (the walk's first step takes the grid radius, no query) gets no hint: deep
// inside a mesh that query would test nearly every triangle, and a path's
// first exact query, far from the site, starts from the grid's own bound.
// The hint only warm-starts a query: the distance found does not depend on it.
template <typename R>
__global__ void slot_hints(Tables tb, int64_t first, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t slot = first + i;
  const DevGeom &g = tb.geoms[tb.slots.geom[slot]];
  int32_t hint = -1;
  if (g.container == kMesh) {
    const double *p = tb.slots.pos + 3 * slot;
    const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
    const V3<R> x{R(p[0]), R(p[1]), R(p[2])};
    // With a distance grid the query starts below the grid's upper bound, at
    // the cell's start node, as the walk's own queries do.
    R init2 = Inf<R>::v();
    int32_t start = 0;
    const float kappa = tb.slots.grid_kappa[slot];
    if (g.grid && kappa > 0.f) {
      R ub;
      const R lb = grid_lower<R>(g, x, &ub);
      const R tau = fmax(R(kappa) * R(g.grid_h), R(2) * R(tb.slots.eps[slot]));
      if (lb >= tau) {
        tb.slots.hint0[slot] = -1;
        return;
      }
      init2 = ub < Inf<R>::v() ? ub * ub : ub;
      start = grid_start_node<R>(g, x, init2);
    }
    V3<R> pt;
    mesh_closest<R, false>(m, x, pt, hint, nullptr, init2, start);
  }
  tb.slots.hint0[slot] = hint;
}

// ---------------------------------------------------------------------------
// Per-site reduction and combine (binary64).  One block per active site folds
// the round's samples into eight power sums {Sa, Sb, Sab, Saa, Sbb, Saab,
// Sabb, Saabb} -- each thread a fixed stride of walk indices, then a wave
// shuffle and an LDS tree -- adds them to the site's resident accumulators
// and recomputes the estimate:
//   score = (Sa Sb - Sab) / (N (N-1))                (wih_twos.cpp:693;
//                                                      study.cpp:117)
//   jackknife SE over leave-one-out estimates        (wih_twos.cpp:698-712)
// from the power sums: the leave-one-out estimate (C - Sa b_i - Sb a_i +
// 2 a_i b_i) / ((N-1)(N-2)) is linear in (a_i, b_i, a_i b_i).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ void site_estimate(const double *s, int64_t count, twos_site_result &r) {
  const double n = double(count);
  r.u = s[0] / n;
  r.v = s[1] / n;
  r.score = (s[0] * s[1] - s[2]) / (n * (n - 1.0));
  r.td = -r.score;
  r.same = s[2] / n;
  if (count < 4) {
    r.score_se = __builtin_nan("");
    return;
  }
  const double sa = s[0], sb = s[1];
  const double z2 = sa * sa * s[4] + sb * sb * s[3] + 4.0 * s[7] + 2.0 * sa * sb * s[2] - 4.0 * sa * s[6] -
                    4.0 * sb * s[5];
  const double z1 = 2.0 * (s[2] - sa * sb);
  const double d = (n - 1.0) * (n - 2.0);
  const double spread = fmax(0.0, z2 - z1 * z1 / n) / (d * d);
  r.score_se = sqrt((n - 1.0) / n * spread);
}

// Coupled finite-cavity estimate from the same sums (both sample slots hold
// d): the mean of d and sqrt(max(0, (sum d^2 - N mean^2) / (N (N-1)))), as
// study.cpp:329-331 computes v.up and v.same.
__device__ __forceinline__ void coupled_estimate(const double *s, int64_t count, twos_site_result &r) {
  const double n = double(count);
  const double up = s[0] / n;
  r.u = r.v = r.score = up;
  r.td = -up;
  r.same = s[2] / n;
  r.score_se = count >= 2 ? sqrt(fmax(0.0, (s[2] - n * up * up) / (n * (n - 1.0)))) : __builtin_nan("");
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void reduce_sites(Tables tb, WalkParams wp) {
  const int32_t k = blockIdx.x;
  if (k >= *tb.round.n_active) return;
  const int32_t slot = tb.round.act_slot[k];
  const int64_t t0 = tb.round.act_prefix[k], t1 = tb.round.act_prefix[k + 1];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long long steps = 0, caps = 0;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const double a = wp.val[2 * t], b = wp.val[2 * t + 1];
    const double ab = a * b;
    s[0] += a; s[1] += b; s[2] += ab; s[3] += a * a; s[4] += b * b;
    s[5] += ab * a; s[6] += ab * b; s[7] += ab * ab;
    const int32_t st0 = wp.steps[2 * t], st1 = wp.steps[2 * t + 1];
    steps += (st0 & 0x3fffffff) + (st1 & 0x3fffffff);
    caps += ((st0 >> 30) & 1) + ((st1 >> 30) & 1);
  }
  __shared__ double red[kBlock / 64][8];
  __shared__ long long redi[kBlock / 64][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 8; ++i) s[i] = wave_sum(s[i]);
  steps = wave_sum_ll(steps);
  caps = wave_sum_ll(caps);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) red[wave][i] = s[i];
    redi[wave][0] = steps;
    redi[wave][1] = caps;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[8];
    long long tsteps = 0, tcaps = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) tot[i] = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
#pragma unroll
      for (int i = 0; i < 8; ++i) tot[i] += red[w][i];
      tsteps += redi[w][0];
      tcaps += redi[w][1];
    }
    double *acc = tb.slots.acc + 8 * size_t(slot);
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] += tot[i];
    const int64_t walks = tb.slots.acc_walks[slot] + (t1 - t0);
    tb.slots.acc_walks[slot] = walks;
    tb.slots.acc_steps[slot] += tsteps;
    tb.slots.acc_caps[slot] += tcaps;
    twos_site_result r;
    if (tb.slots.combine[slot] == kCoupled) coupled_estimate(acc, walks, r);
    else site_estimate(acc, walks, r);
    for (int q = 0; q < 3; ++q) r.grad_u[q] = r.grad_v[q] = __builtin_nan("");
    r.grad_dot = r.grad_dot_se = __builtin_nan("");  // filled by reduce_grad when requested
    r.walks = walks;
    r.steps = tb.slots.acc_steps[slot];
    r.caps = tb.slots.acc_caps[slot];
    const double rel = tb.slots.ci_rel[slot];
    int32_t decision;
    if (!(rel >= 0.0)) {
      decision = 0;  // fixed budget
    } else {
      const bool met = isfinite(r.score_se) && tb.slots.ci_z[slot] * r.score_se <= rel * fabs(r.score);
      decision = met ? 0 : (walks >= tb.slots.walks_max[slot] ? 2 : 1);
    }
    r.status = decision == 2 ? 1 : 0;
    r.pad = 0;
    tb.slots.result[slot] = r;
    tb.round.compact[k] = r;
    tb.round.decision[k] = decision;
  }
}

// Gradient channel of one site from its 28 sums: A = sum a (3), B = sum b (3),
// C = sum a.b, the second moments sum a a^T (6: xx yy zz xy xz yz), sum b b^T
// (6) and sum a b^T (9, row-major).  grad_dot = (A.B - C)/(N(N-1)) counts
// every pair of distinct walks; its SE is the first-order (delta-method) one,
// sqrt((b^T S_aa b + a^T S_bb a + 2 b^T S_ab a) / N) with a, b the means and S
// the sample covariances.
__device__ __forceinline__ void grad_estimate(const double *g, int64_t n, twos_site_result &r) {
  const double N = double(n);
  double a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = g[k] / N;
    b[k] = g[3 + k] / N;
    r.grad_u[k] = a[k];
    r.grad_v[k] = b[k];
  }
  const double ab = g[0] * g[3] + g[1] * g[4] + g[2] * g[5];
  r.grad_dot = n > 1 ? (ab - g[6]) / (N * (N - 1.0)) : __builtin_nan("");
  if (n < 4) {
    r.grad_dot_se = __builtin_nan("");
    return;
  }
  const int ix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  double saa[3][3], sbb[3][3], sab[3][3];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      saa[p][q] = (g[7 + ix[p][q]] - N * a[p] * a[q]) / (N - 1.0);
      sbb[p][q] = (g[13 + ix[p][q]] - N * b[p] * b[q]) / (N - 1.0);
      sab[p][q] = (g[19 + 3 * p + q] - N * a[p] * b[q]) / (N - 1.0);
    }
  double v = 0.0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) v += b[p] * saa[p][q] * b[q] + a[p] * sbb[p][q] * a[q] + 2.0 * b[p] * sab[p][q] * a[q];
  r.grad_dot_se = sqrt(fmax(v, 0.0) / N);
}

// Per-site fold of the gradient samples, after reduce_sites (same work list):
// adds the round's 28 sums to the slot's and completes the result's gradient
// fields (in walk order within each lane, then a fixed tree, like the value
// channel).
template <int kBlock>
__global__ __launch_bounds__(kBlock) void reduce_grad(Tables tb, WalkParams wp) {
  const int32_t k = blockIdx.x;
  if (k >= *tb.round.n_active) return;
  const int32_t slot = tb.round.act_slot[k];
  const int64_t t0 = tb.round.act_prefix[k], t1 = tb.round.act_prefix[k + 1];
  double s[28];
#pragma unroll
  for (int i = 0; i < 28; ++i) s[i] = 0.0;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const double *o = wp.grad + 6 * t;
    const double a[3] = {o[0], o[1], o[2]}, b[3] = {o[3], o[4], o[5]};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      s[q] += a[q];
      s[3 + q] += b[q];
    }
    s[6] += a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    s[7] += a[0] * a[0]; s[8] += a[1] * a[1]; s[9] += a[2] * a[2];
    s[10] += a[0] * a[1]; s[11] += a[0] * a[2]; s[12] += a[1] * a[2];
    s[13] += b[0] * b[0]; s[14] += b[1] * b[1]; s[15] += b[2] * b[2];
    s[16] += b[0] * b[1]; s[17] += b[0] * b[2]; s[18] += b[1] * b[2];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) s[19 + 3 * p + q] += a[p] * b[q];
  }
  __shared__ double red[kBlock / 64][28];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 28; ++i) s[i] = wave_sum(s[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 28; ++i) red[wave][i] = s[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double *acc = tb.slots.acc_grad + 28 * size_t(slot);
    for (int i = 0; i < 28; ++i) {
      double tot = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) tot += red[w][i];
      acc[i] += tot;
    }
    // reduce_sites has already counted this round's walks into acc_walks.
    twos_site_result r = tb.slots.result[slot];
    grad_estimate(acc, tb.slots.acc_walks[slot], r);
    tb.slots.result[slot] = r;
    tb.round.compact[k] = r;
  }
}

}  // namespace twos
