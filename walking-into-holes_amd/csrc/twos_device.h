// Device-side building blocks of the TWoS hot path for gfx950 (CDNA4, wave64).
//
// Everything a walk touches per step lives here: counter-based Philox streams,
// the reference's mt19937_64 replay stream, direction sampling, analytic SDF
// unions, the LBVH closest-point traversal and Ericson's closest point on a
// triangle.  Templates are instantiated twice: binary32 walks (contraction on,
// twos_walk_f32.hip) and binary64 walks (contraction off, twos_walk_f64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "twos_types.h"

namespace twos {

template <typename R>
struct V3 {
  R x, y, z;
};
template <typename R>
__device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R>
__device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R>
__device__ __forceinline__ V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R>
__device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Pointers to device buffers read from a table (the geometry rows) are
// generic to the compiler, which then emits flat loads: those also count on
// the LDS/scalar counter, so each one waits behind the traversal stack's LDS
// traffic.  Every geometry buffer is global memory (hipMalloc or a
// __device__ array), whose generic and global addresses are equal, so the
// pointer is reinterpreted as a global one at the load.
template <typename T>
using GPtr = const __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ GPtr<T> gptr(const T *p) {
  return __builtin_bit_cast(GPtr<T>, p);
}
// 16-byte vector types through a global pointer (their class copy cannot
// bind a global-address-space reference, so the load goes through the
// builtin vector type).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload(const float4 *p) {
  const f32x4 v = *gptr(reinterpret_cast<const f32x4 *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 gload(const int4 *p) {
  const i32x4 v = *gptr(reinterpret_cast<const i32x4 *>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gload(const uint4 *p) {
  const i32x4 v = *gptr(reinterpret_cast<const i32x4 *>(p));
  return make_uint4(uint32_t(v.x), uint32_t(v.y), uint32_t(v.z), uint32_t(v.w));
}

template <typename R> struct Inf;
template <> struct Inf<float> { static __device__ __forceinline__ float v() { return __builtin_huge_valf(); } };
template <> struct Inf<double> { static __device__ __forceinline__ double v() { return __builtin_huge_val(); } };

// ---------------------------------------------------------------------------
// Streams

// splitmix64 finalizer (benchmarks/compiled_bunny/src/wih_twos.cpp:529-534).
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30U)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27U)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31U);
}

// Philox4x32-10 (Salmon et al., SC'11): two 32x32 multiplies per round, Weyl
// key schedule, ten rounds.  Counter = (step, walk, 0, 0).  Each product is
// taken as one 64-bit result (one v_mad_u64_u32 instead of a v_mul_lo_u32 and
// a v_mul_hi_u32: the same bits in half the multiply instructions).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c.x, p1 = uint64_t(0xCD9E8D57u) * c.z;
    const uint32_t lo0 = uint32_t(p0), hi0 = uint32_t(p0 >> 32);
    const uint32_t lo1 = uint32_t(p1), hi1 = uint32_t(p1 >> 32);
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// Stream identity of (seed, site, side), computed on the host per site slot.
inline uint2 philox_key_host(uint64_t seed, int64_t site, int32_t side) {
  const uint64_t k = splitmix64(splitmix64(seed) + ((uint64_t(site) << 1) | uint64_t(side & 1)));
  return make_uint2(uint32_t(k), uint32_t(k >> 32));
}

constexpr double kTwoPi = 6.283185307179586;  // 2 * acos(-1.0) in binary64

// Unit direction from one Philox block.  binary64: 53-bit uniforms from word
// pairs (0,1) and (2,3).  binary32: the top 24 bits of the same two numbers,
// so a binary32 walk and its binary64 rerun follow the same directions to
// within 2^-24.  z = 1 - 2 u1, phi = 2 pi u2 (wih_twos.cpp:536-542).
// Square root of the walk arithmetic: binary64 exact; binary32 the hardware
// v_sqrt_f32 (within 1 ulp) instead of the correctly rounded expansion, which
// the binary32 walk tolerates (a radius off by an ulp; the distance grid's
// 2e-6 relative margin covers its bound).
__device__ __forceinline__ double wsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float wsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ V3<double> philox_direction(uint4 w, double) {
  const double u1 = double(((uint64_t(w.y) << 32) | w.x) >> 11) * 0x1.0p-53;
  const double u2 = double(((uint64_t(w.w) << 32) | w.z) >> 11) * 0x1.0p-53;
  const double z = 1.0 - 2.0 * u1;
  const double phi = kTwoPi * u2;
  const double radial = sqrt(fmax(0.0, 1.0 - z * z));
  double s, c;
  sincos(phi, &s, &c);
  return {radial * c, radial * s, z};
}

// binary32: the hardware sine and cosine of 2 pi u2 (v_sin_f32 / v_cos_f32
// take revolutions, so no range reduction; within a few 1e-7 of the exact
// values, as the rest of the binary32 step).
__device__ __forceinline__ V3<float> philox_direction(uint4 w, float) {
  const float u1 = float(w.y >> 8) * 0x1.0p-24f;
  const float u2 = float(w.w >> 8) * 0x1.0p-24f;
  const float z = 1.0f - 2.0f * u1;
  const float radial = wsqrt(fmaxf(0.0f, (1.0f - z) * (1.0f + z)));
  return {radial * __builtin_amdgcn_cosf(u2), radial * __builtin_amdgcn_sinf(u2), z};
}

// mt19937_64 (Matsumoto-Nishimura; the C++ standard's parameters w=64, n=312,
// m=156, r=31, a=0xB5026F5AA96619E9, tempering u=29 d=0x5555555555555555
// s=17 b=0x71D67FFFEDA60000 t=37 c=0xFFF7EEE000000000 l=43, f=6364136223846793005).
// Used only to replay the reference's own streams in binary64.
struct Mt64 {
  uint64_t mt[312];
  int idx;
  __device__ void seed(uint64_t s) {
    mt[0] = s;
    for (int i = 1; i < 312; ++i) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + uint64_t(i);
    idx = 312;
  }
  __device__ void twist() {
    for (int i = 0; i < 312; ++i) {
      const uint64_t x = (mt[i] & 0xFFFFFFFF80000000ULL) | (mt[i + 1 < 312 ? i + 1 : 0] & 0x7FFFFFFFULL);
      uint64_t xa = x >> 1;
      if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
      const int j = i + 156 < 312 ? i + 156 : i + 156 - 312;
      mt[i] = mt[j] ^ xa;
    }
    idx = 0;
  }
  __device__ uint64_t next() {
    if (idx >= 312) twist();
    uint64_t x = mt[idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
  }
  // uniform_real_distribution<double>(0,1) over a 64-bit engine (libstdc++
  // generate_canonical with one draw): x / 2^64, clamped below 1.
  __device__ double uniform() {
    double u = double(next()) * 0x1.0p-64;
    if (u >= 1.0) u = 0x1.fffffffffffffp-1;
    return u;
  }
  // beta_22: Beta(2,2) by rejection (wih_twos.cpp:569-575).
  __device__ double beta22() {
    while (true) {
      const double value = uniform();
      if (uniform() <= 4.0 * value * (1.0 - value)) return value;
    }
  }
  // Green-ball offset random_unit_vector(rng) * (r * beta_22(rng)); g++
  // evaluates the Beta draw (the right operand) first.
  __device__ V3<double> green(double r) {
    const double t = beta22();
    const V3<double> d = direction();
    return d * (r * t);
  }
  // random_unit_vector (wih_twos.cpp:536-542).
  __device__ V3<double> direction() {
    const double z = 1.0 - 2.0 * uniform();
    const double phi = kTwoPi * uniform();
    const double radial = sqrt(fmax(0.0, 1.0 - z * z));
    return {radial * cos(phi), radial * sin(phi), z};
  }
};

// Per-walk key of the coupled finite-cavity difference (study.cpp:322); its
// start point comes from mt19937_64(splitmix64(key)) (:323-324).
__device__ __forceinline__ uint64_t mt_fd_key(uint64_t seed, int64_t site, int64_t walk) {
  return splitmix64(seed) ^ splitmix64(uint64_t(site) + 734) ^ splitmix64(uint64_t(walk) + 8131);
}

// Per-walk mt19937_64 seed of the reference's key layouts.
__device__ __forceinline__ uint64_t mt_walk_seed(int rng, uint64_t seed, int32_t repeat, int64_t site,
                                                  int64_t walk) {
  if (rng == kRngMtSite) {
    // evaluate_site key (wih_twos.cpp:681-685), seeded through walk() (:638).
    const uint64_t key = splitmix64(seed + 0x2d358dccaa6c78a5ULL) ^
                         splitmix64(uint64_t(repeat) + 0x8bb84b93962eacc9ULL) ^
                         splitmix64(uint64_t(site) + 0x4b33a62ed433d4a3ULL) ^
                         splitmix64(uint64_t(walk) + 0x4d5a2da51de1aa47ULL);
    return splitmix64(key);
  }
  if (rng == kRngMtPaired) {
    // paired_site (benchmarks/submission_study/study.cpp:89).
    return splitmix64(seed) ^ splitmix64(uint64_t(site) + 91347) ^ splitmix64(uint64_t(walk) + 571);
  }
  if (rng == kRngMtFd) {
    // The coupled finite-cavity paths: walk(..., key + 719) (study.cpp:326).
    return splitmix64(mt_fd_key(seed, site, walk) + 719);
  }
  // Explicit key (write_trace: seed + 900000 + site, wih_twos.cpp:1737).
  return splitmix64(seed + uint64_t(site));
}

// a / b.  binary32 walks take the hardware reciprocal (1 ulp) instead of the
// correctly rounded IEEE sequence; binary64 divides exactly.
template <typename R>
__device__ __forceinline__ R ratio(R a, R b) { return a / b; }
template <>
__device__ __forceinline__ float ratio<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// ---------------------------------------------------------------------------
// Closest point on a triangle: Ericson, Real-Time Collision Detection 5.1.5
// (the reference's closest_point_triangle, wih_twos.cpp:247-289).
template <typename R>
__device__ __forceinline__ V3<R> closest_point_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c) {
  const V3<R> ab = b - a;
  const V3<R> ac = c - a;
  const V3<R> ap = p - a;
  const R d1 = dot(ab, ap);
  const R d2 = dot(ac, ap);
  if (d1 <= R(0) && d2 <= R(0)) return a;
  const V3<R> bp = p - b;
  const R d3 = dot(ab, bp);
  const R d4 = dot(ac, bp);
  if (d3 >= R(0) && d4 <= d3) return b;
  const R vc = d1 * d4 - d3 * d2;
  if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) {
    const R v = ratio(d1, d1 - d3);
    return a + ab * v;
  }
  const V3<R> cp = p - c;
  const R d5 = dot(ab, cp);
  const R d6 = dot(ac, cp);
  if (d6 >= R(0) && d5 <= d6) return c;
  const R vb = d5 * d2 - d1 * d6;
  if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) {
    const R w = ratio(d2, d2 - d6);
    return a + ac * w;
  }
  const R va = d3 * d6 - d5 * d4;
  if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) {
    const V3<R> bc = c - b;
    const R w = ratio(d4 - d3, (d4 - d3) + (d5 - d6));
    return b + bc * w;
  }
  const R denom = ratio(R(1), va + vb + vc);
  const R v = vb * denom;
  const R w = vc * denom;
  return a + ab * v + ac * w;
}

// binary32: the same Voronoi-region decisions in the same order, without
// branches: every lane computes the six dot products once, picks its region
// by selects, takes one reciprocal for the region's ratio and forms
// a + ab v + ac w.  (Ericson's early returns cost a wave every region any of
// its lanes lands in.)  The point can differ from the branchy form by an ulp
// of rounding (b as a + ab, for instance); binary64 keeps the branchy form.
template <>
__device__ __forceinline__ V3<float> closest_point_triangle<float>(V3<float> p, V3<float> a, V3<float> b,
                                                                    V3<float> c) {
  const V3<float> ab = b - a, ac = c - a, ap = p - a, bp = p - b, cp = p - c;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cp),
              d6 = dot(ac, cp);
  const float vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
  const float e43 = d4 - d3, e56 = d5 - d6;
  const bool rA = d1 <= 0.f && d2 <= 0.f;
  const bool rB = !rA && d3 >= 0.f && d4 <= d3;
  const bool rAB = !rA && !rB && vc <= 0.f && d1 >= 0.f && d3 <= 0.f;
  const bool done3 = rA || rB || rAB;
  const bool rC = !done3 && d6 >= 0.f && d5 <= d6;
  const bool rAC = !done3 && !rC && vb <= 0.f && d2 >= 0.f && d6 <= 0.f;
  const bool rBC = !done3 && !rC && !rAC && va <= 0.f && e43 >= 0.f && e56 >= 0.f;
  const bool inner = !done3 && !rC && !rAC && !rBC;
  // The region's ratio num / den (vertex regions take none).
  const float num = rAB ? d1 : (rAC ? d2 : (rBC ? e43 : 1.f));
  const float den = rAB ? d1 - d3 : (rAC ? d2 - d6 : (rBC ? e43 + e56 : (inner ? va + vb + vc : 1.f)));
  const float t = num * __builtin_amdgcn_rcpf(den);
  float v = 0.f, w = 0.f;
  v = rB ? 1.f : v;
  v = rAB ? t : v;
  w = rC ? 1.f : w;
  w = rAC ? t : w;
  v = rBC ? 1.f - t : v;
  w = rBC ? t : w;
  v = inner ? vb * t : v;
  w = inner ? vc * t : w;
  return a + ab * v + ac * w;
}

// ---------------------------------------------------------------------------
// LBVH closest-point traversal.
//
// Node i (internal, 0 <= i < n-1) is 64 bytes, four float4:
//   q0 = (L.lo.x, L.lo.y, L.lo.z, L.hi.x)
//   q1 = (L.hi.y, L.hi.z, R.lo.x, R.lo.y)
//   q2 = (R.lo.z, R.hi.x, R.hi.y, R.hi.z)
//   q3 = (child L, child R, parent, -)   as int bits; a child >= 0 is an
//        internal node, a child < 0 is leaf ~k (sorted triangle k).
// One 64-byte fetch yields both child boxes, so a visit costs one line.
// Triangles are stored in sorted order: binary32 as 9 floats, binary64 as 9
// doubles, plus the original triangle index (tie-break for exact queries).

struct MeshView {
  const float4 *nodes;   // binary LBVH (ray queries)
  const float4 *nodes4;  // 4-wide BVH (closest-point queries)
  const float *tf;
  const double *td;
  const int32_t *tid;
  int32_t n_tris;
};

template <typename R>
__device__ __forceinline__ R box_d2(V3<R> p, R lx, R ly, R lz, R hx, R hy, R hz) {
  const R dx = fmax(fmax(lx - p.x, p.x - hx), R(0));
  const R dy = fmax(fmax(ly - p.y, p.y - hy), R(0));
  const R dz = fmax(fmax(lz - p.z, p.z - hz), R(0));
  return dx * dx + dy * dy + dz * dz;
}

template <typename R>
__device__ __forceinline__ void load_tri(const MeshView &m, int32_t k, V3<R> &a, V3<R> &b, V3<R> &c);

template <>
__device__ __forceinline__ void load_tri<float>(const MeshView &m, int32_t k, V3<float> &a, V3<float> &b,
                                                V3<float> &c) {
  const GPtr<float> t = gptr(m.tf + 9 * size_t(k));
  a = {t[0], t[1], t[2]};
  b = {t[3], t[4], t[5]};
  c = {t[6], t[7], t[8]};
}

template <>
__device__ __forceinline__ void load_tri<double>(const MeshView &m, int32_t k, V3<double> &a, V3<double> &b,
                                                 V3<double> &c) {
  const GPtr<double> t = gptr(m.td + 9 * size_t(k));
  a = {t[0], t[1], t[2]};
  b = {t[3], t[4], t[5]};
  c = {t[6], t[7], t[8]};
}

struct TraversalCount {
  uint32_t nodes;
  uint32_t tris;
  uint32_t grid;  // distance-grid lookups
  uint32_t hits;  // lookups whose bound was taken as the radius (no traversal)
  uint32_t top;   // visits to nodes in the top kTopLevels levels of the 4-wide tree
};

// Counting launches report visits to the top three 4-wide levels (1 + 4 + 16
// nodes): the nodes a per-block copy in LDS could serve.
constexpr int kTopLevels = 3;
constexpr int32_t kCellNodeDepth = 1 << 20;  // depth recorded for per-cell nodes

// 4-wide node, 128 bytes, eight float4: the four children's bounds as
// lo.x, lo.y, lo.z, hi.x, hi.y, hi.z rows (one child per lane of the float4),
// then the four child links (>= 0 node, ~k leaf = sorted triangle k,
// kEmptyChild unused slot whose box is empty), then padding.
constexpr int32_t kEmptyChild = int32_t(0x80000000u);
constexpr int kNode4Quads = 8;

// Traversal stack depth: a 4-wide node pushes at most three siblings, and
// the builder rejects trees whose 4-wide depth would overflow it.
constexpr int kStackDepth = 48;

// Start node of an exact query at p whose best squared distance so far is
// ub2 (after the warm-start triangle): the cell's start node when the record
// grid has them, p lies in the grid and ub2 is within their validity, else
// the root.
template <typename R>
__device__ __forceinline__ int32_t grid_start_node(const DevGeom &g, V3<R> p, R ub2) {
  if (!g.grid_start || !(ub2 <= R(g.grid_start_r2))) return 0;
  const R inv_h = R(g.grid_inv_h);
  const int i = int(floor((p.x - R(g.grid_lo[0])) * inv_h));
  const int k = int(floor((p.y - R(g.grid_lo[1])) * inv_h));
  const int l = int(floor((p.z - R(g.grid_lo[2])) * inv_h));
  if (i < 0 || k < 0 || l < 0 || i >= g.grid_n[0] || k >= g.grid_n[1] || l >= g.grid_n[2]) return 0;
  return gptr(g.grid_start)[(size_t(i) * g.grid_n[1] + k) * g.grid_n[2] + l];
}

// Exact closest point on the mesh over the 4-wide BVH.  The query first tests
// `hint` (the previous step's closest triangle: a walk moves by less than
// twice its distance, so this bound is tight), then visits children
// nearest-first, testing leaves as soon as their box is within the bound and
// re-checking each stacked subtree's box distance when it is popped.
// binary64: binary64 distances against outward-rounded binary32 boxes, ties
// broken by the original triangle index, so the result equals a brute-force
// scan.  binary32: the same traversal in binary32.
template <typename R, bool kCount>
__device__ __forceinline__ R mesh_closest(const MeshView &m, V3<R> p, V3<R> &best_pt, int32_t &hint,
                                          TraversalCount *cnt, R init2 = Inf<R>::v(), int32_t start = 0) {
  constexpr bool kExact = sizeof(R) == 8;
  R best = init2;
  int32_t best_id = 0x7fffffff;
  int32_t best_k = -1;
  auto test = [&](int32_t k) {
    V3<R> a, b, c;
    load_tri<R>(m, k, a, b, c);
    if (kCount) cnt->tris += 1;
    const V3<R> q = closest_point_triangle(p, a, b, c);
    const V3<R> d = p - q;
    const R d2 = dot(d, d);
    if constexpr (kExact) {
      const int32_t id = gptr(m.tid)[k];
      if (d2 < best || (d2 == best && id < best_id)) {
        best = d2;
        best_pt = q;
        best_id = id;
        best_k = k;
      }
    } else if (d2 < best) {
      best = d2;
      best_pt = q;
      best_k = k;
    }
  };
  // A non-finite point would pass every box test (fmax drops NaN): no query.
  if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) return best;
  if (m.n_tris == 1) {
    best = Inf<R>::v();
    test(0);
    hint = best_k;
    return best;
  }
  if (hint >= 0 && hint < m.n_tris) test(hint);
  // Relative slack keeps binary64 rounding in a box bound from pruning a
  // triangle that ties the current best.
  const R slack = kExact ? R(1.0 - 1e-12) : R(1);
  auto within = [&](R d) { return kExact ? d * slack <= best : d < best; };
  int32_t stk_node[kStackDepth];
  R stk_d[kStackDepth];
  int sp = 0;
  int32_t node = start;
  int32_t visits = m.n_tris + 8;  // 4-wide nodes < triangles; guards a corrupt tree
  while (--visits >= 0) {
    const float4 *q = m.nodes4 + size_t(kNode4Quads) * node;
    const float4 lx = gload(q), ly = gload(q + 1), lz = gload(q + 2), hx = gload(q + 3), hy = gload(q + 4),
                 hz = gload(q + 5);
    const int4 ch = gload(reinterpret_cast<const int4 *>(q + 6));
    if (kCount) cnt->nodes += 1;
    R d[4] = {box_d2<R>(p, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x), box_d2<R>(p, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y),
              box_d2<R>(p, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z), box_d2<R>(p, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w)};
    int32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
    // Leaves within the bound are tested at once (they tighten it).
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (c[i] < 0 && c[i] != kEmptyChild && within(d[i])) test(~c[i]);
    // Internal children within the (tightened) bound, nearest first.
    R key[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) key[i] = (c[i] >= 0 && within(d[i])) ? d[i] : Inf<R>::v();
    auto cswap = [&](int i, int j) {
      if (key[j] < key[i]) {
        const R tk = key[i]; key[i] = key[j]; key[j] = tk;
        const int32_t tc = c[i]; c[i] = c[j]; c[j] = tc;
      }
    };
    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
    const R inf = Inf<R>::v();
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
      if (within(stk_d[sp])) {
        node = stk_node[sp];
        found = true;
        break;
      }
    }
    if (!found) {
      if (best_k < 0 && best < Inf<R>::v()) {
        // A search started from a bound found nothing below it (never
        // expected: the bound is above the true distance); search the whole tree.
        best = Inf<R>::v();
        node = 0;
        continue;
      }
      break;
    }
  }
  hint = best_k;
  return best;
}

// ---------------------------------------------------------------------------
// Domain: container (ball, box or mesh) minus sphere / cylinder holes minus
// reference-style cavities.  Returns the distance to the boundary, the
// closest boundary point and its class (0-5 container faces, 6 holes,
// 7 cavities).  The mesh is first and later candidates win only when strictly
// closer, as domain_distance / closest_boundary order them
// (wih_twos.cpp:585-593, :602-626).

template <typename R>
struct Vec4 { R x, y, z, w; };

template <typename R>
__device__ __forceinline__ Vec4<R> load4(const DevGeom &g, int kind, int i);

template <>
__device__ __forceinline__ Vec4<float> load4<float>(const DevGeom &g, int kind, int i) {
  const float4 v = gload((kind == 0 ? g.spheres_f : (kind == 1 ? g.cylinders_f : g.cavities_f)) + i);
  return {v.x, v.y, v.z, v.w};
}

template <>
__device__ __forceinline__ Vec4<double> load4<double>(const DevGeom &g, int kind, int i) {
  const GPtr<double> b = gptr(kind == 0 ? g.spheres_d : (kind == 1 ? g.cylinders_d : g.cavities_d));
  return {b[4 * i + 0], b[4 * i + 1], b[4 * i + 2], b[4 * i + 3]};
}

template <typename R>
__device__ __forceinline__ R comp(V3<R> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Distance to an analytic container (ball or box interior), its closest
// boundary point and class.
template <typename R>
__device__ __forceinline__ R container_analytic(const DevGeom &g, V3<R> p, V3<R> &point, int &cls) {
  R best;
  if (g.container == kBall) {
    const V3<R> c{R(g.ball[0]), R(g.ball[1]), R(g.ball[2])};
    const R rad = R(g.ball[3]);
    const V3<R> d = p - c;
    const R len = wsqrt(dot(d, d));
    best = rad - len;
    point = len > R(0) ? c + d * (rad / len) : V3<R>{c.x + rad, c.y, c.z};
    cls = 0;
  } else {
    best = Inf<R>::v();
    cls = 0;
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int k = face >> 1;
      const R x = comp(p, k);
      const R dist = (face & 1) ? R(g.box_hi[k]) - x : x - R(g.box_lo[k]);
      if (dist < best) {
        best = dist;
        cls = face;
      }
    }
    point = p;
    const R plane = (cls & 1) ? R(g.box_hi[cls >> 1]) : R(g.box_lo[cls >> 1]);
    if ((cls >> 1) == 0) point.x = plane; else if ((cls >> 1) == 1) point.y = plane; else point.z = plane;
  }
  return best;
}

// Holes (kKind != kGeomMesh) and cavities after the container: each wins
// only when strictly closer than the best so far.
template <typename R, int kKind>
__device__ __forceinline__ R holes_and_cavities(const DevGeom &g, V3<R> p, R best, V3<R> &point, int &cls,
                                                int32_t n_cavities) {
  if (kKind != kGeomMesh) {
    for (int i = 0; i < g.n_spheres; ++i) {
      const Vec4<R> s = load4<R>(g, 0, i);
      const V3<R> c{s.x, s.y, s.z};
      const V3<R> d = p - c;
      const R len = wsqrt(dot(d, d));
      const R dist = len - s.w;
      if (dist < best && len > R(0)) {
        best = dist;
        point = c + d * (s.w / len);
        cls = 6;
      }
    }
    for (int i = 0; i < g.n_cylinders; ++i) {
      const Vec4<R> s = load4<R>(g, 1, i);
      const int axis = int(s.x);
      const int ka = axis == 0 ? 1 : 0, kb = axis == 2 ? 1 : 2;
      const R da = comp(p, ka) - s.y, db = comp(p, kb) - s.z;
      const R len = wsqrt(da * da + db * db);
      const R dist = len - s.w;
      if (dist < best && len > R(0)) {
        best = dist;
        point = p;
        const R pa = s.y + da * (s.w / len), pb = s.z + db * (s.w / len);
        if (ka == 0) point.x = pa; else point.y = pa;
        if (kb == 1) point.y = pb; else point.z = pb;
        cls = 6;
      }
    }
  }
  for (int i = 0; i < n_cavities; ++i) {
    const Vec4<R> s = load4<R>(g, 2, i);
    const V3<R> c{s.x, s.y, s.z};
    const V3<R> d = p - c;
    const R len = wsqrt(dot(d, d));
    const R cr = R(g.cavity_radius);
    const R dist = len - cr;
    if (dist < best && len > R(0)) {
      best = dist;
      point = c + d * (cr / len);
      cls = 7;
    }
  }
  return best;
}

// Mesh distance from a squared closest distance: binary64 walks replaying the
// reference's recorded outputs round it to binary32 (the float query radius
// of wih_twos.cpp:426).
template <typename R>
__device__ __forceinline__ R mesh_distance(const DevGeom &g, R d2) {
  R best = wsqrt(d2);
  if (sizeof(R) == 8 && g.mesh_radius_f32) best = R(float(best));
  return best;
}

// Conservative radius from the distance grid.  The mesh distance d is
// 1-Lipschitz, so d(p) >= d(c) - |p - c| for every sample point c; a small
// margin covers the binary32 rounding of d(c) and of c itself.
//   centre grid: c = the centre of the cell holding p (clamped to the grid);
//   corner grid: the best of the eight corners of that cell, which is much
//   tighter where d is smooth (the corner most aligned with the gradient
//   loses |p - c| (1 - cos angle) instead of up to |p - c|).
//   corner records: the same eight corners stored per cell in binary16,
//   rounded down (so still lower bounds), read with one 16-byte load.
// Any sample gives a valid bound, so the cell index may round either way at a
// cell face: multiply by 1/h instead of dividing.
__device__ __forceinline__ float f16_lo(uint32_t w) { return float(__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float f16_hi(uint32_t w) { return float(__builtin_bit_cast(_Float16, (unsigned short)(w >> 16))); }

// Upper bound of the mesh distance at p from a sample value v (at most a
// binary16 ulp below the distance at its point, itself within binary32
// rounding) at distance s from p: the distance is 1-Lipschitz.
template <typename R>
__device__ __forceinline__ R grid_upper(R v, R s) {
  return (v * R(1.00390625) + s) * R(1.00001) + R(1e-6);
}

struct NoShadow {
  __device__ void operator()() const {}
};

// shadow() runs after the grid sample loads are issued and before their
// values are used, so independent work there overlaps the loads' latency.
template <typename R, class Shadow = NoShadow>
__device__ __forceinline__ R grid_lower(const DevGeom &g, V3<R> p, R *upper = nullptr, Shadow shadow = {}) {
  const R h = R(g.grid_h), inv_h = R(g.grid_inv_h);
  const int last = g.grid_corners == 1 ? 2 : 1;  // highest first index of a cell (per sample count)
  int i = int(floor((p.x - R(g.grid_lo[0])) * inv_h));
  int k = int(floor((p.y - R(g.grid_lo[1])) * inv_h));
  int l = int(floor((p.z - R(g.grid_lo[2])) * inv_h));
  i = min(max(i, 0), g.grid_n[0] - last);
  k = min(max(k, 0), g.grid_n[1] - last);
  l = min(max(l, 0), g.grid_n[2] - last);
  if (!g.grid_corners) {
    const V3<R> c{R(g.grid_lo[0]) + (R(i) + R(0.5)) * h, R(g.grid_lo[1]) + (R(k) + R(0.5)) * h,
                  R(g.grid_lo[2]) + (R(l) + R(0.5)) * h};
    const R dc = R(gptr(g.grid)[(size_t(i) * g.grid_n[1] + k) * g.grid_n[2] + l]);
    shadow();
    const V3<R> d = p - c;
    const R dist = wsqrt(dot(d, d));
    if (upper) *upper = grid_upper(dc, dist);
    return dc - dist - R(2e-6) * (R(1) + dc);
  }
  const R x0 = p.x - (R(g.grid_lo[0]) + R(i) * h), y0 = p.y - (R(g.grid_lo[1]) + R(k) * h),
          z0 = p.z - (R(g.grid_lo[2]) + R(l) * h);
  if (g.grid_corners == 2) {
    const uint4 rec = gload(reinterpret_cast<const uint4 *>(g.grid) + (size_t(i) * g.grid_n[1] + k) * g.grid_n[2] + l);
    shadow();
    const uint32_t w[4] = {rec.x, rec.y, rec.z, rec.w};
    R best = -Inf<R>::v(), up = Inf<R>::v();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const R dx = x0 - R(a) * h, dy = y0 - R(b) * h;
        const R dxy = dx * dx + dy * dy;
        const R d0 = R(f16_lo(w[2 * a + b])), d1 = R(f16_hi(w[2 * a + b]));
        const R s0 = wsqrt(dxy + z0 * z0), s1 = wsqrt(dxy + (z0 - h) * (z0 - h));
        // d - s - 2e-6 (1 + d) and the upper bound's d (1 + 2^-8) + s as
        // one fused multiply-add per corner; the constant parts after the
        // max and min.
        best = fmax(best, fmax(fma(d0, R(1) - R(2e-6), -s0), fma(d1, R(1) - R(2e-6), -s1)));
        up = fmin(up, fmin(fma(d0, R(1.00390625), s0), fma(d1, R(1.00390625), s1)));
      }
    if (upper) *upper = up * R(1.00001) + R(1e-6);
    return best - R(2e-6);
  }
  const GPtr<float> row = gptr(g.grid + (size_t(i) * g.grid_n[1] + k) * g.grid_n[2] + l);
  shadow();
  const size_t si = size_t(g.grid_n[1]) * g.grid_n[2], sk = size_t(g.grid_n[2]);
  R best = -Inf<R>::v(), up = Inf<R>::v();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const GPtr<float> q = row + a * si + b * sk;
      const R dx = x0 - R(a) * h, dy = y0 - R(b) * h;
      const R dxy = dx * dx + dy * dy;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const R dc = R(q[c]);
        const R dz = z0 - R(c) * h;
        const R s = wsqrt(dxy + dz * dz);
        best = fmax(best, dc - s - R(2e-6) * (R(1) + dc));
        up = fmin(up, grid_upper(dc, s));
      }
    }
  if (upper) *upper = up;
  return best;
}

// tau < inf: a mesh step whose grid lower bound is at least tau takes that
// bound as its radius (any ball inside the domain is a valid Walk on Spheres
// step) and skips the traversal; the point and class are then unused,
// because such a step never ends the path (tau > eps).
template <typename R, int kKind, bool kCount>
__device__ __forceinline__ R closest_boundary(const DevGeom &g, V3<R> p, V3<R> &point, int &cls, int32_t &hint,
                                              TraversalCount *cnt, R tau = Inf<R>::v()) {
  R best;
  if (kKind == kGeomMesh || (kKind == kGeomGeneric && g.container == kMesh)) {
    R ub = Inf<R>::v();
    const R lb = tau < Inf<R>::v() ? grid_lower<R>(g, p, &ub) : R(-1);
    if (lb >= tau) {
      best = lb;
      point = p;
    } else {
      MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
      // With the grid on (tau finite) an exact query searches below the
      // grid's upper bound and may start below the root.
      const R init2 = ub < Inf<R>::v() ? ub * ub : ub;
      best = mesh_distance<R>(g, mesh_closest<R, kCount>(m, p, point, hint, cnt, init2,
                                                         grid_start_node<R>(g, p, init2)));
    }
    cls = 0;
  } else {
    best = container_analytic<R>(g, p, point, cls);
  }
  return holes_and_cavities<R, kKind>(g, p, best, point, cls, g.n_cavities);
}

// ---------------------------------------------------------------------------
// Resumable closest-point query for the wavefront kernel: the same traversal
// as mesh_closest, one 4-wide node per call, so lanes whose query ends early
// go on with their walk while others are still traversing.  The stack lives
// in LDS (kLds entries per lane, column-major with the block's lane count
// as stride so a wave's lanes hit distinct banks) with a private overflow of
// kStackDepth - kLds entries; an entry packs the 4-wide node index (21 bits)
// with a lower bound of its squared box distance (the top 11 bits of the
// binary32 pattern, truncated, so never above the bound).  kLds is a kernel
// variant (16, 24 or 32 entries) chosen by the autotuner.

template <typename R>
struct MeshQuery {
  R best;        // squared distance so far (at first the bound it started from), or the grid radius when `grid`
  V3<R> pt;
  int32_t best_id;
  int32_t best_k;
  int32_t node;
  int32_t sp;
  int32_t pending;  // warm-start triangle still to test (first, in the first visit), or -1
  bool grid;
};

// Largest 4-wide node index an entry holds.  A tree over n triangles has at
// most n - 1 internal nodes, so meshes of up to 2^21 triangles fit.
constexpr int32_t kMaxStackIndex = (1 << 21) - 1;

__device__ __forceinline__ uint32_t stack_pack(int32_t node, float lower) {
  return (uint32_t(node) << 11) | ((__float_as_uint(lower) >> 20) & 0x7FFu);
}
__device__ __forceinline__ int32_t stack_node(uint32_t e) { return int32_t(e >> 11); }
__device__ __forceinline__ float stack_lower(uint32_t e) { return __uint_as_float((e & 0x7FFu) << 20); }

template <typename R>
__device__ __forceinline__ float lower_f32(R d2);
template <>
__device__ __forceinline__ float lower_f32<float>(float d2) { return d2; }
template <>
__device__ __forceinline__ float lower_f32<double>(double d2) { return __double2float_rd(d2); }

template <typename R, bool kCount>
__device__ __forceinline__ void mq_test(const MeshView &m, V3<R> p, MeshQuery<R> &q, int32_t k, TraversalCount *cnt) {
  V3<R> a, b, c;
  load_tri<R>(m, k, a, b, c);
  if (kCount) cnt->tris += 1;
  const V3<R> cp = closest_point_triangle(p, a, b, c);
  const V3<R> d = p - cp;
  const R d2 = dot(d, d);
  if constexpr (sizeof(R) == 8) {
    const int32_t id = gptr(m.tid)[k];
    if (d2 < q.best || (d2 == q.best && id < q.best_id)) {
      q.best = d2;
      q.pt = cp;
      q.best_id = id;
      q.best_k = k;
    }
  } else if (d2 < q.best) {
    q.best = d2;
    q.pt = cp;
    q.best_k = k;
  }
}

template <typename R>
__device__ __forceinline__ bool mq_within(const MeshQuery<R> &q, R d) {
  if constexpr (sizeof(R) == 8) return d * R(1.0 - 1e-12) <= q.best;
  else return d < q.best;
}

// Starts a query; returns true when it is already complete.  init2: a
// squared distance above the true one (or infinity) the search starts from;
// start: the node to start at (the root, or a grid cell's start node valid
// for that bound).  The warm-start triangle is tested first, in the first
// visit, where it shares a pass with other lanes' leaf tests.
template <typename R, bool kCount>
__device__ __forceinline__ bool mq_begin(const MeshView &m, V3<R> p, MeshQuery<R> &q, int32_t hint,
                                         TraversalCount *cnt, R init2 = Inf<R>::v(), int32_t start = 0) {
  q.best = init2;
  q.best_id = 0x7fffffff;
  q.best_k = -1;
  q.node = start;
  q.sp = 0;
  q.grid = false;
  q.pending = hint >= 0 && hint < m.n_tris ? hint : -1;
  if (m.n_tris == 1) {
    q.best = Inf<R>::v();
    q.pending = -1;
    mq_test<R, kCount>(m, p, q, 0, cnt);
    return true;
  }
  return false;
}

// Visits q.node; returns true when the query is complete.
template <typename R, bool kCount, int kLds>
__device__ __forceinline__ bool mq_visit(const MeshView &m, V3<R> p, MeshQuery<R> &q, uint32_t *lds,
                                         uint32_t stride, uint32_t *overflow, TraversalCount *cnt) {
  const float4 *nq = m.nodes4 + size_t(kNode4Quads) * q.node;
  const float4 lx = gload(nq), ly = gload(nq + 1), lz = gload(nq + 2), hx = gload(nq + 3), hy = gload(nq + 4),
               hz = gload(nq + 5);
  const int4 ch = gload(reinterpret_cast<const int4 *>(nq + 6));
  if (kCount) {
    cnt->nodes += 1;
    cnt->top += __float_as_int(gload(nq + 7).z) < kTopLevels ? 1u : 0u;
  }
  R d[4] = {box_d2<R>(p, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x), box_d2<R>(p, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y),
            box_d2<R>(p, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z), box_d2<R>(p, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w)};
  int32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
  // Triangle tests, one per pass: the pending warm-start triangle first,
  // then leaf children (one triangle each) in slot order; the wave runs as
  // many passes as its lanes have triangles to test, not one per slot that
  // any lane uses.  A leaf whose box no longer beats the best distance is
  // skipped; that skip never changes the result (its triangle is at least as
  // far as its box).
  uint32_t leaves = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (c[i] < 0 && c[i] != kEmptyChild && mq_within(q, d[i])) leaves |= 1u << i;
  int32_t pend = q.pending;
  q.pending = -1;
  while (pend >= 0 || leaves) {
    int32_t tri;
    bool go = true;
    if (pend >= 0) {
      tri = pend;
      pend = -1;
    } else {
      const int i = __builtin_ctz(leaves);
      leaves &= leaves - 1;
      tri = ~(i == 0 ? c[0] : (i == 1 ? c[1] : (i == 2 ? c[2] : c[3])));
      go = mq_within(q, i == 0 ? d[0] : (i == 1 ? d[1] : (i == 2 ? d[2] : d[3])));
    }
    if (go) mq_test<R, kCount>(m, p, q, tri, cnt);
  }
  R key[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) key[i] = (c[i] >= 0 && mq_within(q, d[i])) ? d[i] : Inf<R>::v();
  auto cswap = [&](int i, int j) {
    if (key[j] < key[i]) {
      const R tk = key[i]; key[i] = key[j]; key[j] = tk;
      const int32_t tc = c[i]; c[i] = c[j]; c[j] = tc;
    }
  };
  cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
  const R inf = Inf<R>::v();
  if (key[0] < inf) {
#pragma unroll
    for (int i = 3; i >= 1; --i) {
      if (key[i] < inf && q.sp < kStackDepth) {
        const uint32_t e = stack_pack(c[i], lower_f32<R>(key[i]));
        if (q.sp < kLds) lds[q.sp * stride] = e;
        else overflow[q.sp - kLds] = e;
        ++q.sp;
      }
    }
    q.node = c[0];
    return false;
  }
  while (q.sp > 0) {
    --q.sp;
    const uint32_t e = q.sp < kLds ? lds[q.sp * stride] : overflow[q.sp - kLds];
    if (mq_within(q, R(stack_lower(e)))) {
      q.node = stack_node(e);
      return false;
    }
  }
  if (q.best_k < 0 && q.best < Inf<R>::v()) {
    // A search started from a bound found nothing below it (never expected:
    // the bound is above the true distance); search the whole tree.
    q.best = Inf<R>::v();
    q.node = 0;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Boundary data g(y, class) and volume source f(x) of one side.
template <typename R>
__device__ __forceinline__ R side_payoff(const DevSide &s, V3<R> y, int cls) {
  const R *c;
  R value;
  if constexpr (sizeof(R) == 4) {
    c = s.poly_f;
    value = s.cls_const_f[cls & 7];
  } else {
    c = s.poly;
    value = s.cls_const[cls & 7];
  }
  if ((s.cls_poly_mask >> (cls & 7)) & 1U) {
    value += c[0] + c[1] * y.x + c[2] * y.y + c[3] * y.z + c[4] * y.x * y.x + c[5] * y.y * y.y +
             c[6] * y.z * y.z + c[7] * y.x * y.y + c[8] * y.x * y.z + c[9] * y.y * y.z;
  }
  return value;
}

// f = ((s0 + sx (x+ox)) + sy (y+oy)) + sz (z+oz); the reference's affine
// source 1 + slope (y + 0.5) (wih_twos.cpp:546-552) is one instance and gives
// the same binary64 value in the contraction-free binary64 build.
template <typename R>
__device__ __forceinline__ R side_source(const DevSide &s, V3<R> x) {
  const R *c;
  if constexpr (sizeof(R) == 4) c = s.src_f; else c = s.src;
  return ((c[0] + c[1] * (x.x + c[4])) + c[2] * (x.y + c[5])) + c[3] * (x.z + c[6]);
}

// Sampled source of source_kind 2: f(a) = s0 + sum of the side's term rows in
// order (heat bumps, the unit-mass sensor bump, housing Gaussians), each in
// the reference's operation order (study.cpp:29-51, wih_twos.cpp:553-566).
__device__ __forceinline__ float exp_r(float v) { return expf(v); }
__device__ __forceinline__ double exp_r(double v) { return exp(v); }

template <typename R>
__device__ __forceinline__ R side_field(const DevSide &s, V3<R> a) {
  R f = R(s.src[0]);
  const int n = min(s.n_terms, TWOS_MAX_TERMS);
  for (int k = 0; k < n; ++k) {
    const double *t = s.terms[k];
    const int type = int(t[0]);
    if (type == TWOS_TERM_GAUSS) {
      const R u = (a.x - R(t[1])) / R(t[4]), v = (a.y - R(t[2])) / R(t[5]), w = (a.z - R(t[3])) / R(t[6]);
      f += R(t[7]) * exp_r(-(u * u) - v * v - w * w);
    } else {
      const V3<R> d = a - V3<R>{R(t[1]), R(t[2]), R(t[3])};
      const R q = fmax(R(0), R(1) - dot(d, d) / (R(t[4]) * R(t[4])));
      if (type == TWOS_TERM_UNIT) f += q * q / R(t[5]);
      else f += R(t[7]) * (q * q);
    }
  }
  return f;
}

// Exact Green-ball integral of a compact bump A (1 - s^2/R^2)^2 (s the
// distance to its centre) over the ball B(x, r) whose centre is d from the
// bump's: with Phi the bump's Newtonian potential (-Lap Phi = bump, Phi -> 0
// at infinity), the integral of G_B(x, .) times the bump is Phi(d) minus the
// mean of Phi over the sphere |y - x| = r, and that mean is
// (P(d + r) - P(|d - r|)) / (2 d r) with P' (s) = s Phi(s) (the reference's
// exact_bump_green, study.cpp:56-76, evaluates the same quantity in long
// double).  In u = s / R and total mass M = 4 pi A R^3 8/105:
//   Phi = A R^2 (1/6 - u^2/6 + u^4/10 - u^6/42)          (s < R)
//       = A R^3 (8/105) / s                               (s >= R)
//   P   = A R^4 u^2 (1/12 - u^2/24 + u^4/60 - u^6/336)    (s <= R)
//       = P(R) + A R^3 (8/105) (s - R)                    (s > R)
__device__ __forceinline__ double bump_potential(double s, double R, double A) {
  if (s >= R) return A * R * R * R * (8.0 / 105.0) / s;
  const double u2 = (s / R) * (s / R);
  return A * R * R * (1.0 / 6.0 + u2 * (-1.0 / 6.0 + u2 * (1.0 / 10.0 - u2 / 42.0)));
}
__device__ __forceinline__ double bump_primitive(double s, double R, double A) {
  const double u = fmin(s / R, 1.0), u2 = u * u;
  const double inner = A * R * R * R * R * u2 * (1.0 / 12.0 + u2 * (-1.0 / 24.0 + u2 * (1.0 / 60.0 - u2 / 336.0)));
  return s > R ? inner + A * R * R * R * (8.0 / 105.0) * (s - R) : inner;
}
__device__ __forceinline__ double bump_green_ball(double d, double r, double R, double A) {
  if (d >= r + R) return 0.0;  // the bump lies outside the ball
  double mean;
  if (d + R <= r) mean = A * R * R * R * (8.0 / 105.0) / r;  // the sphere encloses the whole bump
  else if (d < 1e-12 * R) mean = bump_potential(r, R, A);    // concentric: Phi is constant on the sphere
  else mean = (bump_primitive(d + r, R, A) - bump_primitive(fabs(d - r), R, A)) / (2.0 * d * r);
  return fmax(0.0, bump_potential(d, R, A) - mean);
}

// source_kind 3: the exact integral of the ball's Green function against s0
// (s0 r^2 / 6) and each BUMP / UNIT row, in binary64.
template <typename R>
__device__ __forceinline__ double side_analytic(const DevSide &s, V3<R> x, R r) {
  const double rr = double(r);
  double sum = s.src[0] * rr * rr / 6.0;
  const int n = min(s.n_terms, TWOS_MAX_TERMS);
  for (int k = 0; k < n; ++k) {
    const double *t = s.terms[k];
    const double dx = double(x.x) - t[1], dy = double(x.y) - t[2], dz = double(x.z) - t[3];
    const double amp = int(t[0]) == TWOS_TERM_UNIT ? 1.0 / t[5] : t[7];
    sum += bump_green_ball(sqrt(dx * dx + dy * dy + dz * dz), rr, t[4], amp);
  }
  return sum;
}

// (r^2/6) f of one step for one side.  x: ball centre, a: Green point.
// paired_form follows paired_site's r*r/6*f (study.cpp:106-108), otherwise
// walk()'s f*r*r/6 (wih_twos.cpp:648-653).
template <typename R>
__device__ __forceinline__ R side_green_term(const DevSide &s, V3<R> x, V3<R> a, R r, bool paired_form) {
  if (s.source_kind == 1) return side_source<R>(s, x) * r * r / R(6);
  if (s.source_kind == 2) return paired_form ? r * r / R(6) * side_field<R>(s, a) : side_field<R>(s, a) * r * r / R(6);
  if (s.source_kind == 3) return R(side_analytic<R>(s, x, r));
  return R(0);
}

// Green-ball offset r t omega from the Philox stream of (key, walk, step):
// omega from block (step, walk, 1, 0) as the move direction is built; t the
// median of three 24-bit uniforms from words 0..2 of block (step, walk, 2, 0)
// (the median of three uniforms is Beta(2,2)); the same t in both precisions.
template <typename R>
__device__ __forceinline__ V3<R> philox_green(uint2 key, int32_t walk, int32_t step, R r) {
  const uint4 w1 = philox4x32_10(make_uint4(uint32_t(step), uint32_t(walk), 1u, 0u), key);
  const uint4 w2 = philox4x32_10(make_uint4(uint32_t(step), uint32_t(walk), 2u, 0u), key);
  const V3<R> d = philox_direction(w1, R(0));
  const R a = R(w2.x >> 8) * R(0x1.0p-24), b = R(w2.y >> 8) * R(0x1.0p-24), c = R(w2.z >> 8) * R(0x1.0p-24);
  const R t = fmax(fmin(a, b), fmin(fmax(a, b), c));
  return d * (r * t);
}

}  // namespace twos
