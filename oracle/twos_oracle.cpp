// CPU oracle for the TWoS estimator -- TEST INFRASTRUCTURE ONLY.
//
// This file restates the reference estimator in plain C++17 + OpenMP, binary64,
// no contraction.  It is the checker for the HIP library, never part of it:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// Reference citations are paths under the hertjo/walking-into-holes tree.
//
// Parity of each part is pinned by tests (see twos_oracle.h for the list).

#include "twos_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x = 0.0, y = 0.0, z = 0.0;
};
inline V3 operator+(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(const V3 &a) { return std::sqrt(dot(a, a)); }
inline V3 load3(const double *p) { return {p[0], p[1], p[2]}; }
inline double comp(const V3 &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

constexpr double kInf = std::numeric_limits<double>::infinity();

// splitmix64 finalizer: benchmarks/compiled_bunny/src/wih_twos.cpp:529-534.
uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30U)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27U)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31U);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with a Weyl key
// schedule.  Multipliers and Weyl increments are the published constants.
void philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream identity of one (seed, site, side): a 64-bit key for Philox.
void philox_key(uint64_t seed, int64_t site, int32_t side, uint32_t key[2]) {
  const uint64_t k = splitmix64(splitmix64(seed) + (uint64_t(site) << 1 | uint64_t(side & 1)));
  key[0] = uint32_t(k);
  key[1] = uint32_t(k >> 32);
}

inline double u53(uint32_t lo, uint32_t hi) {
  return double(((uint64_t(hi) << 32) | lo) >> 11) * 0x1.0p-53;
}

constexpr double kTwoPi = 6.283185307179586;  // == 2 * acos(-1.0) in binary64

// Direction for draw block `step` of walk `walk` under `key`: one Philox block
// per step, u1 from words (0,1), u2 from words (2,3).  The same z / phi / radial
// construction as random_unit_vector (wih_twos.cpp:536-542).
// Counter word 2 selects the draw within a step: 0 the move direction, 1 and
// 2 the Green-ball point of a sampled source.
V3 philox_direction(const uint32_t key[2], uint32_t walk, uint32_t step, uint32_t lane = 0u) {
  const uint32_t ctr[4] = {step, walk, lane, 0u};
  uint32_t w[4];
  philox4x32_10(ctr, key, w);
  const double z = 1.0 - 2.0 * u53(w[0], w[1]);
  const double phi = kTwoPi * u53(w[2], w[3]);
  const double radial = std::sqrt(std::max(0.0, 1.0 - z * z));
  return {radial * std::cos(phi), radial * std::sin(phi), z};
}

// random_unit_vector: wih_twos.cpp:536-542 (libstdc++ uniform_real_distribution).
V3 random_unit_vector(std::mt19937_64 &rng) {
  std::uniform_real_distribution<double> uniform(0.0, 1.0);
  const double z = 1.0 - 2.0 * uniform(rng);
  const double phi = 2.0 * std::acos(-1.0) * uniform(rng);
  const double radial = std::sqrt(std::max(0.0, 1.0 - z * z));
  return {radial * std::cos(phi), radial * std::sin(phi), z};
}

// beta_22: Beta(2,2) by rejection from U(0,1) pairs, wih_twos.cpp:569-575.
double beta_22(std::mt19937_64 &rng) {
  std::uniform_real_distribution<double> uniform(0.0, 1.0);
  while (true) {
    const double value = uniform(rng);
    if (uniform(rng) <= 4.0 * value * (1.0 - value)) return value;
  }
}

// Beta(2,2) for the Philox streams: the median of three uniforms is Beta(2,2)
// (order statistic 2 of 3).  Uniforms carry 24 bits so the binary32 and
// binary64 walks draw the same value: words 0..2 of block (step, walk, 2, 0).
double philox_beta22(const uint32_t key[2], uint32_t walk, uint32_t step) {
  const uint32_t ctr[4] = {step, walk, 2u, 0u};
  uint32_t w[4];
  philox4x32_10(ctr, key, w);
  const double a = double(w[0] >> 8) * 0x1.0p-24, b = double(w[1] >> 8) * 0x1.0p-24,
               c = double(w[2] >> 8) * 0x1.0p-24;
  return std::max(std::min(a, b), std::min(std::max(a, b), c));
}

// Per-path random draws.  dir(step): the move direction; green(step, r): the
// offset r * t * omega of a Green-distributed point in the ball of radius r
// (t ~ Beta(2,2), omega uniform on the sphere), the source point of
// wih_twos.cpp:648-653 and study.cpp:106.
struct PhiloxStream {
  uint32_t key[2];
  uint32_t walk;
  V3 dir(int32_t step) const { return philox_direction(key, walk, uint32_t(step)); }
  V3 green(int32_t step, double radius) const {
    const V3 d = philox_direction(key, walk, uint32_t(step), 1u);
    return d * (radius * philox_beta22(key, walk, uint32_t(step)));
  }
};

// The reference draws the Green point as random_unit_vector(rng) *
// (radius * beta_22(rng)); g++ evaluates the right operand, the Beta draw,
// first.
struct MtStream {
  std::mt19937_64 rng;
  V3 dir(int32_t) { return random_unit_vector(rng); }
  V3 green(int32_t, double radius) {
    const double t = beta_22(rng);
    const V3 d = random_unit_vector(rng);
    return d * (radius * t);
  }
};

// Ericson's closest point on a triangle: wih_twos.cpp:247-289.
V3 closest_point_triangle(const V3 &p, const V3 &a, const V3 &b, const V3 &c) {
  const V3 ab = b - a;
  const V3 ac = c - a;
  const V3 ap = p - a;
  const double d1 = dot(ab, ap);
  const double d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const V3 bp = p - b;
  const double d3 = dot(ab, bp);
  const double d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    return a + ab * v;
  }
  const V3 cp = p - c;
  const double d5 = dot(ab, cp);
  const double d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    return a + ac * w;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const V3 bc = c - b;
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + bc * w;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom;
  const double w = vc * denom;
  return a + ab * v + ac * w;
}

}  // namespace

// ---------------------------------------------------------------------------
// Triangle mesh with the oracle's own BVH (median split, <=4 triangles/leaf).
// The closest-point query returns the exact minimum over all triangles of
// |p - closest_point_triangle|, ties broken by the lowest triangle index, so it
// equals a brute-force scan bit for bit.  The reference's Embree query
// (wih_twos.cpp:339-359, :416-430) only prunes and computes the same binary64
// distance in its callback.
struct og_mesh {
  struct Node {
    double lo[3], hi[3];
    int32_t left, right;  // internal: children; leaf: left = first, right = -count
  };
  std::vector<V3> v;
  std::vector<std::array<uint32_t, 3>> f;
  std::vector<uint32_t> order;
  std::vector<Node> nodes;

  void bounds_of(uint32_t t, double lo[3], double hi[3]) const {
    for (int k = 0; k < 3; ++k) { lo[k] = kInf; hi[k] = -kInf; }
    for (int j = 0; j < 3; ++j) {
      const V3 &q = v[f[t][j]];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], comp(q, k));
        hi[k] = std::max(hi[k], comp(q, k));
      }
    }
  }

  int32_t build(uint32_t begin, uint32_t end) {
    Node node;
    for (int k = 0; k < 3; ++k) { node.lo[k] = kInf; node.hi[k] = -kInf; }
    double clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
    for (uint32_t i = begin; i < end; ++i) {
      double lo[3], hi[3];
      bounds_of(order[i], lo, hi);
      for (int k = 0; k < 3; ++k) {
        node.lo[k] = std::min(node.lo[k], lo[k]);
        node.hi[k] = std::max(node.hi[k], hi[k]);
        const double c = 0.5 * (lo[k] + hi[k]);
        clo[k] = std::min(clo[k], c);
        chi[k] = std::max(chi[k], c);
      }
    }
    const int32_t index = int32_t(nodes.size());
    nodes.push_back(node);
    if (end - begin <= 4) {
      nodes[index].left = int32_t(begin);
      nodes[index].right = -int32_t(end - begin);
      return index;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    const uint32_t mid = begin + (end - begin) / 2;
    std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end,
                     [&](uint32_t a, uint32_t b) {
                       double la[3], ha[3], lb[3], hb[3];
                       bounds_of(a, la, ha);
                       bounds_of(b, lb, hb);
                       const double ca = la[axis] + ha[axis], cb = lb[axis] + hb[axis];
                       return ca < cb || (ca == cb && a < b);
                     });
    const int32_t l = build(begin, mid);
    const int32_t r = build(mid, end);
    nodes[index].left = l;
    nodes[index].right = r;
    return index;
  }

  static double box_d2(const Node &n, const V3 &p) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double x = comp(p, k);
      const double d = x < n.lo[k] ? n.lo[k] - x : (x > n.hi[k] ? x - n.hi[k] : 0.0);
      s += d * d;
    }
    return s;
  }

  void test_triangle(uint32_t t, const V3 &p, double &best, V3 &best_pt, int64_t &best_tri) const {
    const V3 q = closest_point_triangle(p, v[f[t][0]], v[f[t][1]], v[f[t][2]]);
    const V3 d = p - q;
    const double d2 = dot(d, d);
    if (d2 < best || (d2 == best && int64_t(t) < best_tri)) {
      best = d2;
      best_pt = q;
      best_tri = t;
    }
  }

  double closest(const V3 &p, V3 &out, int64_t &tri, bool brute, int64_t *nv, int64_t *nt) const {
    double best = kInf;
    V3 best_pt;
    int64_t best_tri = std::numeric_limits<int64_t>::max();
    int64_t visited = 0, tested = 0;
    if (brute) {
      for (uint32_t t = 0; t < f.size(); ++t) test_triangle(t, p, best, best_pt, best_tri);
      tested = int64_t(f.size());
    } else {
      int32_t stack[128];
      int sp = 0;
      stack[sp++] = 0;
      while (sp) {
        const Node &n = nodes[size_t(stack[--sp])];
        ++visited;
        // Relative slack keeps rounding in the box bound from pruning a triangle
        // whose computed distance ties the current best.
        if (box_d2(n, p) * (1.0 - 1e-12) > best) continue;
        if (n.right < 0) {
          for (int32_t i = 0; i < -n.right; ++i) {
            ++tested;
            test_triangle(order[size_t(n.left + i)], p, best, best_pt, best_tri);
          }
          continue;
        }
        const double dl = box_d2(nodes[size_t(n.left)], p);
        const double dr = box_d2(nodes[size_t(n.right)], p);
        if (dl <= dr) {
          stack[sp++] = n.right;
          stack[sp++] = n.left;
        } else {
          stack[sp++] = n.left;
          stack[sp++] = n.right;
        }
      }
    }
    if (nv) *nv += visited;
    if (nt) *nt += tested;
    out = best_pt;
    tri = best_tri;
    return std::sqrt(best);
  }

  // Closest ray hit with t in [tnear, tfar] (Moller-Trumbore in binary64).
  bool ray_hit(const V3 &o, const V3 &d, double tnear, double tfar, double &t_hit) const {
    double best = tfar;
    bool found = false;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node &n = nodes[size_t(stack[--sp])];
      double t0 = tnear, t1 = best;
      bool miss = false;
      for (int k = 0; k < 3 && !miss; ++k) {
        const double ok = comp(o, k), dk = comp(d, k);
        if (dk == 0.0) {
          if (ok < n.lo[k] || ok > n.hi[k]) miss = true;
          continue;
        }
        double ta = (n.lo[k] - ok) / dk, tb = (n.hi[k] - ok) / dk;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta - 1e-12);
        t1 = std::min(t1, tb + 1e-12);
        if (t0 > t1) miss = true;
      }
      if (miss) continue;
      if (n.right < 0) {
        for (int32_t i = 0; i < -n.right; ++i) {
          const auto &tri = f[order[size_t(n.left + i)]];
          const V3 a = v[tri[0]], e1 = v[tri[1]] - a, e2 = v[tri[2]] - a;
          const V3 pv{d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x};
          const double det = dot(e1, pv);
          if (det == 0.0) continue;
          const double inv = 1.0 / det;
          const V3 tv = o - a;
          const double uu = dot(tv, pv) * inv;
          if (uu < 0.0 || uu > 1.0) continue;
          const V3 qv{tv.y * e1.z - tv.z * e1.y, tv.z * e1.x - tv.x * e1.z, tv.x * e1.y - tv.y * e1.x};
          const double vv = dot(d, qv) * inv;
          if (vv < 0.0 || uu + vv > 1.0) continue;
          const double t = dot(e2, qv) * inv;
          if (t >= tnear && t <= best) {
            best = t;
            found = true;
          }
        }
        continue;
      }
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
    t_hit = best;
    return found;
  }

  // Ray parity along one fixed irrational direction, stepping from hit to hit:
  // EmbreeMesh::inside, wih_twos.cpp:361-394 (tnear 1e-6, tfar 4, step 2e-6,
  // at most 128 crossings).  Returns -1 when the crossing limit is exceeded.
  int inside(const V3 &p, int64_t *rays = nullptr) const {
    V3 direction{1.0, 0.3713906763541037, 0.1570796326794897};
    direction = direction * (1.0 / norm(direction));
    V3 origin = p;
    unsigned hits = 0;
    for (; hits < 128; ++hits) {
      double t;
      if (rays) ++*rays;  // one ray per intersection query, as note_query(1) counts them
      if (!ray_hit(origin, direction, 1.0e-6, 4.0, t)) break;
      origin = origin + direction * (t + 2.0e-6);
    }
    if (hits == 128) return -1;
    return int(hits % 2U == 1U);
  }
};

namespace {

// Closest boundary point of the domain and its boundary class.
// Mesh + cavities follows domain_distance / closest_boundary
// (wih_twos.cpp:585-593, :602-626): the mesh first, then each cavity wins only
// when strictly closer.  Analytic containers and holes are the build's data
// model (sphere, box and axis-aligned cylinder SDFs), evaluated the same way.
double closest_boundary(const og_geom &g, const V3 &p, V3 &point, int32_t &cls) {
  double best;
  if (g.container == OG_MESH) {
    int64_t tri;
    best = g.mesh->closest(p, point, tri, false, nullptr, nullptr);
    // The retained reference outputs (walk-traces.csv, field-*.csv) were written
    // by a build whose distance was the float query radius that
    // point_query_callback stores (wih_twos.cpp:426): replaying them bit for
    // bit needs the binary32-rounded distance.
    if (g.mesh_radius_f32) best = double(float(best));
    cls = 0;
  } else if (g.container == OG_BALL) {
    const V3 c{g.ball[0], g.ball[1], g.ball[2]};
    const V3 d = p - c;
    const double len = norm(d);
    best = g.ball[3] - len;
    point = len > 0.0 ? c + d * (g.ball[3] / len) : V3{c.x + g.ball[3], c.y, c.z};
    cls = 0;
  } else {
    best = kInf;
    for (int face = 0; face < 6; ++face) {
      const int k = face >> 1;
      const double x = comp(p, k);
      const double dist = (face & 1) ? g.box_hi[k] - x : x - g.box_lo[k];
      if (dist < best) {
        best = dist;
        point = p;
        const double plane = (face & 1) ? g.box_hi[k] : g.box_lo[k];
        if (k == 0) point.x = plane; else if (k == 1) point.y = plane; else point.z = plane;
        cls = face;
      }
    }
  }
  for (int32_t i = 0; i < g.n_spheres; ++i) {
    const double *s = g.spheres + 4 * i;
    const V3 c{s[0], s[1], s[2]};
    const V3 d = p - c;
    const double len = norm(d);
    const double dist = len - s[3];
    if (dist < best && len > 0.0) {
      best = dist;
      point = c + d * (s[3] / len);
      cls = 6;
    }
  }
  for (int32_t i = 0; i < g.n_cylinders; ++i) {
    const double *s = g.cylinders + 4 * i;
    const int axis = int(s[0]);
    const int ka = axis == 0 ? 1 : 0, kb = axis == 2 ? 1 : 2;
    const double da = comp(p, ka) - s[1], db = comp(p, kb) - s[2];
    const double len = std::sqrt(da * da + db * db);
    const double dist = len - s[3];
    if (dist < best && len > 0.0) {
      best = dist;
      point = p;
      const double pa = s[1] + da * (s[3] / len), pb = s[2] + db * (s[3] / len);
      if (ka == 0) point.x = pa; else point.y = pa;
      if (kb == 1) point.y = pb; else point.z = pb;
      cls = 6;
    }
  }
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const V3 c = load3(g.cavities + 3 * i);
    const V3 radial = p - c;
    const double len = norm(radial);
    const double dist = len - g.cavity_radius;
    if (dist < best && len > 0.0) {
      best = dist;
      point = c + radial * (g.cavity_radius / len);
      cls = 7;
    }
  }
  return best;
}

double payoff(const og_side &s, const V3 &y, int32_t cls) {
  double value = s.cls_const[cls & 7];
  if (s.cls_poly_mask >> (cls & 7) & 1U) {
    const double *c = s.poly;
    value += c[0] + c[1] * y.x + c[2] * y.y + c[3] * y.z + c[4] * y.x * y.x + c[5] * y.y * y.y +
             c[6] * y.z * y.z + c[7] * y.x * y.y + c[8] * y.x * y.z + c[9] * y.y * y.z;
  }
  return value;
}

// Source of the Green-ball term: the reference's affine source
// f = 1 + slope * (y + 0.5) (wih_twos.cpp:546-552) is the case s0 = 1,
// sy = slope, oy = 0.5, every other coefficient zero, and evaluates to the
// same binary64 value.
double source(const og_side &s, const V3 &x) {
  const double *c = s.src;
  return ((c[0] + c[1] * (x.x + c[4])) + c[2] * (x.y + c[5])) + c[3] * (x.z + c[6]);
}

// Sampled source of source_kind 2: f(y) = s0 + sum of the side's terms, summed
// in order.  The heat bumps of study.cpp:29-36 (s0 = 0, BUMP rows), its
// normalised sensor bump (:44-51, one UNIT row with p1 = 32 pi R^3 / 105) and
// the housing chips of wih_twos.cpp:553-566 (s0 = 0.08, GAUSS rows) evaluate
// to the reference's binary64 values.
double field(const og_side &s, const V3 &y) {
  double f = s.src[0];
  for (int32_t k = 0; k < s.n_terms && k < OG_MAX_TERMS; ++k) {
    const double *t = s.terms[k];
    const int type = int(t[0]);
    if (type == OG_TERM_GAUSS) {
      const double a = (y.x - t[1]) / t[4], b = (y.y - t[2]) / t[5], c = (y.z - t[3]) / t[6];
      f += t[7] * std::exp(-(a * a) - b * b - c * c);
    } else {
      const V3 d = y - V3{t[1], t[2], t[3]};
      const double q = std::max(0.0, 1.0 - dot(d, d) / (t[4] * t[4]));
      if (type == OG_TERM_UNIT) f += q * q / t[5];
      else f += t[7] * (q * q);
    }
  }
  return f;
}

// Analytic source (source_kind 3): the exact integral of the Green function
// of B(x, r) against a compact bump A (1 - s^2/R^2)^2 whose centre is d from
// x, as exact_bump_green computes it (study.cpp:56-76, long double):
// Phi(d) minus the mean of Phi over the sphere |y - x| = r, Phi the bump's
// Newtonian potential and the mean (P(d + r) - P(|d - r|)) / (2 d r) with
// P' = s Phi.  Derivation: -Lap Phi = bump with Phi radial, so
// Phi = A R^2 (1/6 - u^2/6 + u^4/10 - u^6/42) inside (u = s/R) and
// M / (4 pi s) outside, M = 4 pi A R^3 8/105 the bump's mass.
using LD = long double;
LD bump_potential(LD s, LD R, LD A) {
  if (s >= R) return A * R * R * R * (8.0L / 105.0L) / s;
  const LD u = s / R, u2 = u * u;
  return A * R * R * (1.0L / 6.0L - u2 / 6.0L + u2 * u2 / 10.0L - u2 * u2 * u2 / 42.0L);
}
LD bump_primitive(LD s, LD R, LD A) {
  const LD u = std::min(s / R, 1.0L), u2 = u * u;
  const LD inner = A * R * R * R * R * (u2 / 12.0L - u2 * u2 / 24.0L + u2 * u2 * u2 / 60.0L - u2 * u2 * u2 * u2 / 336.0L);
  return s > R ? inner + A * R * R * R * (8.0L / 105.0L) * (s - R) : inner;
}
double bump_green_ball(const V3 &x, double radius, const V3 &c, double support, double amplitude) {
  const V3 dv = x - c;
  const LD d = std::sqrt(LD(dv.x) * dv.x + LD(dv.y) * dv.y + LD(dv.z) * dv.z);
  const LD r = radius, R = support, A = amplitude;
  if (d >= r + R) return 0.0;
  LD mean;
  if (d + R <= r) mean = A * R * R * R * (8.0L / 105.0L) / r;
  else if (d < 1e-12L * R) mean = bump_potential(r, R, A);
  else mean = (bump_primitive(d + r, R, A) - bump_primitive(std::fabs(d - r), R, A)) / (2 * d * r);
  return double(std::max(0.0L, bump_potential(d, R, A) - mean));
}

double analytic_term(const og_side &s, const V3 &x, double radius) {
  double sum = s.src[0] * radius * radius / 6.0;
  for (int32_t k = 0; k < std::min(s.n_terms, int32_t(OG_MAX_TERMS)); ++k) {
    const double *t = s.terms[k];
    const double amp = int(t[0]) == OG_TERM_UNIT ? 1.0 / t[5] : t[7];
    sum += bump_green_ball(x, radius, V3{t[1], t[2], t[3]}, t[4], amp);
  }
  return sum;
}

// Green-ball source term of one step for one side: (r^2/6) f.  `paired_form`
// follows paired_site's association r*r/6*f (study.cpp:106-108), otherwise
// walk()'s f*r*r/6 (wih_twos.cpp:648-653).
double green_term(const og_side &s, const V3 &x, const V3 &a, double radius, bool paired_form) {
  if (s.source_kind == 3) return analytic_term(s, x, radius);
  if (s.source_kind == 1) return source(s, x) * radius * radius / 6.0;
  if (s.source_kind == 2)
    return paired_form ? radius * radius / 6.0 * field(s, a) : field(s, a) * radius * radius / 6.0;
  return 0.0;
}

struct WalkOut {
  double a = 0.0, b = 0.0;
  int32_t steps = 0;
  bool cap = false;
  int32_t caps = 0;  // capped paths of this sample (two paths for CROSS and COUPLED)
  // First step, for the gradient channel: 3 omega0 / r0, r0 (0 if the path
  // ended at its first query) and both sides' values after its source term.
  V3 w0{0.0, 0.0, 0.0};
  double r0 = 0.0, s0 = 0.0, s1 = 0.0;
  V3 ga{0.0, 0.0, 0.0}, gb{0.0, 0.0, 0.0};  // the gradient samples
};

// Gradient sample of one side from a path's first step: (3/r0)(V - s0) omega0
// plus (r0^2/10) grad f for an affine source (the solution of -Lap w = f in the
// ball, w = 0 on its sphere, has gradient r0^2 grad f / 10 at the centre).
V3 grad_sample(const WalkOut &o, double value, double first, const og_side &s) {
  if (!(o.r0 > 0.0)) return {0.0, 0.0, 0.0};
  V3 g = o.w0 * (value - first);
  if (s.source_kind == 1) {
    const double c = o.r0 * o.r0 / 10.0;
    g = g + V3{s.src[1] * c, s.src[2] * c, s.src[3] * c};
  }
  return g;
}

// One Walk on Spheres path: wih_twos.cpp:635-664.  Each step queries the
// domain distance r, stops when !(r > eps), adds the Green-ball source
// (r^2/6) f(x) at the ball centre, and moves to x + r * direction.  A path
// that takes max_steps steps is capped.  The build adds the Dirichlet payoff
// g(closest boundary point) at termination (the reference's terminal value
// is zero; its 2-D harmonic-data experiment projects to the boundary the same
// way, experiments/run_controlled_experiments.py:115-147).  `two` accumulates
// a second side's source and payoff along the same path (study.cpp:91-113).
template <class Stream>
WalkOut walk(const og_geom &g, const og_side &s0, const og_side *s1, V3 x, double eps,
             int32_t max_steps, Stream &stream, bool paired_form, double *trace, int32_t max_rows,
             int32_t *rows) {
  const bool sampled = s0.source_kind == 2 || (s1 && s1->source_kind == 2);
  WalkOut out;
  int32_t row = 0;
  for (int32_t step = 0; step < max_steps; ++step) {
    V3 y;
    int32_t cls;
    const double radius = closest_boundary(g, x, y, cls);
    if (trace && row < max_rows) {
      trace[4 * row + 0] = x.x; trace[4 * row + 1] = x.y;
      trace[4 * row + 2] = x.z; trace[4 * row + 3] = radius;
      ++row;
    }
    if (!(radius > eps)) {
      out.a += payoff(s0, y, cls);
      if (s1) out.b += payoff(*s1, y, cls);
      if (rows) *rows = row;
      return out;
    }
    // Both sides of a paired path share one Green point (study.cpp:104-106).
    const V3 a = sampled ? x + stream.green(step, radius) : x;
    if (s0.source_kind) out.a += green_term(s0, x, a, radius, paired_form);
    if (s1 && s1->source_kind) out.b += green_term(*s1, x, a, radius, paired_form);
    const V3 d = stream.dir(step);
    if (step == 0) {
      out.w0 = d * (3.0 / radius);
      out.r0 = radius;
      out.s0 = out.a;
      out.s1 = out.b;
    }
    x = x + d * radius;
    ++out.steps;
  }
  if (trace && row < max_rows) {
    V3 y;
    int32_t cls;
    const double radius = closest_boundary(g, x, y, cls);
    trace[4 * row + 0] = x.x; trace[4 * row + 1] = x.y;
    trace[4 * row + 2] = x.z; trace[4 * row + 3] = radius;
    ++row;
  }
  if (rows) *rows = row;
  out.cap = true;
  out.caps = 1;
  return out;
}

// Per-walk key of evaluate_site: wih_twos.cpp:681-685.
uint64_t site_key(uint64_t seed, int32_t repeat, int64_t site, int64_t sample) {
  return splitmix64(seed + 0x2d358dccaa6c78a5ULL) ^
         splitmix64(uint64_t(repeat) + 0x8bb84b93962eacc9ULL) ^
         splitmix64(uint64_t(site) + 0x4b33a62ed433d4a3ULL) ^
         splitmix64(uint64_t(sample) + 0x4d5a2da51de1aa47ULL);
}

// Coupled finite-cavity sample of walk `w` at candidate site `s` (centre z):
// study.cpp:317-332.  The start point is uniform in the probe ball (centre c,
// radius R: the single TERM_UNIT row of sides[1]), drawn as cbrt(U) for the
// radial fraction and then random_unit_vector; both paths of side 0 start
// there on the same stream, the second with a cavity of radius
// g.cavity_radius at z appended after the geometry's own (`added`, :320).
// d = w (a - b) / (4 pi r) with w = (1 - rho^2)^2 35/8.
//   MT_FD: key = splitmix64(seed) ^ splitmix64(s + 734) ^ splitmix64(w + 8131)
//   (:322); the probe from mt19937_64(splitmix64(key)) (:323-324), each path
//   from walk()'s mt19937_64(splitmix64(key + 719)) (:326, wih_twos.cpp:638).
//   PHILOX: the side-0 key of (seed, s); the probe from counter block
//   (0xFFFFFFFF, w, 0, 0) (radial fraction, words 0 and 1) and the direction of
//   block (0xFFFFFFFF, w, 1, 0); each path is walk w of that key.
WalkOut coupled_sample(const og_geom &g, const og_side sides[2], const og_params &p, const V3 &z, int64_t s,
                       int64_t w) {
  const double *t = sides[1].terms[0];
  const V3 c{t[1], t[2], t[3]};
  const double R = t[4];
  double rad;
  V3 dir;
  uint64_t key = 0;
  uint32_t pkey[2] = {0u, 0u};
  if (p.rng == OG_RNG_MT_FD) {
    key = splitmix64(p.seed) ^ splitmix64(uint64_t(s) + 734) ^ splitmix64(uint64_t(w) + 8131);
    std::mt19937_64 rng(splitmix64(key));
    std::uniform_real_distribution<double> U(0.0, 1.0);
    rad = std::cbrt(U(rng));
    dir = random_unit_vector(rng);
  } else {
    philox_key(p.seed, s, 0, pkey);
    const uint32_t ctr[4] = {0xFFFFFFFFu, uint32_t(w), 0u, 0u};
    uint32_t b[4];
    philox4x32_10(ctr, pkey, b);
    rad = std::cbrt(u53(b[0], b[1]));
    dir = philox_direction(pkey, uint32_t(w), 0xFFFFFFFFu, 1u);
  }
  const V3 x = c + dir * (R * rad);
  const double weight = std::pow(1.0 - rad * rad, 2) * 35.0 / 8.0;
  std::vector<double> cav(g.cavities, g.cavities + 3 * g.n_cavities);
  cav.push_back(z.x);
  cav.push_back(z.y);
  cav.push_back(z.z);
  og_geom g2 = g;
  g2.n_cavities = g.n_cavities + 1;
  g2.cavities = cav.data();
  WalkOut a, b;
  if (p.rng == OG_RNG_MT_FD) {
    MtStream sa, sb;
    sa.rng.seed(splitmix64(key + 719));
    sb.rng.seed(splitmix64(key + 719));
    a = walk(g, sides[0], nullptr, x, p.eps, p.max_steps, sa, false, nullptr, 0, nullptr);
    b = walk(g2, sides[0], nullptr, x, p.eps, p.max_steps, sb, false, nullptr, 0, nullptr);
  } else {
    PhiloxStream sa{{pkey[0], pkey[1]}, uint32_t(w)};
    a = walk(g, sides[0], nullptr, x, p.eps, p.max_steps, sa, false, nullptr, 0, nullptr);
    b = walk(g2, sides[0], nullptr, x, p.eps, p.max_steps, sa, false, nullptr, 0, nullptr);
  }
  WalkOut o;
  o.a = o.b = weight * (a.a - b.a) / (4 * 3.141592653589793 * g.cavity_radius);
  o.steps = a.steps + b.steps;
  o.cap = a.cap || b.cap;
  o.caps = a.caps + b.caps;
  return o;
}

// One sample (a, b) of walk `w` at site `s`.
WalkOut sample(const og_geom &g, const og_side sides[2], const og_params &p, const V3 &z,
               int64_t s, int64_t w) {
  if (p.combine == OG_COUPLED) return coupled_sample(g, sides, p, z, s, w);
  if (p.rng == OG_RNG_PHILOX) {
    PhiloxStream st0{{0u, 0u}, uint32_t(w)};
    philox_key(p.seed, s, 0, st0.key);
    if (p.combine == OG_SQUARE) {
      WalkOut o = walk(g, sides[0], nullptr, z, p.eps, p.max_steps, st0, false, nullptr, 0, nullptr);
      o.b = o.a;
      o.ga = o.gb = grad_sample(o, o.a, o.s0, sides[0]);
      return o;
    }
    if (p.combine == OG_PAIRED) {
      WalkOut o = walk(g, sides[0], &sides[1], z, p.eps, p.max_steps, st0, false, nullptr, 0, nullptr);
      o.ga = grad_sample(o, o.a, o.s0, sides[0]);
      o.gb = grad_sample(o, o.b, o.s1, sides[1]);
      return o;
    }
    PhiloxStream st1{{0u, 0u}, uint32_t(w)};
    philox_key(p.seed, s, 1, st1.key);
    WalkOut a = walk(g, sides[0], nullptr, z, p.eps, p.max_steps, st0, false, nullptr, 0, nullptr);
    WalkOut b = walk(g, sides[1], nullptr, z, p.eps, p.max_steps, st1, false, nullptr, 0, nullptr);
    a.ga = grad_sample(a, a.a, a.s0, sides[0]);
    a.gb = grad_sample(b, b.a, b.s0, sides[1]);
    a.b = b.a;
    a.steps += b.steps;
    a.cap = a.cap || b.cap;
    a.caps += b.caps;
    return a;
  }
  MtStream st;
  if (p.rng == OG_RNG_MT_SITE) {
    // walk(): std::mt19937_64 rng(splitmix64(seed)) with evaluate_site's key.
    st.rng.seed(splitmix64(site_key(p.seed, p.repeat, s, w)));
  } else {
    // paired_site: rng(splitmix64(seed)^splitmix64(site+91347)^splitmix64(n+571)),
    // benchmarks/submission_study/study.cpp:89.
    st.rng.seed(splitmix64(p.seed) ^ splitmix64(uint64_t(s) + 91347) ^ splitmix64(uint64_t(w) + 571));
  }
  const bool paired_form = p.rng == OG_RNG_MT_PAIRED;
  if (p.combine == OG_PAIRED) {
    WalkOut o = walk(g, sides[0], &sides[1], z, p.eps, p.max_steps, st, paired_form, nullptr, 0, nullptr);
    o.ga = grad_sample(o, o.a, o.s0, sides[0]);
    o.gb = grad_sample(o, o.b, o.s1, sides[1]);
    return o;
  }
  WalkOut o = walk(g, sides[0], nullptr, z, p.eps, p.max_steps, st, paired_form, nullptr, 0, nullptr);
  o.b = o.a;
  o.ga = o.gb = grad_sample(o, o.a, o.s0, sides[0]);
  return o;
}

// Cross-index combine and leave-one-out jackknife over explicit samples,
// evaluated in the reference's order: wih_twos.cpp:687-715 (a == b) and
// study.cpp:114-119 (paired).  out = {u, v, score, score_se, same, n}.
void combine(const double *a, const double *b, int64_t count, bool square, double out[6]) {
  double sa = 0.0, sb = 0.0, sab = 0.0;
  for (int64_t i = 0; i < count; ++i) sa += a[i];
  if (square) {
    for (int64_t i = 0; i < count; ++i) sab += a[i] * a[i];
    sb = sa;
  } else {
    for (int64_t i = 0; i < count; ++i) sb += b[i];
    for (int64_t i = 0; i < count; ++i) sab += a[i] * b[i];
  }
  const double n = double(count);
  out[0] = sa / n;
  out[1] = sb / n;
  out[2] = (sa * sb - sab) / (n * (n - 1.0));
  out[4] = sab / n;
  out[5] = n;
  if (count >= 4) {
    std::vector<double> leave(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) {
      const double bi = square ? a[i] : b[i];
      const double la = sa - a[i];
      const double lb = sb - bi;
      const double lab = sab - a[i] * bi;
      leave[size_t(i)] = (la * lb - lab) / ((n - 1.0) * (n - 2.0));
    }
    double mean = 0.0;
    for (double x : leave) mean += x;
    mean /= n;
    double spread = 0.0;
    for (double x : leave) spread += (x - mean) * (x - mean);
    out[3] = std::sqrt((n - 1.0) / n * spread);
  } else {
    out[3] = std::numeric_limits<double>::quiet_NaN();
  }
}

// ---------------------------------------------------------------------------
// Differential WoS for cavity-shape derivatives (evaluate_differential_site,
// wih_twos.cpp:725-928; study.cpp:303-316 runs it over every point with
// repeat 0).  Mesh container with the reference's cavities only.

struct Hit {
  double distance = kInf;
  V3 point{0.0, 0.0, 0.0}, normal{0.0, 0.0, 0.0};
  int32_t cavity = -1;
};

// closest_boundary with the hit's outward normal and cavity index
// (wih_twos.cpp:602-626): the mesh first (normal towards the mesh point), then
// each cavity when strictly closer (normal into the cavity).
Hit boundary_hit(const og_geom &g, const V3 &p, int64_t *queries) {
  Hit h;
  int64_t tri;
  ++*queries;
  h.distance = g.mesh->closest(p, h.point, tri, false, nullptr, nullptr);
  if (g.mesh_radius_f32) h.distance = double(float(h.distance));
  if (h.distance > 0.0) h.normal = (h.point - p) * (1.0 / h.distance);
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const V3 c = load3(g.cavities + 3 * i);
    const V3 radial = p - c;
    const double len = norm(radial);
    const double dist = len - g.cavity_radius;
    if (dist < h.distance && len > 0.0) {
      const V3 dir = radial * (1.0 / len);
      h.distance = dist;
      h.point = c + dir * g.cavity_radius;
      h.normal = dir * -1.0;
      h.cavity = i;
    }
  }
  return h;
}

// domain_distance (wih_twos.cpp:585-593): one mesh query, then the cavities.
double domain_distance(const og_geom &g, const V3 &p, int64_t *queries) {
  V3 q;
  int64_t tri;
  ++*queries;
  double best = g.mesh->closest(p, q, tri, false, nullptr, nullptr);
  if (g.mesh_radius_f32) best = double(float(best));
  for (int32_t i = 0; i < g.n_cavities; ++i) best = std::min(best, norm(p - load3(g.cavities + 3 * i)) - g.cavity_radius);
  return best;
}

struct DiffCounts {
  int64_t outer_steps = 0, nested_steps = 0, cavity_terms = 0, near_seam = 0, segment_unsafe = 0, unsafe = 0,
          limits = 0, queries = 0, rays = 0;
};

// classify_offset_safety (wih_twos.cpp:764-833), with the enclosure fallback
// the study executable compiles in (study.cpp:8).  Returns (near_seam,
// segment_unsafe).
std::pair<bool, bool> offset_safety(const og_geom &g, int32_t hit, const V3 &bp, const V3 &off, double ell,
                                    double eps, DiffCounts &c) {
  bool near_seam = false, segment_unsafe = false;
  const double r = g.cavity_radius;
  const V3 seg = off - bp;
  const double seg2 = dot(seg, seg);
  const double tol = 1.0e-10 * std::max(1.0, std::fabs(r));
  const double coll2 = (r + tol) * (r + tol);
  const V3 hc = load3(g.cavities + 3 * hit);
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    if (i == hit) continue;
    const V3 ci = load3(g.cavities + 3 * i);
    const V3 delta = ci - hc;
    const double cd = norm(delta);
    if (cd > tol && cd < 2.0 * r) {
      // Distance from the hit point to the seam circle of two equal spheres.
      const V3 axis = delta * (1.0 / cd);
      const V3 seam_c = (hc + ci) * 0.5;
      const double seam_r = std::sqrt(std::max(0.0, r * r - 0.25 * (cd * cd)));
      const V3 rel = bp - seam_c;
      const double ax = dot(rel, axis);
      const V3 rad = rel - axis * ax;
      const double nr = norm(rad) - seam_r;
      near_seam = near_seam || std::sqrt(ax * ax + nr * nr) <= ell;
    }
    if (seg2 > 0.0) {
      const double t = std::clamp(dot(ci - bp, seg) / seg2, 0.0, 1.0);
      const V3 cl = bp + seg * t;
      if (dot(cl - ci, cl - ci) <= coll2) segment_unsafe = true;
    }
  }
  bool inside = g.mesh->inside(off, &c.rays) == 1;
  if (!inside && g.mesh->inside(hc, &c.rays) == 1) {
    V3 q;
    int64_t tri;
    ++c.queries;
    double dc = g.mesh->closest(hc, q, tri, false, nullptr, nullptr);
    if (g.mesh_radius_f32) dc = double(float(dc));
    if (dc > r + ell + eps && norm(off - hc) <= r + ell + eps * 0.5) inside = true;
  }
  if (!inside || !(domain_distance(g, off, &c.queries) > eps)) segment_unsafe = true;
  return {near_seam, segment_unsafe};
}

// One site: per cavity the mean and SE of the center and radius derivative
// samples (8 values: dc x y z, dc se x y z, dR, dR se).
void differential_site(const og_geom &g, const og_side &side, const og_params &p, const V3 &z, int64_t site,
                       double ell, double *out, DiffCounts &c) {
  const int32_t K = g.n_cavities;
  std::vector<V3> sums(size_t(K), V3{0.0, 0.0, 0.0}), sq(size_t(K), V3{0.0, 0.0, 0.0});
  std::vector<double> rs(size_t(K), 0.0), rsq(size_t(K), 0.0);
  // Philox streams (the production form; no reference counterpart): the
  // outer path of sample n draws direction `step` from counter (step, n, 0, 0)
  // under the key of (dseed, site, 0), the nested path is walk n of the key
  // of (dseed, site, 1), where dseed folds the reference's seed and repeat
  // terms of the outer key (wih_twos.cpp:867-870).
  const bool philox = p.rng == OG_RNG_PHILOX;
  uint32_t outer_key[2] = {0u, 0u}, nested_key[2] = {0u, 0u};
  if (philox) {
    const uint64_t dseed = splitmix64(p.seed + 0x6a09e667f3bcc909ULL) ^ splitmix64(uint64_t(p.repeat) + 0xbb67ae8584caa73bULL);
    philox_key(dseed, site, 0, outer_key);
    philox_key(dseed, site, 1, nested_key);
  }
  for (int64_t n = 0; n < p.walks; ++n) {
    const uint64_t key = splitmix64(p.seed + 0x6a09e667f3bcc909ULL) ^
                         splitmix64(uint64_t(p.repeat) + 0xbb67ae8584caa73bULL) ^
                         splitmix64(uint64_t(site) + 0x3c6ef372fe94f82bULL) ^
                         splitmix64(uint64_t(n) + 0xa54ff53a5f1d36f1ULL);
    // walk_to_boundary (wih_twos.cpp:725-740): no source, the last hit kept.
    std::mt19937_64 rng(splitmix64(key));
    V3 x = z;
    Hit h;
    bool capped = true;
    for (int32_t s = 0; s < p.max_steps; ++s) {
      h = boundary_hit(g, x, &c.queries);
      if (!(h.distance > p.eps)) {
        capped = false;
        break;
      }
      x = x + (philox ? philox_direction(outer_key, uint32_t(n), uint32_t(s)) : random_unit_vector(rng)) * h.distance;
      ++c.outer_steps;
    }
    if (capped) {
      ++c.limits;
      continue;
    }
    if (h.cavity < 0) continue;
    ++c.cavity_terms;
    const V3 off = h.point - h.normal * ell;
    const auto safety = offset_safety(g, h.cavity, h.point, off, ell, p.eps, c);
    c.near_seam += safety.first ? 1 : 0;
    c.segment_unsafe += safety.second ? 1 : 0;
    if (safety.first || safety.second) {
      ++c.unsafe;
      continue;
    }
    // The nested walk (walk(), wih_twos.cpp:635-664, keyed nested_key).
    int32_t rows = 0;
    WalkOut w;
    if (philox) {
      PhiloxStream st{{nested_key[0], nested_key[1]}, uint32_t(n)};
      w = walk(g, side, nullptr, off, p.eps, p.max_steps, st, false, nullptr, 0, &rows);
    } else {
      MtStream st;
      st.rng.seed(splitmix64(splitmix64(key ^ 0x510e527fade682d1ULL)));
      w = walk(g, side, nullptr, off, p.eps, p.max_steps, st, false, nullptr, 0, &rows);
    }
    c.nested_steps += w.steps;
    c.queries += w.steps + (w.cap ? 0 : 1);
    c.limits += w.cap ? 1 : 0;
    const double scale = w.a / ell;
    const V3 cs = h.normal * scale;
    const size_t k = size_t(h.cavity);
    sums[k] = sums[k] + cs;
    sq[k].x += cs.x * cs.x;
    sq[k].y += cs.y * cs.y;
    sq[k].z += cs.z * cs.z;
    rs[k] += -scale;
    rsq[k] += scale * scale;
  }
  const double N = double(p.walks);
  for (int32_t k = 0; k < K; ++k) {
    double *o = out + 8 * size_t(k);
    const V3 mean = sums[size_t(k)] * (1.0 / N);
    o[0] = mean.x; o[1] = mean.y; o[2] = mean.z;
    o[6] = rs[size_t(k)] / N;
    if (p.walks > 1) {
      const double den = N * (N - 1.0);
      const V3 &s = sums[size_t(k)], &q = sq[size_t(k)];
      o[3] = std::sqrt(std::max(0.0, (q.x - s.x * s.x / N) / den));
      o[4] = std::sqrt(std::max(0.0, (q.y - s.y * s.y / N) / den));
      o[5] = std::sqrt(std::max(0.0, (q.z - s.z * s.z / N) / den));
      o[7] = std::sqrt(std::max(0.0, (rsq[size_t(k)] - rs[size_t(k)] * rs[size_t(k)] / N) / den));
    } else {
      o[3] = o[4] = o[5] = o[7] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Certified minimum-benefit sampling (benchmarks/compiled_bunny/src/
// minimum_benefit.hpp): fail-closed walks whose every geometric and source
// quantity is rounded toward the safe side.  The distance is the union of the
// triangles' bounding boxes (a lower bound of the surface distance), each box
// distance rounded down; a sample adds a rounded-down lower bound of the
// source over each step's Green point and stops at the value cap.

double down_(double x) { return std::nextafter(x, -kInf); }
double up_(double x) { return std::nextafter(x, kInf); }
double positive_down(double x) { return std::max(0.0, down_(x)); }

// box_distance_lower (minimum_benefit.hpp:18-25).
double cert_box_lower(const double lo[3], const double hi[3], const V3 &x) {
  double sum = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double g = std::max({0.0, down_(lo[a] - comp(x, a)), down_(comp(x, a) - hi[a])});
    sum = positive_down(sum + positive_down(g * g));
  }
  return positive_down(std::sqrt(sum));
}

// point_distance_bound (:26-35).
double cert_point_bound(const V3 &a, const V3 &b, bool upper) {
  double sum = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = std::fabs(comp(a, k) - comp(b, k));
    d = upper ? up_(d) : positive_down(d);
    const double q = upper ? up_(d * d) : positive_down(d * d);
    sum = upper ? up_(sum + q) : positive_down(sum + q);
  }
  return upper ? up_(std::sqrt(sum)) : positive_down(std::sqrt(sum));
}

// The smallest rounded-down box distance over all triangles
// (CertDistanceTree::distance, :39-81).  The box bound is monotone under box
// containment, so pruning a node whose (exact union) box is already no
// closer than the best leaves the minimum unchanged: the result does not
// depend on the tree.
double cert_mesh_distance(const og_mesh &m, const V3 &p) {
  double best = kInf;
  int32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const og_mesh::Node &n = m.nodes[size_t(stack[--sp])];
    if (cert_box_lower(n.lo, n.hi, p) >= best) continue;
    if (n.right < 0) {
      for (int32_t i = 0; i < -n.right; ++i) {
        double lo[3], hi[3];
        m.bounds_of(m.order[size_t(n.left + i)], lo, hi);
        best = std::min(best, cert_box_lower(lo, hi, p));
      }
      continue;
    }
    stack[sp++] = n.right;
    stack[sp++] = n.left;
  }
  return best;
}

// cert_domain_distance (:83-88).
double cert_domain_distance(const og_geom &g, const V3 &p) {
  double d = cert_mesh_distance(*g.mesh, p);
  for (int32_t i = 0; i < g.n_cavities; ++i)
    d = std::min(d, down_(cert_point_bound(p, load3(g.cavities + 3 * i), false) - g.cavity_radius));
  return std::max(0.0, d);
}

// (1 - t/1024)^1024 <= exp(-t), every operation rounded down (:92-96).
double exp_negative_lower(double t_upper) {
  double a = std::max(0.0, down_(1.0 - up_(t_upper / 1024.0)));
  for (int i = 0; i < 10; ++i) a = positive_down(a * a);
  return a;
}

// cert_source_lower (:97-116): the affine source 1 + slope (y + 0.5)
// (slope >= 0) or the housing chips, as lower bounds.
double cert_source_lower(const V3 &p, int32_t source, double slope) {
  if (source == 0) return std::max(0.0, down_(1.0 + down_(slope * down_(p.y + 0.5))));
  static const double centers[3][3] = {{-.48, -.18, -.08}, {.43, .19, -.07}, {.02, -.31, -.02}};
  static const double widths[3][3] = {{.24, .22, .16}, {.21, .20, .15}, {.30, .16, .19}};
  static const double amplitude[3] = {5.2, 4.6, 2.0};
  double value = .08;
  for (int j = 0; j < 3; ++j) {
    double exponent = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double absdiff = up_(std::fabs(comp(p, a) - centers[j][a]));
      const double q = up_(absdiff / widths[j][a]);
      exponent = up_(exponent + up_(q * q));
    }
    value = down_(value + positive_down(amplitude[j] * exp_negative_lower(exponent)));
  }
  return std::max(0.0, value);
}

// Draws of a certified walk: the reference's mt19937_64 stream (the Beta draw,
// then the Green point's direction, then the move's direction, in that order)
// or a Philox stream (step k of sample j: the move's direction from counter
// (k, j, 0, 0), the Green point's direction from (k, j, 1, 0) and its Beta
// from (k, j, 2, 0), as the estimator's walks draw them).
struct CertMt {
  std::mt19937_64 rng;
  V3 green(int32_t, double s) {
    const double t = beta_22(rng);
    return random_unit_vector(rng) * (s * t);
  }
  V3 dir(int32_t) { return random_unit_vector(rng); }
};

// cert_walk (:122-143).  flags: 1 capped, 2 clipped, 4 containment failure
// (the walk stops there; its record must not be used).
template <class Stream>
double cert_walk(const og_geom &g, const og_cert_params &p, V3 point, Stream &stream, int32_t &steps,
                 int32_t &flags) {
  double value = 0.0;
  steps = 0;
  flags = 0;
  for (int32_t k = 0; k < p.max_steps; ++k) {
    const double d = cert_domain_distance(g, point);
    if (d <= p.eps) return value;
    const double s = std::min(p.radius_cap, positive_down(0.95 * d));
    V3 y = point;
    // random_unit_vector(rng) * (s * beta_22(rng)): the Beta draw first.
    if (p.source == 1) y = point + stream.green(k, s);
    const V3 next = point + stream.dir(k) * s;
    if (!(cert_point_bound(point, next, true) < d) || !(cert_point_bound(point, y, true) < d)) {
      flags |= 4;
      return value;
    }
    const double weight = positive_down(positive_down(s * s) / 6.0);
    const double increment = positive_down(weight * cert_source_lower(y, p.source, p.slope));
    value = positive_down(value + increment);
    ++steps;
    if (value >= p.value_cap) {
      flags |= 2;
      return p.value_cap;
    }
    point = next;
  }
  flags |= 1;
  return value;
}

}  // namespace

extern "C" {

int32_t og_differential(const og_geom *g, const og_side *side, const og_params *prm, const double *sites, int64_t n,
                        double ell, double *out, int64_t *counts) {
  if (!g || g->container != OG_MESH || !g->mesh || prm->walks < 1 || !(ell > 0.0) ||
      (prm->rng != OG_RNG_MT_DIFF && prm->rng != OG_RNG_PHILOX))
    return 2;
  const og_params p = *prm;
  std::vector<DiffCounts> per(size_t(std::max<int64_t>(n, 1)));
#ifdef _OPENMP
  if (p.threads > 0) omp_set_num_threads(p.threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int64_t s = 0; s < n; ++s)
    differential_site(*g, *side, p, load3(sites + 3 * s), s, ell, out + 8 * size_t(g->n_cavities) * size_t(s),
                      per[size_t(s)]);
  for (int64_t s = 0; s < n; ++s) {
    const DiffCounts &c = per[size_t(s)];
    const int64_t row[9] = {c.outer_steps, c.nested_steps, c.cavity_terms, c.near_seam, c.segment_unsafe, c.unsafe,
                            c.limits, c.queries, c.rays};
    for (int k = 0; k < 9; ++k) counts[9 * s + k] = row[k];
  }
  return 0;
}

int32_t og_certify(const og_geom *g, const og_cert_params *prm, const double *sites, int64_t n, double *value,
                   int32_t *steps, int32_t *flags) {
  if (!g || g->container != OG_MESH || !g->mesh || prm->walks < 1 || prm->max_steps < 1 ||
      !(prm->radius_cap > 0.0) || !(prm->value_cap > 0.0) || (prm->source != 0 && prm->source != 1) ||
      (prm->source == 0 && prm->slope < 0.0) || (prm->rng != OG_RNG_PHILOX && prm->rng != OG_RNG_MT_KEY) ||
      (prm->rng == OG_RNG_PHILOX && prm->walks > 0x7fffffff))
    return 2;
  const og_cert_params p = *prm;
  const int64_t jobs = n * int64_t(p.walks);
#ifdef _OPENMP
  if (p.threads > 0) omp_set_num_threads(p.threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t t = 0; t < jobs; ++t) {
    const int64_t site = t / p.walks, j = t - site * p.walks;
    if (p.rng == OG_RNG_PHILOX) {
      // Philox: the key of (splitmix64(seed + 0xa1224), site, 0), sample j
      // in the counter.
      PhiloxStream st{{0u, 0u}, uint32_t(j)};
      philox_key(splitmix64(p.seed + 0xa1224ULL), p.site_offset + site, 0, st.key);
      value[t] = cert_walk(*g, p, load3(sites + 3 * site), st, steps[t], flags[t]);
    } else {
      // Key of sample j at site (minimum_benefit.hpp:168).
      const uint64_t key = splitmix64(p.seed) ^ splitmix64(uint64_t(p.site_offset + site) + 0xa1224ULL) ^
                           splitmix64(uint64_t(j) + 0x6242cULL);
      CertMt st{std::mt19937_64(splitmix64(key))};
      value[t] = cert_walk(*g, p, load3(sites + 3 * site), st, steps[t], flags[t]);
    }
  }
  return 0;
}

int32_t og_cert_distance(const og_geom *g, const double *points, int64_t n, double *out) {
  if (!g || g->container != OG_MESH || !g->mesh) return 2;
  for (int64_t i = 0; i < n; ++i) out[i] = cert_domain_distance(*g, load3(points + 3 * i));
  return 0;
}

uint64_t og_splitmix64(uint64_t x) { return splitmix64(x); }

void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}

void og_philox_key(uint64_t seed, int64_t site, int32_t side, uint32_t key[2]) {
  philox_key(seed, site, side, key);
}

void og_mt19937_64(uint64_t seed, uint64_t *out, int32_t n) {
  std::mt19937_64 rng(seed);
  for (int32_t i = 0; i < n; ++i) out[i] = rng();
}

void og_mt_uniforms(uint64_t seed, double *out, int32_t n) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> uniform(0.0, 1.0);
  for (int32_t i = 0; i < n; ++i) out[i] = uniform(rng);
}

og_mesh *og_mesh_create(const double *vertices, int64_t nv, const uint32_t *faces, int64_t nf) {
  if (nv <= 0 || nf <= 0) return nullptr;
  auto *m = new og_mesh;
  m->v.resize(size_t(nv));
  for (int64_t i = 0; i < nv; ++i) m->v[size_t(i)] = load3(vertices + 3 * i);
  m->f.resize(size_t(nf));
  for (int64_t i = 0; i < nf; ++i) {
    for (int k = 0; k < 3; ++k) {
      if (faces[3 * i + k] >= uint64_t(nv)) {
        delete m;
        return nullptr;
      }
      m->f[size_t(i)][size_t(k)] = faces[3 * i + k];
    }
  }
  m->order.resize(size_t(nf));
  for (int64_t i = 0; i < nf; ++i) m->order[size_t(i)] = uint32_t(i);
  m->nodes.reserve(size_t(nf));
  m->build(0, uint32_t(nf));
  return m;
}

void og_mesh_destroy(og_mesh *mesh) { delete mesh; }

double og_mesh_closest(const og_mesh *mesh, const double p[3], double out_point[3], int64_t *out_tri,
                       int32_t brute_force, int64_t *nodes_visited, int64_t *tris_tested) {
  V3 q;
  int64_t tri;
  const double d = mesh->closest(load3(p), q, tri, brute_force != 0, nodes_visited, tris_tested);
  if (out_point) { out_point[0] = q.x; out_point[1] = q.y; out_point[2] = q.z; }
  if (out_tri) *out_tri = tri;
  return d;
}

void og_closest_point_triangle(const double p[3], const double a[3], const double b[3],
                               const double c[3], double out[3]) {
  const V3 q = closest_point_triangle(load3(p), load3(a), load3(b), load3(c));
  out[0] = q.x; out[1] = q.y; out[2] = q.z;
}

int32_t og_mesh_inside(const og_mesh *mesh, const double p[3]) { return mesh->inside(load3(p)); }

int32_t og_mesh_inside_many(const og_mesh *mesh, const double *p, int64_t n, int32_t *out,
                            int32_t threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t i = 0; i < n; ++i) out[i] = mesh->inside(load3(p + 3 * i));
  return 0;
}

double og_closest_boundary(const og_geom *g, const double p[3], double out_point[3], int32_t *out_cls) {
  V3 q;
  int32_t cls = 0;
  const double d = closest_boundary(*g, load3(p), q, cls);
  if (out_point) { out_point[0] = q.x; out_point[1] = q.y; out_point[2] = q.z; }
  if (out_cls) *out_cls = cls;
  return d;
}

void og_combine(const double *a, const double *b, int64_t n, int32_t square, double out[6]) {
  combine(a, b, n, square != 0, out);
}

// Exact leave-one-out jackknife from eight binary64 power sums
// {Sa, Sb, Sab, Saa, Sbb, Saab, Sabb, Saabb}.  The leave-one-out estimate is
// (C - Sa b_i - Sb a_i + 2 a_i b_i) / ((n-1)(n-2)), C = Sa Sb - Sab, which is
// linear in (a_i, b_i, a_i b_i); its spread needs only those second moments.
void og_combine_powersums(const double s[8], int64_t count, double out[6]) {
  const double n = double(count);
  out[0] = s[0] / n;
  out[1] = s[1] / n;
  out[2] = (s[0] * s[1] - s[2]) / (n * (n - 1.0));
  out[4] = s[2] / n;
  out[5] = n;
  if (count < 4) {
    out[3] = std::numeric_limits<double>::quiet_NaN();
    return;
  }
  const double sa = s[0], sb = s[1];
  const double z2 = sa * sa * s[4] + sb * sb * s[3] + 4.0 * s[7] + 2.0 * sa * sb * s[2] -
                    4.0 * sa * s[6] - 4.0 * sb * s[5];
  const double z1 = 2.0 * (s[2] - sa * sb);
  const double d = (n - 1.0) * (n - 2.0);
  const double spread = std::max(0.0, z2 - z1 * z1 / n) / (d * d);
  out[3] = std::sqrt((n - 1.0) / n * spread);
}

// Gradient channel of one site from its 28 sums (A = sum a, B = sum b,
// C = sum a.b, sum a a^T as xx yy zz xy xz yz, sum b b^T, sum a b^T row-major):
// grad_dot = (A.B - C)/(N(N-1)) over pairs of distinct walks, SE to first order
// sqrt((b^T S_aa b + a^T S_bb a + 2 b^T S_ab a)/N) with a, b the means and S
// the sample covariances.
void grad_estimate(const double *g, int64_t n, og_site &r) {
  const double N = double(n);
  double a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = g[k] / N;
    b[k] = g[3 + k] / N;
    r.grad_u[k] = a[k];
    r.grad_v[k] = b[k];
  }
  const double ab = g[0] * g[3] + g[1] * g[4] + g[2] * g[5];
  r.grad_dot = n > 1 ? (ab - g[6]) / (N * (N - 1.0)) : std::nan("");
  if (n < 4) {
    r.grad_dot_se = std::nan("");
    return;
  }
  const int ix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  double v = 0.0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      const double saa = (g[7 + ix[p][q]] - N * a[p] * a[q]) / (N - 1.0);
      const double sbb = (g[13 + ix[p][q]] - N * b[p] * b[q]) / (N - 1.0);
      const double sab = (g[19 + 3 * p + q] - N * a[p] * b[q]) / (N - 1.0);
      v += b[p] * saa * b[q] + a[p] * sbb * a[q] + 2.0 * b[p] * sab * a[q];
    }
  r.grad_dot_se = std::sqrt(std::max(v, 0.0) / N);
}

void grad_accumulate(const V3 &ga, const V3 &gb, double *s) {
  const double a[3] = {ga.x, ga.y, ga.z}, b[3] = {gb.x, gb.y, gb.z};
  for (int q = 0; q < 3; ++q) {
    s[q] += a[q];
    s[3 + q] += b[q];
  }
  s[6] += a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  s[7] += a[0] * a[0]; s[8] += a[1] * a[1]; s[9] += a[2] * a[2];
  s[10] += a[0] * a[1]; s[11] += a[0] * a[2]; s[12] += a[1] * a[2];
  s[13] += b[0] * b[0]; s[14] += b[1] * b[1]; s[15] += b[2] * b[2];
  s[16] += b[0] * b[1]; s[17] += b[0] * b[2]; s[18] += b[1] * b[2];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) s[19 + 3 * p + q] += a[p] * b[q];
}

int32_t og_estimate(const og_geom *g, const og_side sides[2], const og_params *prm,
                    const double *sites, int64_t n, og_site *out, double *dump_a, double *dump_b,
                    int32_t *dump_steps, int64_t dump_stride) {
  const og_params p = *prm;
  if (p.walks < 2 || p.max_steps < 1 || !(p.eps > 0.0) || n < 0) return 2;
  if (p.rng != OG_RNG_PHILOX && p.combine == OG_CROSS) return 2;
  if (p.rng == OG_RNG_MT_KEY || p.rng < 0 || p.rng > OG_RNG_MT_FD) return 2;
  if ((p.rng == OG_RNG_MT_FD) != (p.combine == OG_COUPLED) && p.rng != OG_RNG_PHILOX) return 2;
  if (p.combine == OG_COUPLED &&
      (p.gradient || !(g->cavity_radius > 0.0) || sides[1].source_kind < 2 || sides[1].n_terms != 1 ||
       sides[1].terms[0][0] != OG_TERM_UNIT || !(sides[1].terms[0][4] > 0.0)))
    return 2;
  if (p.gradient && (sides[0].source_kind >= 2 || sides[1].source_kind >= 2)) return 2;
  for (int k = 0; k < 2; ++k)
    if (sides[k].source_kind == 3)
      for (int32_t i = 0; i < sides[k].n_terms; ++i)
        if (sides[k].terms[i][0] == OG_TERM_GAUSS) return 2;
  const int32_t walks_max = std::max(p.walks, p.walks_max);
  int32_t status = 0;
#ifdef _OPENMP
  if (p.threads > 0) omp_set_num_threads(p.threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int64_t s = 0; s < n; ++s) {
    const V3 z = load3(sites + 3 * s);
    std::vector<double> a, b;
    a.reserve(size_t(walks_max));
    b.reserve(size_t(walks_max));
    og_site r{};
    int64_t steps = 0, caps = 0;
    int64_t done = 0, target = p.walks;
    double c[6];
    double gs[28] = {0.0};
    while (true) {
      for (int64_t w = done; w < target; ++w) {
        const WalkOut o = sample(*g, sides, p, z, s, w);
        a.push_back(o.a);
        b.push_back(o.b);
        if (p.gradient) grad_accumulate(o.ga, o.gb, gs);
        steps += o.steps;
        caps += o.caps;
        if (dump_a && w < dump_stride) {
          dump_a[s * dump_stride + w] = o.a;
          dump_b[s * dump_stride + w] = o.b;
          dump_steps[s * dump_stride + w] = o.steps | (o.cap ? int32_t(1u << 30) : 0);
        }
      }
      done = target;
      if (p.combine == OG_COUPLED) {
        // v.up and v.same of study.cpp:329-331: the mean of d and its SE.
        double up = 0.0, sum2 = 0.0;
        for (int64_t i = 0; i < done; ++i) {
          up += a[size_t(i)];
          sum2 += a[size_t(i)] * a[size_t(i)];
        }
        const double nn = double(done);
        up /= nn;
        c[0] = c[1] = c[2] = up;
        c[3] = std::sqrt(std::max(0.0, (sum2 - nn * up * up) / (nn * (nn - 1.0))));
        c[4] = sum2 / nn;
        c[5] = nn;
      } else {
        combine(a.data(), b.data(), done, p.combine == OG_SQUARE, c);
      }
      const bool met = std::isfinite(c[3]) && p.ci_z * c[3] <= p.ci_rel * std::fabs(c[2]);
      if (walks_max <= p.walks || met) {
        r.status = 0;
        break;
      }
      if (done >= walks_max) {
        r.status = 1;
        break;
      }
      target = std::min<int64_t>(2 * done, walks_max);
    }
    r.u = c[0]; r.v = c[1]; r.score = c[2]; r.score_se = c[3];
    r.td = -c[2]; r.same = c[4];
    r.walks = done; r.steps = steps; r.caps = caps;
    if (p.gradient) {
      grad_estimate(gs, done, r);
    } else {
      for (int q = 0; q < 3; ++q) r.grad_u[q] = r.grad_v[q] = std::nan("");
      r.grad_dot = r.grad_dot_se = std::nan("");
    }
    out[s] = r;
  }
  for (int64_t s = 0; s < n; ++s)
    if (out[s].caps) status = 3;
  return status;
}

double og_bump_green_ball(const double x[3], double radius, const double c[3], double support, double amplitude) {
  return bump_green_ball(load3(x), radius, load3(c), support, amplitude);
}

void og_combine_grad(const double *ga, const double *gb, int64_t n, double out[8]) {
  double s[28] = {0.0};
  for (int64_t i = 0; i < n; ++i) grad_accumulate(load3(ga + 3 * i), load3(gb + 3 * i), s);
  og_site r{};
  grad_estimate(s, n, r);
  for (int k = 0; k < 3; ++k) {
    out[k] = r.grad_u[k];
    out[3 + k] = r.grad_v[k];
  }
  out[6] = r.grad_dot;
  out[7] = r.grad_dot_se;
}

int32_t og_trace(const og_geom *g, const og_side *side, const og_params *prm, const double start[3],
                 uint64_t key, int64_t site, int32_t w, double *rows, int32_t max_rows) {
  int32_t count = 0;
  const og_params &p = *prm;
  if (p.rng == OG_RNG_PHILOX) {
    PhiloxStream st{{0u, 0u}, uint32_t(w)};
    philox_key(p.seed, site, 0, st.key);
    walk(*g, *side, nullptr, load3(start), p.eps, p.max_steps, st, false, rows, max_rows, &count);
    return count;
  }
  MtStream st;
  if (p.rng == OG_RNG_MT_KEY) st.rng.seed(splitmix64(key));
  else if (p.rng == OG_RNG_MT_SITE) st.rng.seed(splitmix64(site_key(p.seed, p.repeat, site, w)));
  else st.rng.seed(splitmix64(p.seed) ^ splitmix64(uint64_t(site) + 91347) ^ splitmix64(uint64_t(w) + 571));
  walk(*g, *side, nullptr, load3(start), p.eps, p.max_steps, st, p.rng == OG_RNG_MT_PAIRED, rows, max_rows,
       &count);
  return count;
}

}  // extern "C"
