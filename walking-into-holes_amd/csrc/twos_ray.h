// Ray queries on the binary LBVH: the ray-parity inside test of
// EmbreeMesh::inside (benchmarks/compiled_bunny/src/wih_twos.cpp:361-394).
#pragma once

#include "twos_device.h"

namespace twos {

// Ray parity inside test: closest hit along a fixed irrational direction,
// step 2e-6 past it, count crossings up to 128 (EmbreeMesh::inside,
// wih_twos.cpp:361-394).  Binary64 Moller-Trumbore against outward-rounded
// binary32 boxes.
__device__ inline bool ray_hit(const MeshView &m, V3<double> o, V3<double> d, double tnear, double tfar, double &t_out) {
  double best = tfar;
  bool found = false;
  int32_t stack[kStackDepth];
  int sp = 0;
  int32_t node = 0;
  if (m.n_tris == 1) node = ~0;
  int32_t visits = 2 * m.n_tris + 8;
  // Slab test with the reciprocal direction (one division per axis per ray,
  // not per box).  The boxes are outward-rounded and the slab interval is
  // padded by 1e-12, far above the reciprocal's rounding (about 1e-14 at
  // t = 30), so no box that holds a hit is culled and hits are unchanged.
  const V3<double> inv{d.x != 0.0 ? 1.0 / d.x : 0.0, d.y != 0.0 ? 1.0 / d.y : 0.0, d.z != 0.0 ? 1.0 / d.z : 0.0};
  while (--visits >= 0) {
    if (node >= 0) {
      const float4 *nn = m.nodes + 4 * size_t(node);
      const float4 q0 = gload(nn), q1 = gload(nn + 1), q2 = gload(nn + 2), q3 = gload(nn + 3);
      const float bx[2][6] = {{q0.x, q0.y, q0.z, q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w}};
      const int32_t ch[2] = {__float_as_int(q3.x), __float_as_int(q3.y)};
      bool hit[2];
      for (int c = 0; c < 2; ++c) {
        double t0 = tnear, t1 = best;
        bool miss = false;
        for (int k = 0; k < 3 && !miss; ++k) {
          const double ok = comp(o, k), dk = comp(d, k), ik = comp(inv, k);
          const double lo = bx[c][k], hi = bx[c][3 + k];
          if (dk == 0.0) {
            if (ok < lo || ok > hi) miss = true;
            continue;
          }
          double ta = (lo - ok) * ik, tb = (hi - ok) * ik;
          if (ta > tb) { const double s = ta; ta = tb; tb = s; }
          t0 = fmax(t0, ta - 1e-12);
          t1 = fmin(t1, tb + 1e-12);
          if (t0 > t1) miss = true;
        }
        hit[c] = !miss;
      }
      if (hit[0] && hit[1]) {
        if (sp < kStackDepth) stack[sp++] = ch[1];
        node = ch[0];
        continue;
      }
      if (hit[0]) { node = ch[0]; continue; }
      if (hit[1]) { node = ch[1]; continue; }
    } else {
      V3<double> a, b, c;
      load_tri<double>(m, ~node, a, b, c);
      const V3<double> e1 = b - a, e2 = c - a;
      const V3<double> pv{d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x};
      const double det = dot(e1, pv);
      if (det != 0.0) {
        const double inv = 1.0 / det;
        const V3<double> tv = o - a;
        const double uu = dot(tv, pv) * inv;
        if (uu >= 0.0 && uu <= 1.0) {
          const V3<double> qv{tv.y * e1.z - tv.z * e1.y, tv.z * e1.x - tv.x * e1.z, tv.x * e1.y - tv.y * e1.x};
          const double vv = dot(d, qv) * inv;
          if (vv >= 0.0 && uu + vv <= 1.0) {
            const double t = dot(e2, qv) * inv;
            if (t >= tnear && t <= best) {
              best = t;
              found = true;
            }
          }
        }
      }
    }
    if (sp == 0) break;
    node = stack[--sp];
  }
  t_out = best;
  return found;
}

// Crossings along the fixed direction (1, 0.3713906763541037,
// 0.1570796326794897), normalised: each hit moves the origin 2e-6 past it, at
// most 128 crossings.  Returns the parity (1 inside), -1 past the crossing
// limit; *rays counts the ray queries made (one per hit plus the final miss).
__device__ inline int32_t inside_parity(const MeshView &m, V3<double> o, int32_t *rays) {
  V3<double> dir{1.0, 0.3713906763541037, 0.1570796326794897};
  const double len = sqrt(dot(dir, dir));
  dir = dir * (1.0 / len);
  unsigned hits = 0;
  for (; hits < 128; ++hits) {
    double t;
    ++*rays;
    if (!ray_hit(m, o, dir, 1.0e-6, 4.0, t)) break;
    o = o + dir * (t + 2.0e-6);
  }
  return hits == 128 ? -1 : int32_t(hits & 1u);
}

}  // namespace twos
