// Differential Walk on Spheres, binary64: the reference's mt19937_64
// streams (diff_walk, one thread per job), the Philox binary64 form of
// diff_persistent, and the per-(site, cavity) fold.  The binary32 Philox
// form is built in twos_diff_f32.hip (contraction on).
//
// Keys of the reference's streams (wih_twos.cpp:867-872, 895): key =
// splitmix64(seed + 0x6a09e667f3bcc909) ^ splitmix64(repeat + 0xbb67ae8584caa73b)
// ^ splitmix64(site + 0x3c6ef372fe94f82b) ^ splitmix64(sample + 0xa54ff53a5f1d36f1);
// the outer path draws from mt19937_64(splitmix64(key)), the nested path from
// mt19937_64(splitmix64(splitmix64(key ^ 0x510e527fade682d1))).

#include "twos_diff.h"
#include "twos_launch.h"

namespace twos {

namespace {

struct Hit {
  double distance;
  V3<double> point, normal;
  int32_t cavity;
};

// closest_boundary with the hit's outward normal and cavity index
// (wih_twos.cpp:602-626): the mesh first (normal towards the mesh point),
// then each cavity when strictly closer (normal into the cavity).
__device__ __forceinline__ Hit boundary_hit(const DevGeom &g, V3<double> p, int32_t &hint) {
  Hit h;
  const MeshView m{g.nodes, g.nodes4, g.tris_f, g.tris_d, g.tri_ids, g.n_tris};
  h.distance = mesh_distance<double>(g, mesh_closest<double, false>(m, p, h.point, hint, nullptr));
  h.normal = h.distance > 0.0 ? (h.point - p) * (1.0 / h.distance) : V3<double>{0.0, 0.0, 0.0};
  h.cavity = -1;
  const double r = g.cavity_radius;
  for (int32_t i = 0; i < g.n_cavities; ++i) {
    const Vec4<double> s = load4<double>(g, 2, i);
    const V3<double> c{s.x, s.y, s.z};
    const V3<double> radial = p - c;
    const double len = sqrt(dot(radial, radial));
    const double dist = len - r;
    if (dist < h.distance && len > 0.0) {
      const V3<double> dir = radial * (1.0 / len);
      h.distance = dist;
      h.point = c + dir * r;
      h.normal = dir * -1.0;
      h.cavity = i;
    }
  }
  return h;
}

__device__ __forceinline__ uint64_t diff_key(uint64_t seed, int32_t repeat, int64_t site, int64_t sample) {
  return splitmix64(seed + 0x6a09e667f3bcc909ULL) ^ splitmix64(uint64_t(repeat) + 0xbb67ae8584caa73bULL) ^
         splitmix64(uint64_t(site) + 0x3c6ef372fe94f82bULL) ^ splitmix64(uint64_t(sample) + 0xa54ff53a5f1d36f1ULL);
}

// One job: site = site_base + job / walks, sample = job % walks.  Record:
// val[4 job + 0..2] = the outward normal n, val[4 job + 3] = V / ell;
// rec[6 job + 0..5] = cavity (-1: no sample), flags, outer steps, nested
// steps, distance queries, ray queries.
__global__ __launch_bounds__(64) void diff_walk(const DevGeom *geom, const DevSide *side, DiffParams p) {
  const int64_t job = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (job >= p.jobs) return;
  const DevGeom &g = *geom;
  const DevSide &s = *side;
  const int64_t local = job / p.walks;
  const int64_t sample = job - local * p.walks;
  const int64_t site = p.site_base + local;
  const uint64_t key = diff_key(p.seed, p.repeat, p.site_offset + site, sample);
  V3<double> x{p.sites[3 * site], p.sites[3 * site + 1], p.sites[3 * site + 2]};
  int32_t hint = -1, flags = 0, outer = 0, nested = 0, queries = 0, rays = 0, cavity = -1;
  double value = 0.0;
  Hit h;
  Mt64 mt;
  mt.seed(splitmix64(key));
  bool capped = true;
  for (int32_t step = 0; step < p.max_steps; ++step) {
    h = boundary_hit(g, x, hint);
    ++queries;
    if (!(h.distance > p.eps)) {
      capped = false;
      break;
    }
    x = x + mt.direction() * h.distance;
    ++outer;
  }
  if (capped) {
    flags |= kDiffLimit;
  } else if (h.cavity >= 0) {
    flags |= kDiffCavity;
    const V3<double> off = h.point - h.normal * p.ell;
    const int32_t safety = offset_safety(g, h.cavity, h.point, off, p.ell, p.eps, queries, rays, hint);
    flags |= safety;
    if (safety) {
      flags |= kDiffUnsafe;
    } else {
      // The nested path from the offset point (walk(), wih_twos.cpp:644-663):
      // the cap check precedes the query, a terminated path made steps + 1
      // queries.  A non-finite distance ends the path as capped.
      mt.seed(splitmix64(splitmix64(key ^ 0x510e527fade682d1ULL)));
      const bool sampled = s.source_kind == 2;
      V3<double> y = off;
      bool ncap = true;
      for (int32_t step = 0; step < p.max_steps; ++step) {
        V3<double> bp;
        int cls;
        const double r = closest_boundary<double, kGeomMesh, false>(g, y, bp, cls, hint, nullptr);
        ++queries;
        if (!isfinite(r)) break;
        if (!(r > p.eps)) {
          value += side_payoff<double>(s, bp, cls);
          ncap = false;
          break;
        }
        if (s.source_kind) value += side_green_term<double>(s, y, sampled ? y + mt.green(r) : y, r, false);
        y = y + mt.direction() * r;
        ++nested;
      }
      if (ncap) flags |= kDiffLimit;
      cavity = h.cavity;
      value = value / p.ell;
    }
  }
  diff_record<double>(p, job, h.normal, value, cavity, flags, outer, nested, queries, rays);
}

// One thread per (site, cavity) of the chunk (one per site when there are no
// cavities): the mean and SE of the centre and radius samples of that cavity,
// in sample order (out[8] = dc x y z, their SEs, dR, its SE), and for cavity
// 0 the site's counters.
__global__ __launch_bounds__(64) void diff_fold(DiffParams p, int32_t n_sites, int32_t n_cav) {
  const int32_t width = n_cav > 0 ? n_cav : 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n_sites) * width) return;
  const int32_t local = int32_t(t / width), k = int32_t(t - int64_t(local) * width);
  const int64_t first = int64_t(local) * p.walks;
  double sx = 0.0, sy = 0.0, sz = 0.0, qx = 0.0, qy = 0.0, qz = 0.0, rs = 0.0, rq = 0.0;
  long long c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t n = 0; n < p.walks; ++n) {
    const int32_t *r = p.rec + 6 * (first + n);
    if (k == 0) {
      const int32_t f = r[1];
      c[0] += r[2];
      c[1] += r[3];
      c[2] += (f & kDiffCavity) ? 1 : 0;
      c[3] += (f & kDiffSeam) ? 1 : 0;
      c[4] += (f & kDiffSegment) ? 1 : 0;
      c[5] += (f & kDiffUnsafe) ? 1 : 0;
      c[6] += (f & kDiffLimit) ? 1 : 0;
      c[7] += r[4];
      c[8] += r[5];
    }
    if (r[0] != k) continue;
    const double *v = p.val + 4 * (first + n);
    const double scale = v[3];
    const double cx = v[0] * scale, cy = v[1] * scale, cz = v[2] * scale;
    sx += cx;
    sy += cy;
    sz += cz;
    qx += cx * cx;
    qy += cy * cy;
    qz += cz * cz;
    rs += -scale;
    rq += scale * scale;
  }
  if (k == 0) {
    long long *cnt = reinterpret_cast<long long *>(p.counts) + 9 * size_t(p.site_base + local);
    for (int i = 0; i < 9; ++i) cnt[i] = c[i];
  }
  if (k >= n_cav) return;
  const double N = double(p.walks);
  double *o = p.out + 8 * (size_t(p.site_base + local) * n_cav + k);
  o[0] = sx * (1.0 / N);
  o[1] = sy * (1.0 / N);
  o[2] = sz * (1.0 / N);
  o[6] = rs / N;
  if (p.walks > 1) {
    const double den = N * (N - 1.0);
    o[3] = sqrt(fmax(0.0, (qx - sx * sx / N) / den));
    o[4] = sqrt(fmax(0.0, (qy - sy * sy / N) / den));
    o[5] = sqrt(fmax(0.0, (qz - sz * sz / N) / den));
    o[7] = sqrt(fmax(0.0, (rq - rs * rs / N) / den));
  } else {
    o[3] = o[4] = o[5] = o[7] = 0.0;
  }
}

inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

hipError_t launch_differential(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int32_t n_sites,
                               int32_t n_cav, hipStream_t stream) {
  if (p.jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(diff_walk, dim3(blocks_for(p.jobs, 64)), dim3(64), 0, stream, d_geom, d_side, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_diff_fold(p, n_sites, n_cav, stream);
}

hipError_t launch_diff_fold(const DiffParams &p, int32_t n_sites, int32_t n_cav, hipStream_t stream) {
  hipLaunchKernelGGL(diff_fold, dim3(blocks_for(int64_t(n_sites) * (n_cav > 0 ? n_cav : 1), 64)), dim3(64), 0,
                     stream, p, n_sites, n_cav);
  return hipGetLastError();
}

hipError_t launch_diff_persistent_f64(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int cus,
                                      hipStream_t stream) {
  return launch_diff_persistent<double>(d_geom, d_side, p, cus, stream);
}

}  // namespace twos
