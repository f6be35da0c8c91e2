// Host AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle --
// TEST INFRASTRUCTURE ONLY (built by `make -C oracle sanitize`, run by
// tests/test_oracle_sanitize.py).  Exercises every entry point on small
// inputs: a closed octahedron mesh with a cavity, the ball, both stream
// families, the coupled, differential and certified estimators.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "twos_oracle.h"

namespace {

og_side heat_side() {
  og_side s;
  std::memset(&s, 0, sizeof s);
  s.source_kind = 2;
  s.n_terms = 2;
  const double rows[2][8] = {{OG_TERM_BUMP, 0.2, 0.0, 0.1, 0.3, 0, 0, 2.0}, {OG_TERM_GAUSS, -0.1, 0.1, 0, 0.2, 0.3, 0.25, 1.5}};
  std::memcpy(s.terms, rows, sizeof rows);
  s.cls_const[0] = 1.0;
  return s;
}

og_params params(int combine, int rng, int walks) {
  og_params p;
  std::memset(&p, 0, sizeof p);
  p.combine = combine;
  p.rng = rng;
  p.seed = 7;
  p.max_steps = 256;
  p.eps = 1e-3;
  p.walks = walks;
  p.ci_rel = 0.02;
  p.ci_z = 1.96;
  p.threads = 2;
  return p;
}

}  // namespace

int main() {
  // Octahedron |x| + |y| + |z| <= 1 (outward faces).
  const double v[18] = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  const uint32_t f[24] = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  og_mesh *mesh = og_mesh_create(v, 6, f, 8);
  const double cav[3] = {0.3, 0.0, 0.0};
  og_geom g;
  std::memset(&g, 0, sizeof g);
  g.container = OG_MESH;
  g.mesh = mesh;
  g.n_cavities = 1;
  g.cavities = cav;
  g.cavity_radius = 0.1;
  const double sites[6] = {-0.2, 0.1, 0.05, 0.0, -0.3, 0.1};
  const og_side sides[2] = {heat_side(), heat_side()};
  std::vector<og_site> out(2);
  std::vector<double> da(2 * 64), db(2 * 64);
  std::vector<int32_t> ds(2 * 64);
  int bad = 0;
#define CHECK(cond)                                         \
  do {                                                      \
    if (!(cond)) {                                          \
      std::printf("check failed: %s (line %d)\n", #cond, __LINE__); \
      ++bad;                                                \
    }                                                       \
  } while (0)
  for (int rng : {OG_RNG_PHILOX, OG_RNG_MT_SITE, OG_RNG_MT_PAIRED}) {
    og_params p = params(rng == OG_RNG_PHILOX ? OG_CROSS : (rng == OG_RNG_MT_PAIRED ? OG_PAIRED : OG_SQUARE), rng, 64);
    CHECK(!(og_estimate(&g, sides, &p, sites, 2, out.data(), da.data(), db.data(), ds.data(), 64) > 1));
  }
  og_side probe[2] = {heat_side(), heat_side()};
  probe[1].n_terms = 1;
  const double unit[8] = {OG_TERM_UNIT, -0.2, 0.0, 0.0, 0.15, 32 * 3.141592653589793 * 0.15 * 0.15 * 0.15 / 105, 0, 0};
  std::memcpy(probe[1].terms[0], unit, sizeof unit);
  og_params pf = params(OG_COUPLED, OG_RNG_MT_FD, 32);
  CHECK(!(og_estimate(&g, probe, &pf, sites, 2, out.data(), nullptr, nullptr, nullptr, 0) > 1));
  og_params pa = params(OG_SQUARE, OG_RNG_PHILOX, 16);
  pa.walks_max = 256;
  CHECK(!(og_estimate(&g, sides, &pa, sites, 2, out.data(), nullptr, nullptr, nullptr, 0) > 1));
  double rows[4 * 300];
  CHECK(!(og_trace(&g, &sides[0], &pa, sites, 99, 0, 0, rows, 300) < 1));
  double q[3];
  int64_t tri = 0, nv = 0, nt = 0;
  const double d_bvh = og_mesh_closest(mesh, sites, q, &tri, 0, &nv, &nt);
  const double d_brute = og_mesh_closest(mesh, sites, q, &tri, 1, &nv, &nt);
  CHECK(!(d_bvh != d_brute));
  CHECK(!(og_mesh_inside(mesh, sites) != 1));
  // Differential WoS and the certified sampler.
  og_params pd = params(OG_SQUARE, OG_RNG_MT_DIFF, 32);
  std::vector<double> dv(2 * 8);
  std::vector<int64_t> dc(2 * 9);
  CHECK(!(og_differential(&g, &sides[0], &pd, sites, 2, 0.01, dv.data(), dc.data()) != 0));
  pd.rng = OG_RNG_PHILOX;
  CHECK(!(og_differential(&g, &sides[0], &pd, sites, 2, 0.01, dv.data(), dc.data()) != 0));
  og_cert_params pc;
  std::memset(&pc, 0, sizeof pc);
  pc.seed = 5;
  pc.walks = 32;
  pc.max_steps = 64;
  pc.eps = 1e-4;
  pc.radius_cap = 0.1;
  pc.value_cap = 0.05;
  pc.slope = 0.65;
  pc.threads = 2;
  std::vector<double> cv(64);
  std::vector<int32_t> cs(64), cf(64);
  for (int rng : {int(OG_RNG_MT_KEY), int(OG_RNG_PHILOX)})
    for (int source : {0, 1}) {
      pc.rng = rng;
      pc.source = source;
      CHECK(!(og_certify(&g, &pc, sites, 2, cv.data(), cs.data(), cf.data()) != 0));
    }
  double cd[2];
  CHECK(!(og_cert_distance(&g, sites, 2, cd) != 0));
  CHECK(cd[0] >= 0.0 && cd[0] <= d_bvh);
  // The ball and the box containers.
  og_geom b;
  std::memset(&b, 0, sizeof b);
  b.container = OG_BALL;
  b.ball[3] = 1.0;
  og_params pb = params(OG_CROSS, OG_RNG_PHILOX, 64);
  CHECK(!(og_estimate(&b, sides, &pb, sites, 2, out.data(), nullptr, nullptr, nullptr, 0) > 1));
  og_mesh_destroy(mesh);
  std::printf("sanitized oracle run: %s\n", bad ? "FAILED" : "ok");
  return bad ? 1 : 0;
}
