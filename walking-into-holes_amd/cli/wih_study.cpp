// wih_study: drop-in for the reference's paired-estimator executable
// (benchmarks/submission_study/study.cpp:272-381) running on an MI355X
// through the C-ABI.  One request = one command line: the wih_twos flags plus
// --config (sensor line, then heat bumps), --hot-only, --fd, --diff and
// --sensor-side.  Output site,u,p,score,same_index,steps,caps and a one-line
// JSON summary, as the reference writes them.  `wih_study --serve` reads
// length-prefixed argument lists from stdin (a line with the count n, then
// n lines, the first being the program name) and answers each with
// "DONE <code>", keeping the device context, the mesh's LBVH and its
// geometries resident across requests; a session keeps one outer mesh.
//
// --fd 1 is the coupled finite-cavity difference (study.cpp:317-332): per
// point, walks start uniformly in the sensor ball and run with and without a
// cavity at the point on the same stream (TWOS_COUPLED); the score column is
// the mean of w (a - b) / (4 pi r), same_index its standard error, u and p
// zero, and no point is skipped.  --analytic-source 1 runs the reference's
// analytic-source variant (analytic_query.cpp: study.cpp built with
// WIH_ANALYTIC_SOURCE, :98-103): each step adds the exact Green-ball integral
// of the heat and sensor bumps (source_kind 3) instead of sampling a point;
// like that variant it applies to the paired estimator, not to --fd.
// --diff 1 is Differential WoS (study.cpp:303-316; twos_differential): per
// point and cavity, the derivative of the heat solution with respect to the
// cavity's centre and radius, written as site,x,y,z then per cavity
// dc{k}_x,y,z, dc{k}_se_x,y,z, dR{k}, dR{k}_se (write_differential_field,
// wih_twos.cpp:1576-1602); every point is evaluated, and an unsafe offset or a
// capped path fails the request after the field is written.  The reference's
// multigrid solver (--mode mg) is another solver; a request asking for it
// fails with status 1.
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "wih_io.h"
#include "wih_options.h"

namespace {

struct Session {
  std::unique_ptr<wih::Context> ctx;
  std::string mesh_path;
  int32_t mesh_id = -1;
};

// A request's geometry (outer mesh minus its cavities), released at the end
// of the request; the mesh and its LBVH stay resident for the session.
class Geometry {
 public:
  Geometry(Session &s, const std::vector<double> &holes, double radius) : ctx_(s.ctx->get()) {
    twos_geometry_desc d{};
    d.container = TWOS_MESH;
    d.mesh = s.mesh_id;
    d.n_cavities = int32_t(holes.size() / 3);
    d.cavities = holes.empty() ? nullptr : holes.data();
    d.cavity_radius = radius;
    wih::check(twos_geometry_create(ctx_, &d, &id_), "geometry");
  }
  ~Geometry() {
    if (id_ >= 0) twos_geometry_release(ctx_, id_);
  }
  Geometry(const Geometry &) = delete;
  Geometry &operator=(const Geometry &) = delete;
  int32_t id() const { return id_; }

 private:
  twos_ctx *ctx_;
  int32_t id_ = -1;
};

// Sensor (cx cy cz R) then heat bumps (cx cy cz R A) (study.cpp:37-43): the
// heat side sums A (1 - |a-c|^2/R^2)_+^2, the sensor side is the unit-mass
// bump t^2 / (32 pi R^3 / 105) (:44-51).
void load_config(const std::string &path, twos_side &heat, twos_side &sensor) {
  std::ifstream f(path);
  double c[4];
  if (!(f >> c[0] >> c[1] >> c[2] >> c[3])) throw std::runtime_error("bad sensor config");
  heat = twos_side{};
  sensor = twos_side{};
  heat.source_kind = 2;
  double row[5];
  while (f >> row[0] >> row[1] >> row[2] >> row[3] >> row[4]) {
    if (heat.n_terms == TWOS_MAX_TERMS) throw std::runtime_error("at most 8 heat bumps");
    double *t = heat.terms[heat.n_terms++];
    t[0] = TWOS_TERM_BUMP;
    t[1] = row[0]; t[2] = row[1]; t[3] = row[2]; t[4] = row[3]; t[7] = row[4];
  }
  if (heat.n_terms == 0) throw std::runtime_error("missing compact heat sources");
  sensor.source_kind = 2;
  sensor.n_terms = 1;
  double *t = sensor.terms[0];
  t[0] = TWOS_TERM_UNIT;
  t[1] = c[0]; t[2] = c[1]; t[3] = c[2]; t[4] = c[3];
  t[5] = 32 * 3.141592653589793 * std::pow(c[3], 3) / 105;
}

bool inside_cavity(const double *p, const std::vector<double> &holes, double radius) {
  for (size_t k = 0; k + 2 < holes.size(); k += 3) {
    const double dx = p[0] - holes[k], dy = p[1] - holes[k + 1], dz = p[2] - holes[k + 2];
    if (std::sqrt(dx * dx + dy * dy + dz * dz) <= radius) return true;
  }
  return false;
}

// --diff 1: Differential WoS over every point, repeat 0, the heat sources
// (study.cpp:303-316).
int differential_request(Session &session, const wih::Options &o, int32_t geom, const twos_side &heat,
                         const std::vector<double> &points, const std::vector<double> &holes,
                         std::chrono::steady_clock::time_point start) {
  const size_t n = points.size() / 3, k = holes.size() / 3;
  twos_params p = wih::make_params(o, TWOS_SQUARE, TWOS_RNG_MT_SITE, 0);
  p.precision = 64;  // the reference's mt19937_64 keys, binary64 only
  std::vector<double> out(n * k * 8);
  std::vector<int64_t> counts(n * 9);
  wih::check(twos_differential(session.ctx->get(), geom, &heat, &p, points.data(), int64_t(n), o.diff_offset,
                               out.data(), counts.data()),
             "differential");
  const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  std::ofstream f(o.output_path);
  if (!f) throw std::runtime_error("cannot write differential field output: " + o.output_path);
  f << "site,x,y,z";
  for (size_t c = 0; c < k; ++c)
    f << ",dc" << c << "_x,dc" << c << "_y,dc" << c << "_z,dc" << c << "_se_x,dc" << c << "_se_y,dc" << c
      << "_se_z,dR" << c << ",dR" << c << "_se";
  f << '\n' << std::setprecision(17);
  uint64_t steps = 0, unsafe = 0, caps = 0, queries = 0, rays = 0;
  for (size_t i = 0; i < n; ++i) {
    f << i << ',' << points[3 * i] << ',' << points[3 * i + 1] << ',' << points[3 * i + 2];
    for (size_t c = 0; c < 8 * k; ++c) f << ',' << out[8 * k * i + c];
    f << '\n';
    const int64_t *q = &counts[9 * i];
    steps += uint64_t(q[0] + q[1]);
    unsafe += uint64_t(q[5]);
    caps += uint64_t(q[6]);
    queries += uint64_t(q[7]);
    rays += uint64_t(q[8]);
  }
  f.close();
  if (unsafe || caps)
    throw std::runtime_error("unsafe differential path: unsafe=" + std::to_string(unsafe) +
                             " caps=" + std::to_string(caps));
  std::ofstream s(o.summary_path);
  s << std::setprecision(17) << "{\"total_seconds\":" << seconds << ",\"steps\":" << steps << ",\"caps\":" << caps
    << ",\"peak_rss_megabytes\":" << wih::peak_rss_megabytes() << ",\"distance_queries\":" << queries
    << ",\"parity_rays\":" << rays << "}\n";
  return 0;
}

int process_request(int argc, char **argv, Session &session) {
  try {
    bool hot = false, fd = false, diff = false, analytic = false;
    std::string config;
    std::vector<char *> args{argv[0]};
    for (int i = 1; i < argc; ++i) {
      const std::string key = argv[i];
      if (key == "--config") {
        if (i + 1 == argc) throw std::runtime_error("missing config");
        config = argv[++i];
      } else if (key == "--fd") {
        if (i + 1 == argc) throw std::runtime_error("missing fd flag");
        fd = std::stoi(argv[++i]) != 0;
      } else if (key == "--analytic-source") {
        if (i + 1 == argc) throw std::runtime_error("missing analytic-source flag");
        analytic = std::stoi(argv[++i]) != 0;
      } else if (key == "--diff") {
        if (i + 1 == argc) throw std::runtime_error("missing diff flag");
        diff = std::stoi(argv[++i]) != 0;
      } else if (key == "--sensor-side") {
        if (i + 1 == argc) throw std::runtime_error("missing sensor side");
        const int side = std::stoi(argv[++i]);
        if (side != 1 && side != -1) throw std::runtime_error("sensor side must be -1 or 1");
      } else if (key == "--hot-only") {
        if (i + 1 == argc) throw std::runtime_error("missing hot flag");
        hot = std::stoi(argv[++i]) != 0;
      } else {
        args.push_back(argv[i]);
      }
    }
    const wih::Options o = wih::parse_options(int(args.size()), args.data(), "wih_study");
    twos_side heat, sensor;
    load_config(config, heat, sensor);
    if (analytic && !fd) heat.source_kind = sensor.source_kind = 3;
    if (o.mode != "twos" && o.mode != "mg") throw std::runtime_error("use twos or mg");
    if (o.mode == "mg") throw std::runtime_error("mode mg is not provided by the MI355X backend (twos only)");
    const auto start = std::chrono::steady_clock::now();
    if (!session.ctx) {
      const wih::PlyMesh mesh = wih::load_ascii_ply(o.mesh_path);
      session.ctx = std::make_unique<wih::Context>();
      double ms = 0.0;
      wih::check(twos_mesh_create(session.ctx->get(), mesh.vertices.data(), int64_t(mesh.vertex_count()),
                                  mesh.faces.data(), int64_t(mesh.face_count()), &session.mesh_id, &ms),
                 "mesh");
      session.mesh_path = o.mesh_path;
    }
    if (session.mesh_path != o.mesh_path) throw std::runtime_error("a session must retain the same outer mesh");
    const std::vector<double> points = wih::read_points(o.points_path);
    const std::vector<double> holes = wih::read_optional_points(o.cavities_path);
    const Geometry geometry(session, holes, o.cavity_radius);
    const int32_t geom = geometry.id();
    if (diff) return differential_request(session, o, geom, heat, points, holes, start);
    // Sites inside a cavity are skipped and report zeros (study.cpp:341-342);
    // the finite-cavity difference evaluates every point (:319).
    const size_t n = points.size() / 3;
    std::vector<size_t> live;
    std::vector<double> live_pts;
    for (size_t i = 0; i < n; ++i)
      if (fd || !inside_cavity(&points[3 * i], holes, o.cavity_radius)) {
        live.push_back(i);
        live_pts.insert(live_pts.end(), &points[3 * i], &points[3 * i] + 3);
      }
    std::vector<twos_site_result> res(live.size());
    twos_stats stats{};
    if (!live.empty()) {
      const twos_side sides[2] = {heat, hot && !fd ? twos_side{} : sensor};
      const twos_params p = fd ? wih::make_params(o, TWOS_COUPLED, TWOS_RNG_MT_FD, 0)
                               : wih::make_params(o, TWOS_PAIRED, TWOS_RNG_MT_PAIRED, 0);
      const twos_status st = twos_estimate(session.ctx->get(), geom, sides, &p, live_pts.data(),
                                           int64_t(live.size()), res.data(), &stats);
      if (st != TWOS_OK && st != TWOS_ERR_STEP_CAP) wih::check(st, "estimate");
    }
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::vector<twos_site_result> all(n);
    for (size_t k = 0; k < live.size(); ++k) all[live[k]] = res[k];
    uint64_t steps = 0, caps = 0, paths = 0;
    std::ofstream f(o.output_path);
    if (!f) throw std::runtime_error("cannot write field output: " + o.output_path);
    f << "site,u,p,score,same_index,steps,caps\n" << std::setprecision(17);
    for (size_t i = 0; i < n; ++i) {
      const twos_site_result &r = all[i];
      steps += uint64_t(r.steps);
      caps += uint64_t(r.caps);
      paths += uint64_t(r.walks) * (fd ? 2 : 1);
      if (fd) {
        // PairValue{0, 0, up, same} of the finite-cavity difference (:329-331).
        f << i << ",0,0," << r.score << ',' << r.score_se << ',' << r.steps << ',' << r.caps << '\n';
        continue;
      }
      // Mean-only mode reports the primal mean as its score (study.cpp:117).
      const double score = hot ? r.u : r.score;
      f << i << ',' << r.u << ',' << (hot ? 0.0 : r.v) << ',' << score << ',' << (hot ? 0.0 : r.same) << ','
        << r.steps << ',' << r.caps << '\n';
    }
    f.close();
    if (caps) throw std::runtime_error("walk step limit reached");
    // One closest-point query per step plus the terminating one of each
    // uncapped path.
    const uint64_t queries = steps + paths - caps;
    std::ofstream s(o.summary_path);
    s << std::setprecision(17) << "{\"total_seconds\":" << seconds
      << ",\"peak_rss_megabytes\":" << wih::peak_rss_megabytes() << ",\"distance_queries\":" << queries
      << ",\"parity_rays\":0,\"enclosure_fallbacks\":0,\"relative_residual\":0,\"steps\":" << steps
      << ",\"caps\":" << caps << "}\n";
    return 0;
  } catch (const std::exception &e) {
    std::cerr << e.what() << '\n';
    return 1;
  }
}

}  // namespace

int main(int argc, char **argv) {
  Session session;
  if (argc == 2 && std::string(argv[1]) == "--serve") {
    std::string line;
    while (std::getline(std::cin, line)) {
      int n = 0;
      try {
        n = std::stoi(line);
      } catch (...) {
        return 2;
      }
      if (n <= 0 || n > 200) return 2;
      std::vector<std::string> values(static_cast<size_t>(n));
      for (auto &value : values)
        if (!std::getline(std::cin, value)) return 2;
      std::vector<char *> args;
      for (auto &value : values) args.push_back(value.data());
      const int code = process_request(n, args.data(), session);
      std::cout << "DONE " << code << std::endl;
      if (code) return code;
    }
    return 0;
  }
  return process_request(argc, argv, session);
}
