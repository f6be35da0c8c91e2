// wih_twos: drop-in for the reference's compiled backend
// (benchmarks/compiled_bunny/src/wih_twos.cpp:1799-1864) running the
// estimator on an MI355X through the C-ABI.  Same flags, same files in and
// out, same exit codes: usage errors exit 2, runtime errors print
// "fatal: <what>" and exit 1, a path that hits --max-steps is fatal.
//
// Modes twos, diffwos (Differential WoS cavity derivatives,
// run_differential_wos, wih_twos.cpp:1604-1705; twos_differential) and
// validate (certified minimum-benefit samples, run_minimum_benefit,
// minimum_benefit.hpp:145-188; twos_certify); the reference's voxel and mg
// modes are other solvers.  By default the walks replay the reference's own streams
// (mt19937_64 per walk keyed as evaluate_site keys them, binary64), so the
// field file matches the reference's; --streams philox selects the
// production path (Philox streams, binary32 walks, optional distance grid and
// adaptive budgets).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "wih_io.h"
#include "wih_options.h"

namespace {

using wih::Options;

// The problem's single side: zero Dirichlet data and the source of
// --source-mode (wih_twos.cpp:546-567).
twos_side make_side(const Options &o) {
  twos_side s{};
  if (o.source_mode == "affine") {
    s.source_kind = 1;
    s.src[0] = 1.0;
    s.src[2] = o.source_y_slope;
    s.src[5] = 0.5;
  } else {
    s.source_kind = 2;
    s.src[0] = 0.08;
    const double rows[3][8] = {{TWOS_TERM_GAUSS, -0.48, -0.18, -0.08, 0.24, 0.22, 0.16, 5.2},
                               {TWOS_TERM_GAUSS, 0.43, 0.19, -0.07, 0.21, 0.20, 0.15, 4.6},
                               {TWOS_TERM_GAUSS, 0.02, -0.31, -0.02, 0.30, 0.16, 0.19, 2.0}};
    s.n_terms = 3;
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 8; ++c) s.terms[k][c] = rows[k][c];
  }
  return s;
}

void write_field(const std::string &path, const std::vector<double> &points,
                 const std::vector<twos_site_result> &r) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write field output: " + path);
  out << "site,x,y,z,u,score,score_se,steps,limit_hits\n" << std::setprecision(17);
  for (size_t i = 0; i < r.size(); ++i)
    out << i << ',' << points[3 * i] << ',' << points[3 * i + 1] << ',' << points[3 * i + 2] << ',' << r[i].u << ','
        << r[i].score << ',' << r[i].score_se << ',' << r[i].steps << ',' << r[i].caps << '\n';
}

// write_trace (wih_twos.cpp:1725-1744): trace_sites evenly spaced sites, each
// one path keyed by seed + 900000 + site, rows (position, domain distance).
void write_trace(const std::string &path, twos_ctx *ctx, int32_t geom, const twos_side &side,
                 const std::vector<double> &points, const Options &o) {
  if (path.empty() || o.trace_sites == 0) return;
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write trace output: " + path);
  out << "path,step,x,y,z,radius\n" << std::setprecision(17);
  const size_t n = points.size() / 3;
  const int count = std::min<int>(o.trace_sites, int(n));
  twos_params p = wih::make_params(o, TWOS_SQUARE, TWOS_RNG_MT_SITE, 0);
  p.rng = o.streams == "reference" ? TWOS_RNG_MT_KEY : TWOS_RNG_PHILOX;
  p.seed = o.streams == "reference" ? o.seed + 900000ULL : o.seed;
  std::vector<double> rows(4 * size_t(o.max_steps + 2));
  for (int i = 0; i < count; ++i) {
    const size_t site = size_t(i) * n / size_t(count);
    int32_t nrows = 0;
    wih::check(twos_trace(ctx, geom, &side, &p, &points[3 * site], int64_t(site), 0, rows.data(),
                          int32_t(rows.size() / 4), &nrows),
               "trace");
    for (int32_t k = 0; k < nrows; ++k)
      out << i << ',' << k << ',' << rows[4 * k] << ',' << rows[4 * k + 1] << ',' << rows[4 * k + 2] << ','
          << rows[4 * k + 3] << '\n';
  }
}

void write_summary(const std::string &path, const wih::PlyMesh &mesh, size_t n_cavities, const Options &o,
                   size_t sites, double build_s, const std::vector<double> &times, uint64_t steps,
                   uint64_t limits, double walks_per_site) {
  const double runtime = wih::median(times);
  const double paths = double(sites) * walks_per_site;
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write summary: " + path);
  out << std::setprecision(17) << "{\n"
      << "  \"engine\": \"MI355X HIP LBVH\",\n"
      << "  \"vertices\": " << mesh.vertex_count() << ",\n"
      << "  \"faces\": " << mesh.face_count() << ",\n"
      << "  \"source_mode\": \"" << o.source_mode << "\",\n"
      << "  \"cavities\": " << n_cavities << ",\n"
      << "  \"cavity_radius\": " << o.cavity_radius << ",\n"
      << "  \"sites\": " << sites << ",\n"
      << "  \"walks_per_site\": " << walks_per_site << ",\n"
      << "  \"repeats\": " << o.repeats << ",\n"
      << "  \"threads\": " << o.threads << ",\n"
      << "  \"epsilon\": " << o.epsilon << ",\n"
      << "  \"mesh_build_seconds\": " << build_s << ",\n"
      << "  \"runtime_seconds_median\": " << runtime << ",\n"
      << "  \"runtime_seconds_min\": " << *std::min_element(times.begin(), times.end()) << ",\n"
      << "  \"paths_per_second\": " << paths / runtime << ",\n"
      << "  \"distance_queries_per_second\": " << (double(steps) + paths) / runtime << ",\n"
      << "  \"mean_steps\": " << double(steps) / paths << ",\n"
      << "  \"max_step_limit_hits\": " << limits << ",\n"
      << "  \"mesh_input_megabytes\": " << double(mesh.input_bytes()) / (1024.0 * 1024.0) << ",\n"
      << "  \"peak_rss_megabytes\": " << wih::peak_rss_megabytes() << ",\n"
      << "  \"streams\": \"" << o.streams << "\",\n"
      << "  \"timings_seconds\": [";
  for (size_t i = 0; i < times.size(); ++i) out << (i ? ", " : "") << times[i];
  out << "]\n}\n";
}

// --mode diffwos: every repeat runs the differential estimator with its
// repeat index in the keys, the field and counters are repeat 0's
// (run_differential_wos, wih_twos.cpp:1604-1705).
int run_differential(const Options &o, twos_ctx *ctx, int32_t geom, const twos_side &side, const wih::PlyMesh &mesh,
                     const std::vector<double> &points, size_t n_cavities, double build_s) {
  const size_t n = points.size() / 3;
  std::vector<double> times, out(n * n_cavities * 8), kept;
  std::vector<int64_t> counts(n * 9), kept_counts;
  for (int repeat = 0; repeat < o.repeats; ++repeat) {
    twos_params p = wih::make_params(o, TWOS_SQUARE, TWOS_RNG_MT_SITE, repeat);
    p.precision = 64;  // the reference's mt19937_64 keys, binary64 only
    const auto start = std::chrono::steady_clock::now();
    wih::check(twos_differential(ctx, geom, &side, &p, points.data(), int64_t(n), o.diff_offset, out.data(),
                                 counts.data()),
               "differential");
    times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count());
    if (repeat == 0) {
      kept = out;
      kept_counts = counts;
    }
  }
  uint64_t c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 9; ++k) c[k] += uint64_t(kept_counts[9 * i + size_t(k)]);
  if (c[6] != 0) throw std::runtime_error("one or more differential paths hit --max-steps");
  std::ofstream f(o.output_path);
  if (!f) throw std::runtime_error("cannot write differential field output: " + o.output_path);
  f << "site,x,y,z";
  for (size_t k = 0; k < n_cavities; ++k)
    f << ",dc" << k << "_x,dc" << k << "_y,dc" << k << "_z,dc" << k << "_se_x,dc" << k << "_se_y,dc" << k
      << "_se_z,dR" << k << ",dR" << k << "_se";
  f << '\n' << std::setprecision(17);
  for (size_t i = 0; i < n; ++i) {
    f << i << ',' << points[3 * i] << ',' << points[3 * i + 1] << ',' << points[3 * i + 2];
    for (size_t v = 0; v < 8 * n_cavities; ++v) f << ',' << kept[8 * n_cavities * i + v];
    f << '\n';
  }
  f.close();
  const double paths = double(n) * o.walks, runtime = wih::median(times);
  std::ofstream s(o.summary_path);
  if (!s) throw std::runtime_error("cannot write differential summary: " + o.summary_path);
  s << std::setprecision(17) << "{\n"
    << "  \"engine\": \"MI355X HIP LBVH Differential WoS\",\n"
    << "  \"mode\": \"diffwos\",\n"
    << "  \"vertices\": " << mesh.vertex_count() << ",\n"
    << "  \"faces\": " << mesh.face_count() << ",\n"
    << "  \"source_mode\": \"" << o.source_mode << "\",\n"
    << "  \"cavities\": " << n_cavities << ",\n"
    << "  \"gradient_parameters\": " << 4 * n_cavities << ",\n"
    << "  \"cavity_radius\": " << o.cavity_radius << ",\n"
    << "  \"sites\": " << n << ",\n"
    << "  \"walks_per_site\": " << o.walks << ",\n"
    << "  \"seed\": " << o.seed << ",\n"
    << "  \"epsilon\": " << o.epsilon << ",\n"
    << "  \"normal_offset\": " << o.diff_offset << ",\n"
    << "  \"seam_strip_multiplier\": 1,\n"
    << "  \"runtime_seconds_median\": " << runtime << ",\n"
    << "  \"outer_mean_steps\": " << double(c[0]) / paths << ",\n"
    << "  \"nested_mean_steps_per_outer_path\": " << double(c[1]) / paths << ",\n"
    << "  \"cavity_termination_fraction\": " << double(c[2]) / paths << ",\n"
    << "  \"near_seam_fraction\": " << double(c[3]) / paths << ",\n"
    << "  \"segment_unsafe_fraction\": " << double(c[4]) / paths << ",\n"
    << "  \"unsafe_offset_fraction\": " << double(c[5]) / paths << ",\n"
    << "  \"unsafe_given_cavity_fraction\": " << (c[2] ? double(c[5]) / double(c[2]) : 0.0) << ",\n"
    << "  \"near_seam_terminations\": " << c[3] << ",\n"
    << "  \"segment_unsafe_terminations\": " << c[4] << ",\n"
    << "  \"unsafe_offset_terminations\": " << c[5] << ",\n"
    << "  \"distance_queries\": " << c[7] << ",\n"
    << "  \"parity_rays\": " << c[8] << ",\n"
    << "  \"mesh_build_seconds\": " << build_s << ",\n"
    << "  \"peak_rss_megabytes\": " << wih::peak_rss_megabytes() << ",\n"
    << "  \"timings_seconds\": [";
  for (size_t i = 0; i < times.size(); ++i) s << (i ? ", " : "") << times[i];
  s << "]\n}\n";
  std::cout << "MI355X Differential WoS: " << n << " sites x " << o.walks << " walks, median " << runtime
            << " s\n";
  return 0;
}

// --mode validate: per site, the staged certified samples (1,024, 4,096,
// 16,384, 65,536 below --walks, then --walks), one call per stage; the
// samples file and the staged summary in the reference's formats
// (run_minimum_benefit, minimum_benefit.hpp:145-188).
int run_validation(const Options &o, twos_ctx *ctx, int32_t mesh_id, int32_t geom, const std::vector<double> &points,
                   const std::vector<double> &cavities, double build_s) {
  const size_t n = points.size() / 3;
  std::vector<int32_t> inside(n);
  std::vector<double> clearance(n), dist(n), closest(3 * n);
  std::vector<int32_t> cls(n);
  wih::check(twos_cert_distance(ctx, geom, points.data(), int64_t(n), clearance.data()), "certified distance");
  wih::check(twos_closest(ctx, geom, 64, points.data(), int64_t(n), dist.data(), closest.data(), cls.data()),
             "distance");
  std::ofstream output(o.output_path), summary(o.summary_path);
  if (!output) throw std::runtime_error("cannot write validation output: " + o.output_path);
  if (!summary) throw std::runtime_error("cannot write validation summary: " + o.summary_path);
  output << std::setprecision(17) << "site,sample,value,steps,capped,clipped\n";
  summary << std::setprecision(17)
          << "{\"scope\":\"conditional ideal-sampling certificate; conservative box geometry and source arithmetic\",\n"
          << "\"distance_tree_build_seconds\":" << build_s
          << ",\n\"distance_mode\":\"triangle_AABB_union_lower_bound\",\n"
          << "\"sample_cap\":" << o.validation_value_cap << ",\"radius_cap\":" << o.validation_radius_cap
          << ",\"stages\":[\n";
  std::vector<int> stages;
  for (int s : {1024, 4096, 16384, 65536})
    if (s < o.walks) stages.push_back(s);
  stages.push_back(o.walks);
  twos_cert_params p{};
  p.seed = o.seed;
  p.max_steps = o.max_steps;
  p.eps = o.epsilon;
  p.radius_cap = o.validation_radius_cap;
  p.value_cap = o.validation_value_cap;
  p.slope = o.source_y_slope;
  p.source = o.source_mode == "affine" ? TWOS_CERT_AFFINE : TWOS_CERT_HOUSING;
  uint64_t failures = 0;
  bool first = true;
  for (size_t site = 0; site < n; ++site) {
    const double *z = &points[3 * site];
    int32_t in = 0;
    wih::check(twos_inside(ctx, mesh_id, z, 1, &in), "inside");
    bool in_cavity = false;
    for (size_t k = 0; k + 2 < cavities.size(); k += 3) {
      const double dx = z[0] - cavities[k], dy = z[1] - cavities[k + 1], dz = z[2] - cavities[k + 2];
      in_cavity = in_cavity || std::sqrt(dx * dx + dy * dy + dz * dz) <= o.cavity_radius;
    }
    if (in != 1 || in_cavity)
      throw std::runtime_error("candidate classified outside; exact driver predicate is also required");
    std::vector<double> value(size_t(o.walks));
    std::vector<int32_t> steps(size_t(o.walks)), flags(size_t(o.walks));
    int prev = 0;
    double elapsed = 0.0;
    for (int stage : stages) {
      p.site_offset = int64_t(site);
      p.sample_begin = prev;
      p.walks = stage - prev;
      const auto start = std::chrono::steady_clock::now();
      wih::check(twos_certify(ctx, geom, &p, z, 1, value.data() + prev, steps.data() + prev, flags.data() + prev),
                 "certify");
      elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
      long double sum = 0;
      uint64_t nsteps = 0, clipped = 0, capped = 0;
      for (int j = 0; j < stage; ++j) {
        sum += value[size_t(j)];
        nsteps += uint64_t(steps[size_t(j)]);
        clipped += (flags[size_t(j)] & TWOS_CERT_CLIPPED) ? 1 : 0;
        capped += (flags[size_t(j)] & TWOS_CERT_CAPPED) ? 1 : 0;
      }
      if (!first) summary << ",\n";
      first = false;
      summary << "{\"site\":" << site << ",\"n\":" << stage << ",\"mean\":" << static_cast<double>(sum / stage)
              << ",\"mean_steps\":" << double(nsteps) / stage << ",\"clipped\":" << clipped
              << ",\"path_caps\":" << capped << ",\"elapsed_seconds\":" << elapsed
              << ",\"center_clearance_lower\":" << clearance[site] << ",\"embree_center_distance\":" << dist[site]
              << "}";
      for (int j = prev; j < stage; ++j) {
        const int32_t f = flags[size_t(j)];
        failures += (f & TWOS_CERT_CONTAINMENT) ? 1 : 0;
        output << site << ',' << j << ',' << value[size_t(j)] << ',' << steps[size_t(j)] << ','
               << ((f & TWOS_CERT_CAPPED) ? 1 : 0) << ',' << ((f & TWOS_CERT_CLIPPED) ? 1 : 0) << '\n';
      }
      prev = stage;
    }
  }
  summary << "\n],\"containment_failures\":" << failures << "}\n";
  if (failures) throw std::runtime_error("sample containment guard failed; validation record must be rejected");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    const Options o = wih::parse_options(argc, argv, "wih_twos");
    if (o.mode != "twos" && o.mode != "diffwos" && o.mode != "validate")
      wih::usage("wih_twos", "mode " + o.mode + " is not provided by the MI355X backend (twos, diffwos and validate only)");
    if (o.mode == "diffwos" && o.cavities_path.empty()) wih::usage("wih_twos", "Differential WoS requires --cavities");
    const wih::PlyMesh mesh = wih::load_ascii_ply(o.mesh_path);
    const std::vector<double> points = wih::read_points(o.points_path);
    const std::vector<double> cavities = wih::read_optional_points(o.cavities_path);
    const size_t n = points.size() / 3;
    wih::Context ctx;
    int32_t mesh_id = -1, geom = -1;
    double build_ms = 0.0;
    wih::check(twos_mesh_create(ctx.get(), mesh.vertices.data(), int64_t(mesh.vertex_count()), mesh.faces.data(),
                                int64_t(mesh.face_count()), &mesh_id, &build_ms),
               "mesh");
    twos_geometry_desc d{};
    d.container = TWOS_MESH;
    d.mesh = mesh_id;
    d.mesh_radius_f32 = o.binary32_radius;
    d.n_cavities = int32_t(cavities.size() / 3);
    d.cavities = cavities.empty() ? nullptr : cavities.data();
    d.cavity_radius = o.cavity_radius;
    wih::check(twos_geometry_create(ctx.get(), &d, &geom), "geometry");
    if (o.mode == "validate") return run_validation(o, ctx.get(), mesh_id, geom, points, cavities, build_ms * 1e-3);
    const twos_side side = make_side(o);
    if (o.mode == "diffwos")
      return run_differential(o, ctx.get(), geom, side, mesh, points, cavities.size() / 3, build_ms * 1e-3);
    const twos_side sides[2] = {side, side};
    std::vector<double> times;
    std::vector<twos_site_result> retained(n);
    for (int repeat = 0; repeat < o.repeats; ++repeat) {
      std::vector<twos_site_result> results(n);
      const twos_params p = wih::make_params(o, TWOS_SQUARE, TWOS_RNG_MT_SITE, repeat);
      const auto start = std::chrono::steady_clock::now();
      const twos_status st = twos_estimate(ctx.get(), geom, sides, &p, points.data(), int64_t(n), results.data(),
                                           nullptr);
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count());
      if (st != TWOS_OK && st != TWOS_ERR_STEP_CAP) wih::check(st, "estimate");
      if (repeat == 0) retained = std::move(results);
    }
    uint64_t steps = 0, limits = 0;
    double walks = 0.0;
    for (const auto &r : retained) {
      steps += uint64_t(r.steps);
      limits += uint64_t(r.caps);
      walks += double(r.walks);
    }
    if (limits != 0) throw std::runtime_error("one or more paths hit --max-steps");
    write_field(o.output_path, points, retained);
    write_trace(o.trace_path, ctx.get(), geom, side, points, o);
    const double walks_per_site = n ? walks / double(n) : 0.0;
    write_summary(o.summary_path, mesh, cavities.size() / 3, o, n, build_ms * 1e-3, times, steps, limits,
                  walks_per_site);
    std::cout << "MI355X TWoS: " << n << " sites x " << walks_per_site << " walks, median " << wih::median(times)
              << " s, " << std::fixed << std::setprecision(0) << (double(n) * walks_per_site / wih::median(times))
              << " paths/s\n";
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "fatal: " << e.what() << '\n';
    return 1;
  }
}
