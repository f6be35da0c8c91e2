#include "wih_options.h"

#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <unordered_map>

namespace wih {

void usage(const char *tool, const std::string &message) {
  if (!message.empty()) std::cerr << "error: " << message << "\n\n";
  std::cerr << "Usage: " << tool << " --mesh MESH.ply --points POINTS.csv --output FIELD.csv\n"
            << "                 --summary SUMMARY.json [options]\n\n"
            << "Options (as the reference's wih_twos):\n"
            << "  --mode NAME           twos, diffwos or validate (the MI355X estimators)\n"
            << "  --walks N             paths per evaluation site (default 16)\n"
            << "  --repeats N           timed repetitions; median is reported (default 3)\n"
            << "  --threads N           accepted for compatibility (the GPU sets its own)\n"
            << "  --epsilon X           WoS termination shell (default 5e-4)\n"
            << "  --max-steps N         maximum spheres per path (default 512)\n"
            << "  --seed N              deterministic base seed\n"
            << "  --source-mode NAME    affine or housing (default affine)\n"
            << "  --source-y-slope X    f=1+slope*(y+0.5) (default 0.65)\n"
            << "  --cavities FILE       optional CSV of spherical cavity centers\n"
            << "  --cavity-radius X     common cavity radius (default 0)\n"
            << "  --trace-output FILE   optional CSV of representative paths\n"
            << "  --trace-sites N       number of sites to trace (default 0)\n"
            << "  --diff-offset X       diffwos normal offset (default 0.02)\n"
            << "  --validation-radius-cap X  validate: contained-ball radius cap (default 0.1)\n"
            << "  --validation-value-cap X   validate: sample clipping level (default 0.1)\n"
            << "  --grid, --mg-*: accepted, unused\n"
            << "MI355X options:\n"
            << "  --streams NAME        reference (mt19937_64 per walk, binary64; default) or philox\n"
            << "  --precision N         philox walks in binary 32 (default) or 64\n"
            << "  --walks-max N         philox: double walks per site up to N until the CI target\n"
            << "  --ci-rel X            adaptive target 1.96 SE <= X |score| (default 0.02)\n"
            << "  --distance-grid 0|1   philox: conservative distance-grid radius far from the mesh\n"
            << "  --binary32-radius 0|1 round the mesh distance to binary32 (Embree's query radius)\n";
  std::exit(2);
}

Options parse_options(int argc, char **argv, const char *tool) {
  Options o;
  std::unordered_map<std::string, std::string *> strings = {
      {"--mode", &o.mode},           {"--source-mode", &o.source_mode}, {"--mesh", &o.mesh_path},
      {"--points", &o.points_path},  {"--output", &o.output_path},      {"--summary", &o.summary_path},
      {"--trace-output", &o.trace_path}, {"--cavities", &o.cavities_path}, {"--streams", &o.streams}};
  for (int i = 1; i < argc; ++i) {
    const std::string key = argv[i];
    if (key == "--help" || key == "-h") usage(tool);
    if (i + 1 >= argc) usage(tool, "missing value for " + key);
    const std::string value = argv[++i];
    try {
      if (const auto it = strings.find(key); it != strings.end()) *it->second = value;
      else if (key == "--walks") o.walks = std::stoi(value);
      else if (key == "--repeats") o.repeats = std::stoi(value);
      else if (key == "--max-steps") o.max_steps = std::stoi(value);
      else if (key == "--threads") o.threads = std::stoi(value);
      else if (key == "--trace-sites") o.trace_sites = std::stoi(value);
      else if (key == "--seed") o.seed = std::stoull(value);
      else if (key == "--epsilon") o.epsilon = std::stod(value);
      else if (key == "--source-y-slope") o.source_y_slope = std::stod(value);
      else if (key == "--cavity-radius") o.cavity_radius = std::stod(value);
      else if (key == "--validation-radius-cap") o.validation_radius_cap = std::stod(value);
      else if (key == "--validation-value-cap") o.validation_value_cap = std::stod(value);
      else if (key == "--diff-offset") o.diff_offset = std::stod(value);
      else if (key == "--mg-cycles") o.mg_cycles = std::stoi(value);
      else if (key == "--mg-pre-smooth") o.mg_pre_smooth = std::stoi(value);
      else if (key == "--mg-post-smooth") o.mg_post_smooth = std::stoi(value);
      else if (key == "--mg-coarse-smooth") o.mg_coarse_smooth = std::stoi(value);
      else if (key == "--mg-tolerance") o.mg_tolerance = std::stod(value);
      else if (key == "--mg-omega") o.mg_omega = std::stod(value);
      else if (key == "--grid") {
        std::string parsed = value;
        std::replace(parsed.begin(), parsed.end(), ',', ' ');
        std::istringstream values(parsed);
        if (!(values >> o.grid[0] >> o.grid[1] >> o.grid[2])) usage(tool, "--grid must have the form NX,NY,NZ");
      } else if (key == "--precision") o.precision = std::stoi(value);
      else if (key == "--walks-max") o.walks_max = std::stoi(value);
      else if (key == "--ci-rel") o.ci_rel = std::stod(value);
      else if (key == "--distance-grid") o.distance_grid = std::stoi(value);
      else if (key == "--binary32-radius") o.binary32_radius = std::stoi(value);
      else usage(tool, "unknown option " + key);
    } catch (const std::logic_error &) {
      usage(tool, "invalid value for " + key);
    }
  }
  if (o.mesh_path.empty() || o.points_path.empty() || o.output_path.empty() || o.summary_path.empty())
    usage(tool, "--mesh, --points, --output, and --summary are required");
  // The reference's numeric checks (wih_twos.cpp:208-222).
  if ((o.mode != "twos" && o.mode != "diffwos" && o.mode != "voxel" && o.mode != "mg" && o.mode != "validate") ||
      (o.source_mode != "affine" && o.source_mode != "housing") || o.walks < 2 || o.repeats < 1 ||
      o.max_steps < 1 || o.epsilon <= 0.0 || o.trace_sites < 0 || o.validation_radius_cap <= 0 ||
      o.validation_value_cap <= 0 || o.mg_cycles < 1 || o.mg_pre_smooth < 1 || o.mg_post_smooth < 1 ||
      o.mg_coarse_smooth < 1 || o.diff_offset <= o.epsilon || o.mg_tolerance <= 0.0 || o.mg_omega <= 0.0 ||
      o.mg_omega >= 1.0 || o.cavity_radius < 0.0 || (!o.cavities_path.empty() && o.cavity_radius <= 0.0))
    usage(tool, "invalid numeric option");
  for (int extent : o.grid)
    if (extent < 3) usage(tool, "every --grid extent must be at least three");
  if (o.streams != "reference" && o.streams != "philox") usage(tool, "--streams must be reference or philox");
  if (o.precision != 32 && o.precision != 64) usage(tool, "--precision must be 32 or 64");
  if (o.streams == "reference" && (o.walks_max > 0 || o.distance_grid))
    usage(tool, "--walks-max and --distance-grid need --streams philox");
  return o;
}

twos_params make_params(const Options &o, int combine, int reference_rng, int repeat) {
  twos_params p{};
  p.combine = combine;
  const bool ref = o.streams == "reference";
  p.rng = ref ? reference_rng : TWOS_RNG_PHILOX;
  p.precision = ref ? 64 : o.precision;
  p.schedule = TWOS_PERSISTENT;
  p.seed = o.seed;
  p.repeat = repeat;
  p.max_steps = o.max_steps;
  p.eps = o.epsilon;
  p.walks = o.walks;
  p.walks_max = o.walks_max;
  p.ci_rel = o.ci_rel;
  p.ci_z = 1.96;
  p.use_graph = ref ? 0 : 1;
  p.flags = o.distance_grid ? TWOS_FLAG_DISTANCE_GRID : 0;
  return p;
}

}  // namespace wih
