// Command-line options of the wih_twos / wih_study drop-ins: the reference's
// flags (benchmarks/compiled_bunny/src/wih_twos.cpp:73-240) with the same
// defaults, checks and exit status 2 on a usage error, plus MI355X options.
// Flags of the reference's grid solvers (voxel, multigrid) are accepted and
// checked but have no effect here.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/twos.h"

namespace wih {

struct Options {
  std::string mode = "twos", source_mode = "affine", mesh_path, points_path, output_path, summary_path,
              trace_path, cavities_path;
  int walks = 16, repeats = 3, max_steps = 512, threads = 0, trace_sites = 0;
  uint64_t seed = 20260912ULL;
  double epsilon = 5.0e-4, source_y_slope = 0.65, cavity_radius = 0.0;
  // Differential WoS and validator settings, then flags accepted for
  // compatibility (the reference's grid solvers).
  int grid[3] = {54, 64, 48};
  int mg_cycles = 60, mg_pre_smooth = 3, mg_post_smooth = 3, mg_coarse_smooth = 80;
  double validation_radius_cap = 0.1, validation_value_cap = 0.1, diff_offset = 0.02, mg_tolerance = 1.0e-8,
         mg_omega = 0.72;
  // MI355X options.
  std::string streams = "reference";  // reference: the reference's mt19937_64 streams in binary64
  int precision = 32, walks_max = 0, binary32_radius = 0, distance_grid = 0;
  double ci_rel = 0.02;
};

// `tool` names the executable in the usage text.  Exits with status 2 on a
// usage error, as the reference does.
Options parse_options(int argc, char **argv, const char *tool);

[[noreturn]] void usage(const char *tool, const std::string &message = {});

// twos_params for one estimate: combine, and the stream family of
// o.streams (reference -> `reference_rng` in binary64).
twos_params make_params(const Options &o, int combine, int reference_rng, int repeat);

}  // namespace wih
