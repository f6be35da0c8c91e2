// File formats and helpers shared by the wih_twos / wih_study drop-in
// executables: the reference's ASCII PLY and points readers, its median and
// peak-RSS helpers, and a thin RAII wrapper of the C-ABI context.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/twos.h"

namespace wih {

struct PlyMesh {
  std::vector<double> vertices;  // [nv][3]
  std::vector<uint32_t> faces;   // [nf][3]
  size_t vertex_count() const { return vertices.size() / 3; }
  size_t face_count() const { return faces.size() / 3; }
  // The reference's input_bytes: a binary64 Vec3 per vertex, three uint32 per face.
  size_t input_bytes() const { return vertex_count() * 24 + face_count() * 12; }
};

// ASCII triangle PLY: header `format ascii`, `element vertex N`, `element
// face M`, then the first three numbers of each vertex line and `3 a b c`
// face lines (benchmarks/compiled_bunny/src/wih_twos.cpp:437-493).
PlyMesh load_ascii_ply(const std::string &path);

// Points file: one point per line as x,y,z or x y z, lines starting with '#'
// skipped, lines without three numbers ignored (wih_twos.cpp:508-527).
std::vector<double> read_points(const std::string &path);
std::vector<double> read_optional_points(const std::string &path);

double median(std::vector<double> values);  // wih_twos.cpp:930-935
double peak_rss_megabytes();                 // VmHWM of /proc/self/status (:937-952)

// Throws std::runtime_error with twos_last_error() when status != TWOS_OK.
void check(twos_status status, const char *what);

// One device context (closed on scope exit).
class Context {
 public:
  explicit Context(int device = 0);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  twos_ctx *get() const { return ctx_; }

 private:
  twos_ctx *ctx_ = nullptr;
};

}  // namespace wih
