#include "wih_io.h"

#include <algorithm>
#include <fstream>
#include <sstream>

namespace wih {

PlyMesh load_ascii_ply(const std::string &path) {
  std::ifstream input(path);
  if (!input) throw std::runtime_error("cannot open mesh: " + path);
  std::string line;
  size_t nv = 0, nf = 0;
  bool ascii = false, header_done = false;
  while (std::getline(input, line)) {
    std::istringstream words(line);
    std::string first;
    words >> first;
    if (first == "format") {
      std::string format;
      words >> format;
      ascii = format == "ascii";
    } else if (first == "element") {
      std::string type;
      size_t count = 0;
      words >> type >> count;
      if (type == "vertex") nv = count;
      if (type == "face") nf = count;
    } else if (first == "end_header") {
      header_done = true;
      break;
    }
  }
  if (!ascii || !header_done || nv == 0 || nf == 0)
    throw std::runtime_error("expected a nonempty ASCII triangle PLY mesh");
  PlyMesh mesh;
  mesh.vertices.resize(3 * nv);
  for (size_t i = 0; i < nv; ++i) {
    if (!(input >> mesh.vertices[3 * i] >> mesh.vertices[3 * i + 1] >> mesh.vertices[3 * i + 2]))
      throw std::runtime_error("truncated PLY vertex section");
    std::getline(input, line);
  }
  mesh.faces.resize(3 * nf);
  for (size_t i = 0; i < nf; ++i) {
    unsigned count = 0;
    uint64_t a = 0, b = 0, c = 0;
    if (!(input >> count >> a >> b >> c) || count != 3 || a >= nv || b >= nv || c >= nv)
      throw std::runtime_error("expected valid triangular PLY faces");
    mesh.faces[3 * i] = uint32_t(a);
    mesh.faces[3 * i + 1] = uint32_t(b);
    mesh.faces[3 * i + 2] = uint32_t(c);
    std::getline(input, line);
  }
  return mesh;
}

std::vector<double> read_points(const std::string &path) {
  std::ifstream input(path);
  if (!input) throw std::runtime_error("cannot open points: " + path);
  std::vector<double> points;
  std::string line;
  while (std::getline(input, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::replace(line.begin(), line.end(), ',', ' ');
    std::istringstream values(line);
    double x, y, z;
    if (values >> x >> y >> z) {
      points.push_back(x);
      points.push_back(y);
      points.push_back(z);
    }
  }
  if (points.empty()) throw std::runtime_error("points file contains no points");
  return points;
}

std::vector<double> read_optional_points(const std::string &path) {
  if (path.empty()) return {};
  return read_points(path);
}

double median(std::vector<double> values) {
  std::sort(values.begin(), values.end());
  const size_t middle = values.size() / 2;
  if (values.size() % 2 == 1) return values[middle];
  return 0.5 * (values[middle - 1] + values[middle]);
}

double peak_rss_megabytes() {
  std::ifstream status("/proc/self/status");
  std::string key;
  while (status >> key) {
    if (key == "VmHWM:") {
      double kilobytes = 0.0;
      std::string unit;
      status >> kilobytes >> unit;
      return kilobytes / 1024.0;
    }
    std::string rest;
    std::getline(status, rest);
  }
  return 0.0;
}

void check(twos_status status, const char *what) {
  if (status != TWOS_OK) throw std::runtime_error(std::string(what) + ": " + twos_last_error());
}

Context::Context(int device) { check(twos_create(device, 0, &ctx_), "device"); }

Context::~Context() {
  if (ctx_) twos_destroy(ctx_);
}

}  // namespace wih
