// Host runtime and C-ABI of the TWoS estimator (include/twos.h).
//
// A context owns one device, one stream, the resident meshes and geometries,
// and a round engine.  The engine is a continuous-batching scheduler: sites
// of admitted requests occupy slots with resident binary64 accumulators; each
// round gives every unfinished site its next rung of the doubling ladder,
// runs all of them in one persistent launch (or one launch per site for the
// plain schedule), folds the samples into the accumulators and reads back one
// compact result per site.  Requests wait in FIFO order until enough slots
// are free.  twos_estimate is one request on the context's own engine.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "twos_device.h"
#include "twos_launch.h"
#include "twos_walk.h"

namespace twos {
namespace {

thread_local std::string g_error;

twos_status fail(twos_status code, const std::string &msg) {
  g_error = msg;
  return code;
}

#define TWOS_TRY(x)                              \
  do {                                           \
    const twos_status st_ = (x);                 \
    if (st_ != TWOS_OK) return st_;              \
  } while (0)

// A device failure: also clears the thread's last-error state (see HIP_OK).
twos_status dev_fail(const char *what, hipError_t e) {
  (void)hipGetLastError();
  return fail(TWOS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// A failed runtime call also leaves the thread's last-error state set;
// clearing it keeps the failure from being reported again by the next
// launch check of a later, unrelated call.
#define HIP_OK(x)                                                                             \
  do {                                                                                        \
    const hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                                   \
      (void)hipGetLastError();                                                                \
      return fail(TWOS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));           \
    }                                                                                         \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kMaxClasses = 8;
constexpr int64_t kWalkBlockSites = 64;  // requests of at least twice this many sites are listed in Morton order
constexpr int kCounters = 5;  // traversal counters of a counting launch (nodes, tris, grid lookups, hits, top)
// Distance grid: cells along the longest bounding-box axis, and the radius
// threshold (in cell sizes) above which a mesh step takes the grid bound.
constexpr int kGridCells = 128;
constexpr float kGridKappa = 4.0f;

}  // namespace

struct MeshEntry {
  MeshBuffers mb;
  // Content key: identical vertex and face arrays share one resident mesh.
  uint64_t hash = 0;
  std::vector<double> vertices;
  std::vector<uint32_t> faces;
  int32_t handles = 1;  // creates not yet matched by a release
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // vertex bounds (distance-grid extent)
  int32_t grid_cells = 0;
  int32_t depth = 0;
  int64_t bytes = 0;
  double build_ms = 0.0;
  int32_t users = 0;
};

struct GeomEntry {
  twos_geometry_desc desc{};
  std::vector<double> spheres, cylinders, cavities;
  DevGeom dev{};
  std::vector<void *> owned;
  int64_t owned_bytes = 0;
  int32_t kind = kGeomGeneric;
};

// Device bytes of one site slot: the resident slot arrays plus its entries
// in the round input and output blocks.
constexpr int64_t kSlotBytes = 3 * 8 + 4 + 4 + 4 + 2 * 8 + 8 + 8 + 4 + 4 + 4 + 8 + 8 + 8 + 8 * 8 + 8 + 8 + 8 +
                               int64_t(sizeof(twos_site_result)) + 4 + 4 + 28 * 8 + (4 + 4 + 8) +
                               int64_t(sizeof(twos_site_result)) + 4;

struct Engine;

}  // namespace twos

struct twos_ctx;

namespace twos {
int64_t ctx_device_bytes(const twos_ctx *ctx);
}  // namespace twos

struct twos_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  uint64_t budget = 0;
  std::mutex mu;
  std::map<int32_t, twos::MeshEntry> meshes;
  std::map<int32_t, twos::GeomEntry> geoms;
  int32_t next_mesh = 0, next_geom = 0;
  twos::DevGeom *d_geoms = nullptr;
  int32_t d_geoms_cap = 0;
  std::unique_ptr<twos::Engine> engine;
  // The allocation-free estimate's engine, laid out in the device's static
  // scratch array at twos_create (not listed in `engines`: static memory).
  std::unique_ptr<twos::Engine> arena_engine;
  // Orders device-pointer estimates after the caller's stream and back.
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // Every live engine on this context (its own and each server's): graphs
  // captured by any of them hold d_geoms, and queued or active requests hold
  // geometry ids.
  std::vector<twos::Engine *> engines;
};

namespace twos {

// ---------------------------------------------------------------------------
// Round engine

struct Request {
  int64_t id = 0;
  int32_t geom = 0;
  int32_t problem = -1;
  int32_t status = TWOS_OK;     // server: how the request ended
  twos_params prm{};
  DevSide sides[2];
  std::vector<double> sites;
  int64_t n = 0;
  twos_site_result *out = nullptr;
  int64_t slot0 = -1;
  std::vector<int32_t> done;    // walks done per site (per side)
  std::vector<uint8_t> state;   // 0 needs walks, 1 finished
  int64_t remaining = 0;
  int32_t cls = 0;              // combine class
  bool admitted = false;
  bool capped = false;          // a path of some site hit max_steps
  const double *d_sites = nullptr;  // device [n][3] positions (device entry points), else `sites`
  int64_t steps_sum = 0, paths_sum = 0;  // over the request's finished sites
  // The order in which a round lists the sites: Morton order of their
  // positions (empty: index order), so the walks that run at the same time
  // start near each other and share cache lines of the distance grid and the
  // tree.  A site's result does not depend on it.
  std::vector<int32_t> order;
};

// Morton order of n points (10 bits per axis over their bounding box).
std::vector<int32_t> morton_order(const double *p, int64_t n) {
  double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
  for (int64_t i = 1; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[3 * i + a]);
      hi[a] = std::max(hi[a], p[3 * i + a]);
    }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  auto spread = [](uint64_t v) {  // 10 bits -> every third bit
    v = (v | (v << 16)) & 0x030000FFULL;
    v = (v | (v << 8)) & 0x0300F00FULL;
    v = (v | (v << 4)) & 0x030C30C3ULL;
    v = (v | (v << 2)) & 0x09249249ULL;
    return v;
  };
  const size_t m = static_cast<size_t>(n);
  std::vector<uint32_t> code(m), code2(m);
  std::vector<int32_t> order(m), order2(m);
  for (size_t i = 0; i < m; ++i) {
    uint64_t c = 0;
    for (int a = 0; a < 3; ++a) {
      const double t = (p[3 * i + a] - lo[a]) / ext * 1023.0;
      c |= spread(uint64_t(std::min(1023.0, std::max(0.0, t)))) << a;
    }
    code[i] = uint32_t(c);
    order[i] = int32_t(i);
  }
  // Stable LSD radix sort of the 30-bit codes, 10 bits a pass: ties keep
  // index order, so the order is that of sorting (code, index) pairs.
  for (int shift = 0; shift < 30; shift += 10) {
    size_t start[1025] = {0};
    for (size_t i = 0; i < m; ++i) ++start[((code[i] >> shift) & 1023u) + 1];
    for (int k = 0; k < 1024; ++k) start[k + 1] += start[k];
    for (size_t i = 0; i < m; ++i) {
      const size_t at = start[(code[i] >> shift) & 1023u]++;
      code2[at] = code[i];
      order2[at] = order[i];
    }
    code.swap(code2);
    order.swap(order2);
  }
  return order;
}

struct Engine {
  twos_ctx *ctx = nullptr;
  int64_t cap_sites = 0, cap_walks = 0;
  int32_t cap_problems = 0;
  bool growable = false;
  bool server = false;  // admission by memory; failed requests are reported, not fatal

  SlotArrays s{};
  DevSide *d_sides = nullptr;
  double *d_val = nullptr;
  int32_t *d_steps = nullptr;
  unsigned long long *d_counters = nullptr;  // [kMaxClasses] tickets, then nodes, tris, grid lookups, grid hits, top-level nodes
  // Round input block (one H2D copy) and output block (one D2H copy).
  char *d_in = nullptr, *h_in = nullptr, *d_out = nullptr, *h_out = nullptr;
  // One-site work list of the plain schedule: 64 B in, 96 B out.
  char *d_plain = nullptr, *h_plain = nullptr;
  double *d_grad = nullptr;  // [cap_walks][2][3] gradient samples, allocated on the first gradient request
  // Pinned staging of a request's slot data at admission (grown on demand),
  // so its uploads are asynchronous copies instead of staged pageable ones.
  char *h_stage = nullptr;
  size_t stage_bytes = 0;
  bool stage_pending = false;  // uploads from h_stage may still be in flight on the context's stream
  size_t in_bytes = 0, out_bytes = 0;
  size_t off_slot = 0, off_base = 0, off_prefix = 0, off_nact = 0, off_total = 0;
  size_t off_compact = 0, off_decision = 0;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

  std::deque<Request *> queue;
  std::vector<Request *> active;
  std::vector<Request *> failed;  // server: ended without running (reported by the next step)
  std::map<int64_t, int64_t> free_slots;  // start -> length
  std::vector<int32_t> free_problems;
  int64_t next_id = 1;

  // hipGraph of a single-class persistent round.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_key = ~0ull;  // ~0: nothing captured (no launch shape has this key)

  twos_stats stats{};

  // Server admission by memory: device bytes charged per admitted site, and
  // the engine's fixed buffers (walk samples, round blocks).
  int64_t charged = 0;

  ~Engine() {
    release();
    unbind();
  }

  void bind(twos_ctx *c) {
    ctx = c;
    c->engines.push_back(this);
  }

  void unbind() {
    if (!ctx) return;
    auto &v = ctx->engines;
    v.erase(std::remove(v.begin(), v.end(), this), v.end());
    ctx = nullptr;
  }

  bool uses_geometry(int32_t geom) const {
    for (const Request *r : queue)
      if (r->geom == geom) return true;
    for (const Request *r : active)
      if (r->geom == geom) return true;
    return false;
  }

  // Arena engines (the allocation-free entry point) carve every device
  // buffer from the device's static scratch array and keep their host blocks
  // in host_arena (pageable): nothing to free, no events, no graphs.
  bool arena = false;
  std::vector<char> host_arena;

  void release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_key = ~0ull;
    if (arena) {
      s = SlotArrays{};
      d_sides = nullptr; d_val = nullptr; d_steps = nullptr; d_counters = nullptr;
      d_in = d_out = d_plain = nullptr;
      d_grad = nullptr;
      h_in = h_out = h_plain = nullptr;
      host_arena.clear();
      return;
    }
    dfree(s.pos); dfree(s.geom); dfree(s.problem); dfree(s.combine); dfree(s.keys); dfree(s.site); dfree(s.seed);
    dfree(s.repeat); dfree(s.walks_max); dfree(s.max_steps); dfree(s.eps); dfree(s.ci_rel); dfree(s.ci_z);
    dfree(s.acc); dfree(s.acc_walks); dfree(s.acc_steps); dfree(s.acc_caps); dfree(s.result); dfree(s.hint0); dfree(s.grid_kappa);
    dfree(s.acc_grad); dfree(d_grad);
    dfree(d_sides); dfree(d_val); dfree(d_steps); dfree(d_counters); dfree(d_in); dfree(d_out); dfree(d_plain);
    if (h_in) (void)hipHostFree(h_in);
    if (h_out) (void)hipHostFree(h_out);
    if (h_plain) (void)hipHostFree(h_plain);
    if (stage_pending && ctx) (void)hipStreamSynchronize(ctx->stream);
    stage_pending = false;
    if (h_stage) (void)hipHostFree(h_stage);
    h_in = h_out = h_plain = h_stage = nullptr;
    stage_bytes = 0;
    for (auto &e : ev)
      if (e) { (void)hipEventDestroy(e); e = nullptr; }
  }

  twos_status allocate(int64_t sites, int64_t walks, int32_t problems) {
    release();
    cap_sites = std::max<int64_t>(sites, 1);
    cap_walks = std::max<int64_t>(walks, 64);
    cap_problems = std::max(problems, 1);
    const size_t C = size_t(cap_sites);
    HIP_OK(dalloc(&s.pos, 3 * C)); HIP_OK(dalloc(&s.geom, C)); HIP_OK(dalloc(&s.problem, C));
    HIP_OK(dalloc(&s.combine, C));
    HIP_OK(dalloc(&s.keys, 2 * C)); HIP_OK(dalloc(&s.site, C)); HIP_OK(dalloc(&s.seed, C));
    HIP_OK(dalloc(&s.repeat, C)); HIP_OK(dalloc(&s.walks_max, C)); HIP_OK(dalloc(&s.max_steps, C));
    HIP_OK(dalloc(&s.eps, C)); HIP_OK(dalloc(&s.ci_rel, C)); HIP_OK(dalloc(&s.ci_z, C));
    HIP_OK(dalloc(&s.acc, 8 * C)); HIP_OK(dalloc(&s.acc_walks, C)); HIP_OK(dalloc(&s.acc_steps, C));
    HIP_OK(dalloc(&s.acc_caps, C)); HIP_OK(dalloc(&s.result, C)); HIP_OK(dalloc(&s.hint0, C));
    HIP_OK(dalloc(&s.grid_kappa, C));
    HIP_OK(dalloc(&s.acc_grad, 28 * C));
    HIP_OK(dalloc(&d_sides, 2 * size_t(cap_problems)));
    HIP_OK(dalloc(&d_val, 2 * size_t(cap_walks)));
    HIP_OK(dalloc(&d_steps, 2 * size_t(cap_walks)));
    HIP_OK(dalloc(&d_counters, kMaxClasses + kCounters));
    block_offsets();
    HIP_OK(dalloc(&d_in, in_bytes));
    HIP_OK(dalloc(&d_out, out_bytes));
    HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&h_in), in_bytes, hipHostMallocDefault));
    HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&h_out), out_bytes, hipHostMallocDefault));
    static_assert(64 + sizeof(twos_site_result) + 4 <= 256, "plain-schedule block");
    HIP_OK(dalloc(&d_plain, 256));
    HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&h_plain), 256, hipHostMallocDefault));
    std::memset(h_plain, 0, 256);
    std::memset(h_in, 0, in_bytes);
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    reset_state();
    return TWOS_OK;
  }

  // Offsets of the round input block (one H2D copy) and output block (one
  // D2H copy) for cap_sites slots.
  void block_offsets() {
    const size_t C = size_t(cap_sites);
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    off_slot = 0;
    off_base = align(off_slot + 4 * C);
    off_prefix = align(off_base + 4 * C);
    off_nact = align(off_prefix + 8 * (C + kMaxClasses));
    off_total = align(off_nact + 4 * kMaxClasses);
    in_bytes = align(off_total + 8 * kMaxClasses);
    off_compact = 0;
    off_decision = align(sizeof(twos_site_result) * C);
    out_bytes = align(off_decision + 4 * C);
  }

  void reset_state() {
    queue.clear();
    active.clear();
    failed.clear();
    charged = 0;
    free_slots.clear();
    free_slots[0] = cap_sites;
    free_problems.clear();
    for (int32_t p = cap_problems - 1; p >= 0; --p) free_problems.push_back(p);
  }

  // Lays the engine out in `bytes` of device memory at `dev` (256-byte
  // aligned): `sites` slots without the gradient channel, one problem, and
  // as many walks per round as the rest holds.  Allocates nothing on the
  // device; the host blocks live in host_arena.
  twos_status allocate_arena(char *dev, size_t bytes, int64_t sites) {
    release();
    arena = true;
    cap_sites = std::max<int64_t>(sites, 1);
    cap_problems = 1;
    cap_walks = 0;
    const size_t C = size_t(cap_sites);
    char *cur = dev;
    size_t left = bytes;
    bool ok = true;
    auto carve = [&](auto *&ptr, size_t count) {
      const size_t need = (sizeof(*ptr) * std::max<size_t>(count, 1) + 255) & ~size_t(255);
      if (!ok || need > left) {
        ok = false;
        return;
      }
      ptr = reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(cur);
      cur += need;
      left -= need;
    };
    carve(s.pos, 3 * C); carve(s.geom, C); carve(s.problem, C); carve(s.combine, C); carve(s.keys, 2 * C);
    carve(s.site, C); carve(s.seed, C); carve(s.repeat, C); carve(s.walks_max, C); carve(s.max_steps, C);
    carve(s.eps, C); carve(s.ci_rel, C); carve(s.ci_z, C); carve(s.acc, 8 * C); carve(s.acc_walks, C);
    carve(s.acc_steps, C); carve(s.acc_caps, C); carve(s.result, C); carve(s.hint0, C); carve(s.grid_kappa, C);
    s.acc_grad = nullptr;
    carve(d_sides, 2);
    carve(d_counters, kMaxClasses + kCounters);
    block_offsets();
    carve(d_in, in_bytes);
    carve(d_out, out_bytes);
    carve(d_plain, 256);
    // The rest: per-walk samples (two binary64) and step words (two int32).
    const int64_t walks = ok && left > 2 * 256 ? int64_t((left - 2 * 256) / (2 * (sizeof(double) + sizeof(int32_t)))) : 0;
    if (walks >= 64) {
      carve(d_val, 2 * size_t(walks));
      carve(d_steps, 2 * size_t(walks));
      cap_walks = walks;
    } else {
      ok = false;
    }
    if (!ok) return fail(TWOS_ERR_ADMISSION, "the static scratch array is too small for the engine");
    host_arena.assign(in_bytes + out_bytes + 256, 0);
    h_in = host_arena.data();
    h_out = h_in + in_bytes;
    h_plain = h_out + out_bytes;
    reset_state();
    return TWOS_OK;
  }

  // Kernel timing with the engine's events (arena engines have none).
  twos_status ev_record(int k, hipStream_t st) {
    if (ev[k]) HIP_OK(hipEventRecord(ev[k], st));
    return TWOS_OK;
  }
  float ev_ms(int a, int b) {
    float ms = 0.f;
    if (ev[a] && ev[b] && hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) {
      (void)hipGetLastError();
      ms = 0.f;
    }
    return ms;
  }

  bool idle() const { return queue.empty() && active.empty(); }

  // Device bytes of the buffers every request shares: per-walk samples and
  // step counts (and gradient samples once allocated), round blocks, sides.
  int64_t fixed_bytes() const {
    return cap_walks * 2 * int64_t(sizeof(double) + sizeof(int32_t)) + (d_grad ? cap_walks * 6 * 8 : 0) +
           int64_t(in_bytes + out_bytes) + int64_t(cap_problems) * int64_t(sizeof(DevSide)) + 512;
  }

  int64_t take_slots(int64_t n) {
    for (auto it = free_slots.begin(); it != free_slots.end(); ++it) {
      if (it->second >= n) {
        const int64_t start = it->first, len = it->second;
        free_slots.erase(it);
        if (len > n) free_slots[start + n] = len - n;
        return start;
      }
    }
    return -1;
  }

  void give_slots(int64_t start, int64_t n) {
    auto it = free_slots.emplace(start, n).first;
    auto next = std::next(it);
    if (next != free_slots.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_slots.erase(next);
    }
    if (it != free_slots.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_slots.erase(it);
      }
    }
  }

  Tables plain_tables() const {
    Tables tb{};
    tb.geoms = ctx->d_geoms;
    tb.sides = d_sides;
    tb.slots = s;
    tb.round.act_slot = reinterpret_cast<int32_t *>(d_plain + 0);
    tb.round.walk_base = reinterpret_cast<int32_t *>(d_plain + 4);
    tb.round.act_prefix = reinterpret_cast<int64_t *>(d_plain + 8);
    tb.round.n_active = reinterpret_cast<int32_t *>(d_plain + 24);
    tb.round.total_jobs = reinterpret_cast<int64_t *>(d_plain + 32);
    tb.round.counter = d_counters;
    tb.round.compact = reinterpret_cast<twos_site_result *>(d_plain + 64);
    tb.round.decision = reinterpret_cast<int32_t *>(d_plain + 64 + sizeof(twos_site_result));
    return tb;
  }

  Tables tables(int cls, int64_t e0) const {
    Tables tb{};
    tb.geoms = ctx->d_geoms;
    tb.sides = d_sides;
    tb.slots = s;
    tb.round.act_slot = reinterpret_cast<int32_t *>(d_in + off_slot) + e0;
    tb.round.walk_base = reinterpret_cast<int32_t *>(d_in + off_base) + e0;
    tb.round.act_prefix = reinterpret_cast<int64_t *>(d_in + off_prefix) + e0 + cls;
    tb.round.n_active = reinterpret_cast<int32_t *>(d_in + off_nact) + cls;
    tb.round.total_jobs = reinterpret_cast<int64_t *>(d_in + off_total) + cls;
    tb.round.counter = d_counters + cls;
    tb.round.compact = reinterpret_cast<twos_site_result *>(d_out + off_compact) + e0;
    tb.round.decision = reinterpret_cast<int32_t *>(d_out + off_decision) + e0;
    return tb;
  }

  // Upload one admitted request's per-site slot data and clear its
  // accumulators.
  twos_status admit(Request *r) {
    const int64_t n = r->n, s0 = r->slot0;
    const size_t N = size_t(n);
    hipStream_t st = ctx->stream;
    // Every per-site array is written once into the staging block (pinned;
    // pageable in an arena engine, which allocates nothing) and uploaded from
    // there; the stream is synchronised before the block is reused.
    const size_t need = N * 112 + 16 * 256;
    std::vector<char> pageable;
    char *stage;
    if (arena) {
      pageable.resize(need);
      stage = pageable.data();
    } else {
      // The previous admission's uploads read this block: let them finish.
      if (stage_pending) HIP_OK(hipStreamSynchronize(st));
      stage_pending = false;
      if (stage_bytes < need) {
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        stage_bytes = 0;
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&h_stage), need, hipHostMallocDefault));
        stage_bytes = need;
      }
      stage = h_stage;
      stage_pending = true;  // from the first upload on, also if one fails
    }
    size_t used = 0;
    auto carve = [&](size_t bytes) {
      char *at = stage + used;
      used += (bytes + 255) & ~size_t(255);
      return at;
    };
    auto *gi = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *pi = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *comb = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *keys = reinterpret_cast<uint2 *>(carve(sizeof(uint2) * 2 * N));
    auto *site = reinterpret_cast<int64_t *>(carve(8 * N));
    auto *seed = reinterpret_cast<uint64_t *>(carve(8 * N));
    auto *rep = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *wmax = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *msteps = reinterpret_cast<int32_t *>(carve(4 * N));
    auto *eps = reinterpret_cast<double *>(carve(8 * N));
    auto *rel = reinterpret_cast<double *>(carve(8 * N));
    auto *z = reinterpret_cast<double *>(carve(8 * N));
    auto *kap = reinterpret_cast<float *>(carve(4 * N));
    const float kappa = (r->prm.flags & TWOS_FLAG_DISTANCE_GRID)
                            ? (r->prm.grid_kappa > 0.f ? r->prm.grid_kappa : kGridKappa) : 0.f;
    const bool adaptive = r->prm.walks_max > r->prm.walks;
    for (size_t i = 0; i < N; ++i) {
      const int64_t gsite = int64_t(r->prm.site_offset) + int64_t(i);
      gi[i] = r->geom;
      pi[i] = r->problem;
      comb[i] = r->prm.combine;
      keys[2 * i] = philox_key_host(r->prm.seed, gsite, 0);
      keys[2 * i + 1] = philox_key_host(r->prm.seed, gsite, 1);
      site[i] = gsite;
      seed[i] = r->prm.seed;
      rep[i] = r->prm.repeat;
      wmax[i] = adaptive ? r->prm.walks_max : r->prm.walks;
      msteps[i] = r->prm.max_steps;
      eps[i] = r->prm.eps;
      rel[i] = adaptive ? r->prm.ci_rel : -1.0;
      z[i] = r->prm.ci_z;
      kap[i] = kappa;
    }
    if (r->d_sites) {
      HIP_OK(hipMemcpyAsync(s.pos + 3 * s0, r->d_sites, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
    } else {
      auto *pos = reinterpret_cast<double *>(carve(sizeof(double) * 3 * N));
      std::memcpy(pos, r->sites.data(), sizeof(double) * 3 * N);
      HIP_OK(hipMemcpyAsync(s.pos + 3 * s0, pos, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
    }
    HIP_OK(hipMemcpyAsync(s.geom + s0, gi, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.problem + s0, pi, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.combine + s0, comb, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.keys + 2 * s0, keys, sizeof(uint2) * 2 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.site + s0, site, 8 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.seed + s0, seed, 8 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.repeat + s0, rep, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.walks_max + s0, wmax, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.max_steps + s0, msteps, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.eps + s0, eps, 8 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.ci_rel + s0, rel, 8 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.ci_z + s0, z, 8 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s.grid_kappa + s0, kap, 4 * N, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemsetAsync(s.acc + 8 * s0, 0, sizeof(double) * 8 * n, st));
    HIP_OK(hipMemsetAsync(s.acc_walks + s0, 0, 8 * n, st));
    HIP_OK(hipMemsetAsync(s.acc_steps + s0, 0, 8 * n, st));
    HIP_OK(hipMemsetAsync(s.acc_caps + s0, 0, 8 * n, st));
    if (r->prm.flags & TWOS_FLAG_GRADIENT) {
      HIP_OK(hipMemsetAsync(s.acc_grad + 28 * s0, 0, sizeof(double) * 28 * n, st));
      if (!d_grad) HIP_OK(dalloc(&d_grad, 6 * size_t(cap_walks)));
    }
    HIP_OK(hipMemcpyAsync(d_sides + 2 * r->problem, r->sides, sizeof(DevSide) * 2, hipMemcpyHostToDevice, st));
    // Each new site's first closest triangle seeds its paths' queries.
    if (ctx->d_geoms) HIP_OK(launch_slot_hints(tables(0, 0), s0, n, st));
    // A pageable block dies at return: finish its copies first.  The pinned
    // block stays, and the next admission waits for its copies instead.
    if (arena) HIP_OK(hipStreamSynchronize(st));
    r->done.assign(n, 0);
    r->state.assign(n, 0);
    r->remaining = n;
    r->admitted = true;
    return TWOS_OK;
  }

  // Admits queued requests in FIFO order while slots and side tables are
  // free; a server also needs the request's slot bytes to fit the context's
  // HBM budget beside everything resident (meshes, distance grids,
  // geometries, every engine's buffers and admitted requests).
  twos_status admit_queued() {
    while (!queue.empty()) {
      Request *r = queue.front();
      if (free_problems.empty()) break;
      const int64_t need = r->n * kSlotBytes;
      if (server && ctx_device_bytes(ctx) + need > int64_t(ctx->budget)) {
        if (!active.empty()) break;  // wait until an admitted request frees its memory
        queue.pop_front();
        r->status = fail(TWOS_ERR_ADMISSION, "request does not fit the HBM budget beside the resident geometry");
        failed.push_back(r);
        continue;
      }
      const int64_t slot0 = take_slots(r->n);
      if (slot0 < 0) break;
      r->slot0 = slot0;
      r->problem = free_problems.back();
      free_problems.pop_back();
      queue.pop_front();
      const twos_status st = admit(r);
      if (st != TWOS_OK) {
        if (!server) return st;
        give_slots(r->slot0, r->n);
        free_problems.push_back(r->problem);
        r->status = st;
        failed.push_back(r);
        continue;
      }
      if (server) charged += need;
      active.push_back(r);
    }
    return TWOS_OK;
  }

  int geom_kind_of(int32_t geom) const {
    auto it = ctx->geoms.find(geom);
    return it == ctx->geoms.end() ? int(kGeomGeneric) : it->second.kind;
  }

  int32_t next_inc(const Request *r, int64_t i) const {
    const int32_t d = r->done[i];
    const int32_t wmax = r->prm.walks_max > r->prm.walks ? r->prm.walks_max : r->prm.walks;
    if (d == 0) return r->prm.walks;
    return std::min<int32_t>(d, wmax - d);
  }

  double *grad_of(const Request *r) const { return (r->prm.flags & TWOS_FLAG_GRADIENT) ? d_grad : nullptr; }

  // No side of any request in the class has a volume source.
  template <class E>
  static bool class_bare(const std::vector<E> &entries) {
    for (const E &en : entries)
      if (en.r->sides[0].source_kind != 0 || en.r->sides[1].source_kind != 0) return false;
    return true;
  }

  twos_status launch_walks(const Request *r, int cls, int64_t e0, int64_t total_walks, int kind, bool bare) {
    const twos_params &p = r->prm;
    Tables tb = tables(cls, e0);
    WalkParams wp{};
    wp.combine = p.combine;
    wp.rng = p.rng;
    wp.jobs_per_walk = p.combine == kCross ? 2 : 1;
    wp.job_begin = 0;
    wp.job_end = total_walks * wp.jobs_per_walk;
    wp.val = d_val;
    wp.steps = d_steps;
    wp.grad = grad_of(r);
    wp.counters = d_counters + kMaxClasses;
    wp.refill = p.refill_batch > 0 ? p.refill_batch : kRefillDefault;
    hipStream_t st = ctx->stream;
    const bool count = (p.flags & TWOS_FLAG_COUNT_TRAVERSAL) != 0;
    hipError_t e;
    if (p.rng != kRngPhilox) {
      e = launch_walk_simple_f64(kind, true, tb, wp, 64, st);
    } else if (p.schedule == TWOS_PLAIN) {
      const int block = p.block > 0 ? p.block : 256;
      e = p.precision == 64 ? launch_walk_simple_f64(kind, false, tb, wp, block, st)
                            : launch_walk_simple_f32(kind, tb, wp, block, st);
    } else {
      // Launch shape: block (64..256 lanes), occupancy variant and LDS stack
      // depth (binary32 only); none of them changes a result.
      HIP_OK(launch_zero(tb.round.counter, 1, st));
      int grid = 0;
      const int cus = ctx->prop.multiProcessorCount;
      const int block = p.block > 0 ? p.block : 256;
      e = p.precision == 64
              ? launch_walk_persistent_f64(kind, count, tb, wp, block, cus, st, &grid)
              : launch_walk_persistent_f32(kind, count, p.waves_per_simd, p.lds_stack, p.step_batch, bare, tb, wp,
                                           block, cus, st, &grid);
      stats.grid = grid;
      stats.block = block;
      stats.waves_per_simd = p.waves_per_simd;
    }
    if (e != hipSuccess) return dev_fail("walk launch", e);
    stats.walk_launches += 1;
    return TWOS_OK;
  }

  // One round over all admitted requests.  Completed request ids go to done.
  twos_status step(std::vector<int64_t> &done, double *dump_a = nullptr, double *dump_b = nullptr,
                   int32_t *dump_steps = nullptr, int64_t dump_stride = 0) {
    twos_status st = admit_queued();
    for (Request *r : failed) done.push_back(r->id);
    failed.clear();
    if (st != TWOS_OK) return st;
    if (active.empty()) {
      if (!queue.empty()) return fail(TWOS_ERR_ADMISSION, "request does not fit the engine's slots");
      return TWOS_OK;
    }
    // Work list: each unfinished site gets its next rung; classes group
    // requests that share precision and combine mode.
    struct Entry { Request *r; int64_t i; int32_t inc; };
    std::vector<std::vector<Entry>> by_class;
    std::vector<int64_t> class_key;
    // Sites that do not fit this round's walk budget wait for the next one;
    // a site always receives its whole rung in one round.
    int64_t budget = cap_walks;
    bool full = false;
    for (Request *r : active) {
      if (full) break;
      // Launch-wide choices split classes: precision, combine, the stream
      // family and traversal counting.  Everything else a result depends on
      // (eps, max_steps, the distance-grid threshold, sides, geometry) is
      // per slot, so a site's result is the same alone and in any mix.
      const int64_t key = int64_t(r->prm.precision) | int64_t(r->prm.combine) << 8 | int64_t(r->prm.rng) << 12 |
                          int64_t((r->prm.flags & TWOS_FLAG_COUNT_TRAVERSAL) ? 1 : 0) << 16 |
                          int64_t((r->prm.flags & TWOS_FLAG_GRADIENT) ? 1 : 0) << 17;
      int c = -1;
      for (size_t k = 0; k < class_key.size(); ++k)
        if (class_key[k] == key) c = int(k);
      for (int64_t k = 0; k < r->n; ++k) {
        const int64_t i = r->order.empty() ? k : r->order[size_t(k)];
        if (r->state[i]) continue;
        const int32_t inc = next_inc(r, i);
        if (inc > budget) { full = true; break; }
        if (c < 0) {
          if (int(class_key.size()) == kMaxClasses) { full = true; break; }
          class_key.push_back(key);
          by_class.emplace_back();
          by_class.back().reserve(size_t(r->remaining));
          c = int(class_key.size()) - 1;
        }
        by_class[c].push_back({r, i, inc});
        budget -= inc;
      }
    }
    if (by_class.empty()) return fail(TWOS_ERR_ADMISSION, "walk budget of one round below one site's rung");
    auto *a_slot = reinterpret_cast<int32_t *>(h_in + off_slot);
    auto *a_base = reinterpret_cast<int32_t *>(h_in + off_base);
    auto *a_prefix = reinterpret_cast<int64_t *>(h_in + off_prefix);
    auto *a_nact = reinterpret_cast<int32_t *>(h_in + off_nact);
    auto *a_total = reinterpret_cast<int64_t *>(h_in + off_total);
    std::vector<int64_t> e0s(by_class.size()), totals(by_class.size());
    int64_t e = 0;
    for (size_t c = 0; c < by_class.size(); ++c) {
      e0s[c] = e;
      int64_t run = 0;
      for (size_t k = 0; k < by_class[c].size(); ++k) {
        const Entry &en = by_class[c][k];
        a_slot[e + k] = int32_t(en.r->slot0 + en.i);
        a_base[e + k] = en.r->done[en.i];
        a_prefix[e + c + k] = run;
        run += en.inc;
      }
      a_prefix[e + c + by_class[c].size()] = run;
      a_nact[c] = int32_t(by_class[c].size());
      a_total[c] = run * (by_class[c].front().r->prm.combine == kCross ? 2 : 1);
      totals[c] = run;
      e += int64_t(by_class[c].size());
    }
    const int64_t n_entries = e;
    hipStream_t strm = ctx->stream;
    const bool single = by_class.size() == 1;
    Request *r0 = by_class[0].front().r;
    const bool plain = r0->prm.schedule == TWOS_PLAIN;
    const bool graphable = single && !plain && r0->prm.use_graph && r0->prm.rng == kRngPhilox && !dump_a &&
                           !(r0->prm.flags & TWOS_FLAG_COUNT_TRAVERSAL);
    int kind0 = geom_kind_of(r0->geom);
    for (const Entry &en : by_class[0])
      if (en.r->geom != r0->geom && geom_kind_of(en.r->geom) != kind0) kind0 = kGeomGeneric;

    if (plain) {
      // One launch per site: a one-entry work list, the site's walks one
      // thread per path, one fold, read back, synchronise.
      const Tables tb = plain_tables();
      for (size_t c = 0; c < by_class.size(); ++c) {
        for (size_t k = 0; k < by_class[c].size(); ++k) {
          const Entry &en = by_class[c][k];
          const int jpw = en.r->prm.combine == kCross ? 2 : 1;
          *reinterpret_cast<int32_t *>(h_plain + 0) = int32_t(en.r->slot0 + en.i);
          *reinterpret_cast<int32_t *>(h_plain + 4) = en.r->done[en.i];
          *reinterpret_cast<int64_t *>(h_plain + 8) = 0;
          *reinterpret_cast<int64_t *>(h_plain + 16) = en.inc;
          *reinterpret_cast<int32_t *>(h_plain + 24) = 1;
          *reinterpret_cast<int64_t *>(h_plain + 32) = int64_t(en.inc) * jpw;
          HIP_OK(hipMemcpyAsync(d_plain, h_plain, 64, hipMemcpyHostToDevice, strm));
          TWOS_TRY(ev_record(0, strm));
          WalkParams wp{};
          wp.combine = en.r->prm.combine;
          wp.rng = en.r->prm.rng;
          wp.jobs_per_walk = jpw;
          wp.job_begin = 0;
          wp.job_end = int64_t(en.inc) * jpw;
          wp.val = d_val;
          wp.steps = d_steps;
          wp.grad = grad_of(en.r);
          const int block = en.r->prm.block > 0 ? en.r->prm.block : 256;
          const int kind = geom_kind_of(en.r->geom);
          const hipError_t le = en.r->prm.rng != kRngPhilox ? launch_walk_simple_f64(kind, true, tb, wp, 64, strm)
                                : en.r->prm.precision == 64 ? launch_walk_simple_f64(kind, false, tb, wp, block, strm)
                                                            : launch_walk_simple_f32(kind, tb, wp, block, strm);
          if (le != hipSuccess) return dev_fail("walk launch", le);
          stats.walk_launches += 1;
          TWOS_TRY(ev_record(1, strm));
          if (dump_a) {
            HIP_OK(hipStreamSynchronize(strm));
            const std::vector<Entry> one{en};
            st = dump_round(one, dump_a, dump_b, dump_steps, dump_stride);
            if (st != TWOS_OK) return st;
          }
          HIP_OK(launch_reduce(tb, wp, 1, strm));
          TWOS_TRY(ev_record(2, strm));
          HIP_OK(hipMemcpyAsync(h_plain + 64, d_plain + 64, sizeof(twos_site_result) + 4, hipMemcpyDeviceToHost, strm));
          HIP_OK(hipStreamSynchronize(strm));
          stats.walk_kernel_ms += ev_ms(0, 1);
          stats.reduce_kernel_ms += ev_ms(1, 2);
          twos_site_result res;
          int32_t dec;
          std::memcpy(&res, h_plain + 64, sizeof(res));
          std::memcpy(&dec, h_plain + 64 + sizeof(twos_site_result), 4);
          finish_entry(en.r, en.i, en.inc, res, dec);
        }
      }
    } else if (graphable) {
      // Everything the captured launches depend on: precision, combine,
      // geometry kind, launch shape and the channel flags.
      const twos_params &q = r0->prm;
      const uint64_t key = uint64_t(q.precision) | uint64_t(q.combine) << 8 | uint64_t(kind0) << 12 |
                           uint64_t(q.waves_per_simd) << 16 | uint64_t(q.block) << 24 | uint64_t(q.lds_stack) << 36 |
                           uint64_t(q.flags & 7) << 44 | uint64_t(q.step_batch) << 48 |
                           uint64_t(class_bare(by_class[0])) << 56 | uint64_t(q.refill_batch) << 57;
      if (graph_key != key) {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        HIP_OK(hipStreamBeginCapture(strm, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipMemcpyAsync(d_in, h_in, in_bytes, hipMemcpyHostToDevice, strm);
        twos_status ls = ce == hipSuccess ? launch_walks(r0, 0, 0, 0, kind0, class_bare(by_class[0])) : TWOS_ERR_DEVICE;
        Tables tb = tables(0, 0);
        WalkParams wp{};
        wp.val = d_val;
        wp.steps = d_steps;
        wp.grad = grad_of(r0);
        if (ls == TWOS_OK) ce = launch_reduce(tb, wp, int(cap_sites), strm);
        if (ls == TWOS_OK && ce == hipSuccess)
          ce = hipMemcpyAsync(h_out, d_out, out_bytes, hipMemcpyDeviceToHost, strm);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(strm, &g);
        if (ls != TWOS_OK) return ls;
        HIP_OK(ce);
        HIP_OK(ee);
        graph = g;
        HIP_OK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        graph_key = key;
        stats.walk_launches -= 1;  // counted once per replay below
      }
      TWOS_TRY(ev_record(0, strm));
      HIP_OK(hipGraphLaunch(graph_exec, strm));
      TWOS_TRY(ev_record(1, strm));
      HIP_OK(hipStreamSynchronize(strm));
      stats.walk_launches += 1;
      stats.walk_kernel_ms += ev_ms(0, 1);
      const auto *res = reinterpret_cast<twos_site_result *>(h_out + off_compact);
      const auto *dec = reinterpret_cast<int32_t *>(h_out + off_decision);
      for (size_t k = 0; k < by_class[0].size(); ++k)
        finish_entry(by_class[0][k].r, by_class[0][k].i, by_class[0][k].inc, res[k], dec[k]);
    } else {
      HIP_OK(hipMemcpyAsync(d_in, h_in, in_bytes, hipMemcpyHostToDevice, strm));
      for (size_t c = 0; c < by_class.size(); ++c) {
        Request *rc = by_class[c].front().r;
        int kind = geom_kind_of(rc->geom);
        for (const Entry &en : by_class[c])
          if (en.r->geom != rc->geom && geom_kind_of(en.r->geom) != kind) kind = kGeomGeneric;
        if (rc->prm.flags & TWOS_FLAG_COUNT_TRAVERSAL) HIP_OK(launch_zero(d_counters + kMaxClasses, kCounters, strm));
        TWOS_TRY(ev_record(0, strm));
        st = launch_walks(rc, int(c), e0s[c], totals[c], kind, class_bare(by_class[c]));
        if (st != TWOS_OK) return st;
        TWOS_TRY(ev_record(1, strm));
        if (dump_a) {
          HIP_OK(hipStreamSynchronize(strm));
          st = dump_round(by_class[c], dump_a, dump_b, dump_steps, dump_stride);
          if (st != TWOS_OK) return st;
        }
        Tables tb = tables(int(c), e0s[c]);
        WalkParams wp{};
        wp.val = d_val;
        wp.steps = d_steps;
        wp.grad = grad_of(rc);
        HIP_OK(launch_reduce(tb, wp, int(by_class[c].size()), strm));
        TWOS_TRY(ev_record(2, strm));
        HIP_OK(hipStreamSynchronize(strm));
        stats.walk_kernel_ms += ev_ms(0, 1);
        stats.reduce_kernel_ms += ev_ms(1, 2);
        if (rc->prm.flags & TWOS_FLAG_COUNT_TRAVERSAL) {
          unsigned long long cnt[kCounters];
          HIP_OK(hipMemcpy(cnt, d_counters + kMaxClasses, sizeof(cnt), hipMemcpyDeviceToHost));
          stats.nodes_visited += int64_t(cnt[0]);
          stats.tris_tested += int64_t(cnt[1]);
          stats.grid_lookups += int64_t(cnt[2]);
          stats.grid_hits += int64_t(cnt[3]);
          stats.top_nodes_visited += int64_t(cnt[4]);
        }
      }
      HIP_OK(hipMemcpyAsync(h_out + off_compact, d_out + off_compact, sizeof(twos_site_result) * n_entries,
                            hipMemcpyDeviceToHost, strm));
      HIP_OK(hipMemcpyAsync(h_out + off_decision, d_out + off_decision, 4 * n_entries, hipMemcpyDeviceToHost,
                            strm));
      HIP_OK(hipStreamSynchronize(strm));
      const auto *res = reinterpret_cast<twos_site_result *>(h_out + off_compact);
      const auto *dec = reinterpret_cast<int32_t *>(h_out + off_decision);
      for (size_t c = 0; c < by_class.size(); ++c)
        for (size_t k = 0; k < by_class[c].size(); ++k)
          finish_entry(by_class[c][k].r, by_class[c][k].i, by_class[c][k].inc, res[e0s[c] + k],
                       dec[e0s[c] + k]);
    }
    stats.rounds += 1;
    // Retire completed requests.
    for (size_t k = 0; k < active.size();) {
      Request *r = active[k];
      if (r->remaining == 0) {
        give_slots(r->slot0, r->n);
        free_problems.push_back(r->problem);
        if (server) charged -= r->n * kSlotBytes;
        r->status = r->capped ? TWOS_ERR_STEP_CAP : TWOS_OK;
        done.push_back(r->id);
        active.erase(active.begin() + k);
      } else {
        ++k;
      }
    }
    return TWOS_OK;
  }

  void finish_entry(Request *r, int64_t i, int32_t inc, const twos_site_result &res, int32_t decision) {
    r->done[i] += inc;
    if (decision != 1) {
      r->state[i] = 1;
      r->remaining -= 1;
      if (r->out) r->out[i] = res;
      if (res.caps > 0) r->capped = true;
      r->steps_sum += res.steps;
      r->paths_sum += res.walks * (r->prm.combine == kCross || r->prm.combine == kCoupled ? 2 : 1);
    }
  }

  template <typename E>
  twos_status dump_round(const std::vector<E> &entries, double *da, double *db, int32_t *ds, int64_t stride) {
    int64_t total = 0;
    for (const E &en : entries) total += en.inc;
    std::vector<double> val(2 * size_t(total));
    std::vector<int32_t> stp(2 * size_t(total));
    HIP_OK(hipMemcpy(val.data(), d_val, sizeof(double) * val.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(stp.data(), d_steps, sizeof(int32_t) * stp.size(), hipMemcpyDeviceToHost));
    int64_t t = 0;
    for (const E &en : entries) {
      for (int32_t w = 0; w < en.inc; ++w, ++t) {
        const int64_t walk = en.r->done[en.i] + w;
        if (walk >= stride) continue;
        const int64_t at = en.i * stride + walk;
        da[at] = val[2 * t];
        db[at] = val[2 * t + 1];
        const int32_t s0 = stp[2 * t], s1 = stp[2 * t + 1];
        ds[at] = ((s0 & 0x3fffffff) + (s1 & 0x3fffffff)) | ((s0 | s1) & (1 << 30));
      }
    }
    return TWOS_OK;
  }
};

// ---------------------------------------------------------------------------
// Validation shared by the estimate and server entry points.

twos_status check_params(const twos_params *p, int64_t n) {
  if (!p) return fail(TWOS_ERR_INVALID_ARGUMENT, "params is null");
  if (n < 0) return fail(TWOS_ERR_INVALID_ARGUMENT, "negative site count");
  if (p->walks < 2) return fail(TWOS_ERR_INVALID_ARGUMENT, "walks must be at least 2");
  if (p->max_steps < 1) return fail(TWOS_ERR_INVALID_ARGUMENT, "max_steps must be at least 1");
  if (!(p->eps > 0.0)) return fail(TWOS_ERR_INVALID_ARGUMENT, "eps must be positive");
  if (p->combine < 0 || p->combine > 3) return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown combine mode");
  if (p->rng < 0 || p->rng > 4) return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown stream mode");
  if (p->rng != kRngPhilox && (p->rng == kRngMtFd) != (p->combine == kCoupled))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "COUPLED replays the MT_FD layout, and MT_FD belongs to COUPLED");
  if (p->combine == kCoupled && (p->flags & TWOS_FLAG_GRADIENT))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "the gradient channel does not apply to COUPLED");
  if (p->precision != 32 && p->precision != 64) return fail(TWOS_ERR_INVALID_ARGUMENT, "precision must be 32 or 64");
  if (p->rng != kRngPhilox && p->precision != 64)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "mt19937_64 replay streams run in binary64 only");
  if (p->rng != kRngPhilox && p->combine == kCross)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "mt19937_64 replay streams have no independent adjoint stream");
  if (p->rng == kRngMtKey) return fail(TWOS_ERR_INVALID_ARGUMENT, "MT_KEY streams are for twos_trace");
  if (p->schedule != TWOS_PLAIN && p->schedule != TWOS_PERSISTENT)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown schedule");
  if (p->walks_max > p->walks && !(p->ci_rel >= 0.0))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "adaptive budgets need ci_rel >= 0");
  if (p->block != 0 && (p->block % 64 != 0 || p->block > 256 || p->block < 64))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "block must be 64, 128, 192 or 256");
  if (p->lds_stack != 0 && p->lds_stack != 16 && p->lds_stack != 24 && p->lds_stack != 32)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "lds_stack must be 0, 16, 24 or 32");
  if (!(p->grid_kappa >= 0.f && p->grid_kappa < 1e6f))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "grid_kappa must be a non-negative number of cells");
  if (p->waves_per_simd < 0 || p->waves_per_simd > 16)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "waves_per_simd must be in [0, 16]");
  if (p->step_batch != 0 && p->step_batch != 1 && p->step_batch != 24 && p->step_batch != 32 && p->step_batch != 40)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "step_batch must be 0, 1, 24, 32 or 40");
  if (p->refill_batch < 0 || p->refill_batch > 64)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "refill_batch must be in [0, 64]");
  if (p->step_batch != 0 && ((p->lds_stack != 0 && p->lds_stack != 24) || (p->waves_per_simd != 0 &&
                                                                            p->waves_per_simd != 5)))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "step_batch is built for lds_stack 24 with waves_per_simd 0 or 5");
  return TWOS_OK;
}

twos_status check_side(const twos_side &s) {
  if (s.source_kind < 0 || s.source_kind > 3) return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown source kind");
  if (s.source_kind == 3)
    for (int k = 0; k < s.n_terms && k < TWOS_MAX_TERMS; ++k)
      if (s.terms[k][0] == TWOS_TERM_GAUSS)
        return fail(TWOS_ERR_INVALID_ARGUMENT, "the analytic source integrates BUMP and UNIT rows only");
  if (s.source_kind >= 2) {
    if (s.n_terms < 1 || s.n_terms > TWOS_MAX_TERMS)
      return fail(TWOS_ERR_INVALID_ARGUMENT, "a sampled source needs 1..8 term rows");
    for (int k = 0; k < s.n_terms; ++k) {
      const double *t = s.terms[k];
      if (t[0] != TWOS_TERM_BUMP && t[0] != TWOS_TERM_UNIT && t[0] != TWOS_TERM_GAUSS)
        return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown source term type");
      if (!(t[4] > 0.0) || (t[0] == TWOS_TERM_UNIT && !(t[5] > 0.0)) ||
          (t[0] == TWOS_TERM_GAUSS && !(t[5] > 0.0 && t[6] > 0.0)))
        return fail(TWOS_ERR_INVALID_ARGUMENT, "source term widths must be positive");
    }
  }
  return TWOS_OK;
}

Request *make_request(twos_ctx *ctx, int32_t geom, const twos_side sides[2], const twos_params *p,
                      const double *sites, int64_t n, twos_site_result *out, twos_status *st,
                      const double *d_sites = nullptr) {
  *st = TWOS_OK;
  if (ctx->geoms.find(geom) == ctx->geoms.end()) {
    *st = fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
    return nullptr;
  }
  if (!sides) {
    *st = fail(TWOS_ERR_INVALID_ARGUMENT, "sides is null");
    return nullptr;
  }
  for (int k = 0; k < 2; ++k)
    if ((*st = check_side(sides[k])) != TWOS_OK) return nullptr;
  if ((p->flags & TWOS_FLAG_GRADIENT) && (sides[0].source_kind >= 2 || sides[1].source_kind >= 2)) {
    *st = fail(TWOS_ERR_INVALID_ARGUMENT, "the gradient channel needs an affine source or none");
    return nullptr;
  }
  if (p->combine == kCoupled) {
    const twos_side &q = sides[1];
    if (q.source_kind < 2 || q.n_terms != 1 || q.terms[0][0] != TWOS_TERM_UNIT) {
      *st = fail(TWOS_ERR_INVALID_ARGUMENT, "COUPLED needs sides[1] with one TWOS_TERM_UNIT row (the probe ball)");
      return nullptr;
    }
    if (!(ctx->geoms.at(geom).desc.cavity_radius > 0.0)) {
      *st = fail(TWOS_ERR_INVALID_ARGUMENT, "COUPLED needs a geometry with a positive cavity_radius");
      return nullptr;
    }
  }
  if (d_sites) {
    // Device positions stay on the device: a non-finite one is not caught
    // here, its paths end capped (no query from a non-finite point).
    if (n > 0 && out) {
      *st = fail(TWOS_ERR_INVALID_ARGUMENT, "device requests write their results on the device");
      return nullptr;
    }
  } else {
    if (n > 0 && (!sites || !out)) {
      *st = fail(TWOS_ERR_INVALID_ARGUMENT, "sites and out must be non-null");
      return nullptr;
    }
    for (int64_t i = 0; i < 3 * n; ++i) {
      if (!std::isfinite(sites[i])) {
        *st = fail(TWOS_ERR_INVALID_ARGUMENT, "non-finite site coordinate");
        return nullptr;
      }
    }
  }
  auto *r = new Request;
  r->geom = geom;
  r->prm = *p;
  r->sides[0] = sides[0];
  r->sides[1] = sides[1];
  if (d_sites) r->d_sites = d_sites;
  else r->sites.assign(sites, sites + 3 * n);
  if (!d_sites && n >= 2 * kWalkBlockSites) r->order = morton_order(sites, n);
  r->n = n;
  r->out = out;
  return r;
}

// Context-level helpers.
int64_t ctx_geometry_bytes(const twos_ctx *ctx) {
  int64_t b = int64_t(ctx->d_geoms_cap) * int64_t(sizeof(DevGeom));
  for (const auto &kv : ctx->meshes) b += kv.second.bytes + kv.second.mb.grid_bytes + kv.second.mb.cert_bytes;
  for (const auto &kv : ctx->geoms) b += kv.second.owned_bytes;
  return b;
}

int64_t ctx_engine_bytes(const twos_ctx *ctx) {
  int64_t b = 0;
  for (const Engine *e : ctx->engines)
    b += e->fixed_bytes() + (e->server ? e->charged : e->cap_sites * kSlotBytes);
  return b;
}

int64_t ctx_device_bytes(const twos_ctx *ctx) { return ctx_geometry_bytes(ctx) + ctx_engine_bytes(ctx); }

// Captured rounds hold the geometry table's address and the grids it points
// to: every engine on the context must capture again.
void invalidate_graphs(twos_ctx *ctx) {
  for (Engine *e : ctx->engines) e->graph_key = ~0ull;
}

twos_status upload_geom_table(twos_ctx *ctx) {
  int32_t need = ctx->next_geom;
  if (need == 0) return TWOS_OK;  // no geometry yet (a mesh's grid may be rebuilt before any)
  if (need > ctx->d_geoms_cap) {
    int32_t cap = std::max(16, ctx->d_geoms_cap * 2);
    while (cap < need) cap *= 2;
    DevGeom *p = nullptr;
    HIP_OK(dalloc(&p, size_t(cap)));
    // Rounds of any engine already queued on the stream may read the old table.
    HIP_OK(hipStreamSynchronize(ctx->stream));
    dfree(ctx->d_geoms);
    ctx->d_geoms = p;
    ctx->d_geoms_cap = cap;
    invalidate_graphs(ctx);  // tables() pointer changed
  }
  std::vector<DevGeom> table(size_t(std::max(need, 1)));
  for (auto &kv : ctx->geoms) table[size_t(kv.first)] = kv.second.dev;
  HIP_OK(hipMemcpy(ctx->d_geoms, table.data(), sizeof(DevGeom) * table.size(), hipMemcpyHostToDevice));
  return TWOS_OK;
}

template <typename T>
twos_status upload_owned(GeomEntry &g, const T *host, size_t count, T **dev) {
  *dev = nullptr;
  if (count == 0) return TWOS_OK;
  HIP_OK(dalloc(dev, count));
  g.owned.push_back(*dev);
  g.owned_bytes += int64_t(sizeof(T) * count);
  HIP_OK(hipMemcpy(*dev, host, sizeof(T) * count, hipMemcpyHostToDevice));
  return TWOS_OK;
}

twos_status build_geom(twos_ctx *ctx, GeomEntry &g) {
  const twos_geometry_desc &d = g.desc;
  DevGeom &v = g.dev;
  v = DevGeom{};
  v.container = d.container;
  v.n_spheres = d.n_spheres;
  v.n_cylinders = d.n_cylinders;
  v.n_cavities = d.n_cavities;
  v.mesh_radius_f32 = d.mesh_radius_f32;
  for (int k = 0; k < 4; ++k) v.ball[k] = d.ball[k];
  for (int k = 0; k < 3; ++k) { v.box_lo[k] = d.box_lo[k]; v.box_hi[k] = d.box_hi[k]; }
  v.cavity_radius = d.cavity_radius;
  std::vector<float4> sf(g.spheres.size() / 4), cf(g.cylinders.size() / 4), vf(g.cavities.size() / 3);
  std::vector<double> cav4(4 * vf.size());
  for (size_t i = 0; i < sf.size(); ++i)
    sf[i] = make_float4(float(g.spheres[4 * i]), float(g.spheres[4 * i + 1]), float(g.spheres[4 * i + 2]),
                        float(g.spheres[4 * i + 3]));
  for (size_t i = 0; i < cf.size(); ++i)
    cf[i] = make_float4(float(g.cylinders[4 * i]), float(g.cylinders[4 * i + 1]), float(g.cylinders[4 * i + 2]),
                        float(g.cylinders[4 * i + 3]));
  for (size_t i = 0; i < vf.size(); ++i) {
    vf[i] = make_float4(float(g.cavities[3 * i]), float(g.cavities[3 * i + 1]), float(g.cavities[3 * i + 2]), 0.f);
    for (int k = 0; k < 3; ++k) cav4[4 * i + k] = g.cavities[3 * i + k];
    cav4[4 * i + 3] = 0.0;
  }
  twos_status st;
  float4 *p4 = nullptr;
  double *pd = nullptr;
  if ((st = upload_owned(g, sf.data(), sf.size(), &p4)) != TWOS_OK) return st;
  v.spheres_f = p4;
  if ((st = upload_owned(g, cf.data(), cf.size(), &p4)) != TWOS_OK) return st;
  v.cylinders_f = p4;
  if ((st = upload_owned(g, vf.data(), vf.size(), &p4)) != TWOS_OK) return st;
  v.cavities_f = p4;
  if ((st = upload_owned(g, g.spheres.data(), g.spheres.size(), &pd)) != TWOS_OK) return st;
  v.spheres_d = pd;
  if ((st = upload_owned(g, g.cylinders.data(), g.cylinders.size(), &pd)) != TWOS_OK) return st;
  v.cylinders_d = pd;
  if ((st = upload_owned(g, cav4.data(), cav4.size(), &pd)) != TWOS_OK) return st;
  v.cavities_d = pd;
  if (d.container == kMesh) {
    auto it = ctx->meshes.find(d.mesh);
    if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
    const MeshBuffers &mb = it->second.mb;
    v.nodes = mb.nodes;
    v.tris_f = mb.tris_f;
    v.tris_d = mb.tris_d;
    v.tri_ids = mb.tri_ids;
    v.nodes4 = mb.nodes4;
    v.grid = mb.grid;
    for (int k = 0; k < 3; ++k) {
      v.grid_lo[k] = mb.grid_lo[k];
      v.grid_n[k] = mb.grid_n[k];
    }
    v.grid_h = mb.grid_h;
    v.grid_inv_h = mb.grid_h > 0.f ? 1.0f / mb.grid_h : 0.f;
    v.grid_corners = mb.grid_corners;
    v.grid_start = mb.grid_start;
    v.grid_start_r2 = mb.grid_start_r2;
    v.n_tris = mb.n_tris;
    it->second.users += 1;
  }
  g.kind = d.container == kMesh ? (g.spheres.empty() && g.cylinders.empty() ? int(kGeomMesh) : int(kGeomGeneric))
                                : int(kGeomAnalytic);
  return TWOS_OK;
}

void free_geom(twos_ctx *ctx, GeomEntry &g) {
  for (void *p : g.owned) (void)hipFree(p);
  g.owned.clear();
  if (g.desc.container == kMesh) {
    auto it = ctx->meshes.find(g.desc.mesh);
    if (it != ctx->meshes.end()) it->second.users -= 1;
  }
}

twos_status ensure_engine(twos_ctx *ctx, const twos_params *p, int64_t n) {
  const int64_t walks_max = std::max(p->walks, p->walks_max);
  const int64_t want_sites = std::min<int64_t>(std::max<int64_t>(n, 1), 1 << 20);
  const int64_t limit_walks = int64_t(ctx->budget / 2 / 24);
  int64_t want_walks = std::min<int64_t>(want_sites * walks_max, limit_walks);
  want_walks = std::max<int64_t>(want_walks, walks_max);
  Engine *e = ctx->engine.get();
  if (e && e->cap_sites >= want_sites && e->cap_walks >= want_walks) return TWOS_OK;
  auto ne = std::make_unique<Engine>();
  const int64_t sites = std::max<int64_t>(want_sites, e ? e->cap_sites : 0);
  const int64_t walks = std::max<int64_t>(want_walks, e ? e->cap_walks : 0);
  ctx->engine.reset();
  ne->bind(ctx);
  const twos_status st = ne->allocate(sites, walks, 64);
  if (st != TWOS_OK) return st;
  ctx->engine = std::move(ne);
  return TWOS_OK;
}

// Sites and results in device memory, ordered on the caller's stream.
struct DeviceIo {
  const double *sites = nullptr;    // [n][3]
  twos_site_result *out = nullptr;  // [n]
  hipStream_t stream = nullptr;
};

// One estimate on the context's own engine, or on its arena engine (noalloc:
// every device buffer in the device's static scratch array, nothing
// allocated).  With `dio`, sites and results stay on the device: the
// context's stream first waits for the work queued on dio->stream, each
// chunk's results are copied device to device from the slots, and
// dio->stream then waits for those copies.
twos_status run_estimate(twos_ctx *ctx, int32_t geom, const twos_side sides[2], const twos_params *p,
                         const double *sites, int64_t n, twos_site_result *out, twos_stats *stats, double *da,
                         double *db, int32_t *ds, int64_t stride, const DeviceIo *dio = nullptr,
                         bool noalloc = false) {
  twos_status st = check_params(p, n);
  if (st != TWOS_OK) return st;
  if (dio && n > 0 && (!dio->sites || !dio->out)) return fail(TWOS_ERR_INVALID_ARGUMENT, "device sites and results");
  if (n == 0) {
    if (stats) *stats = twos_stats{};
    return TWOS_OK;
  }
  const double t0 = now_s();
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_OK(hipSetDevice(ctx->device));
  Engine *e;
  if (noalloc) {
    e = ctx->arena_engine.get();
    if (!e) return fail(TWOS_ERR_DEVICE, "no static scratch array on this device");
    if (p->use_graph || (p->flags & TWOS_FLAG_GRADIENT))
      return fail(TWOS_ERR_INVALID_ARGUMENT, "the allocation-free estimate runs without graphs and gradients");
    if (std::max(p->walks, p->walks_max) > e->cap_walks)
      return fail(TWOS_ERR_ADMISSION, "walks per side above the static scratch array's " +
                                          std::to_string(e->cap_walks) + " walks per round");
  } else {
    st = ensure_engine(ctx, p, n);
    if (st != TWOS_OK) return st;
    e = ctx->engine.get();
  }
  if (!e->idle()) return fail(TWOS_ERR_INVALID_ARGUMENT, "context engine busy");
  e->stats = twos_stats{};
  if (dio) {
    HIP_OK(hipEventRecord(ctx->ev_in, dio->stream));
    HIP_OK(hipStreamWaitEvent(ctx->stream, ctx->ev_in, 0));
  }
  // Chunk the sites so each chunk fits the engine's slots; the global site
  // index (site_offset + i) keeps every stream identical to one big call.
  const int64_t chunk = e->cap_sites;
  bool capped = false;
  int64_t steps = 0;
  e->stats.paths = 0;
  for (int64_t b = 0; b < n; b += chunk) {
    const int64_t m = std::min(chunk, n - b);
    twos_params q = *p;
    q.site_offset = p->site_offset + int32_t(b);
    Request *r = dio ? make_request(ctx, geom, sides, &q, nullptr, m, nullptr, &st, dio->sites + 3 * b)
                     : make_request(ctx, geom, sides, &q, sites + 3 * b, m, out + b, &st);
    if (!r) return st;
    std::unique_ptr<Request> hold(r);
    r->id = e->next_id++;
    e->queue.push_back(r);
    std::vector<int64_t> done;
    while (!e->idle()) {
      st = e->step(done, da ? da + b * stride : nullptr, db ? db + b * stride : nullptr, ds ? ds + b * stride : nullptr,
                   stride);
      if (st != TWOS_OK) {
        e->reset_state();
        return st;
      }
    }
    // The slots keep each site's last result until the next chunk is admitted.
    if (dio)
      HIP_OK(hipMemcpyAsync(dio->out + b, e->s.result + r->slot0, sizeof(twos_site_result) * size_t(m),
                            hipMemcpyDeviceToDevice, ctx->stream));
    capped = capped || r->capped;
    steps += r->steps_sum;
    e->stats.paths += r->paths_sum;
  }
  if (dio) {
    HIP_OK(hipEventRecord(ctx->ev_out, ctx->stream));
    HIP_OK(hipStreamWaitEvent(dio->stream, ctx->ev_out, 0));
  }
  e->stats.seconds = now_s() - t0;
  e->stats.steps = steps;
  if (stats) *stats = e->stats;
  if (capped) return fail(TWOS_ERR_STEP_CAP, "one or more paths hit --max-steps");
  return TWOS_OK;
}

}  // namespace twos

namespace twos {
namespace {

// Device scratch of the allocation-free estimate (twos_estimate_device_static):
// every buffer of its engine (site slots, round blocks, per-walk samples) lives
// here.  Each device holds its own copy from module load to process end, and
// every context on a device shares it.  In host code `scratch` names a host
// stand-in: only the address from hipGetSymbolAddress, kept per device in
// g_scratch, may be used.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int64_t kScratchSites = 4096;  // sites per chunk
__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// The device's scratch address; the device must be current.
twos_status scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(TWOS_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(TWOS_ERR_DEVICE, "hipGetSymbolAddress gave no scratch address");
    g_scratch[device] = ptr;
  }
  *out = g_scratch[device];
  return TWOS_OK;
}

}  // namespace
}  // namespace twos

struct twos_server {
  twos_ctx *ctx = nullptr;
  std::unique_ptr<twos::Engine> engine;
  std::map<int64_t, std::unique_ptr<twos::Request>> requests;
};

using namespace twos;

extern "C" {

const char *twos_version(void) { return "twos-mi355x 0.1 (gfx950)"; }

const char *twos_last_error(void) { return g_error.c_str(); }

twos_status twos_create(int32_t device, uint64_t hbm_budget_bytes, twos_ctx **out) {
  if (!out) return fail(TWOS_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(TWOS_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(TWOS_ERR_INVALID_ARGUMENT, "device index out of range");
  auto ctx = std::make_unique<twos_ctx>();
  ctx->device = device;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipGetDeviceProperties(&ctx->prop, device));
  HIP_OK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  size_t free_b = 0, total_b = 0;
  HIP_OK(hipMemGetInfo(&free_b, &total_b));
  ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : uint64_t(free_b / 4);
  HIP_OK(hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming));
  // The allocation-free entry point's engine, laid out once in the device's
  // static scratch array.
  void *arena = nullptr;
  twos_status st = scratch_address(device, &arena);
  if (st != TWOS_OK) return st;
  ctx->arena_engine = std::make_unique<Engine>();
  ctx->arena_engine->ctx = ctx.get();
  st = ctx->arena_engine->allocate_arena(static_cast<char *>(arena), kScratchBytes, kScratchSites);
  if (st != TWOS_OK) return st;
  *out = ctx.release();
  return TWOS_OK;
}

void twos_destroy(twos_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->engine.reset();
  ctx->arena_engine.reset();
  if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
  if (ctx->ev_out) (void)hipEventDestroy(ctx->ev_out);
  // Servers still alive lose their device buffers; their handles only accept
  // twos_server_destroy from now on.
  for (Engine *e : ctx->engines) {
    e->release();
    e->ctx = nullptr;
  }
  ctx->engines.clear();
  for (auto &kv : ctx->geoms) free_geom(ctx, kv.second);
  for (auto &kv : ctx->meshes) {
    MeshBuffers &mb = kv.second.mb;
    (void)hipFree(mb.nodes); (void)hipFree(mb.nodes4); (void)hipFree(mb.grid); (void)hipFree(mb.grid_start); (void)hipFree(mb.tris_f); (void)hipFree(mb.tris_d); (void)hipFree(mb.tri_ids); (void)hipFree(const_cast<float *>(mb.cert.d));
  }
  (void)hipFree(ctx->d_geoms);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

twos_status twos_mesh_create(twos_ctx *ctx, const double *vertices, int64_t nv, const uint32_t *faces, int64_t nf,
                             int32_t *mesh_id, double *build_ms) {
  if (!ctx || !vertices || !faces || !mesh_id) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (nv <= 0 || nf <= 0) return fail(TWOS_ERR_INVALID_ARGUMENT, "expected a nonempty triangle mesh");
  // Traversal stack entries hold 21-bit 4-wide node indices.
  if (nf > int64_t(kMaxStackIndex) + 1)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "meshes are limited to 2^21 triangles");
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (faces[i] >= uint64_t(nv)) return fail(TWOS_ERR_INVALID_ARGUMENT, "expected valid triangular faces");
  for (int64_t i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(vertices[i])) return fail(TWOS_ERR_INVALID_ARGUMENT, "non-finite vertex");
  // Content hash of the exact bit patterns; a hit is confirmed by comparing
  // the arrays themselves.
  uint64_t hash = splitmix64(uint64_t(nv) * 0x9E3779B97F4A7C15ULL ^ uint64_t(nf));
  for (int64_t i = 0; i < 3 * nv; ++i) {
    uint64_t bits;
    std::memcpy(&bits, vertices + i, sizeof bits);
    hash = splitmix64(hash ^ bits);
  }
  for (int64_t i = 0; i < 3 * nf; i += 2)
    hash = splitmix64(hash ^ (uint64_t(faces[i]) | (i + 1 < 3 * nf ? uint64_t(faces[i + 1]) << 32 : 0)));
  std::lock_guard<std::mutex> lock(ctx->mu);
  for (auto &kv : ctx->meshes) {
    MeshEntry &old = kv.second;
    if (old.hash == hash && int64_t(old.vertices.size()) == 3 * nv && int64_t(old.faces.size()) == 3 * nf &&
        std::memcmp(old.vertices.data(), vertices, sizeof(double) * 3 * size_t(nv)) == 0 &&
        std::memcmp(old.faces.data(), faces, sizeof(uint32_t) * 3 * size_t(nf)) == 0) {
      ++old.handles;
      *mesh_id = kv.first;
      if (build_ms) *build_ms = 0.0;
      return TWOS_OK;
    }
  }
  HIP_OK(hipSetDevice(ctx->device));
  double *d_v = nullptr;
  uint32_t *d_f = nullptr;
  HIP_OK(dalloc(&d_v, size_t(3 * nv)));
  HIP_OK(dalloc(&d_f, size_t(3 * nf)));
  HIP_OK(hipMemcpy(d_v, vertices, sizeof(double) * 3 * nv, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_f, faces, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice));
  MeshEntry m;
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, ctx->stream));
  hipError_t e = lbvh_build(d_v, d_f, int32_t(nf), m.mb, &m.depth, ctx->stream);
  if (e == hipSuccess) {
    double lo[3] = {vertices[0], vertices[1], vertices[2]}, hi[3] = {vertices[0], vertices[1], vertices[2]};
    for (int64_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], vertices[3 * i + k]);
        hi[k] = std::max(hi[k], vertices[3 * i + k]);
      }
    e = build_distance_grid(m.mb, lo, hi, kGridCells, 0, ctx->stream);
    for (int k = 0; k < 3; ++k) { m.lo[k] = lo[k]; m.hi[k] = hi[k]; }
    m.grid_cells = kGridCells;
  }
  (void)hipEventRecord(b, ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(d_v);
  (void)hipFree(d_f);
  if (e != hipSuccess) return dev_fail("LBVH build", e);
  m.build_ms = ms;
  m.bytes = int64_t(nf > 1 ? nf - 1 : 1) * (64 + 128) + nf * (36 + 72 + 4);
  // The 4-wide traversal stacks at most three siblings per 4-wide level.
  if (3 * ((m.depth - 1 + 1) / 2) >= kStackDepth) {
    (void)hipFree(m.mb.nodes); (void)hipFree(m.mb.nodes4); (void)hipFree(m.mb.grid); (void)hipFree(m.mb.grid_start); (void)hipFree(m.mb.tris_f); (void)hipFree(m.mb.tris_d); (void)hipFree(m.mb.tri_ids); (void)hipFree(const_cast<float *>(m.mb.cert.d));
    return fail(TWOS_ERR_GEOMETRY, "LBVH deeper than the traversal stack");
  }
  m.hash = hash;
  m.vertices.assign(vertices, vertices + 3 * nv);
  m.faces.assign(faces, faces + 3 * nf);
  const int32_t id = ctx->next_mesh++;
  ctx->meshes[id] = std::move(m);
  *mesh_id = id;
  if (build_ms) *build_ms = ms;
  return TWOS_OK;
}

twos_status twos_mesh_set_distance_grid(twos_ctx *ctx, int32_t mesh_id, int32_t cells, int32_t corners,
                                        double *build_ms) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  if (cells < 8 || cells > 1024) return fail(TWOS_ERR_INVALID_ARGUMENT, "cells must be in [8, 1024]");
  if (corners < 0 || corners > 2) return fail(TWOS_ERR_INVALID_ARGUMENT, "grid mode must be 0, 1 or 2");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->meshes.find(mesh_id);
  if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
  MeshEntry &m = it->second;
  if (m.grid_cells == cells && m.mb.grid_corners == corners && m.mb.grid) {
    if (build_ms) *build_ms = 0.0;  // that grid is resident already
    return TWOS_OK;
  }
  HIP_OK(hipSetDevice(ctx->device));
  // Work already queued may read the old grid: finish it before the swap.
  HIP_OK(hipStreamSynchronize(ctx->stream));
  MeshBuffers nb = m.mb;
  nb.grid = nullptr;
  nb.grid_start = nullptr;
  nb.grid_start_r2 = 0.f;
  nb.grid_bytes = 0;
  nb.cell_nodes = 0;
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, ctx->stream));
  const hipError_t e = build_distance_grid(nb, m.lo, m.hi, cells, corners, ctx->stream);
  (void)hipEventRecord(b, ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) {
    if (nb.grid) (void)hipFree(nb.grid);
    if (nb.grid_start) (void)hipFree(nb.grid_start);
    return dev_fail("distance grid", e);
  }
  (void)hipFree(m.mb.grid);
  if (m.mb.grid_start) (void)hipFree(m.mb.grid_start);
  // A record grid's build copies the 4-wide nodes into an array with its cell
  // nodes appended; the old array goes.
  if (nb.nodes4 != m.mb.nodes4) (void)hipFree(m.mb.nodes4);
  m.mb = nb;
  m.grid_cells = cells;
  for (auto &kv : ctx->geoms) {
    GeomEntry &g = kv.second;
    if (g.desc.container != kMesh || g.desc.mesh != mesh_id) continue;
    g.dev.nodes4 = nb.nodes4;
    g.dev.grid = nb.grid;
    for (int k = 0; k < 3; ++k) {
      g.dev.grid_lo[k] = nb.grid_lo[k];
      g.dev.grid_n[k] = nb.grid_n[k];
    }
    g.dev.grid_h = nb.grid_h;
    g.dev.grid_inv_h = nb.grid_h > 0.f ? 1.0f / nb.grid_h : 0.f;
    g.dev.grid_corners = nb.grid_corners;
    g.dev.grid_start = nb.grid_start;
    g.dev.grid_start_r2 = nb.grid_start_r2;
  }
  invalidate_graphs(ctx);
  if (build_ms) *build_ms = ms;
  return upload_geom_table(ctx);
}

twos_status twos_mesh_info(twos_ctx *ctx, int32_t mesh_id, int32_t *depth, int64_t *nodes, int64_t *bytes) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  auto it = ctx->meshes.find(mesh_id);
  if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
  if (depth) *depth = it->second.depth;
  if (nodes) *nodes = std::max<int64_t>(it->second.mb.n_tris - 1, 1);
  if (bytes) *bytes = it->second.bytes;
  return TWOS_OK;
}

twos_status twos_memory_info(twos_ctx *ctx, int64_t *geometry_bytes, int64_t *engine_bytes, int64_t *budget_bytes,
                             int64_t *slot_bytes) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (geometry_bytes) *geometry_bytes = ctx_geometry_bytes(ctx);
  if (engine_bytes) *engine_bytes = ctx_engine_bytes(ctx);
  if (budget_bytes) *budget_bytes = int64_t(ctx->budget);
  if (slot_bytes) *slot_bytes = kSlotBytes;
  return TWOS_OK;
}

twos_status twos_mesh_release(twos_ctx *ctx, int32_t mesh_id) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->meshes.find(mesh_id);
  if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
  if (it->second.handles > 1) {
    --it->second.handles;  // another create of the same mesh still holds it
    return TWOS_OK;
  }
  if (it->second.users > 0) return fail(TWOS_ERR_INVALID_ARGUMENT, "mesh still used by a geometry");
  MeshBuffers &mb = it->second.mb;
  (void)hipFree(mb.nodes); (void)hipFree(mb.nodes4); (void)hipFree(mb.grid); (void)hipFree(mb.grid_start); (void)hipFree(mb.tris_f); (void)hipFree(mb.tris_d); (void)hipFree(mb.tri_ids); (void)hipFree(const_cast<float *>(mb.cert.d));
  ctx->meshes.erase(it);
  return TWOS_OK;
}

twos_status twos_geometry_create(twos_ctx *ctx, const twos_geometry_desc *desc, int32_t *geom_id) {
  if (!ctx || !desc || !geom_id) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  const twos_geometry_desc &d = *desc;
  if (d.container < 0 || d.container > 2) return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown container");
  if (d.n_spheres < 0 || d.n_cylinders < 0 || d.n_cavities < 0)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "negative primitive count");
  if ((d.n_spheres && !d.spheres) || (d.n_cylinders && !d.cylinders) || (d.n_cavities && !d.cavities))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "missing primitive rows");
  if (d.n_cavities && !(d.cavity_radius > 0.0))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "cavities need a positive cavity_radius");
  if (d.container == kBall && !(d.ball[3] > 0.0)) return fail(TWOS_ERR_INVALID_ARGUMENT, "ball radius must be positive");
  if (d.container == kBox)
    for (int k = 0; k < 3; ++k)
      if (!(d.box_hi[k] > d.box_lo[k])) return fail(TWOS_ERR_INVALID_ARGUMENT, "empty box");
  for (int i = 0; i < d.n_cylinders; ++i) {
    const double axis = d.cylinders[4 * i];
    if (axis != 0.0 && axis != 1.0 && axis != 2.0) return fail(TWOS_ERR_INVALID_ARGUMENT, "cylinder axis must be 0, 1 or 2");
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_OK(hipSetDevice(ctx->device));
  GeomEntry g;
  g.desc = d;
  g.spheres.assign(d.spheres, d.spheres + 4 * size_t(d.n_spheres));
  g.cylinders.assign(d.cylinders, d.cylinders + 4 * size_t(d.n_cylinders));
  g.cavities.assign(d.cavities, d.cavities + 3 * size_t(d.n_cavities));
  g.desc.spheres = g.desc.cylinders = g.desc.cavities = nullptr;
  twos_status st = build_geom(ctx, g);
  if (st != TWOS_OK) {
    free_geom(ctx, g);
    return st;
  }
  const int32_t id = ctx->next_geom++;
  ctx->geoms[id] = std::move(g);
  st = upload_geom_table(ctx);
  if (st != TWOS_OK) return st;
  *geom_id = id;
  return TWOS_OK;
}

twos_status twos_geometry_union_sphere(twos_ctx *ctx, int32_t base, const double center[3], double radius,
                                       int32_t *geom_id) {
  if (!ctx || !center || !geom_id) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (!(radius > 0.0)) return fail(TWOS_ERR_INVALID_ARGUMENT, "radius must be positive");
  twos_geometry_desc d;
  std::vector<double> sp, cy, cv;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto it = ctx->geoms.find(base);
    if (it == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
    d = it->second.desc;
    sp = it->second.spheres;
    cy = it->second.cylinders;
    cv = it->second.cavities;
  }
  // A cavity of another radius than the base's cavities becomes an analytic
  // sphere hole (class 6 data); a base without cavities takes it as its cavity.
  if (d.n_cavities == 0 || d.cavity_radius == radius) {
    cv.insert(cv.end(), center, center + 3);
    d.n_cavities += 1;
    d.cavity_radius = radius;
  } else {
    sp.insert(sp.end(), {center[0], center[1], center[2], radius});
    d.n_spheres += 1;
  }
  d.spheres = sp.data();
  d.cylinders = cy.data();
  d.cavities = cv.data();
  return twos_geometry_create(ctx, &d, geom_id);
}

twos_status twos_geometry_release(twos_ctx *ctx, int32_t geom_id) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->geoms.find(geom_id);
  if (it == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  for (const Engine *e : ctx->engines)
    if (e->uses_geometry(geom_id))
      return fail(TWOS_ERR_INVALID_ARGUMENT, "geometry in use by a queued or active request");
  free_geom(ctx, it->second);
  ctx->geoms.erase(it);
  return upload_geom_table(ctx);
}

twos_status twos_estimate(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                          const double *sites, int64_t n, twos_site_result *out, twos_stats *stats) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  return run_estimate(ctx, geom_id, sides, params, sites, n, out, stats, nullptr, nullptr, nullptr, 0);
}

twos_status twos_estimate_device(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                                 const double *d_sites, int64_t n, twos_site_result *d_out, twos_stats *stats,
                                 void *stream) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  const DeviceIo dio{d_sites, d_out, static_cast<hipStream_t>(stream)};
  return run_estimate(ctx, geom_id, sides, params, nullptr, n, nullptr, stats, nullptr, nullptr, nullptr, 0, &dio);
}

twos_status twos_estimate_device_static(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2],
                                        const twos_params *params, const double *d_sites, int64_t n,
                                        twos_site_result *d_out, twos_stats *stats, void *stream) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  const DeviceIo dio{d_sites, d_out, static_cast<hipStream_t>(stream)};
  return run_estimate(ctx, geom_id, sides, params, nullptr, n, nullptr, stats, nullptr, nullptr, nullptr, 0, &dio,
                      true);
}

twos_status twos_estimate_dump(twos_ctx *ctx, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                               const double *sites, int64_t n, twos_site_result *out, twos_stats *stats,
                               double *dump_a, double *dump_b, int32_t *dump_steps, int64_t stride) {
  if (!ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "null context");
  if (!dump_a || !dump_b || !dump_steps || stride <= 0) return fail(TWOS_ERR_INVALID_ARGUMENT, "dump buffers");
  if (params && params->use_graph) return fail(TWOS_ERR_INVALID_ARGUMENT, "dumps run without graphs");
  return run_estimate(ctx, geom_id, sides, params, sites, n, out, stats, dump_a, dump_b, dump_steps, stride);
}

twos_status twos_inside(twos_ctx *ctx, int32_t mesh_id, const double *points, int64_t n, int32_t *out) {
  if (!ctx || (n > 0 && (!points || !out))) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (n <= 0) return TWOS_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->meshes.find(mesh_id);
  if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
  HIP_OK(hipSetDevice(ctx->device));
  double *d_p = nullptr;
  int32_t *d_o = nullptr;
  HIP_OK(dalloc(&d_p, size_t(3 * n)));
  HIP_OK(dalloc(&d_o, size_t(n)));
  hipError_t e = hipMemcpyAsync(d_p, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_inside(it->second.mb, d_p, n, d_o, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_o, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_p);
  (void)hipFree(d_o);
  if (e != hipSuccess) return dev_fail("inside", e);
  for (int64_t i = 0; i < n; ++i)
    if (out[i] < 0) return fail(TWOS_ERR_GEOMETRY, "ray-parity traversal exceeded 128 intersections");
  return TWOS_OK;
}

twos_status twos_closest(twos_ctx *ctx, int32_t geom_id, int32_t precision, const double *points, int64_t n,
                         double *distance, double *closest, int32_t *cls) {
  if (!ctx || (n > 0 && (!points || !distance || !closest || !cls)))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (precision != 32 && precision != 64) return fail(TWOS_ERR_INVALID_ARGUMENT, "precision must be 32 or 64");
  if (n <= 0) return TWOS_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->geoms.find(geom_id) == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  HIP_OK(hipSetDevice(ctx->device));
  double *d_p = nullptr, *d_d = nullptr, *d_c = nullptr;
  int32_t *d_k = nullptr;
  HIP_OK(dalloc(&d_p, size_t(3 * n)));
  HIP_OK(dalloc(&d_d, size_t(n)));
  HIP_OK(dalloc(&d_c, size_t(3 * n)));
  HIP_OK(dalloc(&d_k, size_t(n)));
  hipError_t e = hipMemcpyAsync(d_p, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_closest(ctx->d_geoms + geom_id, precision, d_p, n, d_d, d_c, d_k, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(distance, d_d, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(closest, d_c, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(cls, d_k, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_p); (void)hipFree(d_d); (void)hipFree(d_c); (void)hipFree(d_k);
  if (e != hipSuccess) return dev_fail("closest", e);
  return TWOS_OK;
}

twos_status twos_distance_bound(twos_ctx *ctx, int32_t geom_id, int32_t precision, const double *points, int64_t n,
                                double *bound) {
  if (!ctx || (n > 0 && (!points || !bound))) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (precision != 32 && precision != 64) return fail(TWOS_ERR_INVALID_ARGUMENT, "precision must be 32 or 64");
  if (n <= 0) return TWOS_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->geoms.find(geom_id) == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  HIP_OK(hipSetDevice(ctx->device));
  double *d_p = nullptr, *d_b = nullptr;
  HIP_OK(dalloc(&d_p, size_t(3 * n)));
  HIP_OK(dalloc(&d_b, size_t(n)));
  hipError_t e = hipMemcpyAsync(d_p, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_bound(ctx->d_geoms + geom_id, precision, d_p, n, d_b, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bound, d_b, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_p); (void)hipFree(d_b);
  if (e != hipSuccess) return dev_fail("distance bound", e);
  return TWOS_OK;
}

twos_status twos_trace(twos_ctx *ctx, int32_t geom_id, const twos_side *side, const twos_params *params,
                       const double start[3], int64_t site, int32_t walk, double *rows, int32_t max_rows,
                       int32_t *n_rows) {
  if (!ctx || !side || !params || !start || !rows || !n_rows || max_rows < 1)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  const twos_params &p = *params;
  if (p.max_steps < 1 || !(p.eps > 0.0)) return fail(TWOS_ERR_INVALID_ARGUMENT, "invalid numeric option");
  if (p.rng != kRngPhilox && p.precision != 64) return fail(TWOS_ERR_INVALID_ARGUMENT, "mt streams need binary64");
  if (twos_status sst = check_side(*side)) return sst;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->geoms.find(geom_id) == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  HIP_OK(hipSetDevice(ctx->device));
  Engine e;
  e.bind(ctx);
  twos_status st = e.allocate(1, 64, 1);
  if (st != TWOS_OK) return st;
  Request r;
  r.geom = geom_id;
  r.prm = p;
  r.prm.walks = 2;
  r.prm.walks_max = 0;
  r.prm.site_offset = int32_t(site);
  r.sides[0] = *side;
  r.sides[1] = *side;
  r.sites.assign(start, start + 3);
  r.n = 1;
  r.slot0 = 0;
  r.problem = 0;
  st = e.admit(&r);
  if (st != TWOS_OK) return st;
  auto *a_slot = reinterpret_cast<int32_t *>(e.h_in + e.off_slot);
  auto *a_base = reinterpret_cast<int32_t *>(e.h_in + e.off_base);
  auto *a_prefix = reinterpret_cast<int64_t *>(e.h_in + e.off_prefix);
  auto *a_nact = reinterpret_cast<int32_t *>(e.h_in + e.off_nact);
  auto *a_total = reinterpret_cast<int64_t *>(e.h_in + e.off_total);
  a_slot[0] = 0;
  a_base[0] = walk;
  a_prefix[0] = 0;
  a_prefix[1] = 1;
  a_nact[0] = 1;
  a_total[0] = 1;
  HIP_OK(hipMemcpy(e.d_in, e.h_in, e.in_bytes, hipMemcpyHostToDevice));
  double *d_rows = nullptr;
  HIP_OK(dalloc(&d_rows, 4 * size_t(max_rows)));
  Tables tb = e.tables(0, 0);
  WalkParams wp{};
  wp.combine = kSquare;
  wp.rng = p.rng;
  wp.jobs_per_walk = 1;
  wp.job_begin = 0;
  wp.job_end = 1;
  wp.val = e.d_val;
  wp.steps = e.d_steps;
  wp.trace = d_rows;
  wp.trace_rows = max_rows;
  hipError_t he = p.precision == 64
                      ? launch_walk_simple_f64(kGeomGeneric, p.rng != kRngPhilox, tb, wp, 64, ctx->stream)
                      : launch_walk_simple_f32(kGeomGeneric, tb, wp, 64, ctx->stream);
  int32_t steps = 0;
  if (he == hipSuccess) he = hipMemcpyAsync(rows, d_rows, sizeof(double) * 4 * max_rows, hipMemcpyDeviceToHost, ctx->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(&steps, e.d_steps, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_rows);
  if (he != hipSuccess) return dev_fail("trace", he);
  *n_rows = std::min((steps & 0x3fffffff) + 1, max_rows);
  return TWOS_OK;
}

twos_status twos_differential(twos_ctx *ctx, int32_t geom_id, const twos_side *side, const twos_params *params,
                              const double *sites, int64_t n, double ell, double *out, int64_t *counts) {
  if (!ctx || !side || !params || (n > 0 && (!sites || !out))) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  const twos_params &p = *params;
  if (n < 0 || p.walks < 1 || p.max_steps < 1 || !(p.eps > 0.0) || !(ell > 0.0) || !std::isfinite(ell))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "invalid numeric option");
  const bool philox = p.rng == TWOS_RNG_PHILOX;
  if (!philox && p.precision != 64)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "the differential estimator's replay streams run in binary64");
  if (p.precision != 32 && p.precision != 64) return fail(TWOS_ERR_INVALID_ARGUMENT, "precision must be 32 or 64");
  if (!(p.grid_kappa >= 0.f && p.grid_kappa < 1e6f))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "grid_kappa must be a non-negative number of cells");
  if (twos_status sst = check_side(*side)) return sst;
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(sites[i])) return fail(TWOS_ERR_INVALID_ARGUMENT, "non-finite site coordinate");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->geoms.find(geom_id);
  if (it == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  const GeomEntry &ge = it->second;
  if (ge.desc.container != kMesh || ge.dev.n_spheres != 0 || ge.dev.n_cylinders != 0)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "the differential estimator needs a mesh container with cavities only");
  const int32_t n_cav = ge.dev.n_cavities;
  if (n_cav > 0 && !(ge.desc.cavity_radius > 0.0))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "the differential estimator needs a positive cavity radius");
  if (n == 0) return TWOS_OK;
  HIP_OK(hipSetDevice(ctx->device));
  // Per-job records take 56 bytes; chunks of sites keep them within 1/4 of
  // the context's budget (and at most 2^22 jobs).
  const int64_t job_cap = std::max<int64_t>(p.walks, std::min<int64_t>(int64_t(1) << 22, int64_t(ctx->budget / 4 / 56)));
  const int64_t chunk = std::max<int64_t>(1, job_cap / p.walks);
  const int64_t jobs_max = std::min(n, chunk) * p.walks;
  struct Buffers {
    double *sites = nullptr, *val = nullptr, *out = nullptr;
    int32_t *rec = nullptr;
    int64_t *counts = nullptr;
    DevSide *side = nullptr;
    unsigned long long *counter = nullptr;
    ~Buffers() { dfree(sites); dfree(val); dfree(out); dfree(rec); dfree(counts); dfree(side); dfree(counter); }
  } b;
  HIP_OK(dalloc(&b.sites, 3 * size_t(n)));
  HIP_OK(dalloc(&b.val, 4 * size_t(jobs_max)));
  HIP_OK(dalloc(&b.rec, 6 * size_t(jobs_max)));
  HIP_OK(dalloc(&b.out, 8 * size_t(n) * size_t(n_cav)));
  HIP_OK(dalloc(&b.counts, 9 * size_t(n)));
  HIP_OK(dalloc(&b.side, 1));
  if (philox) HIP_OK(dalloc(&b.counter, 1));
  HIP_OK(hipMemcpy(b.sites, sites, sizeof(double) * 3 * size_t(n), hipMemcpyHostToDevice));
  const DevSide dside = *side;
  HIP_OK(hipMemcpy(b.side, &dside, sizeof(DevSide), hipMemcpyHostToDevice));
  DiffParams dp{};
  dp.sites = b.sites;
  dp.val = b.val;
  dp.rec = b.rec;
  dp.out = b.out;
  dp.counts = b.counts;
  dp.walks = p.walks;
  dp.site_offset = p.site_offset;
  dp.seed = p.seed;
  dp.repeat = p.repeat;
  dp.max_steps = p.max_steps;
  dp.eps = p.eps;
  dp.ell = ell;
  dp.counter = b.counter;
  dp.pseed = splitmix64(p.seed + 0x6a09e667f3bcc909ULL) ^ splitmix64(uint64_t(p.repeat) + 0xbb67ae8584caa73bULL);
  dp.grid_kappa = (p.flags & TWOS_FLAG_DISTANCE_GRID) ? (p.grid_kappa > 0.f ? p.grid_kappa : kGridKappa) : 0.f;
  dp.safety_batch = 8;
  const int cus = ctx->prop.multiProcessorCount;
  for (int64_t s0 = 0; s0 < n; s0 += chunk) {
    const int32_t cnt = int32_t(std::min(chunk, n - s0));
    dp.site_base = s0;
    dp.jobs = int64_t(cnt) * p.walks;
    if (philox) {
      HIP_OK(hipMemsetAsync(b.counter, 0, sizeof(unsigned long long), ctx->stream));
      HIP_OK(p.precision == 32 ? launch_diff_persistent_f32(ctx->d_geoms + geom_id, b.side, dp, cus, ctx->stream)
                               : launch_diff_persistent_f64(ctx->d_geoms + geom_id, b.side, dp, cus, ctx->stream));
      HIP_OK(launch_diff_fold(dp, cnt, n_cav, ctx->stream));
    } else {
      HIP_OK(launch_differential(ctx->d_geoms + geom_id, b.side, dp, cnt, n_cav, ctx->stream));
    }
  }
  HIP_OK(hipMemcpyAsync(out, b.out, sizeof(double) * 8 * size_t(n) * size_t(n_cav), hipMemcpyDeviceToHost,
                        ctx->stream));
  if (counts)
    HIP_OK(hipMemcpyAsync(counts, b.counts, sizeof(int64_t) * 9 * size_t(n), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return TWOS_OK;
}

namespace twos {
namespace {

// The mesh's certified corner grid with `cells` cells along its longest axis
// (one cell of margin around the vertex bounds), built once per resolution.
twos_status ensure_cert_grid(twos_ctx *ctx, MeshEntry &me, int32_t cells) {
  MeshBuffers &mb = me.mb;
  if (mb.cert.d && mb.cert_cells == cells) return TWOS_OK;
  (void)hipFree(const_cast<float *>(mb.cert.d));
  mb.cert = CertGrid{};
  mb.cert_cells = 0;
  mb.cert_bytes = 0;
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, me.hi[k] - me.lo[k]);
  if (!(ext > 0.0)) return fail(TWOS_ERR_INVALID_ARGUMENT, "degenerate mesh bounds");
  CertGrid cg;
  cg.h = ext / cells;
  cg.inv_h = 1.0 / cg.h;
  for (int k = 0; k < 3; ++k) {
    cg.lo[k] = me.lo[k] - cg.h;
    cg.n[k] = int32_t(std::ceil((me.hi[k] - me.lo[k]) / cg.h)) + 3;  // cells + 2 of margin, as corners
  }
  const size_t count = size_t(cg.n[0]) * cg.n[1] * cg.n[2];
  float *d = nullptr;
  HIP_OK(dalloc(&d, count));
  cg.d = d;
  if (hipError_t e = build_cert_grid(mb, cg, ctx->stream); e != hipSuccess) {
    (void)hipFree(d);
    HIP_OK(e);
  }
  if (hipError_t e = hipStreamSynchronize(ctx->stream); e != hipSuccess) {
    (void)hipFree(d);
    HIP_OK(e);
  }
  mb.cert = cg;
  mb.cert_cells = cells;
  mb.cert_bytes = int64_t(count * sizeof(float));
  return TWOS_OK;
}

// A mesh geometry minus cavities only (the validator's and the differential
// estimator's domain).
twos_status cavity_mesh_geometry(twos_ctx *ctx, int32_t geom_id, const GeomEntry **out) {
  auto it = ctx->geoms.find(geom_id);
  if (it == ctx->geoms.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown geometry id");
  const GeomEntry &ge = it->second;
  if (ge.desc.container != kMesh || ge.dev.n_spheres != 0 || ge.dev.n_cylinders != 0)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "needs a mesh container minus cavities only");
  *out = &ge;
  return TWOS_OK;
}

}  // namespace
}  // namespace twos

twos_status twos_certify(twos_ctx *ctx, int32_t geom_id, const twos_cert_params *params, const double *sites,
                         int64_t n, double *value, int32_t *steps, int32_t *flags) {
  if (!ctx || !params || (n > 0 && (!sites || !value || !steps || !flags)))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  const twos_cert_params &p = *params;
  if (n < 0 || p.walks < 1 || p.max_steps < 1 || !(p.eps >= 0.0) || !(p.radius_cap > 0.0) || !(p.value_cap > 0.0) ||
      p.sample_begin < 0)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "invalid numeric option");
  if (p.source != TWOS_CERT_AFFINE && p.source != TWOS_CERT_HOUSING)
    return fail(TWOS_ERR_INVALID_ARGUMENT, "unknown certified source");
  if (p.source == TWOS_CERT_AFFINE && !(p.slope >= 0.0))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "validator expects nonnegative affine slope");
  if (p.philox != 0 && p.philox != 1) return fail(TWOS_ERR_INVALID_ARGUMENT, "philox must be 0 or 1");
  if (p.grid_cells != 0 && (!p.philox || p.grid_cells < 8 || p.grid_cells > 512))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "grid_cells must be 0, or 8..512 with philox streams");
  if (!(p.grid_kappa >= 0.f && p.grid_kappa < 1e6f))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "grid_kappa must be a non-negative number of cells");
  if (p.philox && p.sample_begin + int64_t(p.walks) > (int64_t(1) << 32))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "Philox sample indices must fit in 32 bits");
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(sites[i])) return fail(TWOS_ERR_INVALID_ARGUMENT, "non-finite site coordinate");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const GeomEntry *ge = nullptr;
  if (twos_status st = cavity_mesh_geometry(ctx, geom_id, &ge)) return st;
  if (n == 0) return TWOS_OK;
  HIP_OK(hipSetDevice(ctx->device));
  CertGrid grid;
  if (p.grid_cells > 0) {
    auto mit = ctx->meshes.find(ge->desc.mesh);
    if (mit == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
    if (twos_status st = ensure_cert_grid(ctx, mit->second, p.grid_cells)) return st;
    grid = mit->second.mb.cert;
    const double kappa = p.grid_kappa > 0.f ? double(p.grid_kappa) : 2.0;
    grid.tau = std::max(kappa * grid.h, 2.0 * p.eps);
  }
  // Chunks of sites keep the per-sample buffers (16 B) within 2^22 samples.
  const int64_t chunk = std::max<int64_t>(1, (int64_t(1) << 22) / p.walks);
  const int64_t rows = std::min(n, chunk);
  struct Buffers {
    double *sites = nullptr, *value = nullptr;
    int32_t *steps = nullptr, *flags = nullptr;
    unsigned long long *counter = nullptr;
    ~Buffers() { dfree(sites); dfree(value); dfree(steps); dfree(flags); dfree(counter); }
  } b;
  HIP_OK(dalloc(&b.sites, 3 * size_t(n)));
  HIP_OK(dalloc(&b.value, size_t(rows) * size_t(p.walks)));
  HIP_OK(dalloc(&b.steps, size_t(rows) * size_t(p.walks)));
  HIP_OK(dalloc(&b.flags, size_t(rows) * size_t(p.walks)));
  HIP_OK(dalloc(&b.counter, 1));
  HIP_OK(hipMemcpy(b.sites, sites, sizeof(double) * 3 * size_t(n), hipMemcpyHostToDevice));
  const int cus = ctx->prop.multiProcessorCount;
  for (int64_t s0 = 0; s0 < n; s0 += chunk) {
    const int64_t cnt = std::min(chunk, n - s0);
    twos_cert_params q = p;
    q.site_offset = p.site_offset + s0;
    const size_t m = size_t(cnt) * size_t(p.walks), o = size_t(s0) * size_t(p.walks);
    HIP_OK(launch_certify(ctx->d_geoms + geom_id, q, b.sites + 3 * s0, cnt, p.sample_begin, p.walks, b.value,
                          b.steps, b.flags, b.counter, cus, grid, ctx->stream));
    HIP_OK(hipMemcpyAsync(value + o, b.value, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(steps + o, b.steps, sizeof(int32_t) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(flags + o, b.flags, sizeof(int32_t) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
  }
  return TWOS_OK;
}

twos_status twos_cert_distance(twos_ctx *ctx, int32_t geom_id, const double *points, int64_t n, double *out) {
  if (!ctx || (n > 0 && (!points || !out))) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 0) return fail(TWOS_ERR_INVALID_ARGUMENT, "invalid point count");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const GeomEntry *ge = nullptr;
  if (twos_status st = cavity_mesh_geometry(ctx, geom_id, &ge)) return st;
  if (n == 0) return TWOS_OK;
  HIP_OK(hipSetDevice(ctx->device));
  double *d_pts = nullptr, *d_out = nullptr;
  HIP_OK(dalloc(&d_pts, 3 * size_t(n)));
  hipError_t e = dalloc(&d_out, size_t(n));
  if (e == hipSuccess) e = hipMemcpy(d_pts, points, sizeof(double) * 3 * size_t(n), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_cert_distance(ctx->d_geoms + geom_id, d_pts, n, d_out, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(d_pts);
  dfree(d_out);
  if (e != hipSuccess) return dev_fail("cert distance", e);
  return TWOS_OK;
}

twos_status twos_cert_grid_lower(twos_ctx *ctx, int32_t mesh_id, int32_t cells, const double *points, int64_t n,
                                 double *out) {
  if (!ctx || (n > 0 && (!points || !out))) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 0 || cells < 8 || cells > 512) return fail(TWOS_ERR_INVALID_ARGUMENT, "cells must be 8..512");
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(points[i])) return fail(TWOS_ERR_INVALID_ARGUMENT, "non-finite point coordinate");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->meshes.find(mesh_id);
  if (it == ctx->meshes.end()) return fail(TWOS_ERR_NOT_FOUND, "unknown mesh id");
  HIP_OK(hipSetDevice(ctx->device));
  if (twos_status st = ensure_cert_grid(ctx, it->second, cells)) return st;
  if (n == 0) return TWOS_OK;
  struct Buffers {
    double *pts = nullptr, *out = nullptr;
    ~Buffers() { dfree(pts); dfree(out); }
  } b;
  HIP_OK(dalloc(&b.pts, 3 * size_t(n)));
  HIP_OK(dalloc(&b.out, size_t(n)));
  HIP_OK(hipMemcpy(b.pts, points, sizeof(double) * 3 * size_t(n), hipMemcpyHostToDevice));
  HIP_OK(launch_cert_grid_lower(it->second.mb.cert, b.pts, n, b.out, ctx->stream));
  HIP_OK(hipMemcpyAsync(out, b.out, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return TWOS_OK;
}

twos_status twos_server_create(twos_ctx *ctx, int64_t max_sites, int64_t max_walks_per_round, twos_server **out) {
  if (!ctx || !out || max_sites < 1 || max_walks_per_round < 64 || max_sites > (int64_t(1) << 30) ||
      max_walks_per_round > (int64_t(1) << 34))
    return fail(TWOS_ERR_INVALID_ARGUMENT, "invalid server capacity");
  std::lock_guard<std::mutex> lock(ctx->mu);
  // The walk buffers and the slot pool must fit beside what is resident.
  // Admission then charges each request's slots against the budget, so
  // geometry built later (a finer distance grid, derived geometries) holds
  // back requests instead of overcommitting the device.
  const int64_t need = max_walks_per_round * 2 * int64_t(sizeof(double) + sizeof(int32_t)) + max_sites * kSlotBytes;
  if (ctx_device_bytes(ctx) + need > int64_t(ctx->budget))
    return fail(TWOS_ERR_ADMISSION, "server capacity exceeds the context's HBM budget beside the resident geometry");
  HIP_OK(hipSetDevice(ctx->device));
  auto srv = std::make_unique<twos_server>();
  srv->ctx = ctx;
  srv->engine = std::make_unique<Engine>();
  srv->engine->server = true;
  srv->engine->bind(ctx);
  const twos_status st = srv->engine->allocate(max_sites, max_walks_per_round, 1024);
  if (st != TWOS_OK) return st;
  *out = srv.release();
  return TWOS_OK;
}

void twos_server_destroy(twos_server *srv) {
  if (!srv) return;
  twos_ctx *ctx = srv->engine ? srv->engine->ctx : nullptr;
  if (!ctx) {  // its context is gone, and with it the device buffers
    delete srv;
    return;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  srv->engine.reset();
  delete srv;
}

twos_status twos_server_submit(twos_server *srv, int32_t geom_id, const twos_side sides[2], const twos_params *params,
                               const double *sites, int64_t n, twos_site_result *out, int64_t *request_id) {
  if (!srv || !request_id) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  twos_status st = check_params(params, n);
  if (st != TWOS_OK) return st;
  if (params->schedule != TWOS_PERSISTENT) return fail(TWOS_ERR_INVALID_ARGUMENT, "the server runs the persistent schedule");
  if (n < 1) return fail(TWOS_ERR_INVALID_ARGUMENT, "a request needs at least one site");
  if (!srv->engine->ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "the server's context was destroyed");
  std::lock_guard<std::mutex> lock(srv->ctx->mu);
  Engine *e = srv->engine.get();
  if (n > e->cap_sites) return fail(TWOS_ERR_ADMISSION, "request has more sites than the server's slots");
  if (std::max(params->walks, params->walks_max) > e->cap_walks)
    return fail(TWOS_ERR_ADMISSION, "request walk budget exceeds one round's capacity");
  Request *r = make_request(srv->ctx, geom_id, sides, params, sites, n, out, &st);
  if (!r) return st;
  r->id = e->next_id++;
  srv->requests[r->id].reset(r);
  e->queue.push_back(r);
  *request_id = r->id;
  return TWOS_OK;
}

twos_status twos_server_step_status(twos_server *srv, int64_t *done_ids, int32_t *done_status, int32_t max_done,
                                    int32_t *n_done) {
  if (!srv || !n_done) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  *n_done = 0;
  if (!srv->engine->ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "the server's context was destroyed");
  std::lock_guard<std::mutex> lock(srv->ctx->mu);
  HIP_OK(hipSetDevice(srv->ctx->device));
  std::vector<int64_t> done;
  const twos_status st = srv->engine->step(done);
  for (int64_t id : done) {
    if (*n_done < max_done) {
      if (done_ids) done_ids[*n_done] = id;
      if (done_status) done_status[*n_done] = srv->requests.at(id)->status;
    }
    *n_done += 1;
    srv->requests.erase(id);
  }
  return st;
}

twos_status twos_server_step(twos_server *srv, int64_t *done_ids, int32_t max_done, int32_t *n_done) {
  return twos_server_step_status(srv, done_ids, nullptr, max_done, n_done);
}

twos_status twos_server_pending(twos_server *srv, int64_t *queued, int64_t *active) {
  if (!srv) return fail(TWOS_ERR_INVALID_ARGUMENT, "null argument");
  if (!srv->engine->ctx) return fail(TWOS_ERR_INVALID_ARGUMENT, "the server's context was destroyed");
  std::lock_guard<std::mutex> lock(srv->ctx->mu);
  if (queued) *queued = int64_t(srv->engine->queue.size());
  if (active) *active = int64_t(srv->engine->active.size());
  return TWOS_OK;
}

}  // extern "C"
