// GPU LBVH over a triangle mesh, plus the mesh point/ray queries that are not
// part of the walk itself (closest-point batches, ray-parity inside tests).
//
// Build (Karras, "Maximizing parallelism in the construction of BVHs, octrees
// and k-d trees", HPG 2012):
//   1. per-triangle binary64 bounds, centroids, outward-rounded binary32 boxes
//   2. centroid bounds (block partials, one finishing block)
//   3. 30-bit Morton codes of the centroids
//   4. LSD radix sort of (code, triangle) pairs: 8-bit digits, one item per
//      thread, per-tile digit counts from wave ballots, one scan, stable scatter
//   5. hierarchy: each internal node finds its key range and split in parallel
//   6. bottom-up refit: the second child to arrive at a node writes its box
//      into the node's parent
// The nodes store both child boxes (64 B per node, see twos_device.h), and
// triangles are reordered into leaf order in binary32 and binary64.

#include <algorithm>
#include <cmath>
#include <vector>

#include "twos_device.h"
#include "twos_launch.h"
#include "twos_ray.h"

namespace twos {

namespace {

constexpr int kSortBlock = 256;

__device__ __forceinline__ float round_down(double v) {
  float f = float(v);
  if (double(f) > v) f = nextafterf(f, -__builtin_huge_valf());
  return f;
}

__device__ __forceinline__ float round_up(double v) {
  float f = float(v);
  if (double(f) < v) f = nextafterf(f, __builtin_huge_valf());
  return f;
}

__global__ void tri_prep(const double *vtx, const uint32_t *faces, int32_t n, double *centroid, float *box) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) { lo[k] = __builtin_huge_val(); hi[k] = -__builtin_huge_val(); }
  for (int c = 0; c < 3; ++c) {
    const double *p = vtx + 3 * size_t(faces[3 * size_t(i) + c]);
    for (int k = 0; k < 3; ++k) { lo[k] = fmin(lo[k], p[k]); hi[k] = fmax(hi[k], p[k]); }
  }
  for (int k = 0; k < 3; ++k) {
    centroid[3 * size_t(i) + k] = 0.5 * (lo[k] + hi[k]);
    box[6 * size_t(i) + k] = round_down(lo[k]);
    box[6 * size_t(i) + 3 + k] = round_up(hi[k]);
  }
}

// Centroid bounds: every block reduces a grid-stride share to a partial; the
// last block to finish (atomic ticket) folds the partials.
__global__ __launch_bounds__(256) void centroid_bounds(const double *centroid, int32_t n, double *partial,
                                                       unsigned int *ticket, double *out) {
  __shared__ double red[256][6];
  double v[6] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val(),
                 -__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int k = 0; k < 3; ++k) {
      v[k] = fmin(v[k], centroid[3 * size_t(i) + k]);
      v[3 + k] = fmax(v[3 + k], centroid[3 * size_t(i) + k]);
    }
  }
  for (int k = 0; k < 6; ++k) red[threadIdx.x][k] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; ++k)
        red[threadIdx.x][k] = k < 3 ? fmin(red[threadIdx.x][k], red[threadIdx.x + s][k])
                                    : fmax(red[threadIdx.x][k], red[threadIdx.x + s][k]);
    __syncthreads();
  }
  __shared__ bool last;
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; ++k) partial[6 * blockIdx.x + k] = red[0][k];
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    double r[6] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val(),
                   -__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
    for (unsigned b = 0; b < gridDim.x; ++b)
      for (int k = 0; k < 6; ++k) {
        const double x = __hip_atomic_load(partial + 6 * b + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        r[k] = k < 3 ? fmin(r[k], x) : fmax(r[k], x);
      }
    for (int k = 0; k < 6; ++k) out[k] = r[k];
  }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void morton(const double *centroid, const double *bounds, int32_t n, uint32_t *keys, uint32_t *vals) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const double ext = bounds[3 + k] - bounds[k];
    const double t = ext > 0.0 ? (centroid[3 * size_t(i) + k] - bounds[k]) / ext : 0.5;
    q[k] = uint32_t(fmin(fmax(t * 1024.0, 0.0), 1023.0));
  }
  keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[i] = uint32_t(i);
}

// Digit counts of one 256-item tile: lanes holding the same digit are found
// with eight ballots; the lowest lane of each group records the group size.
__device__ __forceinline__ void tile_rank(uint32_t digit, bool valid, uint32_t (*wave_hist)[256], int *rank_in_wave) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t match = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t bits = __ballot(valid && ((digit >> b) & 1u));
    match &= ((digit >> b) & 1u) ? bits : ~bits;
  }
  if (valid) {
    *rank_in_wave = __popcll(match & ((1ull << lane) - 1ull));
    if (*rank_in_wave == 0) wave_hist[wave][digit] = uint32_t(__popcll(match));
  }
}

__global__ __launch_bounds__(kSortBlock) void radix_hist(const uint32_t *keys, int32_t n, int shift,
                                                         uint32_t *hist, int32_t n_tiles) {
  __shared__ uint32_t wave_hist[kSortBlock / 64][256];
  for (int i = threadIdx.x; i < (kSortBlock / 64) * 256; i += kSortBlock) (&wave_hist[0][0])[i] = 0;
  __syncthreads();
  const int32_t i = blockIdx.x * kSortBlock + threadIdx.x;
  const bool valid = i < n;
  const uint32_t digit = valid ? (keys[i] >> shift) & 0xFFu : 0u;
  int rank = 0;
  tile_rank(digit, valid, wave_hist, &rank);
  __syncthreads();
  for (int d = threadIdx.x; d < 256; d += kSortBlock) {
    uint32_t c = 0;
    for (int w = 0; w < kSortBlock / 64; ++w) c += wave_hist[w][d];
    hist[size_t(d) * n_tiles + blockIdx.x] = c;
  }
}

// Exclusive scan of the digit-major count matrix in one block.
__global__ __launch_bounds__(1024) void scan_exclusive(uint32_t *data, int32_t n) {
  __shared__ uint32_t part[1024];
  const int32_t per = (n + 1023) / 1024;
  const int32_t begin = threadIdx.x * per, end = min(begin + per, n);
  uint32_t sum = 0;
  for (int32_t i = begin; i < end; ++i) sum += data[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int32_t i = begin; i < end; ++i) {
    const uint32_t v = data[i];
    data[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kSortBlock) void radix_scatter(const uint32_t *keys, const uint32_t *vals, int32_t n,
                                                            int shift, const uint32_t *offsets, int32_t n_tiles,
                                                            uint32_t *keys_out, uint32_t *vals_out) {
  __shared__ uint32_t wave_hist[kSortBlock / 64][256];
  for (int i = threadIdx.x; i < (kSortBlock / 64) * 256; i += kSortBlock) (&wave_hist[0][0])[i] = 0;
  __syncthreads();
  const int32_t i = blockIdx.x * kSortBlock + threadIdx.x;
  const bool valid = i < n;
  const uint32_t key = valid ? keys[i] : 0u;
  const uint32_t digit = (key >> shift) & 0xFFu;
  int rank = 0;
  tile_rank(digit, valid, wave_hist, &rank);
  __syncthreads();
  if (!valid) return;
  const int wave = threadIdx.x >> 6;
  uint32_t before = 0;
  for (int w = 0; w < wave; ++w) before += wave_hist[w][digit];
  const uint32_t dst = offsets[size_t(digit) * n_tiles + blockIdx.x] + before + uint32_t(rank);
  keys_out[dst] = key;
  vals_out[dst] = vals[i];
}

__device__ __forceinline__ int delta(const uint32_t *codes, int32_t n, int32_t i, int32_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = codes[i], b = codes[j];
  if (a == b) return 32 + __clz(uint32_t(i) ^ uint32_t(j));
  return __clz(a ^ b);
}

__global__ void karras(const uint32_t *codes, int32_t n, float4 *nodes, int32_t *leaf_parent, int2 *range) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(codes, n, i, i - d);
  int32_t lmax = 2;
  while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int32_t l = 0;
  for (int32_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int32_t j = i + l * d;
  const int dnode = delta(codes, n, i, j);
  int32_t s = 0, step = l;
  do {
    step = (step + 1) >> 1;
    if (delta(codes, n, i, i + (s + step) * d) > dnode) s += step;
  } while (step > 1);
  const int32_t gamma = i + s * d + min(d, 0);
  const int32_t lo = min(i, j), hi = max(i, j);
  range[i] = make_int2(lo, hi);  // the node's sorted leaves
  const int32_t left = lo == gamma ? ~gamma : gamma;
  const int32_t right = hi == gamma + 1 ? ~(gamma + 1) : gamma + 1;
  float4 q3;
  q3.x = __int_as_float(left);
  q3.y = __int_as_float(right);
  q3.z = __int_as_float(-1);
  q3.w = 0.0f;
  // The parent field is filled by the children's owners below; keep it in a
  // separate pass so no node is written by two threads in one store.
  nodes[4 * size_t(i) + 3] = q3;
  if (left < 0) leaf_parent[~left] = i;
  if (right < 0) leaf_parent[~right] = i;
}

__global__ void link_parents(float4 *nodes, int32_t n_internal) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_internal) return;
  const float4 q3 = nodes[4 * size_t(i) + 3];
  const int32_t l = __float_as_int(q3.x), r = __float_as_int(q3.y);
  float *pz = reinterpret_cast<float *>(nodes);
  if (l >= 0) pz[16 * size_t(l) + 14] = __int_as_float(i);
  if (r >= 0) pz[16 * size_t(r) + 14] = __int_as_float(i);
}

// Writes box (lo, hi) into slot `which` (0 left, 1 right) of node p.
__device__ __forceinline__ void store_child_box(float4 *nodes, int32_t p, int which, const float *b) {
  float *f = reinterpret_cast<float *>(nodes + 4 * size_t(p));
  const int base = which ? 6 : 0;
  for (int k = 0; k < 6; ++k) __hip_atomic_store(f + base + k, b[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void refit(float4 *nodes, const int32_t *leaf_parent, const uint32_t *order, const float *tri_box,
                      int32_t n, unsigned int *flags) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float b[6];
  const uint32_t t = order[k];
  for (int c = 0; c < 6; ++c) b[c] = tri_box[6 * size_t(t) + c];
  int32_t child = ~k;
  int32_t p = leaf_parent[k];
  int guard = 0;
  while (p >= 0 && guard++ < 4096) {
    float *f = reinterpret_cast<float *>(nodes + 4 * size_t(p));
    const int32_t left = __float_as_int(__hip_atomic_load(f + 12, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const int which = left == child ? 0 : 1;
    store_child_box(nodes, p, which, b);
    __threadfence();
    const unsigned prev = __hip_atomic_fetch_add(flags + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0) return;  // the sibling has not arrived; it will continue
    __threadfence();
    float o[12];
    for (int c = 0; c < 12; ++c) o[c] = __hip_atomic_load(f + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int c = 0; c < 3; ++c) {
      b[c] = fminf(o[c], o[6 + c]);
      b[3 + c] = fmaxf(o[3 + c], o[9 + c]);
    }
    child = p;
    p = __float_as_int(__hip_atomic_load(f + 14, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
}

__global__ void reorder_tris(const double *vtx, const uint32_t *faces, const uint32_t *order, int32_t n, float *tf,
                             double *td, int32_t *tid) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t t = order[k];
  for (int c = 0; c < 3; ++c) {
    const double *p = vtx + 3 * size_t(faces[3 * size_t(t) + c]);
    for (int a = 0; a < 3; ++a) {
      td[9 * size_t(k) + 3 * c + a] = p[a];
      tf[9 * size_t(k) + 3 * c + a] = float(p[a]);
    }
  }
  tid[k] = int32_t(t);
}

// 4-wide collapse: binary nodes at even depth become 4-wide nodes whose
// children are the grandchildren through each internal child (a leaf child
// stays a leaf).  flag[i] = 1 + depth / 2 for even depth (the 4-wide depth
// plus one), else 0; an exclusive scan of map[i] = (flag[i] != 0) numbers the
// 4-wide nodes, the root staying 0.
__global__ void even_depth(const float4 *nodes, int32_t n_internal, uint32_t *flag, uint32_t *map) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_internal) return;
  int32_t depth = 0;
  int32_t p = __float_as_int(nodes[4 * size_t(i) + 3].z);
  while (p >= 0 && depth < 4096) {
    ++depth;
    p = __float_as_int(nodes[4 * size_t(p) + 3].z);
  }
  flag[i] = (depth & 1) == 0 ? 1u + uint32_t(depth / 2) : 0u;
  map[i] = flag[i] != 0 ? 1u : 0u;
}

__global__ void collapse4(const float4 *nodes, const uint32_t *flag, const uint32_t *map, const int2 *range,
                          int32_t n_internal, float4 *out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_internal || !flag[i]) return;
  float lo[3][4], hi[3][4];
  int32_t ch[4];
  for (int s = 0; s < 4; ++s) {
    for (int k = 0; k < 3; ++k) { lo[k][s] = __builtin_huge_valf(); hi[k][s] = -__builtin_huge_valf(); }
    ch[s] = kEmptyChild;
  }
  int slot = 0;
  const float *f = reinterpret_cast<const float *>(nodes + 4 * size_t(i));
  for (int side = 0; side < 2; ++side) {
    const int32_t c = __float_as_int(f[12 + side]);
    if (c < 0) {
      for (int k = 0; k < 3; ++k) { lo[k][slot] = f[6 * side + k]; hi[k][slot] = f[6 * side + 3 + k]; }
      ch[slot++] = c;
      continue;
    }
    const float *g = reinterpret_cast<const float *>(nodes + 4 * size_t(c));
    for (int s2 = 0; s2 < 2; ++s2) {
      const int32_t cc = __float_as_int(g[12 + s2]);
      for (int k = 0; k < 3; ++k) { lo[k][slot] = g[6 * s2 + k]; hi[k][slot] = g[6 * s2 + 3 + k]; }
      ch[slot++] = cc < 0 ? cc : int32_t(map[cc]);
    }
  }
  float4 *o = out + size_t(kNode4Quads) * map[i];
  o[0] = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  o[1] = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  o[2] = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  o[3] = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  o[4] = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  o[5] = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  o[6] = make_float4(__int_as_float(ch[0]), __int_as_float(ch[1]), __int_as_float(ch[2]), __int_as_float(ch[3]));
  // Row 7: the node's leaf range (first, last sorted triangle), for the
  // start-node descent of record grids, and its depth in the 4-wide tree
  // (read by counting launches only).
  o[7] = make_float4(__int_as_float(range[i].x), __int_as_float(range[i].y), __int_as_float(int32_t(flag[i]) - 1),
                     0.f);
}

__global__ void leaf_depth(const float4 *nodes, const int32_t *leaf_parent, int32_t n, int32_t *max_depth) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t depth = 1;
  int32_t p = leaf_parent[k];
  while (p >= 0 && depth < 4096) {
    ++depth;
    p = __float_as_int(nodes[4 * size_t(p) + 3].z);
  }
  atomicMax(max_depth, depth);
}

__global__ void inside_kernel(MeshView m, const double *pts, int64_t n, int32_t *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t rays = 0;
  out[i] = inside_parity(m, V3<double>{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, &rays);
}

// Distance at every sample of a uniform grid (binary32 exact query).
// Sample (i, k, l) sits at lo + (index + offset) h: offset 0.5 for cell
// centres, 0 for cell corners.  A thread takes kGridRun consecutive samples
// along x, each query first testing the previous sample's closest triangle
// (one cell away, so its distance is a tight first bound); a wave's lanes
// hold consecutive z samples.  The query returns the minimum over the mesh
// whatever it tests first, so every value is the unhinted query's.
constexpr int kGridRun = 8;

__global__ void distance_grid(MeshView m, float lx, float ly, float lz, float h, float offset, int nx, int ny,
                              int nz, float *out) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int runs = (nx + kGridRun - 1) / kGridRun;
  if (idx >= int64_t(runs) * ny * nz) return;
  const int l = int(idx % nz), k = int((idx / nz) % ny), r = int(idx / (int64_t(ny) * nz));
  int32_t hint = -1;
  for (int i = r * kGridRun; i < min(nx, (r + 1) * kGridRun); ++i) {
    const V3<float> p{lx + (float(i) + offset) * h, ly + (float(k) + offset) * h, lz + (float(l) + offset) * h};
    V3<float> pt;
    out[(size_t(i) * ny + k) * nz + l] = sqrtf(mesh_closest<float, false>(m, p, pt, hint, nullptr));
  }
}

// binary16 rounding of a non-negative distance toward zero, so a bound built
// from it stays a lower bound.
__device__ __forceinline__ uint32_t f16_down(float v) {
  _Float16 h = static_cast<_Float16>(v);
  uint32_t bits = __builtin_bit_cast(unsigned short, h);
  if (static_cast<float>(h) > v && bits > 0u) bits -= 1u;
  return bits;
}

// Per cell of a corner grid (nx x ny x nz samples), its eight corner values in
// binary16 in one 16-byte record: corner (a, b, c) in word 2a + b, half c.
__global__ void pack_corner_records(const float *node, int nx, int ny, int nz, uint4 *out) {
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= int64_t(cx) * cy * cz) return;
  const int l = int(idx % cz), k = int((idx / cz) % cy), i = int(idx / (int64_t(cy) * cz));
  uint32_t w[4];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const float *q = node + (size_t(i + a) * ny + (k + b)) * size_t(nz) + l;
      w[2 * a + b] = f16_down(q[0]) | (f16_down(q[1]) << 16);
    }
  out[idx] = make_uint4(w[0], w[1], w[2], w[3]);
}

// Start node of exact queries per cell of a record grid.  The triangles
// within `reach` of the cell centre c (a conservative float test: distances
// enlarged by 0.1%) form a contiguous-or-not set of sorted leaves; with lo, hi
// the smallest and largest, the start node is the deepest 4-wide node whose
// leaf range holds [lo, hi] (ranges of a Morton-ordered tree nest).  A query
// at p in the cell whose best distance so far is u <= reach - |p - c| only
// needs triangles closer than u, which all lie within reach of c, so the
// traversal may start there.  Cells whose record bounds the centre's
// distance above reach, cells with no such triangle, and a range search that
// overflows its stack keep the root (0).
//
// A cell with one to four such triangles gets instead a node of its own
// (a "cell node", appended after the tree's nodes): a 4-wide node whose
// children are exactly those triangles as leaves, so a query starting there
// tests them in its first visit and ends, the same guarantee making the
// result exact.  Pass 0 writes the start node, or ~start for a cell that
// wants a cell node, and counts those; pass 1 gives each such cell a slot
// below `capacity` (the count of pass 0), writes its node at index base +
// slot of cell_nodes and points the cell at it (base < 0: no cell nodes).

// Triangles within sqrt(r2) of c: the sorted-leaf range [lo, hi], the count
// and the first four.  false when the search overflowed its stack.
__device__ bool cell_candidates(const MeshView &m, V3<float> c, float r2, int32_t &lo, int32_t &hi, int32_t &count,
                                int32_t cand[4]) {
  lo = 0x7fffffff;
  hi = -1;
  count = 0;
  int32_t stack[kStackDepth];
  int sp = 0;
  int32_t node = 0;
  int32_t visits = m.n_tris + 8;
  while (--visits >= 0) {
    const float4 *q = m.nodes4 + size_t(kNode4Quads) * node;
    const float4 bx0 = q[0], by0 = q[1], bz0 = q[2], bx1 = q[3], by1 = q[4], bz1 = q[5];
    const int4 ch = *reinterpret_cast<const int4 *>(q + 6);
    const float d[4] = {box_d2<float>(c, bx0.x, by0.x, bz0.x, bx1.x, by1.x, bz1.x),
                        box_d2<float>(c, bx0.y, by0.y, bz0.y, bx1.y, by1.y, bz1.y),
                        box_d2<float>(c, bx0.z, by0.z, bz0.z, bx1.z, by1.z, bz1.z),
                        box_d2<float>(c, bx0.w, by0.w, bz0.w, bx1.w, by1.w, bz1.w)};
    const int32_t cc[4] = {ch.x, ch.y, ch.z, ch.w};
    for (int s = 0; s < 4; ++s) {
      if (cc[s] == kEmptyChild || !(d[s] <= r2)) continue;
      if (cc[s] < 0) {
        V3<float> a, b, t;
        load_tri<float>(m, ~cc[s], a, b, t);
        const V3<float> e = c - closest_point_triangle(c, a, b, t);
        if (dot(e, e) <= r2) {
          lo = min(lo, ~cc[s]);
          hi = max(hi, ~cc[s]);
          if (count < 4) cand[count] = ~cc[s];
          ++count;
        }
      } else if (sp < kStackDepth) {
        stack[sp++] = cc[s];
      } else {
        return false;
      }
    }
    if (sp == 0) return true;
    node = stack[--sp];
  }
  return false;
}

// The cell node of candidates cand[0..count): leaf children with the
// triangles' binary64 bounds rounded outward, empty slots as in collapse4.
__device__ void write_cell_node(const MeshView &m, const int32_t cand[4], int32_t count, float4 *o) {
  float lo[3][4], hi[3][4];
  int32_t ch[4];
  for (int s = 0; s < 4; ++s) {
    for (int k = 0; k < 3; ++k) { lo[k][s] = __builtin_huge_valf(); hi[k][s] = -__builtin_huge_valf(); }
    ch[s] = kEmptyChild;
    if (s >= count) continue;
    const double *t = m.td + 9 * size_t(cand[s]);
    for (int k = 0; k < 3; ++k) {
      const double a = fmin(fmin(t[k], t[3 + k]), t[6 + k]), b = fmax(fmax(t[k], t[3 + k]), t[6 + k]);
      lo[k][s] = __double2float_rd(a);
      hi[k][s] = __double2float_ru(b);
    }
    ch[s] = ~cand[s];
  }
  o[0] = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  o[1] = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  o[2] = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  o[3] = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  o[4] = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  o[5] = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  o[6] = make_float4(__int_as_float(ch[0]), __int_as_float(ch[1]), __int_as_float(ch[2]), __int_as_float(ch[3]));
  // Not a tree node: an empty leaf range keeps the start-node descent out;
  // its depth is below any tree level.
  o[7] = make_float4(__int_as_float(-1), __int_as_float(-2), __int_as_float(kCellNodeDepth), 0.f);
}

template <int kPass>
__global__ void cell_start_nodes(MeshView m, const uint4 *rec, float lx, float ly, float lz, float h, int nx,
                                 int ny, int nz, float reach, int32_t *out, unsigned int *counter, float4 *cell_nodes,
                                 int32_t base, int32_t capacity) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= int64_t(nx) * ny * nz) return;
  if (kPass == 1 && out[idx] >= 0) return;
  const int l = int(idx % nz), k = int((idx / nz) % ny), i = int(idx / (int64_t(ny) * nz));
  const uint4 r = rec[idx];
  const uint32_t w[4] = {r.x, r.y, r.z, r.w};
  float dmax = 0.f;
  for (int a = 0; a < 4; ++a) dmax = fmaxf(dmax, fmaxf(f16_lo(w[a]), f16_hi(w[a])));
  int32_t start = 0;
  // Every corner is half a diagonal from the centre, and the corner values
  // are rounded down: the centre is at least dmax - half_diag from the mesh.
  const float half_diag = 0.8660254f * h, wide = reach * 1.001f;
  bool cell_node = false;
  int32_t cand[4] = {-1, -1, -1, -1};
  int32_t count = 0;
  if (m.n_tris > 1 && dmax - half_diag <= wide) {
    const V3<float> c{lx + (float(i) + 0.5f) * h, ly + (float(k) + 0.5f) * h, lz + (float(l) + 0.5f) * h};
    int32_t lo, hi;
    const bool ok = cell_candidates(m, c, wide * wide, lo, hi, count, cand);
    if (ok && lo <= hi) {
      int32_t node = 0;
      for (int depth = 0; depth < 64; ++depth) {
        const int4 ch = *reinterpret_cast<const int4 *>(m.nodes4 + size_t(kNode4Quads) * node + 6);
        const int32_t cc[4] = {ch.x, ch.y, ch.z, ch.w};
        int32_t next = -1;
        for (int s = 0; s < 4; ++s) {
          if (cc[s] < 0) continue;
          const float4 rg = m.nodes4[size_t(kNode4Quads) * cc[s] + 7];
          if (__float_as_int(rg.x) <= lo && hi <= __float_as_int(rg.y)) next = cc[s];
        }
        if (next < 0) break;
        node = next;
      }
      start = node;
      cell_node = base >= 0 && count >= 1 && count <= 4;
    }
  }
  if (kPass == 0) {
    if (cell_node) atomicAdd(counter, 1u);
    out[idx] = cell_node ? ~start : start;
    return;
  }
  const unsigned int slot = cell_node ? atomicAdd(counter, 1u) : 0xFFFFFFFFu;
  if (slot < unsigned(capacity)) {
    write_cell_node(m, cand, count, cell_nodes + size_t(kNode4Quads) * (size_t(base) + slot));
    out[idx] = base + int32_t(slot);
  } else {
    out[idx] = ~out[idx];  // no node: back to the start node
  }
}

// Lower bound of the mesh distance from the geometry's distance grid (-inf
// without a grid).
template <typename R>
__global__ void bound_kernel(const DevGeom *g, const double *pts, int64_t n, double *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<R> p{R(pts[3 * i]), R(pts[3 * i + 1]), R(pts[3 * i + 2])};
  out[i] = g->grid ? double(grid_lower<R>(*g, p)) : -__builtin_huge_val();
}

template <typename R>
__global__ void closest_kernel(const DevGeom *g, const double *pts, int64_t n, double *dist, double *cp, int32_t *cls) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<R> p{R(pts[3 * i]), R(pts[3 * i + 1]), R(pts[3 * i + 2])};
  V3<R> q;
  int c = 0;
  int32_t hint = -1;
  const R d = closest_boundary<R, kGeomGeneric, false>(*g, p, q, c, hint, nullptr);
  dist[i] = double(d);
  cp[3 * i] = double(q.x);
  cp[3 * i + 1] = double(q.y);
  cp[3 * i + 2] = double(q.z);
  cls[i] = c;
}

#define TWOS_CHECK(x)                          \
  do {                                         \
    const hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) return e_;           \
  } while (0)

inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

hipError_t lbvh_build(const double *d_vtx, const uint32_t *d_faces, int32_t n, MeshBuffers &mb, int32_t *depth,
                      hipStream_t stream) {
  const int32_t n_tiles = int32_t(blocks_for(n, kSortBlock));
  double *centroid = nullptr, *bounds = nullptr, *partial = nullptr;
  float *tri_box = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr}, *hist = nullptr;
  int32_t *leaf_parent = nullptr, *d_depth = nullptr;
  int2 *range = nullptr;
  unsigned int *flags = nullptr, *ticket = nullptr;
  const int red_blocks = 64;
  hipError_t err = hipSuccess;
  auto cleanup = [&]() {
    (void)hipFree(centroid); (void)hipFree(bounds); (void)hipFree(partial); (void)hipFree(tri_box);
    (void)hipFree(keys[0]); (void)hipFree(keys[1]); (void)hipFree(vals[0]); (void)hipFree(vals[1]); (void)hipFree(hist);
    (void)hipFree(leaf_parent); (void)hipFree(d_depth); (void)hipFree(flags); (void)hipFree(ticket);
    (void)hipFree(range);
  };
#define TRY(x)                    \
  do {                            \
    err = (x);                    \
    if (err != hipSuccess) {      \
      cleanup();                  \
      return err;                 \
    }                             \
  } while (0)
  TRY(hipMalloc(&centroid, sizeof(double) * 3 * size_t(n)));
  TRY(hipMalloc(&bounds, sizeof(double) * 6));
  TRY(hipMalloc(&partial, sizeof(double) * 6 * red_blocks));
  TRY(hipMalloc(&tri_box, sizeof(float) * 6 * size_t(n)));
  for (int b = 0; b < 2; ++b) {
    TRY(hipMalloc(&keys[b], sizeof(uint32_t) * size_t(n)));
    TRY(hipMalloc(&vals[b], sizeof(uint32_t) * size_t(n)));
  }
  TRY(hipMalloc(&hist, sizeof(uint32_t) * 256 * size_t(n_tiles)));
  TRY(hipMalloc(&leaf_parent, sizeof(int32_t) * size_t(n)));
  TRY(hipMalloc(&range, sizeof(int2) * size_t(std::max(n - 1, 1))));
  TRY(hipMalloc(&d_depth, sizeof(int32_t)));
  TRY(hipMalloc(&flags, sizeof(unsigned int) * size_t(std::max(n - 1, 1))));
  TRY(hipMalloc(&ticket, sizeof(unsigned int)));
  TRY(hipMalloc(&mb.nodes, sizeof(float4) * 4 * size_t(std::max(n - 1, 1))));
  TRY(hipMalloc(&mb.nodes4, sizeof(float4) * kNode4Quads * size_t(std::max(n - 1, 1))));
  TRY(hipMalloc(&mb.tris_f, sizeof(float) * 9 * size_t(n)));
  TRY(hipMalloc(&mb.tris_d, sizeof(double) * 9 * size_t(n)));
  TRY(hipMalloc(&mb.tri_ids, sizeof(int32_t) * size_t(n)));
  mb.n_tris = n;

  hipLaunchKernelGGL(tri_prep, dim3(blocks_for(n, 256)), dim3(256), 0, stream, d_vtx, d_faces, n, centroid, tri_box);
  TRY(hipMemsetAsync(ticket, 0, sizeof(unsigned int), stream));
  hipLaunchKernelGGL(centroid_bounds, dim3(red_blocks), dim3(256), 0, stream, centroid, n, partial, ticket, bounds);
  hipLaunchKernelGGL(morton, dim3(blocks_for(n, 256)), dim3(256), 0, stream, centroid, bounds, n, keys[0], vals[0]);
  int cur = 0;
  for (int shift = 0; shift < 32; shift += 8) {
    hipLaunchKernelGGL(radix_hist, dim3(n_tiles), dim3(kSortBlock), 0, stream, keys[cur], n, shift, hist, n_tiles);
    hipLaunchKernelGGL(scan_exclusive, dim3(1), dim3(1024), 0, stream, hist, 256 * n_tiles);
    hipLaunchKernelGGL(radix_scatter, dim3(n_tiles), dim3(kSortBlock), 0, stream, keys[cur], vals[cur], n, shift,
                       hist, n_tiles, keys[cur ^ 1], vals[cur ^ 1]);
    cur ^= 1;
  }
  TRY(hipGetLastError());
  TRY(hipMemsetAsync(leaf_parent, 0xff, sizeof(int32_t) * size_t(n), stream));
  TRY(hipMemsetAsync(mb.nodes, 0, sizeof(float4) * 4 * size_t(std::max(n - 1, 1)), stream));
  if (n > 1) {
    hipLaunchKernelGGL(karras, dim3(blocks_for(n - 1, 256)), dim3(256), 0, stream, keys[cur], n, mb.nodes,
                       leaf_parent, range);
    hipLaunchKernelGGL(link_parents, dim3(blocks_for(n - 1, 256)), dim3(256), 0, stream, mb.nodes, n - 1);
    TRY(hipMemsetAsync(flags, 0, sizeof(unsigned int) * size_t(n - 1), stream));
    hipLaunchKernelGGL(refit, dim3(blocks_for(n, 256)), dim3(256), 0, stream, mb.nodes, leaf_parent, vals[cur],
                       tri_box, n, flags);
    // flags doubles as the parity flags, keys[cur ^ 1] as the 4-wide numbering.
    uint32_t *parity = flags, *map = keys[cur ^ 1];
    hipLaunchKernelGGL(even_depth, dim3(blocks_for(n - 1, 256)), dim3(256), 0, stream, mb.nodes, n - 1, parity, map);
    hipLaunchKernelGGL(scan_exclusive, dim3(1), dim3(1024), 0, stream, map, n - 1);
    hipLaunchKernelGGL(collapse4, dim3(blocks_for(n - 1, 256)), dim3(256), 0, stream, mb.nodes, parity, map, range,
                       n - 1, mb.nodes4);
  }
  hipLaunchKernelGGL(reorder_tris, dim3(blocks_for(n, 256)), dim3(256), 0, stream, d_vtx, d_faces, vals[cur], n,
                     mb.tris_f, mb.tris_d, mb.tri_ids);
  TRY(hipMemsetAsync(d_depth, 0, sizeof(int32_t), stream));
  hipLaunchKernelGGL(leaf_depth, dim3(blocks_for(n, 256)), dim3(256), 0, stream, mb.nodes, leaf_parent, n, d_depth);
  TRY(hipGetLastError());
  TRY(hipMemcpyAsync(depth, d_depth, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  TRY(hipStreamSynchronize(stream));
  cleanup();
#undef TRY
  return hipSuccess;
}

hipError_t build_distance_grid(MeshBuffers &mb, const double lo[3], const double hi[3], int cells, int mode,
                               hipStream_t stream) {
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[k] - lo[k]);
  const double pad = 1e-3 * ext + 1e-9;
  const double h = (ext + 2 * pad) / cells;
  const bool corners = mode != 0;
  int64_t total = 1;
  int n[3];
  for (int k = 0; k < 3; ++k) {
    mb.grid_lo[k] = float(lo[k] - pad);
    n[k] = std::max(1, int(std::ceil((hi[k] - lo[k] + 2 * pad) / h))) + (corners ? 1 : 0);
    total *= n[k];
  }
  mb.grid_h = float(h);
  mb.grid_corners = mode;
  float *samples = nullptr;
  hipError_t e = hipMalloc(&samples, sizeof(float) * size_t(total));
  if (e != hipSuccess) return e;
  MeshView m{mb.nodes, mb.nodes4, mb.tris_f, mb.tris_d, mb.tri_ids, mb.n_tris};
  const int64_t threads = int64_t((n[0] + kGridRun - 1) / kGridRun) * n[1] * n[2];
  hipLaunchKernelGGL(distance_grid, dim3(blocks_for(threads, 128)), dim3(128), 0, stream, m, mb.grid_lo[0],
                     mb.grid_lo[1], mb.grid_lo[2], mb.grid_h, corners ? 0.0f : 0.5f, n[0], n[1], n[2], samples);
  if ((e = hipGetLastError()) != hipSuccess) {
    (void)hipFree(samples);
    return e;
  }
  if (mode != 2) {
    for (int k = 0; k < 3; ++k) mb.grid_n[k] = n[k];
    mb.grid = samples;
    mb.grid_bytes = int64_t(sizeof(float)) * total;
    return hipSuccess;
  }
  // Corner records: one per cell, built from the corner samples.
  const int64_t n_cells = int64_t(n[0] - 1) * (n[1] - 1) * (n[2] - 1);
  uint4 *rec = nullptr;
  if ((e = hipMalloc(&rec, sizeof(uint4) * size_t(n_cells))) != hipSuccess) {
    (void)hipFree(samples);
    return e;
  }
  hipLaunchKernelGGL(pack_corner_records, dim3(blocks_for(n_cells, 128)), dim3(128), 0, stream, samples, n[0], n[1],
                     n[2], rec);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(samples);
  if (e != hipSuccess) {
    (void)hipFree(rec);
    return e;
  }
  for (int k = 0; k < 3; ++k) mb.grid_n[k] = n[k] - 1;
  mb.grid = reinterpret_cast<float *>(rec);
  mb.grid_bytes = int64_t(sizeof(uint4)) * n_cells;
  // Start nodes: reach two cells, so a query whose best distance is up to
  // about 1.1 cells starts below the root (a search from the root costs about
  // eight visits, from a start node two; reaches of 1.5, 1.75, 2.5 and 3 cells
  // measured slower on the SH icosphere at 1024 cells); cells with at most
  // four triangles in reach start at a cell node of their own, appended to a copy of the
  // 4-wide node array (mb.nodes4 then points at the copy; the caller frees
  // the array it replaced).
  int32_t *start = nullptr;
  unsigned int *counter = nullptr;
  if (hipMalloc(&start, sizeof(int32_t) * size_t(n_cells)) == hipSuccess &&
      hipMalloc(&counter, sizeof(unsigned int)) == hipSuccess) {
    const float reach = 2.0f * mb.grid_h;
    const int32_t base = std::max(mb.n_tris - 1, 1);  // the tree's node capacity (lbvh_build)
    auto pass = [&](auto kernel, float4 *cell_nodes, int32_t capacity) {
      hipError_t pe = hipMemsetAsync(counter, 0, sizeof(unsigned int), stream);
      if (pe != hipSuccess) return pe;
      hipLaunchKernelGGL(kernel, dim3(blocks_for(n_cells, 128)), dim3(128), 0, stream, m, rec, mb.grid_lo[0],
                         mb.grid_lo[1], mb.grid_lo[2], mb.grid_h, n[0] - 1, n[1] - 1, n[2] - 1, reach, start,
                         counter, cell_nodes, base, capacity);
      pe = hipGetLastError();
      return pe == hipSuccess ? hipStreamSynchronize(stream) : pe;
    };
    e = pass(cell_start_nodes<0>, nullptr, 0);
    unsigned int wanted = 0;
    if (e == hipSuccess) e = hipMemcpy(&wanted, counter, sizeof(unsigned int), hipMemcpyDeviceToHost);
    float4 *nodes = nullptr;
    if (e == hipSuccess && wanted > 0) {
      // Without room for the cell nodes the cells keep their start nodes.
      if (hipMalloc(&nodes, sizeof(float4) * kNode4Quads * (size_t(base) + wanted)) != hipSuccess) {
        (void)hipGetLastError();
        nodes = nullptr;
      }
      if (nodes) e = hipMemcpyAsync(nodes, mb.nodes4, sizeof(float4) * kNode4Quads * size_t(base),
                                    hipMemcpyDeviceToDevice, stream);
      if (e == hipSuccess) e = pass(cell_start_nodes<1>, nodes, nodes ? int32_t(wanted) : 0);
    }
    (void)hipFree(counter);
    if (e != hipSuccess) {
      (void)hipFree(start);
      if (nodes) (void)hipFree(nodes);
      return e;
    }
    const float valid = 0.999f * (reach - 0.8660254f * mb.grid_h);
    mb.grid_start = start;
    mb.grid_start_r2 = valid * valid;
    mb.grid_bytes += int64_t(sizeof(int32_t)) * n_cells;
    if (nodes) {
      mb.nodes4 = nodes;
      mb.cell_nodes = int32_t(wanted);
      mb.grid_bytes += int64_t(sizeof(float4)) * kNode4Quads * int64_t(wanted);
    }
  } else {
    (void)hipGetLastError();
    if (start) (void)hipFree(start);
  }
  return hipSuccess;
}

hipError_t launch_inside(const MeshBuffers &mb, const double *d_pts, int64_t n, int32_t *d_out, hipStream_t stream) {
  MeshView m{mb.nodes, mb.nodes4, mb.tris_f, mb.tris_d, mb.tri_ids, mb.n_tris};
  hipLaunchKernelGGL(inside_kernel, dim3(blocks_for(n, 128)), dim3(128), 0, stream, m, d_pts, n, d_out);
  return hipGetLastError();
}

hipError_t launch_bound(const DevGeom *d_geom, int precision, const double *d_pts, int64_t n, double *d_out,
                        hipStream_t stream) {
  if (precision == 64)
    hipLaunchKernelGGL(bound_kernel<double>, dim3(blocks_for(n, 128)), dim3(128), 0, stream, d_geom, d_pts, n, d_out);
  else
    hipLaunchKernelGGL(bound_kernel<float>, dim3(blocks_for(n, 128)), dim3(128), 0, stream, d_geom, d_pts, n, d_out);
  return hipGetLastError();
}

hipError_t launch_closest(const DevGeom *d_geom, int precision, const double *d_pts, int64_t n, double *d_dist,
                          double *d_cp, int32_t *d_cls, hipStream_t stream) {
  if (precision == 64)
    hipLaunchKernelGGL(closest_kernel<double>, dim3(blocks_for(n, 128)), dim3(128), 0, stream, d_geom, d_pts, n,
                       d_dist, d_cp, d_cls);
  else
    hipLaunchKernelGGL(closest_kernel<float>, dim3(blocks_for(n, 128)), dim3(128), 0, stream, d_geom, d_pts, n,
                       d_dist, d_cp, d_cls);
  return hipGetLastError();
}

}  // namespace twos
