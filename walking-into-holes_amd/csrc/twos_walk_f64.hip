// binary64 walks (built with -ffp-contract=off so each operation rounds once,
// as the reference build's do), the mt19937_64 replay streams, and the
// binary64 per-site reduction.
#include <algorithm>

#include "twos_launch.h"
#include "twos_walk.h"

namespace twos {

namespace {
inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }
constexpr int kReduceBlock = 256;
}  // namespace

hipError_t launch_walk_simple_f64(int geom_kind, bool mt, const Tables &tb, const WalkParams &wp, int block,
                                  hipStream_t stream) {
  const int64_t n = wp.job_end - wp.job_begin;
  if (n <= 0) return hipSuccess;
  const dim3 grid(blocks_for(n, block)), blk(block);
  if (mt) {
    if (wp.grad) hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, true, kChanGrad>), grid, blk, 0, stream, tb, wp);
    else if (wp.combine == kCoupled)
      hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, true, kChanCoupled>), grid, blk, 0, stream, tb, wp);
    else hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, true>), grid, blk, 0, stream, tb, wp);
    return hipGetLastError();
  }
  if (wp.grad) {  // gradient channel: one generic variant
    hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, false, kChanGrad>), grid, blk, 0, stream, tb, wp);
    return hipGetLastError();
  }
  if (wp.combine == kCoupled) {  // coupled finite-cavity pairs: one generic variant
    hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, false, kChanCoupled>), grid, blk, 0, stream, tb, wp);
    return hipGetLastError();
  }
  switch (geom_kind) {
    case kGeomMesh:
      hipLaunchKernelGGL((walk_simple<double, kGeomMesh, false>), grid, blk, 0, stream, tb, wp);
      break;
    case kGeomAnalytic:
      hipLaunchKernelGGL((walk_simple<double, kGeomAnalytic, false>), grid, blk, 0, stream, tb, wp);
      break;
    default:
      hipLaunchKernelGGL((walk_simple<double, kGeomGeneric, false>), grid, blk, 0, stream, tb, wp);
  }
  return hipGetLastError();
}

hipError_t launch_walk_persistent_f64(int geom_kind, bool count, const Tables &tb, const WalkParams &wp, int block,
                                      int cus, hipStream_t stream, int *grid_out) {
  void (*k)(Tables, WalkParams);
  if (wp.grad) k = walk_persistent<double, kGeomGeneric, false, 1, kLdsDefault, kChanGrad>;
  else if (wp.combine == kCoupled) k = walk_persistent<double, kGeomGeneric, false, 1, kLdsDefault, kChanCoupled>;
  else if (count) k = walk_persistent<double, kGeomGeneric, true, 1, kLdsDefault>;
  else if (geom_kind == kGeomMesh) k = walk_persistent<double, kGeomMesh, false, 1, kLdsDefault>;
  else if (geom_kind == kGeomAnalytic) k = walk_persistent<double, kGeomAnalytic, false, 1, kLdsDefault>;
  else k = walk_persistent<double, kGeomGeneric, false, 1, kLdsDefault>;
  const bool stackless = geom_kind == kGeomAnalytic && !count && !wp.grad && wp.combine != kCoupled;
  const size_t smem = stackless ? 0 : persistent_lds_bytes(kLdsDefault, block);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, smem);
  if (e != hipSuccess) return e;
  const int grid = std::max(1, per_cu) * cus;
  *grid_out = grid;
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), smem, stream, tb, wp);
  return hipGetLastError();
}

namespace {
__global__ void zero_u64(unsigned long long *p, int n) {
  if (int(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}
}  // namespace

// Clears ticket counters with a kernel (graph-safe: no memset nodes).
hipError_t launch_zero(unsigned long long *p, int n, hipStream_t stream) {
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, stream, p, n);
  return hipGetLastError();
}

// The per-site fold; with gradient samples (wp.grad) a second fold of the
// gradient channel follows on the same stream.
hipError_t launch_reduce(const Tables &tb, const WalkParams &wp, int grid, hipStream_t stream) {
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL((reduce_sites<kReduceBlock>), dim3(grid), dim3(kReduceBlock), 0, stream, tb, wp);
  if (wp.grad) hipLaunchKernelGGL((reduce_grad<kReduceBlock>), dim3(grid), dim3(kReduceBlock), 0, stream, tb, wp);
  return hipGetLastError();
}

}  // namespace twos
