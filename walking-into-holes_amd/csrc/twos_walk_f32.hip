// binary32 walks: positions, distance queries and Poisson-kernel weights in
// binary32 (contraction on); samples leave the kernel as binary64.
#include "twos_launch.h"
#include "twos_walk.h"

namespace twos {

namespace {
inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }
}  // namespace

hipError_t launch_walk_simple_f32(int geom_kind, const Tables &tb, const WalkParams &wp, int block,
                                  hipStream_t stream) {
  const int64_t n = wp.job_end - wp.job_begin;
  if (n <= 0) return hipSuccess;
  const dim3 grid(blocks_for(n, block)), blk(block);
  switch (geom_kind) {
    case kGeomMesh:
      hipLaunchKernelGGL((walk_simple<float, kGeomMesh, false>), grid, blk, 0, stream, tb, wp);
      break;
    case kGeomAnalytic:
      hipLaunchKernelGGL((walk_simple<float, kGeomAnalytic, false>), grid, blk, 0, stream, tb, wp);
      break;
    default:
      hipLaunchKernelGGL((walk_simple<float, kGeomGeneric, false>), grid, blk, 0, stream, tb, wp);
  }
  return hipGetLastError();
}

hipError_t launch_walk_persistent_f32(int geom_kind, bool count, const Tables &tb, const WalkParams &wp, int grid,
                                      int block, hipStream_t stream) {
  const dim3 g(grid), b(block);
  if (count) {
    hipLaunchKernelGGL((walk_persistent<float, kGeomGeneric, true>), g, b, 0, stream, tb, wp);
    return hipGetLastError();
  }
  switch (geom_kind) {
    case kGeomMesh:
      hipLaunchKernelGGL((walk_persistent<float, kGeomMesh, false>), g, b, 0, stream, tb, wp);
      break;
    case kGeomAnalytic:
      hipLaunchKernelGGL((walk_persistent<float, kGeomAnalytic, false>), g, b, 0, stream, tb, wp);
      break;
    default:
      hipLaunchKernelGGL((walk_persistent<float, kGeomGeneric, false>), g, b, 0, stream, tb, wp);
  }
  return hipGetLastError();
}

}  // namespace twos
