// binary32 walks: positions, distance queries and Poisson-kernel weights in
// binary32 (contraction on); samples leave the kernel as binary64.
#include <algorithm>

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
  if (wp.grad) {  // gradient channel: one generic variant
    hipLaunchKernelGGL((walk_simple<float, kGeomGeneric, false, kChanGrad>), grid, blk, 0, stream, tb, wp);
    return hipGetLastError();
  }
  if (wp.combine == kCoupled) {  // coupled finite-cavity pairs: one generic variant
    hipLaunchKernelGGL((walk_simple<float, kGeomGeneric, false, kChanCoupled>), grid, blk, 0, stream, tb, wp);
    return hipGetLastError();
  }
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

namespace {
template <int kWaves, int kLds, int kBatch = 0>
hipError_t persistent_f32(int geom_kind, bool count, bool bare, const Tables &tb, const WalkParams &wp, int block,
                          int cus, hipStream_t stream, int *grid_out) {
  void (*k)(Tables, WalkParams);
  int lds = kLds;
  if (count) k = walk_persistent<float, kGeomGeneric, true, kWaves, kLds>;
  else if (geom_kind == kGeomMesh && bare) k = walk_persistent<float, kGeomMesh, false, kWaves, kLds, kChanBare, kBatch>;
  else if (geom_kind == kGeomMesh) k = walk_persistent<float, kGeomMesh, false, kWaves, kLds, kChanNone, kBatch>;
  else if (geom_kind == kGeomAnalytic) {
    // no traversal stack in use
    k = walk_persistent<float, kGeomAnalytic, false, kWaves, kLdsDefault, kChanNone, kBatch>;
    lds = 0;
  } else k = walk_persistent<float, kGeomGeneric, false, kWaves, kLds, kChanNone, kBatch>;
  const size_t smem = persistent_lds_bytes(lds, block);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, smem);
  if (e != hipSuccess) return e;
  const int grid = std::max(1, per_cu) * cus;
  *grid_out = grid;
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), smem, stream, tb, wp);
  return hipGetLastError();
}
}  // namespace

// waves: occupancy variant (5, 6 or 8 cap the registers for that many waves
// per SIMD; anything else uses the compiler's allocation, 4 waves at the
// default 24-entry stack).  lds: per-lane LDS stack entries (16, 24 or 32;
// 16 and 32 are built for the compiler's allocation and 5 waves).  The grid
// is one resident wave set: blocks-per-CU from the occupancy query times CUs.
hipError_t launch_walk_persistent_f32(int geom_kind, bool count, int waves, int lds, int batch, bool bare,
                                      const Tables &tb, const WalkParams &wp, int block, int cus, hipStream_t stream,
                                      int *grid_out) {
  if (wp.grad || wp.combine == kCoupled) {
    // Gradient channel or coupled finite-cavity pairs: the compiler's
    // allocation, 24-entry stack, any geometry.
    void (*k)(Tables, WalkParams) = wp.grad ? walk_persistent<float, kGeomGeneric, false, 1, kLdsDefault, kChanGrad>
                                            : walk_persistent<float, kGeomGeneric, false, 1, kLdsDefault, kChanCoupled>;
    const size_t smem = persistent_lds_bytes(kLdsDefault, block);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, smem);
    if (e != hipSuccess) return e;
    *grid_out = std::max(1, per_cu) * cus;
    hipLaunchKernelGGL(k, dim3(*grid_out), dim3(block), smem, stream, tb, wp);
    return hipGetLastError();
  }
  if (batch && !count && (lds == 24 || lds == 0)) {
    const bool w5 = waves == 5;
    if (batch == 1) return w5 ? persistent_f32<5, 24, 1>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out)
                              : persistent_f32<1, 24, 1>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
    if (batch == 24) return w5 ? persistent_f32<5, 24, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out)
                               : persistent_f32<1, 24, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
    if (batch == 32) return w5 ? persistent_f32<5, 24, 32>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out)
                               : persistent_f32<1, 24, 32>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
    if (batch == 40) return w5 ? persistent_f32<5, 24, 40>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out)
                               : persistent_f32<1, 24, 40>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  }
  if (lds == 16) {
    if (waves == 5) return persistent_f32<5, 16>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
    return persistent_f32<1, 16>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  }
  if (lds == 32) {
    if (waves == 5) return persistent_f32<5, 32>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
    return persistent_f32<1, 32>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  }
  if (waves == 8) return persistent_f32<8, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  if (waves == 6) return persistent_f32<6, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  if (waves == 5) return persistent_f32<5, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
  return persistent_f32<1, 24>(geom_kind, count, bare, tb, wp, block, cus, stream, grid_out);
}

hipError_t launch_slot_hints(const Tables &tb, int64_t first, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(slot_hints<float>, dim3(blocks_for(n, 128)), dim3(128), 0, stream, tb, first, n);
  return hipGetLastError();
}

}  // namespace twos
