// Differential Walk on Spheres, binary32 walks on Philox streams (the
// production form of twos_differential; contraction on).  The kernel is
// diff_persistent in twos_diff.h; its offset-safety test stays binary64.

#include <algorithm>

#include "twos_diff.h"
#include "twos_launch.h"

namespace twos {

hipError_t launch_diff_persistent_f32(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int cus,
                                      hipStream_t stream) {
  return launch_diff_persistent<float>(d_geom, d_side, p, cus, stream);
}

}  // namespace twos
