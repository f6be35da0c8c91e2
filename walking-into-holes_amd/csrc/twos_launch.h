// Host-side launch entry points of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "twos_types.h"

namespace twos {

struct Tables;

struct MeshBuffers {
  float4 *nodes = nullptr;
  float4 *nodes4 = nullptr;
  float *tris_f = nullptr;
  double *tris_d = nullptr;
  int32_t *tri_ids = nullptr;
  int32_t n_tris = 0;
  float *grid = nullptr;
  float grid_lo[3] = {0.f, 0.f, 0.f};
  float grid_h = 0.f;
  int32_t grid_n[3] = {0, 0, 0};  // samples (modes 0, 1) or cells (mode 2) per axis
  int32_t grid_corners = 0;       // mode: 0 centres, 1 corners, 2 binary16 corner records
  int32_t *grid_start = nullptr;  // mode 2: per cell, the start node of exact queries
  float grid_start_r2 = 0.f;      // squared best distance up to which a start node is valid
  int64_t grid_bytes = 0;         // device bytes of grid + start nodes + cell nodes
  int32_t cell_nodes = 0;         // cell nodes appended to nodes4 after the tree's max(n_tris - 1, 1) nodes
  CertGrid cert;                  // certified corner grid (null until a certified call asks for one)
  int32_t cert_cells = 0;
  int64_t cert_bytes = 0;
};

hipError_t lbvh_build(const double *d_vtx, const uint32_t *d_faces, int32_t n, MeshBuffers &mb, int32_t *depth,
                      hipStream_t stream);
// cells along the longest axis; mode 0 samples the distance at cell centres,
// 1 at cell corners (cells + 1 samples per axis, binary32), 2 stores per cell
// its eight corner values rounded down to binary16 in one 16-byte record.
hipError_t build_distance_grid(MeshBuffers &mb, const double lo[3], const double hi[3], int cells, int mode,
                               hipStream_t stream);
hipError_t launch_inside(const MeshBuffers &mb, const double *d_pts, int64_t n, int32_t *d_out, hipStream_t stream);
hipError_t launch_bound(const DevGeom *d_geom, int precision, const double *d_pts, int64_t n, double *d_out,
                        hipStream_t stream);
hipError_t launch_closest(const DevGeom *d_geom, int precision, const double *d_pts, int64_t n, double *d_dist,
                          double *d_cp, int32_t *d_cls, hipStream_t stream);

// geom_kind: kGeomGeneric / kGeomMesh / kGeomAnalytic.
hipError_t launch_walk_simple_f32(int geom_kind, const Tables &tb, const WalkParams &wp, int block,
                                  hipStream_t stream);
// batch: 0, or 24 / 32 / 40 lanes of a wave ready before it takes walk steps,
// or 1: at least as many ready as traversing (built for the 24-entry stack
// with the compiler's allocation or 5 waves).
// bare: no side of any request in the launch has a volume source (mesh
// geometries take a kernel built without the source code).
hipError_t launch_walk_persistent_f32(int geom_kind, bool count, int waves, int lds, int batch, bool bare,
                                      const Tables &tb, const WalkParams &wp, int block, int cus, hipStream_t stream,
                                      int *grid_out);
hipError_t launch_slot_hints(const Tables &tb, int64_t first, int64_t n, hipStream_t stream);
hipError_t launch_walk_simple_f64(int geom_kind, bool mt, const Tables &tb, const WalkParams &wp, int block,
                                  hipStream_t stream);
hipError_t launch_walk_persistent_f64(int geom_kind, bool count, const Tables &tb, const WalkParams &wp, int block,
                                      int cus, hipStream_t stream, int *grid_out);
hipError_t launch_reduce(const Tables &tb, const WalkParams &wp, int grid, hipStream_t stream);
hipError_t launch_zero(unsigned long long *p, int n, hipStream_t stream);
// Certified sampling of n_sites sites, samples sample_begin .. + count - 1.
// prm.philox: Philox streams in one persistent launch over the jobs (counter: one
// device word for its tickets; cus: the device's compute units); otherwise
// the reference's streams, one thread per job.
hipError_t launch_certify(const DevGeom *d_geom, const twos_cert_params &prm, const double *d_sites, int64_t n_sites,
                          int64_t sample_begin, int64_t count, double *d_value, int32_t *d_steps, int32_t *d_flags,
                          unsigned long long *counter, int cus, const CertGrid &grid, hipStream_t stream);
// Fills grid.d (corners grid.n[0] x grid.n[1] x grid.n[2], allocated by the
// caller) with each corner's certified box-union distance rounded down.
hipError_t build_cert_grid(const MeshBuffers &mb, const CertGrid &grid, hipStream_t stream);
hipError_t launch_cert_grid_lower(const CertGrid &grid, const double *d_pts, int64_t n, double *d_out,
                                  hipStream_t stream);
hipError_t launch_cert_distance(const DevGeom *d_geom, const double *d_pts, int64_t n, double *d_out,
                                hipStream_t stream);
// Differential WoS over one chunk of sites (p.site_base, n_sites sites) on
// the reference's streams: the per-job walks, then the per-(site, cavity) fold.
hipError_t launch_differential(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int32_t n_sites,
                               int32_t n_cav, hipStream_t stream);
// The Philox form: one persistent launch over p.jobs tickets (p.counter zeroed
// before it), binary64 or binary32 walks; then launch_diff_fold.
hipError_t launch_diff_persistent_f64(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int cus,
                                      hipStream_t stream);
hipError_t launch_diff_persistent_f32(const DevGeom *d_geom, const DevSide *d_side, const DiffParams &p, int cus,
                                      hipStream_t stream);
hipError_t launch_diff_fold(const DiffParams &p, int32_t n_sites, int32_t n_cav, hipStream_t stream);

}  // namespace twos
