// nr_bvh_gpu.h — the mesh tree of nr_bvh.h built on the GPU.
//
// Same binned-SAH decisions as the host builder (hittable/include/bvh/bvh.h:128-320:
// 16 centroid bins per axis, cost A_L*N_L + A_R*N_R, axes with centroid extent
// < 1e-4 skipped, longest-axis median split when no axis qualifies, one item per
// leaf), made level by level: every split of a level at once for nodes of more
// than SMALL items, then one wave per remaining subtree in LDS.  Every split
// decision depends only on the set of items under a node (bin counts and boxes
// are sums, minima and maxima); the items are moved in std::partition's own
// order, and a node's place in the output follows from its item range, so the
// node array, the leaf order and the depth equal the host build's.  The one
// exception is a median-split cluster (all centroids within 1e-4) of more than
// 16 items whose sort keys tie: the host's std::sort is an unstable introsort
// there, while this build keeps tied items in their current order (which is
// what std::sort's insertion sort does for up to 16 items).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "nr_types.h"

namespace nrg {

struct Result {
    int32_t root_ref = 0;
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    int32_t depth = 0;
};

// Tree over n >= 1 item boxes (device SoA: lo x, y, z, hi x, y, z, indexed by
// item).  Writes n - 1 child-pair nodes to d_nodes and the slot -> item order
// to d_order (both device, caller-allocated).  Synchronises `st`.  Returns 0,
// or a nectar_hip error code with a message in err.
int build(hipStream_t st, int32_t n, const float* const boxes[6], GNode* d_nodes, int32_t* d_order, Result& r,
          std::string& err);

// Item boxes of n triangles (mesh.h:320-336): vertices are 11 floats each
// (position first), tris 3 indices each; writes the 6 SoA arrays.
void triangle_boxes(hipStream_t st, const float* d_vertices, const uint32_t* d_tris, int32_t n, float* const boxes[6]);

// Triangles in leaf-slot order (GTri: positions + original index) and their
// vertex indices, from the build's slot -> triangle order.
void gather_triangles(hipStream_t st, const float* d_vertices, const uint32_t* d_tris, const int32_t* d_order,
                      int32_t n, GTri* d_out, uint32_t* d_idx);

}  // namespace nrg
