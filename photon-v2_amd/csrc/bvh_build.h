// Host build of the 4-wide BVH that the gfx950 traversal kernels walk.
//
// Takes over the work of TBvhBuilder<2>::buildInformativeBvh + TLinearDepthFirstBinaryBvh::build
// (Math/Algorithm/BVH/TBvhBuilder.ipp:34-86, TLinearDepthFirstBinaryBvh.ipp:16-41): same
// input (triangle bounds), same contract (every triangle reachable through boxes that
// contain it), different tree: binned SAH over FP64 bounds, collapsed to 4-wide nodes of
// exactly 128 bytes.
#pragma once

#include <cstdint>
#include <vector>

namespace phg {

// One 128-byte node: child boxes as planes (lo/hi per axis, 4 lanes each) followed by
// the child references.  child[i] >= 0: index of an inner node; child[i] < 0 and not
// kEmptyChild: a leaf, encoded by leaf_ref(); kEmptyChild: unused slot (its box is empty).
struct alignas(128) WideNode {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node must fill one 128-byte line");

constexpr int32_t kEmptyChild = 0x7fffffff;
constexpr uint32_t kMaxLeafTris = 8;
constexpr uint32_t kLeafFirstBits = 28;

inline int32_t leaf_ref(uint32_t first, uint32_t count) {
    return (int32_t)(0x80000000u | ((count - 1u) << kLeafFirstBits) | first);
}

struct BvhBuildParams {
    uint32_t max_leaf_tris = 4;
    uint32_t sah_bins = 16;
    uint32_t nthreads = 0;
    uint32_t width = 4;              // children per wide node (2..4; unused slots stay empty)
};

struct BvhBuildResult {
    std::vector<WideNode> nodes;     // nodes[0] is the root
    std::vector<uint32_t> tri_order; // position k of the leaf order holds input triangle tri_order[k]
    uint32_t depth = 0;              // wide-tree depth (root = 1)
    uint32_t stack_need = 0;         // worst-case traversal stack occupancy
    uint32_t binary_nodes = 0;
};

// verts: 9 floats per triangle.  Throws std::runtime_error on invalid parameters.
BvhBuildResult build_wide_bvh(const float* verts, uint32_t ntri, const BvhBuildParams& params);

}  // namespace phg
