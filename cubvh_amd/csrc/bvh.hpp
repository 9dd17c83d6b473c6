// bvh.hpp -- in-memory layout of the 4-wide BVH shared by the host builder and the
// HIP traversal kernels.
//
// HBM layout (one array each, 128-byte aligned):
//   nodes : WideNode[n_nodes], 128 bytes = one cache line per node.  The six
//           bound vectors are structure-of-arrays over the 4 child slots, so one
//           16-byte load yields one bound of all four children; the box tests of a
//           node are four independent lane-local computations.
//   tris  : float4[3 * n_faces] in leaf order.  Triangle t occupies
//           tris[3t+0] = (a.x, a.y, a.z, bits of the original face index),
//           tris[3t+1] = (b.x, b.y, b.z, 0), tris[3t+2] = (c.x, c.y, c.z, 0).
//           A leaf is a contiguous run of at most 8 triangles.
//   tbnd  : float4[2 * n_faces] in leaf order, a bounding disk per face:
//           tbnd[2t+0] = (centre, radius), tbnd[2t+1] = (unit normal, thickness).
//           Every point of the face lies within `radius` of the centre in the plane
//           through the centre with that normal, and within `thickness` of that plane.
//           It gives a cheap lower bound on the squared distance to the face.
//   lbnd  : float4[2 * n_faces], the same kind of disk around all faces of a leaf,
//           stored at the leaf's first slot: lbnd[2f+0..1] for the leaf starting at
//           triangle f (other slots unused).  One test can skip a whole leaf.
//
// Child references (WideNode::child[k]):
//   ref >= 0                  internal node index
//   ref == CHILD_EMPTY        unused slot (node with fewer than 4 children)
//   otherwise (bit 31 set)    leaf: first triangle = (ref & 0x7fffffff) >> 3,
//                             count = (ref & 7) + 1
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cubvh_amd {

constexpr int32_t CHILD_EMPTY = -1;
constexpr int32_t LEAF_BIT = (int32_t)0x80000000u;
constexpr int64_t MAX_LEAF_FIRST = (1ll << 28) - 2;  // keeps every leaf ref != CHILD_EMPTY
constexpr int MAX_LEAF_SIZE = 8;

inline int32_t make_leaf_ref(int64_t first, int count) {
    return (int32_t)(0x80000000u | ((uint32_t)first << 3) | (uint32_t)(count - 1));
}

struct alignas(16) WideNode {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    int32_t child[4];
    int32_t n_children;
    int32_t pad[3];
};
static_assert(sizeof(WideNode) == 128, "wide node must be one 128-byte line");

struct HostBvh {
    std::vector<WideNode> nodes;
    std::vector<float> tris;  // 12 floats per triangle, leaf order, see header comment
    std::vector<float> tbnd;  // 8 floats per triangle, leaf order, see header comment
    std::vector<float> lbnd;  // 8 floats per triangle slot, set at leaf starts
    int64_t n_faces = 0;
    int32_t depth = 0;        // root = 1
    int32_t stack_need = 0;   // entries a nearest-first traversal may hold
    int32_t leaf_size = 4;
    int64_t face0_slot = 0;   // leaf-order slot of original face 0
};

// Builds the tree over triangles given as 9 floats each (a, b, c) with original ids.
// Returns an empty string on success, else an error message.
// split_mode: 0 object median on the widest centroid axis, 1 binned SAH.
std::string build_bvh(const float* tri9, const int64_t* ids, int64_t n_faces, int leaf_size, HostBvh& out,
                      int split_mode = 0);

// Fills bvh.tbnd and bvh.lbnd from bvh.tris and the leaf references of bvh.nodes
// (bounding disks of every face and of every leaf, rounded outward).
void compute_face_bounds(HostBvh& bvh);

// Recomputes depth / stack_need / face0_slot of an imported tree and validates every
// reference.  Returns an empty string on success.
std::string validate_bvh(HostBvh& bvh);

}  // namespace cubvh_amd
