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
//           A leaf is a contiguous run of at most MAX_LEAF_SIZE (16) triangles.
//   tbnd  : float4[2 * n_faces] in leaf order, a bounding disk per face:
//           tbnd[2t+0] = (centre, radius), tbnd[2t+1] = (unit normal, thickness).
//           Every point of the face lies within `radius` of the centre in the plane
//           through the centre with that normal, and within `thickness` of that plane.
//           It gives a cheap lower bound on the squared distance to the face.
//   lbnd  : float4[2 * n_faces], the same kind of disk around all faces of a leaf,
//           stored at the leaf's first slot: lbnd[2f+0..1] for the leaf starting at
//           triangle f (other slots unused).  One test can skip a whole leaf.
//   cbnd  : float4[8 * n_nodes], the same kind of disk around all faces below each
//           internal child: cbnd[8i + 2k + 0..1] for child slot k of node i (zero
//           for leaf and empty slots).  Tested with the child's box at the parent.
//   tpre  : float4[9 * n_faces] in leaf order, the per-face constants of the distance
//           formula, each computed with the formula's own fp32 operations:
//           [0] (a, id bits)   [1] (b, |v21|^2)   [2] (c, |v32|^2)   [3] (v21, |v13|^2)
//           [4] (v32, |n|^2)   [5] (v13, n.x)     [6] (v21 x n, n.y) [7] (v32 x n, n.z)
//           [8] (v13 x n, 0)   with v21 = b - a, v32 = c - b, v13 = a - c, n = v21 x v13.
//
// Child references (WideNode::child[k]):
//   ref >= 0                  internal node index
//   ref == CHILD_EMPTY        unused slot (node with fewer than 4 children)
//   otherwise (bit 31 set)    leaf: first triangle = (ref & 0x7fffffff) >> 4,
//                             count = (ref & 15) + 1 (leaf_ref_first / leaf_ref_count)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cubvh_amd {

constexpr int32_t CHILD_EMPTY = -1;
constexpr int32_t LEAF_BIT = (int32_t)0x80000000u;
constexpr int LEAF_COUNT_BITS = 4;
constexpr int64_t MAX_LEAF_FIRST = (1ll << (31 - LEAF_COUNT_BITS)) - 2;  // keeps every leaf ref != CHILD_EMPTY
constexpr int MAX_LEAF_SIZE = 1 << LEAF_COUNT_BITS;

// constexpr: usable in host and device code
constexpr int32_t make_leaf_ref(int64_t first, int count) {
    return (int32_t)(0x80000000u | ((uint32_t)first << LEAF_COUNT_BITS) | (uint32_t)(count - 1));
}
constexpr int32_t leaf_ref_first(int32_t ref) { return (int32_t)(((uint32_t)ref & 0x7fffffffu) >> LEAF_COUNT_BITS); }
constexpr int32_t leaf_ref_count(int32_t ref) { return (ref & (MAX_LEAF_SIZE - 1)) + 1; }

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
    std::vector<float> cbnd;  // 32 floats per node, internal child slots only
    std::vector<float> tpre;  // 36 floats per triangle, leaf order, see header comment
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

// Fills the derived arrays from bvh.tris and bvh.nodes: bvh.tbnd, bvh.lbnd and
// bvh.cbnd (bounding disks of every face, every leaf and every internal subtree,
// rounded outward) and bvh.tpre (per-face constants of the distance formula).  The node
// references must already be valid (build_bvh output or validate_bvh passed).
void compute_face_bounds(HostBvh& bvh);

// Recomputes depth / stack_need / face0_slot of an imported tree, validates every
// reference and sets the box of every empty child slot to +inf .. -inf.  Returns an
// empty string on success.
std::string validate_bvh(HostBvh& bvh);

}  // namespace cubvh_amd
