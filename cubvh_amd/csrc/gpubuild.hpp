// gpubuild.hpp -- host interface of the GPU tree build (gpubuild.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh.hpp"
#include "query.hpp"

namespace cubvh_amd {

// Binary tree node: children >= 0 are nodes, < 0 the leaf ~j (sorted face j; the
// radix tree only); [first, last] is the node's run of sorted faces.  prefix orders the
// collapse (smaller opens first): in the radix tree the length of the common prefix of
// the node's codes (shorter = larger Morton cell), in the SAH tree a key of the node's
// box area (smaller = larger area).
struct BinNode {
    int32_t left, right, first, last, parent, prefix;
};

// Topology of a GPU build.
enum GpuTopology : int {
    kTopologyMedian = 0,  // object median on the axis of largest centroid extent
    kTopologyLinear = 1,  // Morton radix tree (linear BVH)
    kTopologySah = 2,     // surface-area heuristic over every slot of the sorted lists
};

struct GpuBuildWorkspace {
    int32_t* counters = nullptr;  // [0..5] centroid bounds (ordered ints), [6] max |coordinate| bits,
                                  // [7] input errors (1 bad index, 2 non-finite), [8] wide nodes,
                                  // [9] jobs above 4,096 faces, [10] leaves, [11] subtree jobs, [12] depth, [13] slot of
                                  // face 0, [14] structure error, [16 + k] entries made by
                                  // collapse level k; then the gather partials
    float* tri9 = nullptr;        // [9n] gathered vertices, original face order
    uint64_t* keys0 = nullptr;
    uint64_t* keys1 = nullptr;    // sorted Morton codes
    uint32_t* vals0 = nullptr;
    uint32_t* face_of = nullptr;  // [n] original face of sorted position (= leaf slot)
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    BinNode* bn = nullptr;        // [n-1] radix tree; [2n] every node of the SAH tree
    int32_t* leaf_parent = nullptr;
    WideNode* wide = nullptr;     // [n-1] wide nodes (the first `n_nodes` used)
    int4* queue0 = nullptr;
    int4* queue1 = nullptr;
    int4* leaves = nullptr;       // (first, count, wide node, slot) per leaf child
    int4* jobs = nullptr;         // (wide node, slot, first, count) per internal child
    // median topology
    float* cen = nullptr;         // [3n] centroids, axis-major
    uint32_t* lists = nullptr;    // [6n] three centroid-sorted face lists, double-buffered
    uint8_t* side = nullptr;      // [n] lower (0) or upper (1) half of the face's run
    void* flag_scan = nullptr;    // [n] scanned lower-half flags of the three lists
    void* flag_tmp = nullptr;
    size_t flag_tmp_bytes = 0;
    // SAH topology
    float* fbox = nullptr;          // [6n] face boxes
    int32_t* rid = nullptr;         // [n] active run (binary node) of each slot, -1 once a leaf
    uint32_t* split = nullptr;      // [2n] chosen split per node: axis << 28 | slot
    int32_t* act0 = nullptr;        // [n] active nodes of the level, double-buffered
    int32_t* act1 = nullptr;
    int32_t* level_count = nullptr; // per level: active runs, longest, runs above 1,024 slots
    int32_t* bigslot = nullptr;     // [2n] bin slot of a run above 1,024 slots
    int32_t* bigrun = nullptr;      // its node
    int32_t* bins = nullptr;        // per bin slot, list and bin: count and box
    float* grow = nullptr;          // box growth of the area terms
    // subtrees above 4,096 faces, reduced in chunks
    int32_t* bigjobs = nullptr;   // job indices, counted in counters[9]
    double* big_part = nullptr;   // chunk partials
    void* big_frames = nullptr;   // disk frame per big job
    int64_t n = 0;
};

struct GpuBuildResult {
    int32_t input_errors = 0;
    int32_t n_nodes = 0, n_leaves = 0, n_jobs = 0, depth = 0;
    bool structure_ok = false;  // every face in exactly one leaf
    bool too_deep = false;      // deeper than a walk's stack accepts
};

size_t gpu_build_workspace_bytes(int64_t n, int topology);
GpuBuildWorkspace carve_gpu_build_workspace(void* ws, int64_t n, int topology);

// Steps 1-4 (n > leaf_size faces): wide nodes (without boxes) in w.wide, leaf order in
// w.face_of.  kTopologyMedian: every run of faces is halved at the median centroid on
// its axis of largest centroid extent (the host's object-median split);
// kTopologyLinear: the Morton radix tree (fastest to build, slowest to query);
// kTopologySah: every run split where the surface-area cost is least (slowest to build,
// fastest to query; w must be carved for it).  res.input_errors is set when an index
// or coordinate is bad (the tree is then meaningless, but every access stayed in
// bounds); res.too_deep when the tree is deeper than a walk's stack accepts.
// Synchronizes the stream once (the SAH topology a few times more).
hipError_t gpu_build_tree(const float* V, int64_t nv, const int32_t* F, int64_t n, int32_t leaf_size, int topology,
                          const GpuBuildWorkspace& w, GpuBuildResult& res, hipStream_t s);

// Step 5: triangles and derived arrays into `out`, child boxes into w.wide (the caller
// copies w.wide to the tree's nodes afterwards).
hipError_t gpu_build_arrays(int64_t n, const GpuBuildWorkspace& w, const TreeArrays& out, hipStream_t s);

// Leaf-order slot of original face 0 (after gpu_build_arrays).  Synchronizes.
hipError_t gpu_build_face0_slot(const GpuBuildWorkspace& w, int64_t& slot, hipStream_t s);

}  // namespace cubvh_amd
