// forest.hpp -- launch interface of the batched (many-BVH) distance query.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh.hpp"
#include "query.hpp"

namespace cubvh_amd {

struct ForestWorkspace {
    int32_t* bounds = nullptr;   // 6 ints: ordered-int query bounds
    uint32_t* counter = nullptr; // persistent-grid tile counter
    int64_t* pbase = nullptr;    // [M+1] padded segment bases (multiples of 64)
    uint32_t* order = nullptr;   // [npad_max] query per slot, 0xffffffff for padding
    uint32_t* bin_of = nullptr;  // [n]
    uint32_t* rank = nullptr;    // [n]
    uint32_t* hist = nullptr;    // [nbins] counts, then exclusive scan inside chunks
    uint32_t* sums = nullptr;    // [nchunks]
    int64_t npad_max = 0, nbins = 0, nchunks = 0;
};

size_t forest_workspace_bytes(int64_t n, int32_t n_meshes, int32_t bits);
ForestWorkspace carve_forest_workspace(void* ws, int64_t n, int32_t n_meshes, int32_t bits);

hipError_t launch_bounds(int64_t n, const float* positions, int32_t* bounds, hipStream_t s);

// Fills ws.order / ws.pbase for the queries (CSR offsets on the device, n_meshes + 1).
hipError_t forest_order(int64_t n, const float* positions, const int64_t* offsets, int32_t n_meshes, int32_t bits,
                        const ForestWorkspace& ws, hipStream_t s);

struct ForestArgs {
    int64_t n = 0;
    int32_t n_meshes = 0;
    const float* positions = nullptr;
    const int64_t* offsets = nullptr;
    float* distances = nullptr;
    int64_t* face_id = nullptr;
    float* uvw = nullptr;
    TreeArrays tree;
    const int4* mesh_info = nullptr;  // per mesh: root node, face-0 slot, n_faces, 0
    int32_t persistent = 0;
    int32_t waves_per_cu = 8;
    int32_t waves_per_block = 4;  // 1, 2, 4 or 8 (workgroups of 64..512 lanes)
    int32_t xcd_chunk = 0;        // one-wave-per-tile grid: consecutive workgroups per XCD (0 = dispatch order)
    int32_t n_cu = 256;
};

hipError_t launch_forest_distance(const ForestArgs& a, const ForestWorkspace& ws, hipStream_t s);

// Rebases the child references of n nodes copied into a forest (forest.hip).
hipError_t launch_rebase_nodes(WideNode* nodes, int64_t n, int32_t node_base, int32_t tri_base, hipStream_t s);

}  // namespace cubvh_amd
