// query.hpp -- host-side launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh.hpp"

namespace cubvh_amd {

// Device arrays of one tree (or of a forest's concatenated trees); layouts in bvh.hpp.
struct TreeArrays {
    const WideNode* nodes = nullptr;
    const float4* tris = nullptr;  // leaf-ordered triangles
    const float4* tbnd = nullptr;  // per-face bounding disks
    const float4* lbnd = nullptr;  // per-leaf bounding disks at leaf-start slots
    const float4* cbnd = nullptr;  // per-child bounding disks of internal subtrees
    const float4* tpre = nullptr;  // per-face constants of the distance formula
};

struct QueryArgs {
    int64_t n = 0;
    const float* in0 = nullptr;  // positions or ray origins [n,3]
    const float* in1 = nullptr;  // ray directions [n,3]
    const uint32_t* order = nullptr;  // optional query permutation [n] (lane i handles query order[i])
    float* out_f = nullptr;      // distances or depth [n]
    int64_t* out_id = nullptr;   // face ids [n]
    float* out_v = nullptr;      // uvw or hit positions [n,3]
    TreeArrays tree;
    int64_t face0_slot = 0;
    int32_t stack_need = 0;
    bool packet = true;  // walk as wave packets (queries in Morton order) or one lane per query
    void* stream = nullptr;
};

// Stack capacity instantiated for a tree that needs `need` entries (-1 if none).
int stack_capacity_for(int need);

// mode 0 unsigned distance, 1 signed (watertight), 2 signed (raystab)
hipError_t launch_distance(const QueryArgs& q, int mode);
hipError_t launch_ray(const QueryArgs& q);
// Negates q.out_f[i] of every query whose 64 raystab rays all hit (distances from mode 0),
// walking the queries of q.order (required, n < 2^32) in passes over the undecided ones;
// ws holds stab_workspace_bytes(q.n) bytes.
size_t stab_workspace_bytes(int64_t n);
hipError_t launch_stab_sign(const QueryArgs& q, void* ws);

// The VALU issue-ceiling probe (probe.hip): `blocks` workgroups of 256 threads, each
// thread kValuProbeChains independent fp32 fma chains of `iters` steps; form 0 = VGPR
// and literal operands, 1 = one SGPR operand; out holds blocks x 256 floats.
constexpr int kValuProbeChains = 16;
hipError_t launch_valu_probe(int form, int blocks, int32_t iters, float* out, hipStream_t s);

// Morton permutation of point queries (order.hip).
size_t morton_workspace_bytes(int64_t n);
hipError_t morton_order(int64_t n, const float* positions, void* ws, uint32_t* order_out, hipStream_t s);
// Permutation of rays that puts rays of close origins and directions together (order.hip);
// ws holds morton_workspace_bytes(n) bytes.
hipError_t ray_order(int64_t n, const float* rays_o, const float* rays_d, void* ws, uint32_t* order_out, hipStream_t s);

// Static per-device scratch of the allocation-free distance call (order.hip): room for
// the Morton order and ordering workspace of kNoAllocMaxQueries queries.
constexpr int64_t kNoAllocMaxQueries = (int64_t)1 << 20;
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocDevices = 64;
// Device address of the scratch on `device` (the current device); looked up once per device.
hipError_t noalloc_scratch(int device, void** ptr);

hipError_t launch_floodfill(const uint8_t* grid, int32_t B, int32_t H, int32_t W, int32_t D, int32_t* labels,
                            hipStream_t stream);

// Backward of the unsigned distance (grad.hip).  With grad_vertices, `ws` holds
// udf_backward_workspace_bytes(...) bytes of device scratch (sorts and per-face sums);
// point gradients alone need none.  corners / corner_begin / corner_end: the mesh's
// vertex -> corners table (launch_vertex_corners), or all null to build it in `ws`.
size_t udf_backward_workspace_bytes(int64_t n, int64_t n_vertices, int64_t n_faces, bool have_grad_points,
                                    bool have_tables);
hipError_t launch_udf_backward(int64_t n, const float* grad_dist, const float* positions, const int64_t* face_id,
                               const float* uvw, const float* vertices, int64_t n_vertices, const int32_t* faces,
                               int64_t n_faces, float* grad_points, float* grad_vertices, const uint32_t* corners,
                               const uint32_t* corner_begin, const uint32_t* corner_end, void* ws, hipStream_t stream);
// The vertex -> corners table of a mesh: corners (3 n_faces) = the corners 3 f + k grouped
// by vertex, ascending inside a vertex (corners of out-of-range indices last); vertex v
// owns corners[begin[v] .. end[v]).  ws holds vertex_corners_workspace_bytes(...) bytes.
size_t vertex_corners_workspace_bytes(int64_t n_vertices, int64_t n_faces);
hipError_t launch_vertex_corners(const int32_t* faces, int64_t n_faces, int64_t n_vertices, uint32_t* corners,
                                 uint32_t* begin, uint32_t* end, void* ws, hipStream_t stream);

}  // namespace cubvh_amd
