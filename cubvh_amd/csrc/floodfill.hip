// floodfill.hip -- 6-connected labelling of the free voxels of [B,H,W,D] grids.
//
// Replaces the reference's host-driven loop of hook and pointer-jump kernels
// (include/gpu/floodfill.cuh:157-272: up to ~500 iterations at 256^3, each with a
// device synchronization and a 4-byte copy back to the host) by a fixed sequence of
// three kernels with no host round trip, working in place in the output array:
//
//   1. init   : label[i] = occupied ? -1 : i
//   2. merge  : each free voxel unions itself with its free neighbours at -1 along
//               each axis.  Union-find over the label array: a node's parent is
//               always a smaller index of the same component, roots are linked
//               smaller-under-larger with atomicMin and the link is retried until
//               it lands on a root, so concurrent unions never lose an edge.
//   3. flatten: label[i] = root(i).  Since parents only ever point to smaller
//               indices, the root of each component is its minimum linear index.
//
// The result equals the reference's converged output (the minimum linear index of
// each free component, -1 for occupied voxels, indices global across the batch)
// for cubic grids.  Connectivity is tensor-order (last axis fastest) for every grid
// shape; the reference decodes coordinates as if H == W == D.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "query.hpp"

namespace cubvh_amd {
namespace dev {

__device__ __forceinline__ int32_t ld_parent(const int32_t* L, int32_t x) {
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t find_root(const int32_t* L, int32_t x) {
    int32_t p = ld_parent(L, x);
    while (p != x) {
        x = p;
        p = ld_parent(L, x);
    }
    return x;
}

__device__ __forceinline__ void unite(int32_t* L, int32_t a, int32_t b) {
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a > b) {
            int32_t t = a; a = b; b = t;
        }
        // link root b under a; succeeds only if b is still a root
        int32_t old = atomicMin(L + b, a);
        if (old == b) return;
        b = old;
    }
}

__global__ void ff_init(const uint8_t* __restrict__ grid, int32_t* __restrict__ L, int32_t ntot) {
    int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < ntot) L[i] = grid[i] ? -1 : i;
}

__global__ void ff_merge(const uint8_t* __restrict__ grid, int32_t* L, int32_t ntot, int32_t H, int32_t W,
                         int32_t D) {
    int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= ntot || grid[i]) return;
    int32_t d = i % D;
    int32_t w = (i / D) % W;
    int32_t h = (i / (D * W)) % H;
    if (d > 0 && !grid[i - 1]) unite(L, i, i - 1);
    if (w > 0 && !grid[i - D]) unite(L, i, i - D);
    if (h > 0 && !grid[i - D * W]) unite(L, i, i - D * W);
}

__global__ void ff_flatten(int32_t* L, int32_t ntot) {
    int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= ntot) return;
    int32_t p = L[i];
    if (p < 0 || p == i) return;
    L[i] = find_root(L, p);
}

}  // namespace dev

hipError_t launch_floodfill(const uint8_t* grid, int32_t B, int32_t H, int32_t W, int32_t D, int32_t* labels,
                            hipStream_t stream) {
    const int64_t ntot64 = (int64_t)B * H * W * D;
    if (ntot64 <= 0) return hipSuccess;
    const int32_t ntot = (int32_t)ntot64;
    const int block = 256;
    const unsigned grid_n = (unsigned)((ntot64 + block - 1) / block);
    hipLaunchKernelGGL(dev::ff_init, dim3(grid_n), dim3(block), 0, stream, grid, labels, ntot);
    hipLaunchKernelGGL(dev::ff_merge, dim3(grid_n), dim3(block), 0, stream, grid, labels, ntot, H, W, D);
    hipLaunchKernelGGL(dev::ff_flatten, dim3(grid_n), dim3(block), 0, stream, labels, ntot);
    return hipGetLastError();
}

}  // namespace cubvh_amd
