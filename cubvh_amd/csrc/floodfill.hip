// floodfill.hip -- 6-connected labelling of the free voxels of [B,H,W,D] grids.
//
// Replaces the reference's host-driven loop of hook and pointer-jump kernels
// (include/gpu/floodfill.cuh:157-272: up to ~500 iterations at 256^3, each with a
// device synchronization and a 4-byte copy back to the host) by a fixed sequence of
// three kernels with no host round trip, working in place in the output array:
//
//   1. tiles  : each workgroup labels one 16^3 tile in LDS (runs along the last
//               axis, then union-find with LDS atomics between overlapping runs of
//               neighbouring rows) and writes, for every free voxel, the global index
//               of its tile-local root; occupied voxels get -1.
//   2. merge  : voxels on a tile's low faces union themselves with their free
//               neighbour in the adjacent tile (one launch per axis over the face
//               voxels only; a pair with the same two labels as the pair before it
//               on the face is skipped).  Union-find over the label array: a
//               node's parent is always a smaller index of the same component, roots
//               are linked smaller-under-larger with atomicMin and the link is
//               retried until it lands on a root, so concurrent unions never lose an
//               edge.
//   3. flatten: label[i] = root(i).  Since parents only ever point to smaller
//               indices, the root of each component is its minimum linear index.
// Only the ~19% of voxels on tile faces touch global atomics.
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

constexpr int TILE = 16;          // tile edge (voxels)
constexpr int TILE_N = TILE * TILE * TILE;
constexpr int TILE_THREADS = 256;

__device__ __forceinline__ int32_t lds_parent(const int32_t* L, int32_t x) {
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int32_t lds_root(const int32_t* L, int32_t x) {
    int32_t p = lds_parent(L, x);
    while (p != x) {
        x = p;
        p = lds_parent(L, x);
    }
    return x;
}
__device__ __forceinline__ void lds_unite(int32_t* L, int32_t a, int32_t b) {
    for (;;) {
        a = lds_root(L, a);
        b = lds_root(L, b);
        if (a == b) return;
        if (a > b) {
            int32_t t = a; a = b; b = t;
        }
        int32_t old = atomicMin(L + b, a);
        if (old == b) return;
        b = old;
    }
}

// One workgroup of 256 threads per 16^3 tile; tiles numbered d fastest, then w, h, b.
// In LDS a voxel (lh, lw, ld) sits at (lh * 16 + lw) * ROW + ld: rows of 16 voxels
// along d, padded to ROW = 17 words so that one row per lane touches distinct banks.
// The padded index orders voxels like the tile's linear index, so the smallest
// index of a component is also its smallest voxel.
//   0. all threads load the tile's occupancy (coalesced, 16 consecutive d per row);
//   1. thread t labels the runs of free voxels of row t by their first voxel;
//   2. it unites each run-overlap segment of its row with the rows at w - 1 and
//      h - 1 (one union per segment: two runs meet in one contiguous segment);
//   3. all threads write the global index of every free voxel's tile-local root.
constexpr int ROW = TILE + 1;
__device__ __forceinline__ int pad_index(int k) { return (k >> 4) * ROW + (k & 15); }

__global__ __launch_bounds__(TILE_THREADS) void ff_tiles(const uint8_t* __restrict__ grid, int32_t* __restrict__ out,
                                                         int32_t H, int32_t W, int32_t D) {
    __shared__ int32_t L[TILE * TILE * ROW];
    __shared__ uint8_t occ[TILE * TILE * ROW];
    const int32_t nd = (D + TILE - 1) / TILE, nw = (W + TILE - 1) / TILE, nh = (H + TILE - 1) / TILE;
    const int64_t tile = blockIdx.x;
    const int32_t d0 = (int32_t)(tile % nd) * TILE;
    const int32_t w0 = (int32_t)((tile / nd) % nw) * TILE;
    const int64_t rest = tile / ((int64_t)nd * nw);
    const int32_t h0 = (int32_t)(rest % nh) * TILE;
    const int32_t b = (int32_t)(rest / nh);
    const int64_t base = (int64_t)b * H * W * D;
    for (int k = threadIdx.x; k < TILE_N; k += TILE_THREADS) {
        const int lh = k >> 8, lw = (k >> 4) & 15, ld = k & 15;
        const int h = h0 + lh, w = w0 + lw, d = d0 + ld;
        uint8_t o = 1;  // outside the grid counts as occupied
        if (h < H && w < W && d < D) o = grid[base + ((int64_t)h * W + w) * D + d] ? 1 : 0;
        occ[pad_index(k)] = o;
    }
    __syncthreads();
    const int row = (int)threadIdx.x * ROW;
    {
        int32_t start = -1;
        for (int ld = 0; ld < TILE; ++ld) {
            const int k = row + ld;
            if (occ[k]) {
                L[k] = -1;
                start = -1;
            } else {
                if (start < 0) start = k;
                L[k] = start;
            }
        }
    }
    __syncthreads();
    const int lh = (int)threadIdx.x >> 4, lw = (int)threadIdx.x & 15;
    for (int nb = 0; nb < 2; ++nb) {
        if (nb == 0 ? lw == 0 : lh == 0) continue;
        const int off = nb == 0 ? ROW : TILE * ROW;  // row at w - 1, or at h - 1
        bool prev = false;  // previous position was a free/free pair
        for (int ld = 0; ld < TILE; ++ld) {
            const int k = row + ld;
            const bool pair = !occ[k] && !occ[k - off];
            if (pair && !prev) lds_unite(L, lds_parent(L, k), lds_parent(L, k - off));
            prev = pair;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TILE_N; k += TILE_THREADS) {
        const int lhk = k >> 8, lwk = (k >> 4) & 15, ld = k & 15;
        const int h = h0 + lhk, w = w0 + lwk, d = d0 + ld;
        if (h >= H || w >= W || d >= D) continue;
        const int pk = pad_index(k);
        int32_t lab = -1;
        if (!occ[pk]) {
            const int32_t r = lds_root(L, pk);
            const int rrow = r / ROW, rd = r - rrow * ROW;
            lab = (int32_t)(base + ((int64_t)(h0 + (rrow >> 4)) * W + (w0 + (rrow & 15))) * D + (d0 + rd));
        }
        out[base + ((int64_t)h * W + w) * D + d] = lab;
    }
}

// Unions across the tile faces normal to one axis: thread t takes voxel v on a face
// (coordinate along the axis a positive multiple of TILE) and its neighbour one step
// back along the axis.  Both start from their tile-local labels; a pair whose labels
// equal those of the pair one step back along the face's fastest in-plane axis is
// the same union and is skipped, so only the first pair of every run unites (the
// comparison is safe whatever concurrent unions have changed: equal starting labels
// give the identical union).  AXIS 0 = last tensor axis (stride 1), 1 = middle (stride
// D), 2 = first spatial axis (stride D * W).
template <int AXIS>
__global__ __launch_bounds__(256) void ff_merge_axis(const uint8_t* __restrict__ grid, int32_t* L, int32_t B,
                                                     int32_t H, int32_t W, int32_t D) {
    const int32_t dims[3] = {D, W, H};
    const int32_t na = (dims[AXIS] - 1) / TILE;  // faces along the axis
    // in-face axes: the fastest one (`u`) and the other (`v`), plus the batch
    const int32_t ua = AXIS == 0 ? 1 : 0, va = AXIS == 2 ? 1 : 2;
    const int32_t nu = dims[ua], nv = dims[va];
    const int64_t total = (int64_t)B * na * nu * nv;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int32_t c[3];
    c[ua] = (int32_t)(t % nu);
    c[AXIS] = TILE * (1 + (int32_t)((t / nu) % na));
    const int64_t rest = t / ((int64_t)nu * na);
    c[va] = (int32_t)(rest % nv);
    const int32_t b = (int32_t)(rest / nv);
    const int64_t strides[3] = {1, D, (int64_t)D * W};
    const int64_t i = (int64_t)b * H * W * D + c[2] * strides[2] + c[1] * strides[1] + c[0];
    const int64_t step = strides[AXIS];
    if (grid[i] || grid[i - step]) return;
    const int32_t li = L[i], lj = L[i - step];
    if (c[ua] > 0) {
        const int64_t k = i - strides[ua];  // the same pair one step back along u
        if (!grid[k] && !grid[k - step] && L[k] == li && L[k - step] == lj) return;
    }
    unite(L, (int32_t)i, (int32_t)(i - step));
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
    const int64_t n_tiles = (int64_t)B * ((H + dev::TILE - 1) / dev::TILE) * ((W + dev::TILE - 1) / dev::TILE) *
                            ((D + dev::TILE - 1) / dev::TILE);
    hipLaunchKernelGGL(dev::ff_tiles, dim3((unsigned)n_tiles), dim3(dev::TILE_THREADS), 0, stream, grid, labels, H, W,
                       D);
    // unions across tile faces normal to each axis (no launch when the axis has one tile)
    const int64_t nd = (int64_t)B * ((D - 1) / dev::TILE) * W * H;
    const int64_t nw = (int64_t)B * ((W - 1) / dev::TILE) * D * H;
    const int64_t nh = (int64_t)B * ((H - 1) / dev::TILE) * D * W;
    if (nd > 0)
        hipLaunchKernelGGL(dev::ff_merge_axis<0>, dim3((unsigned)((nd + block - 1) / block)), dim3(block), 0, stream,
                           grid, labels, B, H, W, D);
    if (nw > 0)
        hipLaunchKernelGGL(dev::ff_merge_axis<1>, dim3((unsigned)((nw + block - 1) / block)), dim3(block), 0, stream,
                           grid, labels, B, H, W, D);
    if (nh > 0)
        hipLaunchKernelGGL(dev::ff_merge_axis<2>, dim3((unsigned)((nh + block - 1) / block)), dim3(block), 0, stream,
                           grid, labels, B, H, W, D);
    hipLaunchKernelGGL(dev::ff_flatten, dim3(grid_n), dim3(block), 0, stream, labels, ntot);
    return hipGetLastError();
}

}  // namespace cubvh_amd
