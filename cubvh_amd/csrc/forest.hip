// forest.hip -- one launch over many BVHs (no counterpart in the reference).
//
// The trees of M meshes are concatenated in HBM (node references and leaf ranges
// rebased on the host).  Queries arrive concatenated with CSR offsets
// [M + 1] and are walked with the packet traversal of query.hip, one wave per
// 64-query tile, every tile belonging to a single mesh.
//
// Ordering (graph-capture safe: kernels only, no memset, no library calls):
//   1. bounds of all queries                      (order.hip kernels)
//   2. per-query mesh (binary search in the CSR offsets) and bin =
//      mesh * 2^B + top B bits of the Morton code; atomicAdd on the bin's count
//      returns the query's rank inside its bin
//   3. exclusive scan of the bin counts (block scans + scan of block sums)
//   4. each mesh's segment is padded to a multiple of 64 slots; the query goes to
//      padded_base[mesh] + (scan[bin] - scan[first bin of mesh]) + rank
// With B = 0 the queries keep their input order inside each mesh.
//
// Tiles are processed either one wave per tile (grid = all tiles) or by a
// persistent grid whose waves take tiles from an atomic counter.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "forest.hpp"
#include "query.hpp"

namespace cubvh_amd {
namespace dev {

__device__ __forceinline__ int32_t mesh_of(const int64_t* __restrict__ off, int32_t m, int64_t i) {
    // largest k with off[k] <= i, for k in [0, m)
    int32_t lo = 0, hi = m - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// padded segment bases: pbase[m] = sum_{k<m} ceil(count_k / 64) * 64, pbase[M] = total.
// One block of 1024 threads, serial chunks.
__global__ void padded_bases(const int64_t* __restrict__ off, int32_t m, int64_t* __restrict__ pbase) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    int64_t s = 0;
    // a negative count (offsets that decrease) counts as an empty mesh
    for (int k = t * per; k < min(m, (t + 1) * per); ++k) s += ((max(off[k + 1] - off[k], (int64_t)0) + 63) / 64) * 64;
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t base = t > 0 ? part[t - 1] : 0;
    for (int k = t * per; k < min(m, (t + 1) * per); ++k) {
        pbase[k] = base;
        base += ((max(off[k + 1] - off[k], (int64_t)0) + 63) / 64) * 64;
    }
    if (t == 1023) pbase[m] = part[1023];
}

__device__ __forceinline__ uint32_t spread10f(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ float unordered_f(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ __launch_bounds__(256) void bin_count(int64_t n, const float* __restrict__ p, const int32_t* __restrict__ bnd,
                                                 const int64_t* __restrict__ off, int32_t m, int32_t bits,
                                                 uint32_t* __restrict__ hist, uint32_t* __restrict__ bin_of,
                                                 uint32_t* __restrict__ rank) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t mesh = mesh_of(off, m, i);
    uint32_t code = 0;
    if (bits > 0) {
        uint32_t q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float lo = unordered_f(bnd[k]), hi = unordered_f(bnd[3 + k]);
            float ext = hi - lo;
            float s = ext > 0.0f ? 1023.0f / ext : 0.0f;
            float t = fminf(fmaxf((p[3 * i + k] - lo) * s, 0.0f), 1023.0f);
            q[k] = (uint32_t)t;
        }
        code = ((spread10f(q[0]) << 2) | (spread10f(q[1]) << 1) | spread10f(q[2])) >> (30 - bits);
    }
    const uint32_t bin = ((uint32_t)mesh << bits) | code;
    bin_of[i] = bin;
    rank[i] = atomicAdd(hist + bin, 1u);
}

// In-place exclusive scan of 1024-element chunks; chunk totals to sums[].
__global__ __launch_bounds__(1024) void scan_chunks(uint32_t* __restrict__ x, int64_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 1024 + t;
    uint32_t v = i < n ? x[i] : 0u;
    s[t] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint32_t w = t >= d ? s[t - d] : 0u;
        __syncthreads();
        s[t] += w;
        __syncthreads();
    }
    if (i < n) x[i] = s[t] - v;
    if (t == 1023) sums[blockIdx.x] = s[1023];
}

// Exclusive scan of up to 1024 * 16 chunk totals in one block.
__global__ __launch_bounds__(1024) void scan_sums(uint32_t* __restrict__ sums, int32_t nchunks) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int per = (nchunks + 1023) / 1024;
    uint32_t acc = 0;
    for (int k = t * per; k < min(nchunks, (t + 1) * per); ++k) acc += sums[k];
    s[t] = acc;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint32_t w = t >= d ? s[t - d] : 0u;
        __syncthreads();
        s[t] += w;
        __syncthreads();
    }
    uint32_t base = t > 0 ? s[t - 1] : 0u;
    for (int k = t * per; k < min(nchunks, (t + 1) * per); ++k) {
        uint32_t v = sums[k];
        sums[k] = base;
        base += v;
    }
}

// Slots outside [0, npad_max) are skipped: with offsets that are not a CSR of n
// queries (offsets[0] != 0, decreasing, offsets[M] != n) the answers are wrong but
// every write stays inside the order array.
__global__ __launch_bounds__(256) void scatter_order(int64_t n, const int64_t* __restrict__ off, int32_t m,
                                                     int32_t bits, const uint32_t* __restrict__ scan,
                                                     const uint32_t* __restrict__ sums,
                                                     const int64_t* __restrict__ pbase,
                                                     const uint32_t* __restrict__ bin_of,
                                                     const uint32_t* __restrict__ rank, uint32_t* __restrict__ order,
                                                     int64_t npad_max) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bin = bin_of[i];
    const uint32_t mesh = bin >> bits;
    const uint32_t first = mesh << bits;
    const uint32_t pos_bin = scan[bin] + sums[bin >> 10];
    const uint32_t pos_first = scan[first] + sums[first >> 10];
    const int64_t slot = pbase[mesh] + (int64_t)(pos_bin - pos_first) + rank[i];
    if (slot >= 0 && slot < npad_max) order[slot] = (uint32_t)i;
}

__global__ void reset_counter(uint32_t* c) {
    if (threadIdx.x == 0) c[0] = 0u;
}

}  // namespace dev

namespace dev {

// Child references of n nodes of one mesh moved into the forest: internal nodes by
// node_base, leaf ranges by tri_base.
__global__ __launch_bounds__(256) void rebase_nodes(WideNode* nodes, int64_t n, int32_t node_base, int32_t tri_base) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WideNode& w = nodes[i];
    for (int k = 0; k < 4; ++k) {
        const int32_t r = w.child[k];
        if (r == CHILD_EMPTY) continue;
        if (r >= 0) {
            w.child[k] = r + node_base;
        } else {
            w.child[k] = make_leaf_ref((int64_t)leaf_ref_first(r) + tri_base, leaf_ref_count(r));
        }
    }
}

}  // namespace dev

hipError_t launch_rebase_nodes(WideNode* nodes, int64_t n, int32_t node_base, int32_t tri_base, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::rebase_nodes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nodes, n, node_base,
                       tri_base);
    return hipGetLastError();
}

size_t forest_workspace_bytes(int64_t n, int32_t n_meshes, int32_t bits) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t npad = n + 64 * (int64_t)n_meshes;
    const int64_t nbins = (int64_t)n_meshes << bits;
    const int64_t nchunks = (nbins + 1023) / 1024;
    return al(64) + al(64) + al(8 * (size_t)(n_meshes + 1)) + al(4 * (size_t)npad) + 2 * al(4 * (size_t)n) +
           al(4 * (size_t)nbins) + al(4 * (size_t)nchunks);
}

ForestWorkspace carve_forest_workspace(void* ws, int64_t n, int32_t n_meshes, int32_t bits) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t npad = n + 64 * (int64_t)n_meshes;
    const int64_t nbins = (int64_t)n_meshes << bits;
    const int64_t nchunks = (nbins + 1023) / 1024;
    char* w = (char*)ws;
    ForestWorkspace f;
    f.bounds = (int32_t*)w;
    w += al(64);
    f.counter = (uint32_t*)w;
    w += al(64);
    f.pbase = (int64_t*)w;
    w += al(8 * (size_t)(n_meshes + 1));
    f.order = (uint32_t*)w;
    w += al(4 * (size_t)npad);
    f.bin_of = (uint32_t*)w;
    w += al(4 * (size_t)n);
    f.rank = (uint32_t*)w;
    w += al(4 * (size_t)n);
    f.hist = (uint32_t*)w;
    w += al(4 * (size_t)nbins);
    f.sums = (uint32_t*)w;
    f.npad_max = npad;
    f.nbins = nbins;
    f.nchunks = nchunks;
    return f;
}

hipError_t forest_order(int64_t n, const float* positions, const int64_t* offsets, int32_t n_meshes, int32_t bits,
                        const ForestWorkspace& w, hipStream_t s) {
    const unsigned g256 = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(dev::padded_bases, dim3(1), dim3(1024), 0, s, offsets, n_meshes, w.pbase);
    hipLaunchKernelGGL(dev::fill_u32, dim3(2048), dim3(256), 0, s, w.order, w.npad_max, 0xffffffffu);
    hipLaunchKernelGGL(dev::fill_u32, dim3(std::max<int64_t>(1, std::min<int64_t>(2048, (w.nbins + 255) / 256))),
                       dim3(256), 0, s, w.hist, w.nbins, 0u);
    hipError_t e = launch_bounds(n, positions, w.bounds, s);
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(dev::bin_count, dim3(g256), dim3(256), 0, s, n, positions, w.bounds, offsets, n_meshes, bits,
                           w.hist, w.bin_of, w.rank);
    hipLaunchKernelGGL(dev::scan_chunks, dim3((unsigned)w.nchunks), dim3(1024), 0, s, w.hist, w.nbins, w.sums);
    hipLaunchKernelGGL(dev::scan_sums, dim3(1), dim3(1024), 0, s, w.sums, (int32_t)w.nchunks);
    if (n > 0)
        hipLaunchKernelGGL(dev::scatter_order, dim3(g256), dim3(256), 0, s, n, offsets, n_meshes, bits, w.hist, w.sums,
                           w.pbase, w.bin_of, w.rank, w.order, w.npad_max);
    hipLaunchKernelGGL(dev::reset_counter, dim3(1), dim3(64), 0, s, w.counter);
    return hipGetLastError();
}

}  // namespace cubvh_amd
