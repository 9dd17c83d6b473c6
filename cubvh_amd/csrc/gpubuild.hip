// gpubuild.hip -- construction of the 4-wide BVH on the GPU.
//
// Builds the same tree layout as the host builder (build.cpp, which replaces the
// reference's src/bvh.cu:515-620) from a mesh already in device memory:
//   1. faces    : gather each face's vertices, check indices and coordinates, bounds
//                 of the centroids and the largest |coordinate| (order-preserving
//                 int atomics);
//   2-3. topology, one of
//        median : three centroid-sorted face lists, every run halved level by level
//                 at the median of its axis of largest centroid extent (below);
//        SAH    : the same lists, every run split at the slot of least surface-area
//                 cost, from segmented scans of the face boxes (below);
//        linear : 63-bit Morton code of each centroid (21 bits per axis), a radix
//                 sort of (code, face) pairs, and the binary radix tree over the
//                 sorted codes (Karras, "Maximizing parallelism in the construction
//                 of BVHs, octrees and k-d trees", HPG 2012), equal codes ordered by
//                 position; one thread per internal node;
//   4. collapse : top-down, one launch per level: a wide node takes its binary
//                 node's two children and opens the child that comes first (the
//                 largest Morton cell, the largest box area, or the longer median
//                 run) until it has 4; subtrees of <= leaf_size faces become leaves;
//   5. arrays   : leaf-ordered triangles; then one pass over each child's run of
//                 faces gives its box (written into its wide node) and its bounding
//                 disk; face disks and the per-face constants of the distance
//                 formula -- the layouts and rules of the host build (bvh.hpp).
// A subtree of the radix tree covers a contiguous run of sorted faces, so child boxes
// and disks are reductions over ranges of the leaf order, and no bottom-up pass (with
// its cross-workgroup synchronization) is needed.  Child boxes are widened by the host
// build's margin, so query results are the same whichever builder made the tree.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "gpubuild.hpp"

namespace cubvh_amd {
namespace gb {

__device__ __forceinline__ int32_t ordered(float f) {
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float unordered(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// the counter layout is listed at GpuBuildWorkspace::counters
__global__ void init_counters(int32_t* c) {
    if (threadIdx.x < 3) c[threadIdx.x] = 0x7fffffff;                  // centroid lo (ordered)
    else if (threadIdx.x < 6) c[threadIdx.x] = (int32_t)0x80000000;    // centroid hi (ordered)
    else if (threadIdx.x < 64) c[threadIdx.x] = 0;                     // the rest
}

// Gathers face f's vertices into tri9[9f..] and checks them.  Each block writes its
// partial centroid bounds, largest |coordinate| and error bits to part[8 * block]
// (ordered ints; no atomics on shared counters); fold_partials combines them.
constexpr int kGatherBlocks = 2048;

// Internal children with more faces than this are reduced in kChunks chunks by separate
// workgroups (their indices listed in bigjobs, counted in counters[9]); the rest by one
// wave each.
constexpr int32_t kWaveJobMax = 4096;
constexpr int kChunks = 32;

__global__ __launch_bounds__(256) void gather_faces(int64_t n, const float* __restrict__ V, int64_t nv,
                                                    const int32_t* __restrict__ F, float* __restrict__ tri9,
                                                    int32_t* __restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float amax = 0.0f;
    int32_t err = 0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n; f += (int64_t)gridDim.x * blockDim.x) {
        float t[9];
        for (int k = 0; k < 3; ++k) {
            const int32_t vi = F[3 * f + k];
            if (vi < 0 || (int64_t)vi >= nv) {
                err |= 1;
                t[3 * k] = t[3 * k + 1] = t[3 * k + 2] = 0.0f;
                continue;
            }
            for (int a = 0; a < 3; ++a) t[3 * k + a] = V[3 * (int64_t)vi + a];
        }
        for (int k = 0; k < 9; ++k) {
            if (!isfinite(t[k])) err |= 2;
            amax = fmaxf(amax, fabsf(t[k]));
            tri9[9 * f + k] = t[k];
        }
        for (int a = 0; a < 3; ++a) {
            const float cen = (t[a] + t[3 + a] + t[6 + a]) * (1.0f / 3.0f);
            lo[a] = fminf(lo[a], cen);
            hi[a] = fmaxf(hi[a], cen);
        }
    }
    int32_t v[8] = {ordered(lo[0]), ordered(lo[1]), ordered(lo[2]), ordered(hi[0]), ordered(hi[1]), ordered(hi[2]),
                    __float_as_int(amax), err};
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 8; ++k) {
            const int32_t x = __shfl_xor(v[k], o);
            v[k] = k < 3 ? min(v[k], x) : k < 7 ? max(v[k], x) : (v[k] | x);
        }
    __shared__ int32_t sh[4][8];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 8; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        int32_t r = sh[0][k];
        for (int w = 1; w < 4; ++w) r = k < 3 ? min(r, sh[w][k]) : k < 7 ? max(r, sh[w][k]) : (r | sh[w][k]);
        part[8 * blockIdx.x + k] = r;
    }
}

// One block: folds the gather partials into counters [0..7].
__global__ __launch_bounds__(256) void fold_partials(const int32_t* __restrict__ part, int32_t nparts, int32_t* c) {
    __shared__ int32_t sh[256];
    for (int k = 0; k < 8; ++k) {
        int32_t r = k < 3 ? 0x7fffffff : k < 7 ? (int32_t)0x80000000 : 0;
        if (k == 6) r = 0;
        for (int i = threadIdx.x; i < nparts; i += 256) {
            const int32_t x = part[8 * i + k];
            r = k < 3 ? min(r, x) : k < 7 ? max(r, x) : (r | x);
        }
        sh[threadIdx.x] = r;
        __syncthreads();
        for (int step = 128; step > 0; step >>= 1) {
            if ((int)threadIdx.x < step) {
                const int32_t x = sh[threadIdx.x + step];
                sh[threadIdx.x] = k < 3 ? min(sh[threadIdx.x], x) : k < 7 ? max(sh[threadIdx.x], x) : (sh[threadIdx.x] | x);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) c[k] = sh[0];
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(256) void morton_keys(int64_t n, const float* __restrict__ tri9,
                                                   const int32_t* __restrict__ c, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    uint64_t code = 0;
    for (int a = 0; a < 3; ++a) {
        const float lo = unordered(c[a]), hi = unordered(c[3 + a]);
        const float cen = (tri9[9 * f + a] + tri9[9 * f + 3 + a] + tri9[9 * f + 6 + a]) * (1.0f / 3.0f);
        const float ext = hi - lo;
        float u = ext > 0.0f ? (cen - lo) / ext : 0.0f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        const uint64_t q = (uint64_t)fminf(u * 2097152.0f, 2097151.0f);
        code |= spread21(q) << (2 - a);
    }
    keys[f] = code;
    vals[f] = (uint32_t)f;
}

// Common-prefix length of sorted keys i and j (-1 outside [0, n)); equal keys are
// told apart by their positions.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clzll((long long)((uint64_t)i ^ (uint64_t)j));
}

// Internal node i of the radix tree over n sorted keys (n >= 2).  Children: >= 0
// internal node, < 0 leaf ~j (sorted position j).
__global__ __launch_bounds__(256) void radix_tree(int64_t n, const uint64_t* __restrict__ k, BinNode* __restrict__ bn,
                                                  int32_t* __restrict__ leaf_parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int64_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int64_t s = 0;
    for (int64_t div = 2;; div *= 2) {
        const int64_t t = (l + div - 1) / div;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    BinNode& b = bn[i];
    b.first = (int32_t)lo;
    b.last = (int32_t)hi;
    b.prefix = dnode;
    b.left = lo == g ? ~(int32_t)g : (int32_t)g;
    b.right = hi == g + 1 ? ~(int32_t)(g + 1) : (int32_t)(g + 1);
    if (b.left < 0) leaf_parent[g] = (int32_t)i;
    else bn[g].parent = (int32_t)i;
    if (b.right < 0) leaf_parent[g + 1] = (int32_t)i;
    else bn[g + 1].parent = (int32_t)i;
    if (i == 0) b.parent = -1;
}

// ---- median-split topology
// The tree of a median build depends only on n and the leaf size: every run of more
// than leaf_size slots [b, e) is halved at b + (e - b) / 2.  The faces are placed so that
// the lower half holds the faces whose centroids come first, by (centroid, face index),
// on the run's axis of largest centroid extent: the host build's object-median split
// (build.cpp), done level by level over three centroid-sorted face lists.

// Run [b, e) that slot p lies in after `level` halvings (runs of <= leaf slots stop).
__device__ __forceinline__ void run_at(int64_t p, int64_t n, int level, int32_t leaf, int64_t& b, int64_t& e) {
    b = 0;
    e = n;
    for (int l = 0; l < level && e - b > leaf; ++l) {
        const int64_t m = b + (e - b) / 2;
        if (p < m)
            e = m;
        else
            b = m;
    }
}

// Centroids, axis-major (cen[a * n + f]), as the gather computes them, and the sort
// keys of axis 0: the centroid's bits made order-preserving (-0 counts as +0).
__device__ __forceinline__ uint32_t float_key(float x) {
    if (x == 0.0f) x = 0.0f;
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__global__ __launch_bounds__(256) void centroids(int64_t n, const float* __restrict__ tri9, float* __restrict__ cen) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    for (int a = 0; a < 3; ++a)
        cen[a * n + f] = (tri9[9 * f + a] + tri9[9 * f + 3 + a] + tri9[9 * f + 6 + a]) * (1.0f / 3.0f);
}
__global__ __launch_bounds__(256) void axis_keys(int64_t n, const float* __restrict__ c, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    keys[f] = float_key(c[f]);
    vals[f] = (uint32_t)f;
}

// Axis of largest centroid extent of run [b, e) (the first on ties, as the host).
__device__ __forceinline__ int run_axis(const uint32_t* const* L, const float* __restrict__ cen, int64_t n, int64_t b,
                                        int64_t e) {
    int axis = 0;
    float ext = cen[L[0][e - 1]] - cen[L[0][b]];
    for (int a = 1; a < 3; ++a) {
        const float x = cen[a * n + L[a][e - 1]] - cen[a * n + L[a][b]];
        if (x > ext) {
            ext = x;
            axis = a;
        }
    }
    return axis;
}

struct Lists {
    const uint32_t* l[3];
};

// side[f] = 0 for the faces of the lower half of each run still being split, 1 above.
__global__ __launch_bounds__(256) void median_mark(int64_t n, int level, int32_t leaf, Lists L,
                                                   const float* __restrict__ cen, uint8_t* __restrict__ side) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t b, e;
    run_at(p, n, level, leaf, b, e);
    if (e - b <= leaf) return;
    const int axis = run_axis(L.l, cen, n, b, e);
    side[L.l[axis][p]] = p < b + (e - b) / 2 ? 0 : 1;
}

// Lower-half flags of the three lists at slot p, scanned together.
struct U3 {
    uint32_t x, y, z;
};
struct U3Sum {
    __host__ __device__ U3 operator()(const U3& a, const U3& b) const { return U3{a.x + b.x, a.y + b.y, a.z + b.z}; }
};
struct LowerFlags {
    Lists L;
    const uint8_t* side;
    __host__ __device__ U3 operator()(int64_t p) const {
        return U3{side[L.l[0][p]] == 0 ? 1u : 0u, side[L.l[1][p]] == 0 ? 1u : 0u, side[L.l[2][p]] == 0 ? 1u : 0u};
    }
};

// Stable partition of each run of each list by side (S: exclusive scan of the flags);
// runs no longer split are copied.
__global__ __launch_bounds__(256) void median_scatter(int64_t n, int level, int32_t leaf, Lists L, uint32_t* N0,
                                                      uint32_t* N1, uint32_t* N2, const U3* __restrict__ S,
                                                      const uint8_t* __restrict__ side) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t b, e;
    run_at(p, n, level, leaf, b, e);
    uint32_t* out[3] = {N0, N1, N2};
    if (e - b <= leaf) {
        for (int a = 0; a < 3; ++a) out[a][p] = L.l[a][p];
        return;
    }
    const int64_t m = b + (e - b) / 2;
    const U3 sp = S[p], sb = S[b];
    const uint32_t below[3] = {sp.x - sb.x, sp.y - sb.y, sp.z - sb.z};
    for (int a = 0; a < 3; ++a) {
        const uint32_t f = L.l[a][p];
        out[a][side[f] == 0 ? b + below[a] : m + (p - b - below[a])] = f;
    }
}

// The remaining halvings of one run of at most kFinishRun slots, in LDS: one workgroup
// per run, the same marks, scans and stable scatters as the global levels, on local
// face numbers; writes the run's final slot order into face_of.  loc: a scratch word
// per face (local number).
constexpr int kFinishRun = 2048;
constexpr int kFinishThreads = 1024;

__device__ __forceinline__ void local_run(int i, int k, int level, int32_t leaf, int& b, int& e) {
    b = 0;
    e = k;
    for (int l = 0; l < level && e - b > leaf; ++l) {
        const int m = b + (e - b) / 2;
        if (i < m)
            e = m;
        else
            b = m;
    }
}

// Exclusive scan over the workgroup of one 64-bit value per element, two elements per
// thread (elements 2t and 2t + 1); returns the scans of both.
__device__ __forceinline__ void block_scan2(uint64_t x0, uint64_t x1, uint64_t& s0, uint64_t& s1, uint64_t* sh) {
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const uint64_t t = x0 + x1;
    uint64_t v = t;
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    uint64_t base = 0;
    for (int w = 0; w < wave; ++w) base += sh[w];
    s0 = base + v - t;
    s1 = s0 + x0;
    __syncthreads();
}

__global__ __launch_bounds__(kFinishThreads) void median_finish(int64_t n, int global_levels, int32_t leaf, Lists L,
                                                                const float* __restrict__ cen,
                                                                uint32_t* __restrict__ loc,
                                                                uint32_t* __restrict__ face_of) {
    __shared__ uint32_t face[kFinishRun];
    __shared__ float c[3][kFinishRun];
    __shared__ uint16_t lst[2][3][kFinishRun];
    __shared__ uint8_t side[kFinishRun];
    __shared__ uint64_t sh[kFinishThreads / 64];
    // this run: follow the bits of the block index down the halvings
    int64_t b = 0, e = n;
    for (int l = 0; l < global_levels; ++l) {
        const int64_t m = b + (e - b) / 2;
        if ((blockIdx.x >> (global_levels - 1 - l)) & 1u)
            b = m;
        else
            e = m;
    }
    const int k = (int)(e - b);
    for (int i = (int)threadIdx.x; i < k; i += kFinishThreads) {
        const uint32_t f = L.l[0][b + i];
        face[i] = f;
        loc[f] = (uint32_t)i;
        for (int a = 0; a < 3; ++a) c[a][i] = cen[a * n + f];
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < k; i += kFinishThreads) {
        lst[0][0][i] = (uint16_t)i;
        lst[0][1][i] = (uint16_t)loc[L.l[1][b + i]];
        lst[0][2][i] = (uint16_t)loc[L.l[2][b + i]];
    }
    __syncthreads();
    int levels = 0;
    for (int run = k; run > leaf; run -= run / 2) ++levels;
    int cur = 0;
    for (int level = 0; level < levels; ++level) {
        // marks
        for (int i = (int)threadIdx.x; i < k; i += kFinishThreads) {
            int rb, re;
            local_run(i, k, level, leaf, rb, re);
            if (re - rb <= leaf) continue;
            int axis = 0;
            float ext = c[0][lst[cur][0][re - 1]] - c[0][lst[cur][0][rb]];
            for (int a = 1; a < 3; ++a) {
                const float x = c[a][lst[cur][a][re - 1]] - c[a][lst[cur][a][rb]];
                if (x > ext) {
                    ext = x;
                    axis = a;
                }
            }
            side[lst[cur][axis][i]] = i < rb + (re - rb) / 2 ? 0 : 1;
        }
        __syncthreads();
        // lower-half flags of the three lists, 16 bits each, scanned together
        const int i0 = 2 * (int)threadIdx.x, i1 = i0 + 1;
        uint64_t x0 = 0, x1 = 0;
        for (int a = 0; a < 3; ++a) {
            if (i0 < k && side[lst[cur][a][i0]] == 0) x0 |= 1ull << (16 * a);
            if (i1 < k && side[lst[cur][a][i1]] == 0) x1 |= 1ull << (16 * a);
        }
        uint64_t s0, s1;
        block_scan2(x0, x1, s0, s1, sh);
        __shared__ uint64_t scan[kFinishRun];
        if (i0 < k) scan[i0] = s0;
        if (i1 < k) scan[i1] = s1;
        __syncthreads();
        for (int i = (int)threadIdx.x; i < k; i += kFinishThreads) {
            int rb, re;
            local_run(i, k, level, leaf, rb, re);
            for (int a = 0; a < 3; ++a) {
                const uint16_t j = lst[cur][a][i];
                int pos = i;
                if (re - rb > leaf) {
                    const int rm = rb + (re - rb) / 2;
                    const int below = (int)(((scan[i] - scan[rb]) >> (16 * a)) & 0xffffu);
                    pos = side[j] == 0 ? rb + below : rm + (i - rb - below);
                }
                lst[cur ^ 1][a][pos] = j;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    for (int i = (int)threadIdx.x; i < k; i += kFinishThreads) face_of[b + i] = face[lst[cur][0][i]];
}

// ---- SAH topology
// Every run of more than leaf_size slots [f, l] is split where the surface-area cost
// A(left faces) * |left| + A(right faces) * |right| is least, A the surface area of
// the faces' bounding box grown by `grow` (1% of the diagonal of the centroid bounds)
// on every axis, over three candidate orders: the run's slots of each centroid-sorted
// list.  A split is a slot k of one list (faces at [f, k) go left), so the stable
// partition of the three lists is the median topology's.
//   * runs of at most kSahWaveRun slots: every slot is a candidate; one wave per run
//     (one to 16 slots per lane) takes the box of every prefix and suffix from
//     registers and wave scans;
//   * longer runs: kSahBins centroid bins per list; a bin is a contiguous range of the
//     sorted list, so each wave reduces its slots by (run, bin) and updates the bin's
//     count and box with one set of atomics per segment, and one wave per run sweeps
//     the bins as above.
// The binary nodes (every node, leaves included) go to `bn`, which the radix-tree
// collapse then turns into wide nodes; a node's prefix field holds its own box area
// made into an ordering key (smaller = larger box), so that the collapse opens the
// child of largest surface area first, as the host collapse does.  Below level
// kSahLevels the splits fall back to the median on the axis of largest centroid
// extent, so every run ends within a bounded number of levels.
constexpr float kSahDilate = 0.01f;
constexpr int kSahLevels = 40;
constexpr int kSahMaxLevels = 96;
constexpr int kSahWaveRun = 1024;
constexpr int kSahBins = 256;
constexpr int kSahBinWords = 7;  // count, then the box as order-preserving ints

struct Box6 {
    float lx, ly, lz, hx, hy, hz;
};
__device__ __forceinline__ Box6 empty_box() { return Box6{INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}; }
__device__ __forceinline__ Box6 box_union(const Box6& a, const Box6& b) {
    return Box6{fminf(a.lx, b.lx), fminf(a.ly, b.ly), fminf(a.lz, b.lz),
                fmaxf(a.hx, b.hx), fmaxf(a.hy, b.hy), fmaxf(a.hz, b.hz)};
}
__device__ __forceinline__ Box6 box_shfl_up(const Box6& b, int off, int width = 64) {
    return Box6{__shfl_up(b.lx, off, width), __shfl_up(b.ly, off, width), __shfl_up(b.lz, off, width),
                __shfl_up(b.hx, off, width), __shfl_up(b.hy, off, width), __shfl_up(b.hz, off, width)};
}
__device__ __forceinline__ Box6 box_shfl_down(const Box6& b, int off, int width = 64) {
    return Box6{__shfl_down(b.lx, off, width), __shfl_down(b.ly, off, width), __shfl_down(b.lz, off, width),
                __shfl_down(b.hx, off, width), __shfl_down(b.hy, off, width), __shfl_down(b.hz, off, width)};
}
__device__ __forceinline__ float grown_area(const Box6& b, float g) {
    const float d0 = b.hx - b.lx + g, d1 = b.hy - b.ly + g, d2 = b.hz - b.lz + g;
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

// Face boxes; every slot starts in the root run (node 0).
__global__ __launch_bounds__(256) void sah_boxes(int64_t n, const float* __restrict__ tri9, Box6* __restrict__ fbox,
                                                 int32_t* __restrict__ rid) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const float* t = tri9 + 9 * f;
    fbox[f] = Box6{fminf(fminf(t[0], t[3]), t[6]), fminf(fminf(t[1], t[4]), t[7]), fminf(fminf(t[2], t[5]), t[8]),
                   fmaxf(fmaxf(t[0], t[3]), t[6]), fmaxf(fmaxf(t[1], t[4]), t[7]), fmaxf(fmaxf(t[2], t[5]), t[8])};
    rid[f] = 0;
}

// Per level, 4 words of `stats`: active runs, the longest, the runs longer than
// kSahWaveRun (bin slot s: the run at active index bigrun[s]), unused.  All zero on
// entry.  Node ids: the root 0, the children of the run at active index e of a level
// base + 2e and base + 2e + 1, base = 1 + twice the active runs of the levels above.
__global__ void sah_root(int64_t n, BinNode* bn, int32_t* act, int32_t* stats, int32_t* bigslot, int32_t* bigrun,
                         float* grow, int32_t* c) {
    if (threadIdx.x != 0) return;
    bn[0] = BinNode{-1, -1, 0, (int32_t)(n - 1), -1, 0};
    act[0] = 0;
    stats[0] = 1;
    stats[1] = (int32_t)n;
    if (n > kSahWaveRun) {
        stats[2] = 1;
        bigslot[0] = 0;
        bigrun[0] = 0;
    }
    const float dx = unordered(c[3]) - unordered(c[0]), dy = unordered(c[4]) - unordered(c[1]),
                dz = unordered(c[5]) - unordered(c[2]);
    *grow = kSahDilate * sqrtf(dx * dx + dy * dy + dz * dz);
}

// Records the split of active run r = [f, l] at active index e (axis, first slot of
// the upper side k, its own box area) and its two children, in every lane with
// `want` (all 64 lanes of the wave call it); the children longer than leaf go to the
// next level's list, with one atomic per wave, or per block when `blk` (LDS, 33
// words) is given and the whole block calls it, and the longer than kSahWaveRun get a
// bin slot.  Past kSahLevels, or without a valid k, the median of the axis of largest
// centroid extent.
__device__ __forceinline__ void sah_emit(bool want, int level, int32_t node_base, int32_t e, int32_t r, int32_t f,
                                         int32_t l, unsigned long long kb, int axis, float area,
                                         int32_t* __restrict__ stats, int32_t* __restrict__ act_out,
                                         BinNode* __restrict__ bn, int32_t* __restrict__ bigslot,
                                         int32_t* __restrict__ bigrun, uint32_t* __restrict__ split, const Lists& L,
                                         const float* __restrict__ cen, int64_t n, int32_t leaf,
                                         int32_t* blk = nullptr) {
    const int lane = (int)(threadIdx.x & 63);
    int32_t k = f + (int32_t)(uint32_t)kb;  // kb: (cost, left count)
    const int32_t id = node_base + 2 * e;
    if (want) {
        if (level >= kSahLevels || kb == ~0ull || k <= f || k > l) {
            axis = run_axis(L.l, cen, n, f, (int64_t)l + 1);
            k = f + (l + 1 - f) / 2;
        }
        split[r] = ((uint32_t)axis << 28) | (uint32_t)k;
        bn[r].prefix = 0x3fffffff - (int32_t)(__float_as_uint(area) >> 1);
        bn[r].left = id;
        bn[r].right = id + 1;
        bn[id] = BinNode{-1, -1, f, k - 1, r, 0};
        bn[id + 1] = BinNode{-1, -1, k, l, r, 0};
    }
    const int32_t s0 = want ? k - f : 0, s1 = want ? l + 1 - k : 0;
    const bool c0 = s0 > leaf, c1 = s1 > leaf;
    const uint64_t b0 = __ballot(c0), b1 = __ballot(c1);
    int32_t* next = stats + 4 * (level + 1);
    int32_t base = 0;
    const int32_t total = __builtin_popcountll(b0) + __builtin_popcountll(b1);
    int32_t longest = max(c0 ? s0 : 0, c1 ? s1 : 0);
    for (int off = 32; off > 0; off >>= 1) longest = max(longest, __shfl_xor(longest, off));
    if (blk) {
        const int wave = (int)(threadIdx.x >> 6), nwaves = (int)(blockDim.x >> 6);
        if (lane == 0) {
            blk[wave] = total;
            blk[16 + wave] = longest;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t sum = 0, most = 0;
            for (int v = 0; v < nwaves; ++v) {
                const int32_t t = blk[v];
                blk[v] = sum;
                sum += t;
                most = max(most, blk[16 + v]);
            }
            blk[32] = sum > 0 ? atomicAdd(&next[0], sum) : 0;
            if (most > 0) atomicMax(&next[1], most);
        }
        __syncthreads();
        base = blk[32] + blk[wave];
        __syncthreads();  // blk is reused by the next call
    } else {
        if (lane == 0 && total > 0) base = atomicAdd(&next[0], total);
        base = __shfl(base, 0);
        if (lane == 0 && longest > 0) atomicMax(&next[1], longest);
    }
    const uint64_t below = (1ull << lane) - 1ull;
    const int32_t pos = base + __builtin_popcountll(b0 & below) + __builtin_popcountll(b1 & below);
    if (c0) act_out[pos] = id;
    if (c1) act_out[pos + (c0 ? 1 : 0)] = id + 1;
    if (s0 > kSahWaveRun) {
        const int32_t slot = atomicAdd(&next[2], 1);
        bigslot[id] = slot;
        bigrun[slot] = pos;
    }
    if (s1 > kSahWaveRun) {
        const int32_t slot = atomicAdd(&next[2], 1);
        bigslot[id + 1] = slot;
        bigrun[slot] = pos + (c0 ? 1 : 0);
    }
}

// Least SAH split of an ordered sequence of `items` items, computed by a group of W
// lanes (W = 16, 32 or 64; at most W * K items, lane j of the group holds items
// jK .. jK + K - 1), item i with box and count from load(i, box, count).  The split
// after item i puts items 0..i left; both sides must hold a face.  Returns (cost bits
// << 32 | faces on the left) of the least cost (ties to fewer on the left), or ~0;
// `all` = the union of every box.  The same in every lane of the group.
template <int K, int W, class Load>
__device__ __forceinline__ unsigned long long wave_least_split(int items, float g, Load load, Box6& all) {
    const int lane = (int)(threadIdx.x & (W - 1));
    const int i0 = lane * K;
    constexpr bool kKeep = K <= 4;  // keep the loaded items for the second pass
    Box6 sbox[K], ibox[kKeep ? K : 1];
    int32_t scnt[K], icnt[kKeep ? K : 1];
    Box6 acc = empty_box();
    int32_t accn = 0;
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
        Box6 b = empty_box();
        int32_t m = 0;
        if (i0 + i < items) {
            load(i0 + i, b, m);
            if (m > 0) acc = box_union(acc, b);
            accn += m;
        }
        if constexpr (kKeep) {
            ibox[i] = b;
            icnt[i] = m;
        }
        sbox[i] = acc;
        scnt[i] = accn;
    }
    // boxes and counts of the lanes below and above this one
    Box6 below = box_shfl_up(acc, 1, W), above = box_shfl_down(acc, 1, W);
    int32_t nbelow = __shfl_up(accn, 1, W), nabove = __shfl_down(accn, 1, W);
    if (lane == 0) {
        below = empty_box();
        nbelow = 0;
    }
    if (lane == W - 1) {
        above = empty_box();
        nabove = 0;
    }
    for (int off = 1; off < W; off <<= 1) {
        const Box6 bu = box_shfl_up(below, off, W), ad = box_shfl_down(above, off, W);
        const int32_t nu = __shfl_up(nbelow, off, W), nd = __shfl_down(nabove, off, W);
        if (lane >= off) {
            below = box_union(below, bu);
            nbelow += nu;
        }
        if (lane + off < W) {
            above = box_union(above, ad);
            nabove += nd;
        }
    }
    all = box_union(below, box_union(acc, above));
    Box6 pre = below;
    int32_t npre = nbelow;
    unsigned long long key = ~0ull;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (i0 + i < items) {
            Box6 b;
            int32_t m;
            if constexpr (kKeep) {
                b = ibox[i];
                m = icnt[i];
            } else {
                load(i0 + i, b, m);
            }
            if (m > 0) pre = box_union(pre, b);
            npre += m;
            const Box6 right = i + 1 < K ? box_union(sbox[i + 1], above) : above;
            const int32_t nright = (i + 1 < K ? scnt[i + 1] : 0) + nabove;
            if (npre > 0 && nright > 0) {
                const float cost = grown_area(pre, g) * (float)npre + grown_area(right, g) * (float)nright;
                const unsigned long long x = ((unsigned long long)__float_as_uint(cost) << 32) | (uint32_t)npre;
                if (x < key) key = x;
            }
        }
    }
    for (int off = W / 2; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)key, off, W), hi = __shfl_xor((uint32_t)(key >> 32), off, W);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        if (o < key) key = o;
    }
    return key;
}

// Runs of LO < size <= W * K slots, one group of W lanes each (W = 16 or 32: several
// runs per wave): every slot of each list a candidate.  The loop is block-uniform
// (appends to the next level's list take one atomic per block); groups without a run
// of their class compute nothing (items = 0).
template <int K, int W, int LO, int BLOCK>
__global__ __launch_bounds__(BLOCK) void sah_wave_runs(int level, int32_t node_base, int32_t* __restrict__ stats,
                                                     const int32_t* __restrict__ act_in, int32_t* __restrict__ act_out,
                                                     BinNode* __restrict__ bn, int32_t* __restrict__ bigslot,
                                                     int32_t* __restrict__ bigrun, uint32_t* __restrict__ split,
                                                     const Box6* __restrict__ fbox, const float* __restrict__ grow,
                                                     Lists L, const float* __restrict__ cen, int64_t n, int32_t leaf) {
    constexpr int kPerBlock = BLOCK / W;  // runs per block and round
    __shared__ int32_t blk[33];
    const int lane = (int)(threadIdx.x & (W - 1));
    const int32_t m = stats[4 * level];
    const float g = *grow;
    for (int32_t e0 = (int32_t)blockIdx.x * kPerBlock; e0 < m; e0 += (int32_t)gridDim.x * kPerBlock) {
        const int32_t e = e0 + (int32_t)(threadIdx.x / W);
        int32_t r = 0, f = 0, l = 0;
        bool valid = e < m;
        if (valid) {
            r = act_in[e];
            f = bn[r].first;
            l = bn[r].last;
            valid = l - f + 1 > LO && l - f + 1 <= W * K;
        }
        const int32_t items = valid ? l - f + 1 : 0;
        unsigned long long kb = ~0ull;
        int axis = 0;
        Box6 all;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t* list = L.l[a] + f;
            auto load = [&](int i, Box6& b, int32_t& cnt) {
                b = fbox[list[i]];
                cnt = 1;
            };
            const unsigned long long x = wave_least_split<K, W>(items, g, load, all);
            if (x < kb) {
                kb = x;
                axis = a;
            }
        }
        sah_emit(valid && lane == 0, level, node_base, e, r, f, l, kb, axis, grown_area(all, g), stats, act_out, bn,
                 bigslot, bigrun, split, L, cen, n, leaf, blk);
    }
}

// Empty bins (count 0, box +inf / -inf) for the level's `nbig` long runs.
__global__ __launch_bounds__(256) void sah_bin_clear(const int32_t* __restrict__ stats, int level,
                                                     int32_t* __restrict__ bins) {
    const int64_t nb = (int64_t)stats[4 * level + 2] * 3 * kSahBins;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t* w = bins + kSahBinWords * i;
        w[0] = 0;
        w[1] = w[2] = w[3] = ordered(INFINITY);
        w[4] = w[5] = w[6] = ordered(-INFINITY);
    }
}

// Bins of list a: slot p of a long run falls in bin floor((c - lo) * kSahBins / (hi - lo))
// of its centroid c, lo and hi the run's first and last centroids along a (the list is
// sorted), so the bins of a run are contiguous and ascending along the list.
__global__ __launch_bounds__(256) void sah_bin(int64_t n, int a, const int32_t* __restrict__ rid,
                                               const BinNode* __restrict__ bn, const int32_t* __restrict__ bigslot,
                                               const uint32_t* __restrict__ list, const float* __restrict__ cen,
                                               const Box6* __restrict__ fbox, int32_t* __restrict__ bins) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    const int32_t r = p < n ? rid[p] : -1;
    int32_t seg = -1, cnt = 0;
    Box6 b = empty_box();
    if (r >= 0) {
        const int32_t f = bn[r].first, l = bn[r].last;
        if (l - f + 1 > kSahWaveRun) {
            const float* ca = cen + a * n;
            const float lo = ca[list[f]], hi = ca[list[l]], ext = hi - lo;
            const uint32_t face = list[p];
            int bin = 0;
            if (ext > 0.0f) bin = min(max((int)((ca[face] - lo) * ((float)kSahBins / ext)), 0), kSahBins - 1);
            seg = bigslot[r] * kSahBins + bin;
            b = fbox[face];
            cnt = 1;
        }
    }
    for (int off = 1; off < 64; off <<= 1) {
        const Box6 o = box_shfl_down(b, off);
        const int32_t os = __shfl_down(seg, off), on = __shfl_down(cnt, off);
        if (lane + off < 64 && os == seg) {
            b = box_union(b, o);
            cnt += on;
        }
    }
    const int32_t sprev = __shfl_up(seg, 1);
    if (seg >= 0 && (lane == 0 || sprev != seg)) {
        int32_t* w = bins + (int64_t)kSahBinWords * (((int64_t)(seg / kSahBins) * 3 + a) * kSahBins + seg % kSahBins);
        atomicAdd(w, cnt);
        atomicMin(w + 1, ordered(b.lx));
        atomicMin(w + 2, ordered(b.ly));
        atomicMin(w + 3, ordered(b.lz));
        atomicMax(w + 4, ordered(b.hx));
        atomicMax(w + 5, ordered(b.hy));
        atomicMax(w + 6, ordered(b.hz));
    }
}

// One long run per wave: the least split over the bins of its three lists.
__global__ __launch_bounds__(256) void sah_bin_choose(int level, int32_t node_base, int32_t* __restrict__ stats,
                                                      const int32_t* __restrict__ bins,
                                                      const int32_t* __restrict__ act_in, int32_t* __restrict__ act_out,
                                                      BinNode* __restrict__ bn, int32_t* __restrict__ bigslot,
                                                      int32_t* __restrict__ bigrun, uint32_t* __restrict__ split,
                                                      const float* __restrict__ grow, Lists L,
                                                      const float* __restrict__ cen, int64_t n, int32_t leaf) {
    const int lane = (int)(threadIdx.x & 63);
    const int32_t nbig = stats[4 * level + 2];
    const int32_t waves = (int32_t)((gridDim.x * blockDim.x) >> 6);
    for (int32_t slot = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); slot < nbig; slot += waves) {
        const int32_t e = bigrun[slot];
        const int32_t r = act_in[e];
        const int32_t f = bn[r].first, l = bn[r].last;
        const float g = *grow;
        unsigned long long kb = ~0ull;
        int axis = 0;
        Box6 all;
        for (int a = 0; a < 3; ++a) {
            const int32_t* base = bins + (int64_t)kSahBinWords * (((int64_t)slot * 3 + a) * kSahBins);
            auto load = [&](int i, Box6& b, int32_t& cnt) {
                const int32_t* w = base + kSahBinWords * i;
                cnt = w[0];
                b = Box6{unordered(w[1]), unordered(w[2]), unordered(w[3]), unordered(w[4]), unordered(w[5]),
                         unordered(w[6])};
            };
            const unsigned long long x = wave_least_split<kSahBins / 64, 64>(kSahBins, g, load, all);
            if (x < kb) {
                kb = x;
                axis = a;
            }
        }
        sah_emit(lane == 0, level, node_base, e, r, f, l, kb, axis, grown_area(all, g), stats, act_out, bn, bigslot,
                 bigrun, split, L, cen, n, leaf);
    }
}

// side[f] = 0 for the faces before the split slot of the run's chosen list, 1 after.
__global__ __launch_bounds__(256) void sah_mark(int64_t n, const int32_t* __restrict__ rid,
                                                const uint32_t* __restrict__ split, Lists L,
                                                uint8_t* __restrict__ side) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = rid[p];
    if (r < 0) return;
    const uint32_t s = split[r];
    side[L.l[s >> 28][p]] = p < (int64_t)(s & 0x0fffffffu) ? 0 : 1;
}

// Stable partition of every active run of each list by side (S: exclusive scan of the
// flags), and every slot's run for the next level (-1 once its run is a leaf).
__global__ __launch_bounds__(256) void sah_scatter(int64_t n, int32_t* __restrict__ rid,
                                                   const BinNode* __restrict__ bn,
                                                   const uint32_t* __restrict__ split, Lists L, uint32_t* N0,
                                                   uint32_t* N1, uint32_t* N2, const U3* __restrict__ S,
                                                   const uint8_t* __restrict__ side, int32_t leaf, int32_t* c) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t* out[3] = {N0, N1, N2};
    const int32_t r = rid[p];
    if (r < 0) {
        for (int a = 0; a < 3; ++a) out[a][p] = L.l[a][p];
        return;
    }
    const BinNode nd = bn[r];
    const int64_t b = nd.first, k = (int64_t)(split[r] & 0x0fffffffu);
    const U3 sp = S[p], sb = S[b];
    const uint32_t below[3] = {sp.x - sb.x, sp.y - sb.y, sp.z - sb.z};
    for (int a = 0; a < 3; ++a) {
        const uint32_t f = L.l[a][p];
        const int64_t d = side[f] == 0 ? b + below[a] : k + (p - b - below[a]);
        if (d < b || d > nd.last) {  // the lists disagree on the run's faces: never for a sound build
            c[14] = 1;
            continue;
        }
        out[a][d] = f;
    }
    const bool lower = p < k;
    const int64_t size = lower ? k - b : (int64_t)nd.last + 1 - k;
    rid[p] = size > leaf ? (lower ? nd.left : nd.right) : -1;
}

// Outward rounding by a relative and an absolute margin (the host build's rule).
__device__ __forceinline__ float pad_down(float v, double abs_pad) {
    return __double2float_rd((double)v - fabs((double)v) * 1.0e-6 - abs_pad);
}
__device__ __forceinline__ float pad_up(float v, double abs_pad) {
    return __double2float_ru((double)v + fabs((double)v) * 1.0e-6 + abs_pad);
}

// Exclusive prefix sum of x over the lanes of the wave, and the wave's total.
__device__ __forceinline__ int32_t wave_scan(int32_t x, int32_t& total) {
    const int lane = (int)(threadIdx.x & 63);
    int32_t v = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    total = __shfl(v, 63);
    return v - x;
}

// One wide node per thread: entry (binary node, wide node, depth).  Writes the
// node's child references; the child boxes come later from the range passes.  The
// output slots (wide node ids, next-level queue, leaf and subtree job lists) are
// taken with one atomic per wave and counter.
__device__ __forceinline__ void collapse_entry(const int4* __restrict__ in, int32_t e, int32_t n_in,
                                               int4* __restrict__ out, int32_t* n_out,
                                               const BinNode* __restrict__ bn, int32_t leaf_size,
                                               WideNode* __restrict__ wn, int4* __restrict__ leaves,
                                               int4* __restrict__ jobs, int32_t* __restrict__ bigjobs, int32_t* c) {
    const bool valid = e < n_in;
    int4 ent = make_int4(0, 0, 0, 0);
    int32_t kids[4] = {0, 0, 0, 0};
    int nk = 0;
    auto size_of = [&](int32_t r) -> int32_t { return r < 0 ? 1 : bn[r].last - bn[r].first + 1; };
    if (valid) {
        ent = in[e];
        const BinNode& root = bn[ent.x];
        kids[0] = root.left;
        kids[1] = root.right;
        nk = 2;
        while (nk < 4) {
            int best = -1, best_prefix = 1 << 30, best_size = 0;
            for (int q = 0; q < nk; ++q) {
                const int32_t r = kids[q];
                if (r < 0) continue;
                const int32_t sz = size_of(r);
                if (sz <= leaf_size) continue;
                const int32_t pf = bn[r].prefix;
                if (pf < best_prefix || (pf == best_prefix && sz > best_size)) {
                    best = q;
                    best_prefix = pf;
                    best_size = sz;
                }
            }
            if (best < 0) break;
            const int32_t o = kids[best];
            for (int q = nk; q > best + 1; --q) kids[q] = kids[q - 1];
            kids[best] = bn[o].left;
            kids[best + 1] = bn[o].right;
            ++nk;
        }
    }
    int32_t n_leaf = 0, n_int = 0;
    for (int q = 0; q < nk; ++q) {
        if (size_of(kids[q]) <= leaf_size) ++n_leaf;
        else ++n_int;
    }
    int32_t tot_leaf, tot_int;
    const int32_t off_leaf = wave_scan(n_leaf, tot_leaf);
    const int32_t off_int = wave_scan(n_int, tot_int);
    int32_t b_leaf = 0, b_wide = 0, b_queue = 0, b_job = 0;
    if ((threadIdx.x & 63) == 0) {
        if (tot_leaf) b_leaf = atomicAdd(&c[10], tot_leaf);
        if (tot_int) {
            b_wide = atomicAdd(&c[8], tot_int);
            b_queue = atomicAdd(n_out, tot_int);
            b_job = atomicAdd(&c[11], tot_int);
        }
    }
    b_leaf = __shfl(b_leaf, 0) + off_leaf;
    b_wide = __shfl(b_wide, 0) + off_int;
    b_queue = __shfl(b_queue, 0) + off_int;
    b_job = __shfl(b_job, 0) + off_int;
    int32_t depth = valid ? ent.z : 0;
    for (int o = 32; o > 0; o >>= 1) depth = max(depth, __shfl_xor(depth, o));
    if ((threadIdx.x & 63) == 0) atomicMax(&c[12], depth);
    if (!valid) return;
    WideNode w;
    for (int q = 0; q < 4; ++q) {
        w.lo_x[q] = w.lo_y[q] = w.lo_z[q] = INFINITY;
        w.hi_x[q] = w.hi_y[q] = w.hi_z[q] = -INFINITY;
        w.child[q] = CHILD_EMPTY;
    }
    w.n_children = nk;
    w.pad[0] = w.pad[1] = w.pad[2] = 0;
    for (int q = 0; q < nk; ++q) {
        const int32_t r = kids[q];
        const int32_t first = r < 0 ? ~r : bn[r].first;
        const int32_t cnt = size_of(r);
        if (cnt <= leaf_size) {
            w.child[q] = make_leaf_ref(first, cnt);
            leaves[b_leaf++] = make_int4(first, cnt, ent.y, q);
        } else {
            const int32_t id = b_wide++;
            w.child[q] = id;
            out[b_queue++] = make_int4(r, id, ent.z + 1, 0);
            if (cnt > kWaveJobMax) bigjobs[atomicAdd(&c[9], 1)] = b_job;
            jobs[b_job++] = make_int4(ent.y, q, first, cnt);
        }
    }
    wn[ent.y] = w;
}

// Median topology: one wide node per thread, entry (first slot, end slot, wide node,
// depth).  Its children are the run's halves, each halved once more when it is still
// longer than leaf_size; runs of at most leaf_size slots become leaves.
__device__ __forceinline__ void collapse_run_entry(const int4* __restrict__ in, int32_t e, int32_t n_in,
                                                   int4* __restrict__ out, int32_t* n_out, int32_t leaf_size,
                                                   WideNode* __restrict__ wn, int4* __restrict__ leaves,
                                                   int4* __restrict__ jobs, int32_t* __restrict__ bigjobs, int32_t* c) {
    const bool valid = e < n_in;
    int4 ent = make_int4(0, 0, 0, 0);
    int32_t kb[4] = {0, 0, 0, 0}, ke[4] = {0, 0, 0, 0};
    int nk = 0;
    if (valid) {
        ent = in[e];
        const int32_t b = ent.x, end = ent.y, m = b + (end - b) / 2;
        const int32_t hb[2] = {b, m}, he[2] = {m, end};
        for (int h = 0; h < 2; ++h) {
            if (he[h] - hb[h] > leaf_size) {
                const int32_t q = hb[h] + (he[h] - hb[h]) / 2;
                kb[nk] = hb[h];
                ke[nk++] = q;
                kb[nk] = q;
                ke[nk++] = he[h];
            } else {
                kb[nk] = hb[h];
                ke[nk++] = he[h];
            }
        }
    }
    int32_t n_leaf = 0, n_int = 0;
    for (int q = 0; q < nk; ++q) {
        if (ke[q] - kb[q] <= leaf_size) ++n_leaf;
        else ++n_int;
    }
    int32_t tot_leaf, tot_int;
    const int32_t off_leaf = wave_scan(n_leaf, tot_leaf);
    const int32_t off_int = wave_scan(n_int, tot_int);
    int32_t b_leaf = 0, b_wide = 0, b_queue = 0, b_job = 0;
    if ((threadIdx.x & 63) == 0) {
        if (tot_leaf) b_leaf = atomicAdd(&c[10], tot_leaf);
        if (tot_int) {
            b_wide = atomicAdd(&c[8], tot_int);
            b_queue = atomicAdd(n_out, tot_int);
            b_job = atomicAdd(&c[11], tot_int);
        }
    }
    b_leaf = __shfl(b_leaf, 0) + off_leaf;
    b_wide = __shfl(b_wide, 0) + off_int;
    b_queue = __shfl(b_queue, 0) + off_int;
    b_job = __shfl(b_job, 0) + off_int;
    int32_t depth = valid ? ent.z : 0;
    for (int o = 32; o > 0; o >>= 1) depth = max(depth, __shfl_xor(depth, o));
    if ((threadIdx.x & 63) == 0) atomicMax(&c[12], depth);
    if (!valid) return;
    WideNode w;
    for (int q = 0; q < 4; ++q) {
        w.lo_x[q] = w.lo_y[q] = w.lo_z[q] = INFINITY;
        w.hi_x[q] = w.hi_y[q] = w.hi_z[q] = -INFINITY;
        w.child[q] = CHILD_EMPTY;
    }
    w.n_children = nk;
    w.pad[0] = w.pad[1] = w.pad[2] = 0;
    for (int q = 0; q < nk; ++q) {
        const int32_t first = kb[q], cnt = ke[q] - kb[q];
        if (cnt <= leaf_size) {
            w.child[q] = make_leaf_ref(first, cnt);
            leaves[b_leaf++] = make_int4(first, cnt, ent.w, q);
        } else {
            const int32_t id = b_wide++;
            w.child[q] = id;
            out[b_queue++] = make_int4(first, first + cnt, ent.z + 1, id);
            if (cnt > kWaveJobMax) bigjobs[atomicAdd(&c[9], 1)] = b_job;
            jobs[b_job++] = make_int4(ent.w, q, first, cnt);
        }
    }
    wn[ent.w] = w;
}

__global__ __launch_bounds__(256) void collapse_run_level(const int4* __restrict__ in, const int32_t* n_in_ptr,
                                                          int4* __restrict__ out, int32_t* n_out, int32_t leaf_size,
                                                          WideNode* __restrict__ wn, int4* __restrict__ leaves,
                                                          int4* __restrict__ jobs, int32_t* __restrict__ bigjobs,
                                                          int32_t* c) {
    const int32_t n_in = n_in_ptr ? *n_in_ptr : 1;  // the first level holds the root alone
    for (int32_t e0 = (int32_t)(blockIdx.x * blockDim.x); e0 < n_in; e0 += (int32_t)(gridDim.x * blockDim.x))
        collapse_run_entry(in, e0 + (int32_t)threadIdx.x, n_in, out, n_out, leaf_size, wn, leaves, jobs, bigjobs, c);
}

__global__ __launch_bounds__(256) void collapse_level(const int4* __restrict__ in, const int32_t* n_in_ptr,
                                                      int4* __restrict__ out, int32_t* n_out,
                                                      const BinNode* __restrict__ bn, int32_t leaf_size,
                                                      WideNode* __restrict__ wn, int4* __restrict__ leaves,
                                                      int4* __restrict__ jobs, int32_t* __restrict__ bigjobs,
                                                      int32_t* c) {
    const int32_t n_in = n_in_ptr ? *n_in_ptr : 1;  // the first level holds the root alone
    for (int32_t e0 = (int32_t)(blockIdx.x * blockDim.x); e0 < n_in; e0 += (int32_t)(gridDim.x * blockDim.x))
        collapse_entry(in, e0 + (int32_t)threadIdx.x, n_in, out, n_out, bn, leaf_size, wn, leaves, jobs, bigjobs, c);
}

// Writes child slot q of wide node w: the box of its faces widened by the margin.
__device__ __forceinline__ void set_child_box(WideNode* __restrict__ wn, int32_t w, int32_t q, const double* lo,
                                              const double* hi, double abs_pad) {
    WideNode& n = wn[w];
    n.lo_x[q] = pad_down((float)lo[0], abs_pad);
    n.lo_y[q] = pad_down((float)lo[1], abs_pad);
    n.lo_z[q] = pad_down((float)lo[2], abs_pad);
    n.hi_x[q] = pad_up((float)hi[0], abs_pad);
    n.hi_y[q] = pad_up((float)hi[1], abs_pad);
    n.hi_z[q] = pad_up((float)hi[2], abs_pad);
}

__device__ __forceinline__ double margin_of(const int32_t* c) {
    return (double)(fmaxf(__int_as_float(c[6]), 1e-30f) * 4.0e-6f);
}

// Leaf-ordered triangles (bvh.hpp) and the slot of original face 0.
__global__ __launch_bounds__(256) void emit_tris(int64_t n, const uint32_t* __restrict__ face_of,
                                                 const float* __restrict__ tri9, float4* __restrict__ tris,
                                                 int32_t* c) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t f = face_of[s];
    const float* t = tri9 + 9 * (int64_t)f;
    tris[3 * s + 0] = make_float4(t[0], t[1], t[2], __int_as_float((int32_t)f));
    tris[3 * s + 1] = make_float4(t[3], t[4], t[5], 0.0f);
    tris[3 * s + 2] = make_float4(t[6], t[7], t[8], 0.0f);
    if (f == 0) c[13] = (int32_t)s;
}

// Bounding disk of the vertices of slots [first, first + count): centre of their box,
// area-weighted mean normal, radius and thickness measured in double from the rounded
// centre along the rounded normal and rounded outward (the host build's rule).
struct DiskAcc {
    double lo[3], hi[3], n[3];
};

__device__ __forceinline__ void acc_init(DiskAcc& a) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = INFINITY;
        a.hi[k] = -INFINITY;
        a.n[k] = 0.0;
    }
}

__device__ __forceinline__ void acc_face(DiskAcc& a, const float4* __restrict__ tris, int64_t s) {
    double v[3][3];
    for (int i = 0; i < 3; ++i) {
        const float4 r = tris[3 * s + i];
        v[i][0] = r.x;
        v[i][1] = r.y;
        v[i][2] = r.z;
    }
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = v[1][k] - v[0][k];
        e2[k] = v[2][k] - v[0][k];
        for (int i = 0; i < 3; ++i) {
            a.lo[k] = fmin(a.lo[k], v[i][k]);
            a.hi[k] = fmax(a.hi[k], v[i][k]);
        }
    }
    a.n[0] += e1[1] * e2[2] - e1[2] * e2[1];
    a.n[1] += e1[2] * e2[0] - e1[0] * e2[2];
    a.n[2] += e1[0] * e2[1] - e1[1] * e2[0];
}

struct DiskFrame {
    float c[3], n[3];
};

__device__ __forceinline__ DiskFrame frame_of(const double* centre, const double* dir) {
    DiskFrame f;
    for (int k = 0; k < 3; ++k) {
        f.c[k] = (float)centre[k];
        f.n[k] = 0.0f;
    }
    const double nl = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (nl > 0.0 && isfinite(nl))
        for (int k = 0; k < 3; ++k) f.n[k] = (float)(dir[k] / nl);
    return f;
}

// Largest in-plane distance (x) and height (y) of slot s's vertices in frame f.
__device__ __forceinline__ double2 face_extent(const DiskFrame& f, const float4* __restrict__ tris, int64_t s) {
    double r = 0.0, th = 0.0;
    for (int i = 0; i < 3; ++i) {
        const float4 v = tris[3 * s + i];
        const double d[3] = {(double)v.x - f.c[0], (double)v.y - f.c[1], (double)v.z - f.c[2]};
        const double h = d[0] * f.n[0] + d[1] * f.n[1] + d[2] * f.n[2];
        const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        r = fmax(r, sqrt(fmax(dd - h * h, 0.0)));
        th = fmax(th, fabs(h));
    }
    return make_double2(r, th);
}

__device__ __forceinline__ void store_disk(const DiskFrame& f, double r, double th, float4* __restrict__ o) {
    r = r * (1.0 + 1e-6) + 1e-30;
    th = th * (1.0 + 1e-6) + 1e-30;
    o[0] = make_float4(f.c[0], f.c[1], f.c[2], __double2float_ru(r));
    o[1] = make_float4(f.n[0], f.n[1], f.n[2], __double2float_ru(th));
}

// fp32 operations of the distance formula, with its fused operations only.
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float vdot(V3 a, V3 b) {
    return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return V3{__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
              __builtin_fmaf(a.x, b.y, -(a.y * b.x))};
}

// Per face: bounding disk (centroid, plane normal) and the distance-formula constants.
__global__ __launch_bounds__(256) void face_records(int64_t n, const float4* __restrict__ tris,
                                                    float4* __restrict__ tbnd, float4* __restrict__ tpre) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    DiskAcc acc;
    acc_init(acc);
    acc_face(acc, tris, s);
    double cen[3];
    for (int k = 0; k < 3; ++k) {
        double sum = 0.0;
        for (int i = 0; i < 3; ++i) sum += (double)((const float*)&tris[3 * s + i])[k];
        cen[k] = sum / 3.0;
    }
    const DiskFrame fr = frame_of(cen, acc.n);
    const double2 ext = face_extent(fr, tris, s);
    store_disk(fr, ext.x, ext.y, tbnd + 2 * s);
    const float4 r0 = tris[3 * s], r1 = tris[3 * s + 1], r2 = tris[3 * s + 2];
    const V3 a{r0.x, r0.y, r0.z}, b{r1.x, r1.y, r1.z}, c{r2.x, r2.y, r2.z};
    const V3 v21 = vsub(b, a), v32 = vsub(c, b), v13 = vsub(a, c);
    const V3 nn = vcross(v21, v13);
    const V3 g1 = vcross(v21, nn), g2 = vcross(v32, nn), g3 = vcross(v13, nn);
    float4* o = tpre + 9 * s;
    o[0] = make_float4(a.x, a.y, a.z, r0.w);
    o[1] = make_float4(b.x, b.y, b.z, vdot(v21, v21));
    o[2] = make_float4(c.x, c.y, c.z, vdot(v32, v32));
    o[3] = make_float4(v21.x, v21.y, v21.z, vdot(v13, v13));
    o[4] = make_float4(v32.x, v32.y, v32.z, vdot(nn, nn));
    o[5] = make_float4(v13.x, v13.y, v13.z, nn.x);
    o[6] = make_float4(g1.x, g1.y, g1.z, nn.y);
    o[7] = make_float4(g2.x, g2.y, g2.z, nn.z);
    o[8] = make_float4(g3.x, g3.y, g3.z, 0.0f);
}

// One thread per leaf child (first, count, wide node, slot): its box into the wide
// node and its disk at lbnd[2 * first].
__global__ __launch_bounds__(256) void leaf_records(const int4* __restrict__ leaves, const int32_t* c,
                                                    const float4* __restrict__ tris, float4* __restrict__ lbnd,
                                                    WideNode* __restrict__ wn) {
    const int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= c[10]) return;
    const int4 lf = leaves[e];
    DiskAcc acc;
    acc_init(acc);
    for (int j = 0; j < lf.y; ++j) acc_face(acc, tris, lf.x + j);
    set_child_box(wn, lf.z, lf.w, acc.lo, acc.hi, margin_of(c));
    double cen[3];
    for (int k = 0; k < 3; ++k) cen[k] = 0.5 * (acc.lo[k] + acc.hi[k]);
    const DiskFrame fr = frame_of(cen, acc.n);
    double r = 0.0, th = 0.0;
    for (int j = 0; j < lf.y; ++j) {
        const double2 x = face_extent(fr, tris, lf.x + j);
        r = fmax(r, x.x);
        th = fmax(th, x.y);
    }
    store_disk(fr, r, th, lbnd + 2 * lf.x);
}

// Structural check of the collapsed tree: every sorted face lies in exactly one leaf.
__global__ __launch_bounds__(256) void cover_leaves(const int4* __restrict__ leaves, const int32_t* c,
                                                    int32_t* cover) {
    const int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= c[10]) return;
    const int4 lf = leaves[e];
    for (int j = 0; j < lf.y; ++j) atomicAdd(&cover[lf.x + j], 1);
}
__global__ __launch_bounds__(256) void check_cover(int64_t n, const int32_t* __restrict__ cover, int32_t* c) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && cover[s] != 1) atomicOr(&c[14], 1);
}


enum { kSum = 0, kMin = 1, kMax = 2 };

// Workgroup-wide reduction of nv values per thread with one operation: waves reduce
// with shuffles, then wave 0 combines the per-wave results (sh: 3 * 16 doubles).
__device__ __forceinline__ double op_of(double a, double b, int op) {
    return op == kSum ? a + b : op == kMin ? fmin(a, b) : fmax(a, b);
}
template <int BLOCK>
__device__ __forceinline__ void block_reduce(double* sh, double* v, int nv, int op) {
    constexpr int NW = BLOCK / 64;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    for (int k = 0; k < nv; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x = op_of(x, __shfl_xor(x, o), op);
        if (lane == 0) sh[k * NW + wave] = x;
    }
    __syncthreads();
    for (int k = 0; k < nv; ++k) {
        double x = sh[k * NW];
        for (int w = 1; w < NW; ++w) x = op_of(x, sh[k * NW + w], op);
        v[k] = x;
    }
    __syncthreads();
}


__device__ __forceinline__ double wave_reduce(double v, int op) {
    for (int o = 32; o > 0; o >>= 1) {
        const double x = __shfl_xor(v, o);
        v = op == kSum ? v + x : op == kMin ? fmin(v, x) : fmax(v, x);
    }
    return v;
}

// The box of an internal child's run of faces into its wide node and its disk at
// cbnd[8 * wide + 2 * slot], job = (wide node, slot, first, count): one wave per job of
// at most kWaveJobMax faces (shuffle reductions); larger jobs are the big_job_* passes'.
__global__ __launch_bounds__(256) void subtree_records(const int4* __restrict__ jobs, const int32_t* c,
                                                       const float4* __restrict__ tris, float4* __restrict__ cbnd,
                                                       WideNode* __restrict__ wn) {
    const double abs_pad = margin_of(c);
    const int lane = (int)(threadIdx.x & 63);
    for (int32_t e0 = blockIdx.x * 4; e0 < c[11]; e0 += gridDim.x * 4) {
        const int32_t e = e0 + (int32_t)(threadIdx.x >> 6);
        if (e >= c[11]) continue;
        const int4 job = jobs[e];
        if (job.w > kWaveJobMax) continue;  // uniform over the wave
        const int64_t first = job.z, end = (int64_t)job.z + job.w;
        DiskAcc acc;
        acc_init(acc);
        // four faces per lane in flight: the loads do not wait on each other
        for (int64_t s = first + lane; s < end; s += 4 * 64) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (s + u * 64 < end) acc_face(acc, tris, s + u * 64);
        }
        for (int k = 0; k < 3; ++k) {
            acc.lo[k] = wave_reduce(acc.lo[k], kMin);
            acc.hi[k] = wave_reduce(acc.hi[k], kMax);
            acc.n[k] = wave_reduce(acc.n[k], kSum);
        }
        double cen[3];
        for (int k = 0; k < 3; ++k) cen[k] = 0.5 * (acc.lo[k] + acc.hi[k]);
        const DiskFrame fr = frame_of(cen, acc.n);
        double ext[2] = {0.0, 0.0};
        for (int64_t s = first + lane; s < end; s += 4 * 64) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (s + u * 64 < end) {
                    const double2 x = face_extent(fr, tris, s + u * 64);
                    ext[0] = fmax(ext[0], x.x);
                    ext[1] = fmax(ext[1], x.y);
                }
        }
        ext[0] = wave_reduce(ext[0], kMax);
        ext[1] = wave_reduce(ext[1], kMax);
        if (lane == 0) {
            store_disk(fr, ext[0], ext[1], cbnd + 8 * (int64_t)job.x + 2 * job.y);
            set_child_box(wn, job.x, job.y, acc.lo, acc.hi, abs_pad);
        }
    }
}

// Children of more than kWaveJobMax faces: chunk y of big job j reduces its faces'
// box and normal sum (pass 1) and, with the job's frame, their extents (pass 2) into
// fixed slots; one wave per job folds the kChunks partials in a fixed order.
__device__ __forceinline__ void chunk_range(const int4& job, int y, int64_t& b, int64_t& e) {
    const int64_t per = ((int64_t)job.w + kChunks - 1) / kChunks;
    b = (int64_t)job.z + y * per;
    e = min((int64_t)job.z + job.w, b + per);
}

__global__ __launch_bounds__(256) void big_job_sums(const int4* __restrict__ jobs, const int32_t* __restrict__ big,
                                                    const int32_t* c, const float4* __restrict__ tris,
                                                    double* __restrict__ part) {
    __shared__ double sh[3 * 4];
    for (int32_t j = (int32_t)blockIdx.x; j < c[9]; j += (int32_t)gridDim.x) {
        const int4 job = jobs[big[j]];
        int64_t b, e;
        chunk_range(job, (int)blockIdx.y, b, e);
        DiskAcc acc;
        acc_init(acc);
        for (int64_t s = b + threadIdx.x; s < e; s += 4 * 256) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (s + u * 256 < e) acc_face(acc, tris, s + u * 256);
        }
        block_reduce<256>(sh, acc.lo, 3, kMin);
        block_reduce<256>(sh, acc.hi, 3, kMax);
        block_reduce<256>(sh, acc.n, 3, kSum);
        if (threadIdx.x == 0) {
            double* o = part + 9 * ((int64_t)j * kChunks + blockIdx.y);
            for (int k = 0; k < 3; ++k) {
                o[k] = acc.lo[k];
                o[3 + k] = acc.hi[k];
                o[6 + k] = acc.n[k];
            }
        }
    }
}

__global__ __launch_bounds__(256) void big_job_frames(const int4* __restrict__ jobs, const int32_t* __restrict__ big,
                                                      const int32_t* c, const double* __restrict__ part,
                                                      DiskFrame* __restrict__ frames, WideNode* __restrict__ wn) {
    const int lane = (int)(threadIdx.x & 63);
    const double abs_pad = margin_of(c);
    for (int32_t j = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6)); j < c[9]; j += (int32_t)(gridDim.x * 4)) {
        const int4 job = jobs[big[j]];
        double lo[3], hi[3], nn[3];
        const double* q = part + 9 * ((int64_t)j * kChunks + lane);
        for (int k = 0; k < 3; ++k) {
            lo[k] = lane < kChunks ? q[k] : INFINITY;
            hi[k] = lane < kChunks ? q[3 + k] : -INFINITY;
            nn[k] = lane < kChunks ? q[6 + k] : 0.0;
            lo[k] = wave_reduce(lo[k], kMin);
            hi[k] = wave_reduce(hi[k], kMax);
            nn[k] = wave_reduce(nn[k], kSum);
        }
        if (lane == 0) {
            double cen[3];
            for (int k = 0; k < 3; ++k) cen[k] = 0.5 * (lo[k] + hi[k]);
            frames[j] = frame_of(cen, nn);
            set_child_box(wn, job.x, job.y, lo, hi, abs_pad);
        }
    }
}

__global__ __launch_bounds__(256) void big_job_extents(const int4* __restrict__ jobs, const int32_t* __restrict__ big,
                                                       const int32_t* c, const float4* __restrict__ tris,
                                                       const DiskFrame* __restrict__ frames,
                                                       double* __restrict__ part) {
    __shared__ double sh[3 * 4];
    for (int32_t j = (int32_t)blockIdx.x; j < c[9]; j += (int32_t)gridDim.x) {
        const int4 job = jobs[big[j]];
        const DiskFrame fr = frames[j];
        int64_t b, e;
        chunk_range(job, (int)blockIdx.y, b, e);
        double ext[2] = {0.0, 0.0};
        for (int64_t s = b + threadIdx.x; s < e; s += 4 * 256) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (s + u * 256 < e) {
                    const double2 x = face_extent(fr, tris, s + u * 256);
                    ext[0] = fmax(ext[0], x.x);
                    ext[1] = fmax(ext[1], x.y);
                }
        }
        block_reduce<256>(sh, ext, 2, kMax);
        if (threadIdx.x == 0) {
            double* o = part + 2 * ((int64_t)j * kChunks + blockIdx.y);
            o[0] = ext[0];
            o[1] = ext[1];
        }
    }
}

__global__ __launch_bounds__(256) void big_job_disks(const int4* __restrict__ jobs, const int32_t* __restrict__ big,
                                                     const int32_t* c, const double* __restrict__ part,
                                                     const DiskFrame* __restrict__ frames, float4* __restrict__ cbnd) {
    const int lane = (int)(threadIdx.x & 63);
    for (int32_t j = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6)); j < c[9]; j += (int32_t)(gridDim.x * 4)) {
        const int4 job = jobs[big[j]];
        const double* q = part + 2 * ((int64_t)j * kChunks + lane);
        double r = lane < kChunks ? q[0] : 0.0, th = lane < kChunks ? q[1] : 0.0;
        r = wave_reduce(r, kMax);
        th = wave_reduce(th, kMax);
        if (lane == 0) store_disk(frames[j], r, th, cbnd + 8 * (int64_t)job.x + 2 * job.y);
    }
}

}  // namespace gb

namespace {

constexpr int kMaxDepth = 21;  // 3 * 21 + 1 = 64 stack entries

size_t sort_tmp_bytes(int64_t n) {
    size_t a = 0, b = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 63);
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
    return std::max(a, b);
}

using FlagIter = hipcub::TransformInputIterator<gb::U3, gb::LowerFlags, hipcub::CountingInputIterator<int64_t>>;


size_t flag_scan_bytes(int64_t n) {
    size_t b = 0;
    FlagIter it(hipcub::CountingInputIterator<int64_t>(0), gb::LowerFlags{});
    hipcub::DeviceScan::ExclusiveScan(nullptr, b, it, (gb::U3*)nullptr, gb::U3Sum(), gb::U3{0, 0, 0}, (int)n);
    return b;
}

// The workspace layout, measured (ws == nullptr) or carved.
size_t lay_out_workspace(void* ws, int64_t n, int topology, GpuBuildWorkspace& w) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const int64_t ni = n > 1 ? n - 1 : 1;
    char* base = (char*)ws;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* q = base ? base + off : nullptr;
        off += al(bytes);
        return q;
    };
    w.counters = (int32_t*)take((64 + 8 * gb::kGatherBlocks) * sizeof(int32_t));
    w.tri9 = (float*)take(9 * n * sizeof(float));
    w.keys0 = (uint64_t*)take(n * sizeof(uint64_t));
    w.keys1 = (uint64_t*)take(n * sizeof(uint64_t));
    w.vals0 = (uint32_t*)take(n * sizeof(uint32_t));
    w.face_of = (uint32_t*)take(n * sizeof(uint32_t));
    w.sort_tmp_bytes = sort_tmp_bytes(n);
    w.sort_tmp = take(w.sort_tmp_bytes);
    const bool sah = topology == kTopologySah;
    w.bn = (BinNode*)take((sah ? 2 * n : ni) * sizeof(BinNode));
    w.leaf_parent = (int32_t*)take(n * sizeof(int32_t));
    w.wide = (WideNode*)take(ni * sizeof(WideNode));
    w.queue0 = (int4*)take(ni * sizeof(int4));
    w.queue1 = (int4*)take(ni * sizeof(int4));
    w.leaves = (int4*)take(n * sizeof(int4));
    w.jobs = (int4*)take(ni * sizeof(int4));
    w.cen = (float*)take(3 * n * sizeof(float));
    w.lists = (uint32_t*)take(6 * n * sizeof(uint32_t));
    w.side = (uint8_t*)take(n);
    w.flag_scan = take(n * sizeof(gb::U3));
    w.flag_tmp_bytes = flag_scan_bytes(n);
    w.flag_tmp = take(w.flag_tmp_bytes);
    if (sah) {
        // runs longer than kSahWaveRun are disjoint: at most n / (kSahWaveRun + 1) + 1
        const int64_t nbig = n / (gb::kSahWaveRun + 1) + 1;
        w.fbox = (float*)take(6 * n * sizeof(float));
        w.rid = (int32_t*)take(n * sizeof(int32_t));
        w.split = (uint32_t*)take(2 * n * sizeof(uint32_t));
        w.act0 = (int32_t*)take(n * sizeof(int32_t));
        w.act1 = (int32_t*)take(n * sizeof(int32_t));
        w.level_count = (int32_t*)take(4 * (gb::kSahMaxLevels + 1) * sizeof(int32_t));
        w.bigslot = (int32_t*)take(2 * n * sizeof(int32_t));
        w.bigrun = (int32_t*)take(nbig * sizeof(int32_t));
        w.bins = (int32_t*)take(nbig * 3 * gb::kSahBins * gb::kSahBinWords * sizeof(int32_t));
        w.grow = (float*)take(sizeof(float));
    }
    // children above kWaveJobMax faces: disjoint at each depth, so at most
    // kMaxDepth * (n / (kWaveJobMax + 1) + 1) of them
    const int64_t nbig = kMaxDepth * (n / (gb::kWaveJobMax + 1) + 1);
    w.bigjobs = (int32_t*)take(nbig * sizeof(int32_t));
    w.big_part = (double*)take(nbig * gb::kChunks * 9 * sizeof(double));
    w.big_frames = take(nbig * sizeof(gb::DiskFrame));
    w.n = n;
    return off;
}

}  // namespace

size_t gpu_build_workspace_bytes(int64_t n, int topology) {
    GpuBuildWorkspace w;
    return lay_out_workspace(nullptr, n, topology, w);
}

GpuBuildWorkspace carve_gpu_build_workspace(void* ws, int64_t n, int topology) {
    GpuBuildWorkspace w;
    lay_out_workspace(ws, n, topology, w);
    return w;
}

namespace {
unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }
}  // namespace

// The centroids and the three centroid-sorted face lists (by (centroid, face index))
// in cur[0..2].
hipError_t sorted_lists(int64_t n, const GpuBuildWorkspace& w, uint32_t* const* cur, hipStream_t s) {
    hipLaunchKernelGGL(gb::centroids, dim3(blocks_for(n)), dim3(256), 0, s, n, w.tri9, w.cen);
    uint32_t* keys = reinterpret_cast<uint32_t*>(w.keys0);
    uint32_t* keys_sorted = reinterpret_cast<uint32_t*>(w.keys1);
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(gb::axis_keys, dim3(blocks_for(n)), dim3(256), 0, s, n, w.cen + a * n, keys, w.vals0);
        size_t tmp = w.sort_tmp_bytes;
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tmp, keys, keys_sorted, w.vals0, cur[a], (int)n,
                                                          0, 32, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Median topology, steps 2-3: the three centroid-sorted face lists, then one stable
// partition of every run per level until all runs are leaves; the first list is the
// leaf order.
hipError_t median_order(int64_t n, int32_t leaf_size, const GpuBuildWorkspace& w, hipStream_t s) {
    uint32_t* cur[3] = {w.lists, w.lists + n, w.lists + 2 * n};
    uint32_t* nxt[3] = {w.lists + 3 * n, w.lists + 4 * n, w.lists + 5 * n};
    hipError_t e0 = sorted_lists(n, w, cur, s);
    if (e0 != hipSuccess) return e0;
    // global levels until every run fits one workgroup, then the rest in LDS
    int levels = 0;
    for (int64_t run = n; run > gb::kFinishRun; run -= run / 2) ++levels;
    gb::U3* S = reinterpret_cast<gb::U3*>(w.flag_scan);
    for (int level = 0; level < levels; ++level) {
        const gb::Lists L{{cur[0], cur[1], cur[2]}};
        hipLaunchKernelGGL(gb::median_mark, dim3(blocks_for(n)), dim3(256), 0, s, n, level, leaf_size, L, w.cen,
                           w.side);
        FlagIter flags(hipcub::CountingInputIterator<int64_t>(0), gb::LowerFlags{L, w.side});
        size_t tmp = w.flag_tmp_bytes;
        hipError_t e = hipcub::DeviceScan::ExclusiveScan(w.flag_tmp, tmp, flags, S, gb::U3Sum(), gb::U3{0, 0, 0},
                                                         (int)n, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(gb::median_scatter, dim3(blocks_for(n)), dim3(256), 0, s, n, level, leaf_size, L, nxt[0],
                           nxt[1], nxt[2], S, w.side);
        for (int a = 0; a < 3; ++a) std::swap(cur[a], nxt[a]);
    }
    const gb::Lists L{{cur[0], cur[1], cur[2]}};
    hipLaunchKernelGGL(gb::median_finish, dim3(1u << levels), dim3(gb::kFinishThreads), 0, s, n, levels, leaf_size, L,
                       w.cen, w.vals0, w.face_of);
    return hipGetLastError();
}

// SAH topology, steps 2-3: the sorted lists, then one level of binary nodes per round
// until no run is longer than leaf_size: runs longer than kSahWaveRun by bins
// (sah_bin_clear, sah_bin per list, sah_bin_choose), shorter ones by sah_wave_runs
// (one launch per size class: one, four or 16 slots per lane); then one stable
// partition of the three lists.  The binary nodes in w.bn, the first list is the leaf
// order.  The host reads each level's counts (one synchronization per level) to size
// its launches.  finished = false if runs are still active after kSahMaxLevels levels
// (cannot happen: splits fall back to the median below kSahLevels).
hipError_t sah_order(int64_t n, int32_t leaf_size, const GpuBuildWorkspace& w, bool& finished, hipStream_t s) {
    finished = false;
    uint32_t* cur[3] = {w.lists, w.lists + n, w.lists + 2 * n};
    uint32_t* nxt[3] = {w.lists + 3 * n, w.lists + 4 * n, w.lists + 5 * n};
    hipError_t e = sorted_lists(n, w, cur, s);
    if (e != hipSuccess) return e;
    gb::Box6* fbox = reinterpret_cast<gb::Box6*>(w.fbox);
    hipLaunchKernelGGL(gb::sah_boxes, dim3(blocks_for(n)), dim3(256), 0, s, n, w.tri9, fbox, w.rid);
    hipLaunchKernelGGL(gb::fill_i32, dim3(2), dim3(256), 0, s, w.level_count, 4 * ((int64_t)gb::kSahMaxLevels + 1), 0);
    hipLaunchKernelGGL(gb::sah_root, dim3(1), dim3(64), 0, s, n, w.bn, w.act0, w.level_count, w.bigslot, w.bigrun,
                       w.grow, w.counters);
    int32_t* act_in = w.act0;
    int32_t* act_out = w.act1;
    gb::U3* S = reinterpret_cast<gb::U3*>(w.flag_scan);
    int32_t stats[4];  // the current level's counts (the copy completes before the synchronization)
    int32_t node_base = 1;  // ids of the children made at this level
    for (int level = 0;; ++level) {
        e = hipMemcpyAsync(stats, w.level_count + 4 * level, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        const int32_t active = stats[0], longest = stats[1], nbig = stats[2];
        if (active == 0) break;
        if (level >= gb::kSahMaxLevels) return hipGetLastError();
        const gb::Lists L{{cur[0], cur[1], cur[2]}};
        if (nbig > 0) {
            hipLaunchKernelGGL(gb::sah_bin_clear, dim3(std::min<unsigned>(blocks_for((int64_t)nbig * 3 * gb::kSahBins), 4096)),
                               dim3(256), 0, s, w.level_count, level, w.bins);
            for (int a = 0; a < 3; ++a)
                hipLaunchKernelGGL(gb::sah_bin, dim3(blocks_for(n)), dim3(256), 0, s, n, a, w.rid, w.bn, w.bigslot,
                                   cur[a], w.cen, fbox, w.bins);
            hipLaunchKernelGGL(gb::sah_bin_choose, dim3(blocks_for(nbig, 4)), dim3(256), 0, s, level, node_base,
                               w.level_count, w.bins, act_in, act_out, w.bn, w.bigslot, w.bigrun, w.split, w.grow, L,
                               w.cen, n, leaf_size);
        }
        // short runs by size class: groups of 16, 32 and 64 lanes with one slot each, then
        // 64 lanes with 4 and 16 slots each
#define CUBVH_SAH_RUNS(K, W, LO, B)                                                                               \
    hipLaunchKernelGGL((gb::sah_wave_runs<K, W, LO, B>), dim3(std::min<unsigned>(blocks_for(active, B / W), 16384)),  \
                       dim3(B), 0, s, level, node_base, w.level_count, act_in, act_out, w.bn, w.bigslot, w.bigrun,   \
                       w.split, fbox, w.grow, L, w.cen, n, leaf_size)
        CUBVH_SAH_RUNS(1, 16, 0, 1024);
        if (longest > 16) CUBVH_SAH_RUNS(1, 32, 16, 1024);
        if (longest > 32) CUBVH_SAH_RUNS(1, 64, 32, 1024);
        if (longest > 64) CUBVH_SAH_RUNS(4, 64, 64, 1024);
        if (longest > 256) CUBVH_SAH_RUNS(16, 64, 256, 256);
#undef CUBVH_SAH_RUNS
        hipLaunchKernelGGL(gb::sah_mark, dim3(blocks_for(n)), dim3(256), 0, s, n, w.rid, w.split, L, w.side);
        FlagIter flags(hipcub::CountingInputIterator<int64_t>(0), gb::LowerFlags{L, w.side});
        size_t tmp = w.flag_tmp_bytes;
        e = hipcub::DeviceScan::ExclusiveScan(w.flag_tmp, tmp, flags, S, gb::U3Sum(), gb::U3{0, 0, 0}, (int)n, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(gb::sah_scatter, dim3(blocks_for(n)), dim3(256), 0, s, n, w.rid, w.bn, w.split, L,
                           nxt[0], nxt[1], nxt[2], S, w.side, leaf_size, w.counters);
        for (int a = 0; a < 3; ++a) std::swap(cur[a], nxt[a]);
        std::swap(act_in, act_out);
        node_base += 2 * active;
    }
    finished = true;
    e = hipMemcpyAsync(w.face_of, cur[0], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t gpu_build_tree(const float* V, int64_t nv, const int32_t* F, int64_t n, int32_t leaf_size, int topology,
                          const GpuBuildWorkspace& w, GpuBuildResult& res, hipStream_t s) {
    const bool median = topology == kTopologyMedian;
    const bool sah = topology == kTopologySah;
    bool sah_finished = true;
    hipError_t e = hipSuccess;
    hipLaunchKernelGGL(gb::init_counters, dim3(1), dim3(64), 0, s, w.counters);
    const unsigned gblocks = std::min<unsigned>(blocks_for(n), gb::kGatherBlocks);
    hipLaunchKernelGGL(gb::gather_faces, dim3(gblocks), dim3(256), 0, s, n, V, nv, F, w.tri9, w.counters + 64);
    hipLaunchKernelGGL(gb::fold_partials, dim3(1), dim3(256), 0, s, w.counters + 64, (int32_t)gblocks, w.counters);
    if (median) {
        e = median_order(n, leaf_size, w, s);
        if (e != hipSuccess) return e;
    } else if (sah) {
        e = sah_order(n, leaf_size, w, sah_finished, s);
        if (e != hipSuccess) return e;
        if (!sah_finished) {  // no collapse of an unfinished tree
            int32_t c[16];
            e = hipMemcpyAsync(c, w.counters, sizeof c, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            res.input_errors = c[7];
            res.too_deep = true;
            return hipGetLastError();
        }
    } else {
        hipLaunchKernelGGL(gb::morton_keys, dim3(blocks_for(n)), dim3(256), 0, s, n, w.tri9, w.counters, w.keys0,
                           w.vals0);
        size_t tmp = w.sort_tmp_bytes;
        e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tmp, w.keys0, w.keys1, w.vals0, w.face_of, (int)n, 0, 63,
                                               s);
        if (e != hipSuccess) return e;
        const int64_t ni = n - 1;
        hipLaunchKernelGGL(gb::radix_tree, dim3(blocks_for(ni)), dim3(256), 0, s, n, w.keys1, w.bn, w.leaf_parent);
    }
    // collapse: one launch per level for the deepest tree a walk accepts (depth
    // kMaxDepth, 3 * depth + 1 <= 64 stack entries), without host synchronization: level
    // k reads its entry count from counters[16 + k - 1] (the root alone at level 0) and
    // counts its outputs in counters[16 + k]; a tree still growing after the last
    // level is reported as too deep.  The root entry is (binary root 0, wide node 0,
    // depth 1) for the radix and SAH trees, (slots 0 .. n, depth 1, wide node 0) for the median
    // topology.
    const int4 root = median ? make_int4(0, (int32_t)n, 1, 0) : make_int4(0, 0, 1, 0);
    const int32_t one = 1;
    e = hipMemcpyAsync(w.queue0, &root, sizeof root, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(w.counters + 8, &one, sizeof one, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const unsigned cblocks = std::min<unsigned>(blocks_for(n), 1024);
    // the median tree's depth follows from n: two halvings per wide level
    int collapse_levels = kMaxDepth;
    if (median) {
        int halvings = 0;
        for (int64_t run = n; run > leaf_size; run -= run / 2) ++halvings;
        collapse_levels = std::min(kMaxDepth, (halvings + 1) / 2 + 1);
    }
    for (int level = 0; level < collapse_levels; ++level) {
        const int4* qin = (level & 1) ? w.queue1 : w.queue0;
        int4* qout = (level & 1) ? w.queue0 : w.queue1;
        const int32_t* n_in = level == 0 ? (const int32_t*)nullptr : w.counters + 16 + level - 1;
        if (median)
            hipLaunchKernelGGL(gb::collapse_run_level, dim3(cblocks), dim3(256), 0, s, qin, n_in, qout,
                               w.counters + 16 + level, leaf_size, w.wide, w.leaves, w.jobs, w.bigjobs, w.counters);
        else
            hipLaunchKernelGGL(gb::collapse_level, dim3(cblocks), dim3(256), 0, s, qin, n_in, qout,
                               w.counters + 16 + level, w.bn, leaf_size, w.wide, w.leaves, w.jobs, w.bigjobs,
                               w.counters);
    }
    // every sorted face in exactly one leaf (leaf_parent is free again: reused as counts)
    hipLaunchKernelGGL(gb::fill_i32, dim3(std::min<unsigned>(blocks_for(n), 4096)), dim3(256), 0, s, w.leaf_parent, n,
                       0);
    hipLaunchKernelGGL(gb::cover_leaves, dim3(blocks_for(n)), dim3(256), 0, s, w.leaves, w.counters, w.leaf_parent);
    hipLaunchKernelGGL(gb::check_cover, dim3(blocks_for(n)), dim3(256), 0, s, n, w.leaf_parent, w.counters);
    int32_t c[64];
    e = hipMemcpyAsync(c, w.counters, sizeof c, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    res.input_errors = c[7];
    res.too_deep = c[16 + collapse_levels - 1] != 0 || !sah_finished;
    res.structure_ok = c[14] == 0;
    res.n_nodes = c[8];
    res.n_leaves = c[10];
    res.n_jobs = c[11];
    res.depth = c[12];
    return hipGetLastError();
}

hipError_t gpu_build_face0_slot(const GpuBuildWorkspace& w, int64_t& slot, hipStream_t s) {
    int32_t v = -1;
    hipError_t e = hipMemcpyAsync(&v, w.counters + 13, sizeof v, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    slot = v;
    return e;
}

hipError_t gpu_build_arrays(int64_t n, const GpuBuildWorkspace& w, const TreeArrays& out, hipStream_t s) {
    float4* tris = const_cast<float4*>(out.tris);
    hipLaunchKernelGGL(gb::emit_tris, dim3(blocks_for(n)), dim3(256), 0, s, n, w.face_of, w.tri9, tris, w.counters);
    hipLaunchKernelGGL(gb::face_records, dim3(blocks_for(n)), dim3(256), 0, s, n, out.tris,
                       const_cast<float4*>(out.tbnd), const_cast<float4*>(out.tpre));
    hipLaunchKernelGGL(gb::leaf_records, dim3(blocks_for(n)), dim3(256), 0, s, w.leaves, w.counters, out.tris,
                       const_cast<float4*>(out.lbnd), w.wide);
    hipLaunchKernelGGL(gb::subtree_records, dim3(4096), dim3(256), 0, s, w.jobs, w.counters, out.tris,
                       const_cast<float4*>(out.cbnd), w.wide);
    const dim3 chunks(256, gb::kChunks);
    hipLaunchKernelGGL(gb::big_job_sums, chunks, dim3(256), 0, s, w.jobs, w.bigjobs, w.counters, out.tris, w.big_part);
    hipLaunchKernelGGL(gb::big_job_frames, dim3(64), dim3(256), 0, s, w.jobs, w.bigjobs, w.counters, w.big_part,
                       reinterpret_cast<gb::DiskFrame*>(w.big_frames), w.wide);
    hipLaunchKernelGGL(gb::big_job_extents, chunks, dim3(256), 0, s, w.jobs, w.bigjobs, w.counters, out.tris,
                       reinterpret_cast<const gb::DiskFrame*>(w.big_frames), w.big_part);
    hipLaunchKernelGGL(gb::big_job_disks, dim3(64), dim3(256), 0, s, w.jobs, w.bigjobs, w.counters, w.big_part,
                       reinterpret_cast<const gb::DiskFrame*>(w.big_frames), const_cast<float4*>(out.cbnd));
    return hipGetLastError();
}

}  // namespace cubvh_amd
