// smc.hip -- sparse marching cubes.
//
// Output contract of the reference's GPU extraction (include/gpu/spcumc.cuh:631-868):
//   * one vertex per lattice edge that carries a sign change in some cell, numbered in
//     ascending order of (lowest lattice point x, y, z, axis); its position is linear
//     interpolation of the iso level along the edge, from the values of the first cell
//     (input order) that has the edge;
//   * per cell in input order, the case's triangles in table order with the second and
//     third vertex exchanged.
// With ensure_consistency the corner values are first replaced by the mean over every
// cell that shares the lattice point, summed in double in cell order (the reference's
// CPU path, include/cpu/spmc.h:408-433; its GPU path sums in fp32 in an order that its
// library reduction leaves open).
//
// Pipeline: bounds (per-block partials, one fold) -> [corner keys, radix sort, run
// means] -> classify (case, counts) -> exclusive scan -> one synchronization (bounds and
// totals) -> edge keys in cell order -> stable radix sort -> inclusive scan of run
// starts (vertex ids) -> one synchronization (vertex count) -> [caller's buffers]
// vertices from the first cell of each run and a raw -> vertex map -> triangles.  Edge
// and corner keys pack the lattice point relative to the minimum with B bits per axis
// (B from the measured span), so the sorts run over 3B + 2 and 3B bits, with 32-bit
// edge keys whenever 3B + 2 <= 32 (spans up to 1022).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

#include "mc_table.hpp"
#include "smc.hpp"

namespace cubvh_amd {

namespace {

__constant__ uint64_t c_triangles[256] = {CUBVH_MC_TRIANGLES};

constexpr int kBlock = 256;

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

__device__ __forceinline__ uint32_t edge_mask_of(uint32_t c) {
    const uint32_t lo = c & 15u, hi = c >> 4;
    const uint32_t rlo = ((lo >> 1) | (lo << 3)) & 15u, rhi = ((hi >> 1) | (hi << 3)) & 15u;
    return (lo ^ rlo) | ((hi ^ rhi) << 4) | ((lo ^ hi) << 8);
}

// info: [0..2] min x y z, [3..5] max x y z (int32).

// Bounds in two steps without same-address atomics: each block writes its six extremes
// to part[6 * block], one block folds them into info.
__device__ __forceinline__ void block_extremes(int32_t (&lo)[3], int32_t (&hi)[3]) {
    __shared__ int32_t sh[6][kBlock / 64];
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], __shfl_xor(lo[k], o));
            hi[k] = max(hi[k], __shfl_xor(hi[k], o));
        }
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            sh[k][w] = lo[k];
            sh[3 + k][w] = hi[k];
        }
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        lo[k] = sh[k][0];
        hi[k] = sh[3 + k][0];
        for (int j = 1; j < kBlock / 64; ++j) {
            lo[k] = min(lo[k], sh[k][j]);
            hi[k] = max(hi[k], sh[3 + k][j]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void smc_bounds(const int32_t* __restrict__ coords, int64_t n,
                                                     int32_t* __restrict__ part) {
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        for (int k = 0; k < 3; ++k) {
            const int32_t v = coords[3 * i + k];
            lo[k] = min(lo[k], v);
            hi[k] = max(hi[k], v);
        }
    block_extremes(lo, hi);
    if (threadIdx.x < 3) {
        part[6 * blockIdx.x + threadIdx.x] = lo[threadIdx.x];
        part[6 * blockIdx.x + 3 + threadIdx.x] = hi[threadIdx.x];
    }
}

__global__ __launch_bounds__(kBlock) void smc_fold_bounds(const int32_t* __restrict__ part, int n_part,
                                                          int32_t* __restrict__ info) {
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int b = (int)threadIdx.x; b < n_part; b += kBlock)
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], part[6 * b + k]);
            hi[k] = max(hi[k], part[6 * b + 3 + k]);
        }
    block_extremes(lo, hi);
    if (threadIdx.x < 3) {
        info[threadIdx.x] = lo[threadIdx.x];
        info[3 + threadIdx.x] = hi[threadIdx.x];
    }
}

// Lattice point packed relative to the minimum, B bits per axis.
__device__ __forceinline__ uint64_t pack_point(int64_t x, int64_t y, int64_t z, const int32_t* mn, int B) {
    return ((uint64_t)(x - mn[0]) << (2 * B)) | ((uint64_t)(y - mn[1]) << B) | (uint64_t)(z - mn[2]);
}

// ---- corner consistency
template <typename K>
__global__ __launch_bounds__(kBlock) void smc_corner_keys(const int32_t* __restrict__ coords, int64_t n,
                                                          const int32_t* __restrict__ info, int B,
                                                          K* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= 8 * n) return;
    const int64_t i = j >> 3;
    const int k = (int)(j & 7);
    keys[j] = (K)pack_point((int64_t)coords[3 * i] + mc::corner_dx(k), (int64_t)coords[3 * i + 1] + mc::corner_dy(k),
                            (int64_t)coords[3 * i + 2] + mc::corner_dz(k), info, B);
    vals[j] = (uint32_t)j;
}

// Run of equal keys starting at s: mean of the corner values in cell order (the sort is
// stable and the values were laid out in cell order), written to every member.
template <typename K>
__global__ __launch_bounds__(kBlock) void smc_corner_mean(const K* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals, int64_t m,
                                                          const float* __restrict__ corners,
                                                          float* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= m || (s > 0 && keys[s - 1] == keys[s])) return;
    const K key = keys[s];
    double sum = 0.0;
    int64_t e = s;
    for (; e < m && keys[e] == key; ++e) sum += (double)corners[vals[e]];
    const float mean = (float)(sum / (double)(e - s));
    for (int64_t j = s; j < e; ++j) out[vals[j]] = mean;
}

// ---- classification
__global__ __launch_bounds__(kBlock) void smc_classify(const float* __restrict__ corners, int64_t n, float iso,
                                                       uint8_t* __restrict__ cases, uint64_t* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c |= (corners[8 * i + k] < iso ? 1u : 0u) << k;
    cases[i] = (uint8_t)c;
    const uint32_t nv = (uint32_t)__popc(edge_mask_of(c));
    const uint32_t nt = (uint32_t)mc::triangle_count(c_triangles[c]);
    counts[i] = ((uint64_t)nv << 32) | nt;
}

// offsets = exclusive scan of counts; totals = last offset + last count.
__global__ void smc_totals(const uint64_t* __restrict__ counts, const uint64_t* __restrict__ offsets, int64_t n,
                           uint64_t* __restrict__ totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) totals[0] = offsets[n - 1] + counts[n - 1];
}

// ---- edge vertices
// Edge keys (lowest lattice point relative to the minimum, then the axis) and
// (cell, edge) values, laid out in cell order at each cell's first raw slot.
template <typename K>
__global__ __launch_bounds__(kBlock) void smc_edge_keys(const int32_t* __restrict__ coords,
                                                        const uint8_t* __restrict__ cases,
                                                        const uint64_t* __restrict__ offsets, int64_t n,
                                                        const int32_t* __restrict__ info, int B,
                                                        K* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t m = edge_mask_of(cases[i]);
    if (!m) return;
    uint64_t r = offsets[i] >> 32;
    const int64_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
    while (m) {
        const int e = __ffs(m) - 1;
        m &= m - 1;
        const int a = mc::edge_a(e), b = mc::edge_b(e);
        const uint64_t p = pack_point(x + min(mc::corner_dx(a), mc::corner_dx(b)),
                                      y + min(mc::corner_dy(a), mc::corner_dy(b)),
                                      z + min(mc::corner_dz(a), mc::corner_dz(b)), info, B);
        keys[r] = (K)((p << 2) | (uint64_t)mc::edge_axis(e));
        vals[r] = (uint32_t)(i * 16 + e);
        ++r;
    }
}

// 1 where sorted position s starts a run of equal keys (read by the id scan).
template <typename K>
struct RunStart {
    const K* keys;
    __host__ __device__ uint32_t operator()(int64_t s) const { return (s == 0 || keys[s - 1] != keys[s]) ? 1u : 0u; }
};

// Raw vertex slot of edge e in cell i: the cell's first slot + the crossing edges below e.
__device__ __forceinline__ uint32_t raw_slot(uint64_t offset, uint32_t mask, int e) {
    return (uint32_t)(offset >> 32) + (uint32_t)__popc(mask & ((1u << e) - 1u));
}

template <typename K>
__global__ __launch_bounds__(kBlock) void smc_vertices(const K* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ ids, int64_t m,
                                                       const int32_t* __restrict__ coords,
                                                       const float* __restrict__ corners,
                                                       const uint8_t* __restrict__ cases,
                                                       const uint64_t* __restrict__ offsets, float iso,
                                                       float* __restrict__ verts, int32_t* __restrict__ map) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= m) return;
    const uint32_t id = ids[s] - 1u;
    const uint32_t v = vals[s];
    const int64_t i = v >> 4;
    const int e = (int)(v & 15u);
    map[raw_slot(offsets[i], edge_mask_of(cases[i]), e)] = (int32_t)id;
    if (s > 0 && keys[s - 1] == keys[s]) return;
    // include/gpu/spcumc.cuh:455-471: t = (iso - va) / (vb - va) clamped to [0, 1] (0.5
    // on a flat edge), p = pA + t (pB - pA); t (pB - pA) is exact for lattice points
    const int a = mc::edge_a(e), b = mc::edge_b(e);
    const float va = corners[8 * i + a], vb = corners[8 * i + b];
    const float denom = vb - va;
    float t;
    if (fabsf(denom) < 1e-30f) {
        t = 0.5f;
    } else {
        t = (iso - va) / denom;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
    }
    const int64_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
    const float ax = (float)(x + mc::corner_dx(a)), ay = (float)(y + mc::corner_dy(a)),
                az = (float)(z + mc::corner_dz(a));
    const float bx = (float)(x + mc::corner_dx(b)), by = (float)(y + mc::corner_dy(b)),
                bz = (float)(z + mc::corner_dz(b));
    verts[3 * (int64_t)id] = ax + t * (bx - ax);
    verts[3 * (int64_t)id + 1] = ay + t * (by - ay);
    verts[3 * (int64_t)id + 2] = az + t * (bz - az);
}

__global__ __launch_bounds__(kBlock) void smc_triangles(const uint8_t* __restrict__ cases,
                                                        const uint64_t* __restrict__ offsets, int64_t n,
                                                        const int32_t* __restrict__ map,
                                                        int32_t* __restrict__ tris) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cases[i];
    const uint64_t packed = c_triangles[c];
    const int nt = mc::triangle_count(packed);
    if (!nt) return;
    const uint32_t mask = edge_mask_of(c);
    const uint64_t off = offsets[i];
    int32_t* o = tris + 3 * (int64_t)(uint32_t)off;
    for (int t = 0; t < nt; ++t) {
        const int e0 = mc::triangle_edge(packed, 3 * t), e1 = mc::triangle_edge(packed, 3 * t + 1),
                  e2 = mc::triangle_edge(packed, 3 * t + 2);
        o[3 * t] = map[raw_slot(off, mask, e0)];
        o[3 * t + 1] = map[raw_slot(off, mask, e2)];
        o[3 * t + 2] = map[raw_slot(off, mask, e1)];
    }
}

int bits_for(int64_t span) {  // bits to hold 0..span
    int b = 1;
    while (((int64_t)1 << b) <= span) ++b;
    return b;
}

}  // namespace

// Scratch of the extractions on one device, kept between calls and grown when a larger
// one arrives, so that steady-state calls allocate nothing: region A holds the per-cell
// arrays, region B the per-edge arrays (sized once the edge count is known).  One
// extraction at a time owns it (from smc_extract to smc_release); a concurrent one
// allocates its own blocks from the stream-ordered pool.  The next owner's stream waits
// for the event the previous owner recorded at release.
struct Region {
    char* base = nullptr;
    size_t cap = 0;
};

struct DeviceScratch {
    std::mutex mu;
    bool busy = false;
    Region a, b;
    hipEvent_t done = nullptr;
    uint64_t* readback = nullptr;  // pinned host words for the small readbacks
};

namespace {

constexpr int kScratchDevices = 64;
DeviceScratch g_scratch[kScratchDevices];

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

// Bump allocation over one block; with base == nullptr it only measures.
struct Carve {
    char* base = nullptr;
    size_t used = 0;
    template <typename T>
    T* take(int64_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += align_up((size_t)std::max<int64_t>(count, 1) * sizeof(T));
        return p;
    }
};

hipError_t reserve(Region& r, size_t need) {
    if (need <= r.cap) return hipSuccess;
    if (r.base) {
        hipError_t e = hipFree(r.base);  // synchronizes the device: nothing uses the old block
        r.base = nullptr;
        r.cap = 0;
        if (e != hipSuccess) return e;
    }
    need += need / 4;
    hipError_t e = hipMalloc((void**)&r.base, need);
    if (e != hipSuccess) {
        r.base = nullptr;
        return e;
    }
    r.cap = need;
    return hipSuccess;
}

}  // namespace

// State of one extraction between its two phases.
struct SmcState {
    int64_t n = 0, m_raw = 0, m = 0, t = 0;
    float iso = 0.0f;
    bool wide = false;  // 64-bit edge keys (more than 32 key bits)
    const int32_t* coords = nullptr;
    const float* corners = nullptr;  // the caller's, or the averaged copy
    int device = -1;
    DeviceScratch* shared = nullptr;  // the device's scratch, or null for private blocks
    std::vector<void*> blocks;        // private blocks (stream-ordered)
    int32_t* info = nullptr;          // [0..5] bounds, [8..9] totals (uint64)
    uint8_t* cases = nullptr;
    uint64_t* offsets = nullptr;
    void* keys = nullptr;      // sorted edge keys (uint32 or uint64)
    uint32_t* vals = nullptr;  // sorted (cell, edge)
    uint32_t* ids = nullptr;   // vertex id + 1 per sorted position
    int32_t* map = nullptr;    // raw slot -> vertex id
};

namespace {

SmcStatus hip_status(hipError_t e, const char* what) {
    SmcStatus st;
    if (e != hipSuccess) {
        st.code = -2;
        st.message = std::string(what) + ": " + hipGetErrorString(e);
    }
    return st;
}

// A block of `bytes` for region `which` (0 = A, 1 = B): the shared scratch, or a private
// stream-ordered allocation.
hipError_t region(SmcState* st, int which, size_t bytes, hipStream_t s, char** out) {
    if (st->shared) {
        Region& r = which == 0 ? st->shared->a : st->shared->b;
        hipError_t e = reserve(r, bytes);
        *out = r.base;
        return e;
    }
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, std::max<size_t>(bytes, 1), s);
    if (e != hipSuccess) return e;
    st->blocks.push_back(q);
    *out = (char*)q;
    return hipSuccess;
}

template <typename K>
size_t sort_bytes(int64_t m) {
    size_t b = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const K*)nullptr, (K*)nullptr, (const uint32_t*)nullptr,
                                       (uint32_t*)nullptr, (int)m, 0, (int)(8 * sizeof(K)));
    return b;
}

template <typename K>
using StartIter = hipcub::TransformInputIterator<uint32_t, RunStart<K>, hipcub::CountingInputIterator<int64_t>>;

template <typename K>
size_t id_scan_bytes(int64_t m) {
    size_t b = 0;
    StartIter<K> it(hipcub::CountingInputIterator<int64_t>(0), RunStart<K>{nullptr});
    hipcub::DeviceScan::InclusiveSum(nullptr, b, it, (uint32_t*)nullptr, (int)m);
    return b;
}

// Corner means: keys of the 8 n corners, stable sort, run means into `mean`.
template <typename K>
hipError_t average_corners(SmcState* st, Carve& c, int B, float* mean, hipStream_t s) {
    const int64_t m = 8 * st->n;
    K* ck = c.take<K>(m);
    uint32_t* cv = c.take<uint32_t>(m);
    K* ck2 = c.take<K>(m);
    uint32_t* cv2 = c.take<uint32_t>(m);
    size_t bytes = sort_bytes<K>(m);
    char* tmp = c.take<char>((int64_t)bytes);
    hipLaunchKernelGGL(smc_corner_keys<K>, dim3(blocks_for(m)), dim3(kBlock), 0, s, st->coords, st->n, st->info, B,
                       ck, cv);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, ck, ck2, cv, cv2, (int)m, 0, 3 * B, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(smc_corner_mean<K>, dim3(blocks_for(m)), dim3(kBlock), 0, s, ck2, cv2, m, st->corners, mean);
    return hipGetLastError();
}

// Region B: edge keys in cell order, their stable sort, the vertex ids (inclusive scan
// of run starts) and the raw -> vertex map used by the triangles.
template <typename K>
hipError_t order_edges(SmcState* st, int B, hipStream_t s) {
    const int64_t m = st->m_raw;
    const size_t sbytes = sort_bytes<K>(m), ibytes = id_scan_bytes<K>(m);
    auto lay = [&](Carve& c, K** k1, K** k2, uint32_t** v1, uint32_t** v2, char** t1, char** t2) {
        *k1 = c.take<K>(m);
        *k2 = c.take<K>(m);
        *v1 = c.take<uint32_t>(m);
        *v2 = c.take<uint32_t>(m);
        st->ids = c.take<uint32_t>(m);
        st->map = c.take<int32_t>(m);
        *t1 = c.take<char>((int64_t)sbytes);
        *t2 = c.take<char>((int64_t)ibytes);
    };
    K *k1, *k2;
    uint32_t *v1, *v2;
    char *t1, *t2;
    Carve size;
    lay(size, &k1, &k2, &v1, &v2, &t1, &t2);
    Carve c;
    hipError_t e = region(st, 1, size.used, s, &c.base);
    if (e != hipSuccess) return e;
    lay(c, &k1, &k2, &v1, &v2, &t1, &t2);
    hipLaunchKernelGGL(smc_edge_keys<K>, dim3(blocks_for(st->n)), dim3(kBlock), 0, s, st->coords, st->cases,
                       st->offsets, st->n, st->info, B, k1, v1);
    size_t bytes = sbytes;
    e = hipcub::DeviceRadixSort::SortPairs(t1, bytes, k1, k2, v1, v2, (int)m, 0, 3 * B + 2, s);
    if (e != hipSuccess) return e;
    st->keys = k2;
    st->vals = v2;
    StartIter<K> starts(hipcub::CountingInputIterator<int64_t>(0), RunStart<K>{k2});
    bytes = ibytes;
    return hipcub::DeviceScan::InclusiveSum(t2, bytes, starts, st->ids, (int)m, s);
}

template <typename K>
void launch_vertices(const SmcState* st, float* verts, hipStream_t s) {
    hipLaunchKernelGGL(smc_vertices<K>, dim3(blocks_for(st->m_raw)), dim3(kBlock), 0, s, (const K*)st->keys,
                       st->vals, st->ids, st->m_raw, st->coords, st->corners, st->cases, st->offsets, st->iso, verts,
                       st->map);
}

// Copies `bytes` from the device once the stream's work is done (through pinned memory
// when the extraction owns the shared scratch).
hipError_t read_back(SmcState* st, const void* src, size_t bytes, void* dst, hipStream_t s) {
    void* host = dst;
    if (st->shared && st->shared->readback) host = st->shared->readback;
    hipError_t e = hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && host != dst) std::memcpy(dst, host, bytes);
    return e;
}

}  // namespace

void smc_release(SmcState* st, hipStream_t s) {
    if (!st) return;
    for (void* p : st->blocks) (void)hipFreeAsync(p, s);
    if (st->shared) {
        std::lock_guard<std::mutex> lock(st->shared->mu);
        if (!st->shared->done) (void)hipEventCreateWithFlags(&st->shared->done, hipEventDisableTiming);
        if (st->shared->done) (void)hipEventRecord(st->shared->done, s);
        st->shared->busy = false;
    }
    delete st;
}

SmcStatus smc_extract(const int32_t* coords, const float* corners, int64_t n, float iso, bool consistent,
                      hipStream_t s, SmcState** out, int64_t* n_verts, int64_t* n_tris) {
    *out = nullptr;
    *n_verts = *n_tris = 0;
    SmcState* st = new SmcState();
    st->n = n;
    st->iso = iso;
    st->coords = coords;
    st->corners = corners;
    auto bail = [&](SmcStatus r) {
        smc_release(st, s);
        return r;
    };
    auto arg_error = [&](const std::string& msg) {
        SmcStatus r;
        r.code = -1;
        r.message = msg;
        return bail(r);
    };
    if (n == 0) {
        *out = st;
        return SmcStatus();
    }
    hipError_t e = hipGetDevice(&st->device);
    if (e != hipSuccess) return bail(hip_status(e, "hipGetDevice"));
    if (st->device >= 0 && st->device < kScratchDevices) {
        DeviceScratch& d = g_scratch[st->device];
        std::lock_guard<std::mutex> lock(d.mu);
        if (!d.busy) {
            d.busy = true;
            st->shared = &d;
            if (!d.readback) (void)hipHostMalloc((void**)&d.readback, 64, hipHostMallocDefault);
            if (d.done) e = hipStreamWaitEvent(s, d.done, 0);
        }
    }
    if (e != hipSuccess) return bail(hip_status(e, "hipStreamWaitEvent"));

    // region A: per-cell arrays (and the corner averaging, sized for 64-bit keys)
    const unsigned nb = std::min<unsigned>(blocks_for(n), 1024);
    size_t scan_bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)n);
    int32_t* part = nullptr;
    uint64_t* counts = nullptr;
    float* mean = nullptr;
    char* scan_tmp = nullptr;
    Carve corner_area;
    auto lay_a = [&](Carve& c) {
        st->info = c.take<int32_t>(16);
        part = c.take<int32_t>(6 * (int64_t)nb);
        st->cases = c.take<uint8_t>(n);
        counts = c.take<uint64_t>(n);
        st->offsets = c.take<uint64_t>(n);
        scan_tmp = c.take<char>((int64_t)scan_bytes);
        if (consistent) {
            mean = c.take<float>(8 * n);
            corner_area.base = c.base ? c.base + c.used : nullptr;
            const size_t need = 2 * align_up(8 * n * sizeof(uint64_t)) + 2 * align_up(8 * n * sizeof(uint32_t)) +
                                align_up(std::max(sort_bytes<uint64_t>(8 * n), sort_bytes<uint32_t>(8 * n)));
            c.take<char>((int64_t)need);
        }
    };
    Carve size_a;
    lay_a(size_a);
    Carve ca;
    e = region(st, 0, size_a.used, s, &ca.base);
    if (e != hipSuccess) return bail(hip_status(e, "scratch allocation"));
    lay_a(ca);

    // bounds of the cell coordinates
    hipLaunchKernelGGL(smc_bounds, dim3(nb), dim3(kBlock), 0, s, coords, n, part);
    hipLaunchKernelGGL(smc_fold_bounds, dim3(1), dim3(kBlock), 0, s, part, (int)nb, st->info);
    int32_t info[10];
    int B = 0;
    auto read_bounds = [&]() -> SmcStatus {
        int64_t span = 0;
        for (int k = 0; k < 3; ++k) span = std::max<int64_t>(span, (int64_t)info[3 + k] - info[k]);
        if (span > kSmcMaxSpan) {
            SmcStatus r;
            r.code = -1;
            r.message = "cell coordinates span " + std::to_string(span) + " lattice steps on one axis; at most " +
                        std::to_string(kSmcMaxSpan) + " are supported";
            return r;
        }
        B = bits_for(span + 1);  // lattice points min .. max + 1
        return SmcStatus();
    };
    if (consistent) {
        // the corner sort needs the key width before the cells are classified
        e = read_back(st, st->info, 6 * sizeof(int32_t), info, s);
        if (e != hipSuccess) return bail(hip_status(e, "cell bounds"));
        SmcStatus r = read_bounds();
        if (r.code) return bail(r);
        e = 3 * B <= 32 ? average_corners<uint32_t>(st, corner_area, B, mean, s)
                        : average_corners<uint64_t>(st, corner_area, B, mean, s);
        if (e != hipSuccess) return bail(hip_status(e, "corner averaging"));
        st->corners = mean;
    }

    // cases and per-cell counts, their exclusive scan and the totals
    hipLaunchKernelGGL(smc_classify, dim3(blocks_for(n)), dim3(kBlock), 0, s, st->corners, n, iso, st->cases, counts);
    e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, counts, st->offsets, (int)n, s);
    if (e != hipSuccess) return bail(hip_status(e, "count scan"));
    uint64_t* totals = reinterpret_cast<uint64_t*>(st->info + 8);
    hipLaunchKernelGGL(smc_totals, dim3(1), dim3(64), 0, s, counts, st->offsets, n, totals);
    e = read_back(st, st->info, sizeof info, info, s);
    if (e != hipSuccess) return bail(hip_status(e, "bounds and totals"));
    if (!consistent) {
        SmcStatus r = read_bounds();
        if (r.code) return bail(r);
    }
    uint64_t tot = 0;
    std::memcpy(&tot, info + 8, sizeof tot);
    st->m_raw = (int64_t)(tot >> 32);
    st->t = (int64_t)(uint32_t)tot;
    if (st->m_raw >= (int64_t)INT32_MAX) return arg_error("more than 2^31 - 1 edge vertices");
    if (st->m_raw == 0) {
        *out = st;
        *n_tris = st->t;
        return SmcStatus();
    }

    // region B: edge keys in cell order, stable sort, vertex ids from run starts
    st->wide = 3 * B + 2 > 32;
    e = st->wide ? order_edges<uint64_t>(st, B, s) : order_edges<uint32_t>(st, B, s);
    if (e != hipSuccess) return bail(hip_status(e, "edge ordering"));
    uint32_t m = 0;
    e = read_back(st, st->ids + (st->m_raw - 1), sizeof m, &m, s);
    if (e != hipSuccess) return bail(hip_status(e, "vertex count"));
    st->m = m;
    *out = st;
    *n_verts = st->m;
    *n_tris = st->t;
    return SmcStatus();
}

SmcStatus smc_write(SmcState* st, float* verts, int32_t* tris, hipStream_t s) {
    SmcStatus r;
    if (st->m_raw > 0) {
        if (st->wide)
            launch_vertices<uint64_t>(st, verts, s);
        else
            launch_vertices<uint32_t>(st, verts, s);
        if (st->t > 0)
            hipLaunchKernelGGL(smc_triangles, dim3(blocks_for(st->n)), dim3(kBlock), 0, s, st->cases, st->offsets,
                               st->n, st->map, tris);
        r = hip_status(hipGetLastError(), "marching cubes kernel launch");
    }
    return r;
}

}  // namespace cubvh_amd
