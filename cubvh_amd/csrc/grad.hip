// grad.hip -- backward of the unsigned distance (no counterpart in the reference).
//
// d = |p - c| with c the closest point on the nearest face.  By the envelope
// theorem the derivative of the minimum over the face equals the derivative at the
// minimiser, so with n = (p - c) / d:
//   dd/dp   = n
//   dd/dv_k = -w_k n   for the face's vertex k with barycentric weight w_k
// c is rebuilt from the stored face id and barycentrics, against the current
// vertex values.  d = 0 gives the zero subgradient.
//
// Vertex gradients are gathered, not scattered with atomics, so they are the same
// bit for bit on every run (and equal whatever other meshes or queries share the
// call, as long as a mesh's own queries keep their relative order):
//   1. query_terms  : g n per query (the point gradient), and its face as a sort key
//   2. radix sort   : queries grouped by face, stable (ascending query index inside
//                     a face)
//   3. face_ranges  : each face's range of the sorted queries, and its number of
//                     chunks of kChunk queries counted from the face's first query
//   4. scan         : the chunks' positions (exclusive sum over the faces)
//   5. chunk_sums   : one lane per chunk adds -w_k g n over its queries in order
//   6. face_totals  : one lane per face adds its chunks' sums in order
//   7. radix sort   : the face corners (3 f + k) grouped by vertex, stable
//   8. vertex_sums  : one lane per vertex adds its corners' face totals in corner
//                     order and adds the total to grad_vertices
// Queries per face are very uneven (far points all reach the few faces that stick out
// most: on the jittered level-6 sphere with 2^20 N(0, 0.4^2) points, 45% of the faces get
// none and one gets 2,867), so a face's queries are split into chunks that lanes sum
// side by side.  Device-wide float atomics on this multi-XCD part are resolved past the
// per-XCD L2 caches; nine per query made the scatter form 0.6 ms for 2^20 queries.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "query.hpp"

namespace cubvh_amd {
namespace dev {

// Per query: grad_p = g n (zero for an out-of-range face or vertex, or d = 0); the
// sort key is the face, or n_faces (sorted after every face, never summed) when the
// query adds nothing to the vertices.
__global__ __launch_bounds__(256) void query_terms(int64_t n, const float* __restrict__ g,
                                                   const float* __restrict__ p,
                                                   const int64_t* __restrict__ face_id,
                                                   const float* __restrict__ uvw, const float* __restrict__ V,
                                                   const int32_t* __restrict__ F, int64_t n_vertices,
                                                   int64_t n_faces, float* __restrict__ gn,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t f = face_id[i];
    // out-of-range face or vertex indices contribute nothing
    bool ok = f >= 0 && f < n_faces;
    const int32_t ia = ok ? F[3 * f] : 0, ib = ok ? F[3 * f + 1] : 0, ic = ok ? F[3 * f + 2] : 0;
    ok = ok && ia >= 0 && ia < n_vertices && ib >= 0 && ib < n_vertices && ic >= 0 && ic < n_vertices;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f, s = 0.0f;
    if (ok) {
        const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
        float d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = __builtin_fmaf(w, V[3 * ic + k], __builtin_fmaf(v, V[3 * ib + k], u * V[3 * ia + k]));
            d[k] = p[3 * i + k] - c;
        }
        const float len = sqrtf(__builtin_fmaf(d[2], d[2], __builtin_fmaf(d[1], d[1], d[0] * d[0])));
        s = len > 0.0f ? g[i] / len : 0.0f;
        nx = d[0] * s;  // g n
        ny = d[1] * s;
        nz = d[2] * s;
    }
    if (gn) {
        gn[3 * i] = nx;
        gn[3 * i + 1] = ny;
        gn[3 * i + 2] = nz;
    }
    if (key) {
        key[i] = (ok && s != 0.0f) ? (uint32_t)f : (uint32_t)n_faces;
        val[i] = (uint32_t)i;
    }
}

// First position in sorted[0, n) whose value is >= x.
__device__ __forceinline__ int64_t lower_bound(const uint32_t* __restrict__ sorted, int64_t n, uint32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (sorted[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

constexpr int kChunk = 32;  // queries per chunk of a face's sorted queries

// Per face: its range [start, start + count) of the sorted queries and its chunk
// count.  One lane per face, two binary searches in the sorted keys.  nchunk has
// n_faces + 1 entries, the last one 0 (the scan's total).
__global__ __launch_bounds__(256) void face_ranges(int64_t n, int64_t n_faces, const uint32_t* __restrict__ skey,
                                                   uint32_t* __restrict__ start, uint32_t* __restrict__ nchunk) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f > n_faces) return;
    if (f == n_faces) {
        nchunk[f] = 0;
        return;
    }
    const int64_t lo = lower_bound(skey, n, (uint32_t)f);
    const int64_t hi = lower_bound(skey, n, (uint32_t)f + 1u);
    start[f] = (uint32_t)lo;
    nchunk[f] = (uint32_t)((hi - lo + kChunk - 1) / kChunk);
}

// Lane per chunk t < cbase[n_faces]: the face f with cbase[f] <= t < cbase[f + 1],
// chunk k = t - cbase[f] covers the face's queries kChunk k .. kChunk (k + 1) - 1 (of
// its count); their -w_k g n summed in order go to part[9 t ..].
__global__ __launch_bounds__(256) void chunk_sums(int64_t n, int64_t n_faces, int64_t max_chunks,
                                                  const uint32_t* __restrict__ skey,
                                                  const uint32_t* __restrict__ sval, const uint32_t* __restrict__ start,
                                                  const uint32_t* __restrict__ cbase, const float* __restrict__ gn,
                                                  const float* __restrict__ uvw, float* __restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = cbase[n_faces];
    if (t >= total || t >= max_chunks) return;
    // face: last f with cbase[f] <= t (faces without queries have cbase[f] == cbase[f + 1])
    int64_t lo = 0, hi = n_faces - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (cbase[mid] <= (uint32_t)t) lo = mid;
        else hi = mid - 1;
    }
    const int64_t f = lo;
    const int64_t k = t - cbase[f];
    const int64_t j0 = (int64_t)start[f] + kChunk * k;
    float a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j = j0; j < j0 + kChunk && j < n && skey[j] == (uint32_t)f; ++j) {
        const int64_t q = sval[j];
        const float x = gn[3 * q], y = gn[3 * q + 1], z = gn[3 * q + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float wk = uvw[3 * q + c];
            a[3 * c] += -wk * x;
            a[3 * c + 1] += -wk * y;
            a[3 * c + 2] += -wk * z;
        }
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) part[9 * t + c] = a[c];
}

// Lane per face: the face's chunk sums added in chunk order, written as
// acc[3 (3 f + k) + axis] (zero for a face without queries).
__global__ __launch_bounds__(256) void face_totals(int64_t n_faces, int64_t max_chunks,
                                                   const uint32_t* __restrict__ cbase, const float* __restrict__ part,
                                                   float* __restrict__ acc) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces) return;
    float a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t b = cbase[f], e = min((int64_t)cbase[f + 1], max_chunks);
    for (int64_t t = b; t < e; ++t)
#pragma unroll
        for (int c = 0; c < 9; ++c) a[c] += part[9 * t + c];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[9 * f + c] = a[c];
}

// Corner keys: corner s = 3 f + k of every face has key F[s] (its vertex), or
// n_vertices (never summed) for an out-of-range index.
__global__ __launch_bounds__(256) void corner_keys(int64_t n_corners, const int32_t* __restrict__ F,
                                                   int64_t n_vertices, uint32_t* __restrict__ key,
                                                   uint32_t* __restrict__ val) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_corners) return;
    const int32_t v = F[s];
    key[s] = (v >= 0 && v < n_vertices) ? (uint32_t)v : (uint32_t)n_vertices;
    val[s] = (uint32_t)s;
}

// Lane per vertex: adds the sums of its corners (corner index ascending) to gv.
__global__ __launch_bounds__(256) void vertex_sums(int64_t n_corners, int64_t n_vertices,
                                                   const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
                                                   const float* __restrict__ acc, float* __restrict__ gv) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vertices) return;
    const int64_t lo = lower_bound(skey, n_corners, (uint32_t)v);
    const int64_t hi = lower_bound(skey, n_corners, (uint32_t)v + 1u);
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (int64_t j = lo; j < hi; ++j) {
        const int64_t s = sval[j];
        x += acc[3 * s];
        y += acc[3 * s + 1];
        z += acc[3 * s + 2];
    }
    gv[3 * v] += x;
    gv[3 * v + 1] += y;
    gv[3 * v + 2] += z;
}

}  // namespace dev

namespace {

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bits of the largest key (the sentinel `top`).
int key_bits(int64_t top) {
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= top) ++b;
    return b;
}

size_t sort_temp_bytes(int64_t m, int bits) {
    size_t tmp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)m, 0, bits);
    return tmp;
}

size_t scan_temp_bytes(int64_t m) {
    size_t tmp = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)m);
    return tmp;
}

// Chunks of one call: at most n / kChunk full ones plus one partial chunk per face
// that has queries.
int64_t max_chunks(int64_t n, int64_t n_faces) { return n / dev::kChunk + std::min(n, n_faces); }

struct BwdLayout {
    size_t gn = 0, k0 = 0, v0 = 0, k1 = 0, v1 = 0, acc = 0, c0 = 0, cv0 = 0, c1 = 0, cv1 = 0, tmp = 0, total = 0;
    size_t start = 0, nchunk = 0, cbase = 0, part = 0;
    size_t tmp_bytes = 0;
};

BwdLayout bwd_layout(int64_t n, int64_t n_vertices, int64_t n_faces, bool need_gn) {
    BwdLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += al256(bytes);
        return at;
    };
    const int64_t nc = 3 * n_faces;
    if (need_gn) L.gn = take(12 * (size_t)n);
    L.k0 = take(4 * (size_t)n);
    L.v0 = take(4 * (size_t)n);
    L.k1 = take(4 * (size_t)n);
    L.v1 = take(4 * (size_t)n);
    L.acc = take(36 * (size_t)n_faces);
    L.start = take(4 * (size_t)n_faces);
    L.nchunk = take(4 * (size_t)(n_faces + 1));
    L.cbase = take(4 * (size_t)(n_faces + 1));
    L.part = take(36 * (size_t)max_chunks(n, n_faces));
    L.c0 = take(4 * (size_t)nc);
    L.cv0 = take(4 * (size_t)nc);
    L.c1 = take(4 * (size_t)nc);
    L.cv1 = take(4 * (size_t)nc);
    L.tmp_bytes = std::max(std::max(sort_temp_bytes(n, key_bits(n_faces)), sort_temp_bytes(nc, key_bits(n_vertices))),
                           scan_temp_bytes(n_faces + 1));
    L.tmp = take(L.tmp_bytes);
    L.total = o;
    return L;
}

}  // namespace

size_t udf_backward_workspace_bytes(int64_t n, int64_t n_vertices, int64_t n_faces, bool have_grad_points) {
    return bwd_layout(n, n_vertices, n_faces, !have_grad_points).total;
}

hipError_t launch_udf_backward(int64_t n, const float* grad_dist, const float* positions, const int64_t* face_id,
                               const float* uvw, const float* vertices, int64_t n_vertices, const int32_t* faces,
                               int64_t n_faces, float* grad_points, float* grad_vertices, void* ws,
                               hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const unsigned gq = (unsigned)((n + 255) / 256);
    if (!grad_vertices) {
        if (grad_points)
            hipLaunchKernelGGL(dev::query_terms, dim3(gq), dim3(256), 0, stream, n, grad_dist, positions, face_id, uvw,
                               vertices, faces, n_vertices, n_faces, grad_points, (uint32_t*)nullptr,
                               (uint32_t*)nullptr);
        return hipGetLastError();
    }
    const BwdLayout L = bwd_layout(n, n_vertices, n_faces, grad_points == nullptr);
    char* w = (char*)ws;
    float* gn = grad_points ? grad_points : (float*)(w + L.gn);
    uint32_t *k0 = (uint32_t*)(w + L.k0), *v0 = (uint32_t*)(w + L.v0);
    uint32_t *k1 = (uint32_t*)(w + L.k1), *v1 = (uint32_t*)(w + L.v1);
    float* acc = (float*)(w + L.acc);
    uint32_t *c0 = (uint32_t*)(w + L.c0), *cv0 = (uint32_t*)(w + L.cv0);
    uint32_t *c1 = (uint32_t*)(w + L.c1), *cv1 = (uint32_t*)(w + L.cv1);
    void* tmp = w + L.tmp;
    const int64_t nc = 3 * n_faces;
    hipLaunchKernelGGL(dev::query_terms, dim3(gq), dim3(256), 0, stream, n, grad_dist, positions, face_id, uvw,
                       vertices, faces, n_vertices, n_faces, gn, k0, v0);
    size_t tb = L.tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)n, 0, key_bits(n_faces), stream);
    if (e != hipSuccess) return e;
    uint32_t* start = (uint32_t*)(w + L.start);
    uint32_t* nchunk = (uint32_t*)(w + L.nchunk);
    uint32_t* cbase = (uint32_t*)(w + L.cbase);
    float* part = (float*)(w + L.part);
    const int64_t mc = max_chunks(n, n_faces);
    hipLaunchKernelGGL(dev::face_ranges, dim3((unsigned)((n_faces + 1 + 255) / 256)), dim3(256), 0, stream, n, n_faces, k1,
                       start, nchunk);
    tb = L.tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, nchunk, cbase, (int)(n_faces + 1), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::chunk_sums, dim3((unsigned)((mc + 255) / 256)), dim3(256), 0, stream, n, n_faces, mc, k1, v1,
                       start, cbase, gn, uvw, part);
    hipLaunchKernelGGL(dev::face_totals, dim3((unsigned)((n_faces + 255) / 256)), dim3(256), 0, stream, n_faces, mc, cbase,
                       part, acc);
    hipLaunchKernelGGL(dev::corner_keys, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, nc, faces, n_vertices,
                       c0, cv0);
    tb = L.tmp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, c0, c1, cv0, cv1, (int)nc, 0, key_bits(n_vertices), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::vertex_sums, dim3((unsigned)((n_vertices + 255) / 256)), dim3(256), 0, stream, nc,
                       n_vertices, c1, cv1, acc, grad_vertices);
    return hipGetLastError();
}

}  // namespace cubvh_amd
