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
//   3. run_bounds   : each face's range [begin, end) of the sorted queries, from the
//                     run boundaries of the sorted keys (one pass; absent faces 0, 0)
//   4. face_chunks  : each face's number of chunks of kChunk queries, counted from the
//                     face's first query; a scan gives the chunks' positions, and
//                     chunk_marks + a max-scan give each chunk its face
//   5. chunk_sums   : one lane per chunk adds -w_k g n over its queries in order
//   6. a sum tree of fan-in kChunk over each face's chunk sums (chunk_tree_level
//                     passes, in place): a level adds runs of kChunk values in order,
//                     starting from zero, and leaves each run's sum at the run's first
//                     chunk, so a face with k queries costs about log_32(k) short passes
//                     instead of k / kChunk additions on one lane, and its total ends at
//                     its first chunk.  A face of at most kChunk chunks is one run of
//                     its chunk sums; passes beyond a face's own depth leave its total
//                     unchanged (0 + x == x for a sum that began at +0, never -0), so a
//                     pass runs only when chunk_marks found a face with more chunks
//                     than its stride
//   7. vertex_sums  : one lane per vertex adds its corners' face totals in corner
//                     order (skipping faces without queries, whose +0 would change
//                     nothing) and adds the total to grad_vertices
// Step 7 reads the vertex -> corners table (the corners 3 f + k of every face grouped
// by vertex, ascending inside a vertex: a stable radix sort of the corners by vertex
// and run_bounds over it).  It depends only on the faces, so a caller that runs many
// backward passes over one mesh builds it once (launch_vertex_corners); otherwise
// each call builds it in its workspace.
// Queries per face are very uneven (far points all reach the few faces that stick out
// most: on the jittered level-6 sphere with 2^20 N(0, 0.4^2) points, 45% of the faces get
// none and one gets 2,867), so a face's queries are split into chunks that lanes sum
// side by side.  Device-wide float atomics on this multi-XCD part are resolved past the
// per-XCD L2 caches; nine per query made the scatter form 0.6 ms for 2^20 queries.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "query.hpp"
#include "radix.hpp"

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

constexpr int kChunk = 32;  // queries per chunk of a face's sorted queries
// Passes of the faces' sum trees: strides kChunk^0 .. kChunk^(kTreeLevels - 1) cover
// every face of a call below 2^31 queries (at most 2^26 chunks per face).
constexpr int kTreeLevels = 6;

__global__ __launch_bounds__(256) void grad_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Runs of equal keys in sorted[0, n): for each key k < n_keys present, begin[k] = its
// first position and end[k] = one past its last (begin and end must hold 0 for keys
// that are absent; keys >= n_keys are sentinels and ignored).
__global__ __launch_bounds__(256) void run_bounds(int64_t n, const uint32_t* __restrict__ sorted, int64_t n_keys,
                                                  uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sorted[i];
    if ((int64_t)k >= n_keys) return;
    if (i == 0 || sorted[i - 1] != k) begin[k] = (uint32_t)i;
    if (i == n - 1 || sorted[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

// Per face: its number of chunks of kChunk sorted queries.  nchunk has n_faces + 1
// entries, the last one 0 (the scan's total).
__global__ __launch_bounds__(256) void face_chunks(int64_t n_faces, const uint32_t* __restrict__ begin,
                                                   const uint32_t* __restrict__ end, uint32_t* __restrict__ nchunk) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f > n_faces) return;
    nchunk[f] = f == n_faces ? 0u : (end[f] - begin[f] + kChunk - 1) / kChunk;
}

// Per face with queries: marks its first chunk with f + 1 (mark holds zeros elsewhere);
// an inclusive max-scan of the marks then gives every chunk its face + 1, whatever the
// number of chunks of one face.  need[j] (zeroed beforehand) is set when some face has
// more than kChunk^j chunks: sum-tree pass j changes nothing otherwise.
__global__ __launch_bounds__(256) void chunk_marks(int64_t n_faces, int64_t max_chunks,
                                                   const uint32_t* __restrict__ nchunk,
                                                   const uint32_t* __restrict__ cbase, uint32_t* __restrict__ mark,
                                                   uint32_t* __restrict__ need) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_faces || nchunk[f] == 0u || (int64_t)cbase[f] >= max_chunks) return;
    mark[cbase[f]] = (uint32_t)f + 1u;
    const uint32_t c = nchunk[f];
    uint32_t stride = 1u;
    for (int j = 0; j < kTreeLevels && c > stride; ++j, stride *= (uint32_t)kChunk)
        if (need[j] == 0u) need[j] = 1u;  // every writer stores the same value
}

// Lane per chunk t < cbase[n_faces]: chunk k = t - cbase[f] of face f = cface[t] - 1 covers
// the face's sorted queries begin[f] + kChunk k .. (up to end[f]); their -w_k g n summed
// in order go to part[9 t ..].
__global__ __launch_bounds__(256) void chunk_sums(int64_t n_faces, int64_t max_chunks, const uint32_t* __restrict__ sval,
                                                  const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end,
                                                  const uint32_t* __restrict__ cbase, const uint32_t* __restrict__ cface,
                                                  const float* __restrict__ gn, const float* __restrict__ uvw,
                                                  float* __restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = cbase[n_faces];
    if (t >= total || t >= max_chunks) return;
    const int64_t f = (int64_t)cface[t] - 1;
    const int64_t j0 = (int64_t)begin[f] + kChunk * (t - (int64_t)cbase[f]);
    const int64_t j1 = min(j0 + kChunk, (int64_t)end[f]);
    float a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t j = j0; j < j1; ++j) {
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

// One level of each face's sum tree, in place: chunk t = cbase[f] + r of face f with r
// a multiple of kChunk * stride receives the sum, from zero and in order, of the values
// at t + i stride (i = 0 .. kChunk - 1, inside the face's chunks).  With stride =
// kChunk^(L-1) the values read are level L-1's and the ones written level L's; a chunk
// reads only below t + kChunk * stride, where the face's next written chunk is.  The
// chunks are taken grid-stride, so that a pass with nothing to do is one small launch.
__global__ __launch_bounds__(256) void chunk_tree_level(int64_t n_faces, int64_t max_chunks, int64_t stride,
                                                        const uint32_t* __restrict__ need_level,
                                                        const uint32_t* __restrict__ cbase,
                                                        const uint32_t* __restrict__ cface, float* __restrict__ part) {
    if (*need_level == 0u) return;  // no face has more chunks than `stride`
    const int64_t total = min((int64_t)cbase[n_faces], max_chunks);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = (int64_t)cface[t] - 1;
        if ((t - (int64_t)cbase[f]) % (stride * kChunk) != 0) continue;
        const int64_t e = min((int64_t)cbase[f + 1], max_chunks);
        float a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t i = 0, u = t; i < kChunk && u < e; ++i, u += stride)
#pragma unroll
            for (int c = 0; c < 9; ++c) a[c] += part[9 * u + c];
#pragma unroll
        for (int c = 0; c < 9; ++c) part[9 * t + c] = a[c];
    }
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

// Lane per vertex: adds the totals of its corners corners[begin[v] .. end[v]) (corner
// index ascending) to gv; the total of corner s = 3 f + k is part[9 cbase[f] + 3 k ..],
// and a face without queries adds nothing.  A table passed in by the caller is device
// data the host does not check: ranges and corner indices outside the n_corners corners
// are skipped.
__global__ __launch_bounds__(256) void vertex_sums(int64_t n_vertices, int64_t n_corners, int64_t max_chunks,
                                                   const uint32_t* __restrict__ corners,
                                                   const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end,
                                                   const uint32_t* __restrict__ cbase, const float* __restrict__ part,
                                                   float* __restrict__ gv) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vertices) return;
    const int64_t lo = begin[v], hi = min((int64_t)end[v], n_corners);
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (int64_t j = lo; j < hi; ++j) {
        const int64_t s = corners[j];
        if (s >= n_corners) continue;
        const int64_t f = s / 3;
        const int64_t c = cbase[f];
        if (c == (int64_t)cbase[f + 1] || c >= max_chunks) continue;
        const float* a = part + 9 * c + 3 * (s - 3 * f);
        x += a[0];
        y += a[1];
        z += a[2];
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
    radix_sort_pairs_u32(nullptr, tmp, nullptr, nullptr, nullptr, nullptr, m, bits, 0);
    return tmp;
}

size_t max_scan_temp_bytes(int64_t m) {
    size_t tmp = 0;
    hipcub::DeviceScan::InclusiveScan(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr, hipcub::Max(), (int)m);
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

// Workspace of launch_vertex_corners: corner keys and values before the sort, the
// sorted keys, and the sort's temporary storage.
struct CornerLayout {
    size_t c0 = 0, cv0 = 0, c1 = 0, tmp = 0, total = 0, tmp_bytes = 0;
};

CornerLayout corner_layout(int64_t n_vertices, int64_t n_faces) {
    CornerLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += al256(bytes);
        return at;
    };
    const int64_t nc = 3 * n_faces;
    L.c0 = take(4 * (size_t)nc);
    L.cv0 = take(4 * (size_t)nc);
    L.c1 = take(4 * (size_t)nc);
    L.tmp_bytes = sort_temp_bytes(nc, key_bits(n_vertices));
    L.tmp = take(L.tmp_bytes);
    L.total = o;
    return L;
}

struct BwdLayout {
    size_t gn = 0, k0 = 0, v0 = 0, k1 = 0, v1 = 0, need = 0, tmp = 0, total = 0;
    size_t fbeg = 0, fend = 0, nchunk = 0, cbase = 0, mark = 0, cface = 0, part = 0;
    size_t corners = 0, vbeg = 0, vend = 0, corner_ws = 0;
    size_t tmp_bytes = 0;
};

BwdLayout bwd_layout(int64_t n, int64_t n_vertices, int64_t n_faces, bool need_gn, bool need_tables) {
    BwdLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += al256(bytes);
        return at;
    };
    if (need_gn) L.gn = take(12 * (size_t)n);
    L.k0 = take(4 * (size_t)n);
    L.v0 = take(4 * (size_t)n);
    L.k1 = take(4 * (size_t)n);
    L.v1 = take(4 * (size_t)n);
    L.need = take(4 * (size_t)dev::kTreeLevels);
    L.fbeg = take(4 * (size_t)n_faces);
    L.fend = take(4 * (size_t)n_faces);
    L.nchunk = take(4 * (size_t)(n_faces + 1));
    L.cbase = take(4 * (size_t)(n_faces + 1));
    L.mark = take(4 * (size_t)max_chunks(n, n_faces));
    L.cface = take(4 * (size_t)max_chunks(n, n_faces));
    L.part = take(36 * (size_t)max_chunks(n, n_faces));
    if (need_tables) {
        L.corners = take(12 * (size_t)n_faces);
        L.vbeg = take(4 * (size_t)n_vertices);
        L.vend = take(4 * (size_t)n_vertices);
        L.corner_ws = take(corner_layout(n_vertices, n_faces).total);
    }
    L.tmp_bytes = std::max(std::max(sort_temp_bytes(n, key_bits(n_faces)), scan_temp_bytes(n_faces + 1)),
                           max_scan_temp_bytes(max_chunks(n, n_faces)));
    L.tmp = take(L.tmp_bytes);
    L.total = o;
    return L;
}

}  // namespace

size_t vertex_corners_workspace_bytes(int64_t n_vertices, int64_t n_faces) {
    return corner_layout(n_vertices, n_faces).total;
}

hipError_t launch_vertex_corners(const int32_t* faces, int64_t n_faces, int64_t n_vertices, uint32_t* corners,
                                 uint32_t* begin, uint32_t* end, void* ws, hipStream_t stream) {
    if (n_vertices <= 0) return hipSuccess;
    const int64_t nc = 3 * n_faces;
    const unsigned gv = (unsigned)((n_vertices + 255) / 256);
    hipLaunchKernelGGL(dev::grad_fill_u32, dim3(gv), dim3(256), 0, stream, begin, n_vertices, 0u);
    hipLaunchKernelGGL(dev::grad_fill_u32, dim3(gv), dim3(256), 0, stream, end, n_vertices, 0u);
    if (nc <= 0) return hipGetLastError();
    const CornerLayout L = corner_layout(n_vertices, n_faces);
    char* w = (char*)ws;
    uint32_t *c0 = (uint32_t*)(w + L.c0), *cv0 = (uint32_t*)(w + L.cv0), *c1 = (uint32_t*)(w + L.c1);
    hipLaunchKernelGGL(dev::corner_keys, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, nc, faces, n_vertices,
                       c0, cv0);
    size_t tb = L.tmp_bytes;
    hipError_t e = radix_sort_pairs_u32((void*)(w + L.tmp), tb, c0, c1, cv0, corners, nc, key_bits(n_vertices), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::run_bounds, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, nc, c1, n_vertices,
                       begin, end);
    return hipGetLastError();
}

size_t udf_backward_workspace_bytes(int64_t n, int64_t n_vertices, int64_t n_faces, bool have_grad_points,
                                    bool have_tables) {
    return bwd_layout(n, n_vertices, n_faces, !have_grad_points, !have_tables).total;
}

hipError_t launch_udf_backward(int64_t n, const float* grad_dist, const float* positions, const int64_t* face_id,
                               const float* uvw, const float* vertices, int64_t n_vertices, const int32_t* faces,
                               int64_t n_faces, float* grad_points, float* grad_vertices, const uint32_t* corners,
                               const uint32_t* corner_begin, const uint32_t* corner_end, void* ws, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const unsigned gq = (unsigned)((n + 255) / 256);
    if (!grad_vertices) {
        if (grad_points)
            hipLaunchKernelGGL(dev::query_terms, dim3(gq), dim3(256), 0, stream, n, grad_dist, positions, face_id, uvw,
                               vertices, faces, n_vertices, n_faces, grad_points, (uint32_t*)nullptr,
                               (uint32_t*)nullptr);
        return hipGetLastError();
    }
    const bool have_tables = corners && corner_begin && corner_end;
    const BwdLayout L = bwd_layout(n, n_vertices, n_faces, grad_points == nullptr, !have_tables);
    char* w = (char*)ws;
    float* gn = grad_points ? grad_points : (float*)(w + L.gn);
    uint32_t *k0 = (uint32_t*)(w + L.k0), *v0 = (uint32_t*)(w + L.v0);
    uint32_t *k1 = (uint32_t*)(w + L.k1), *v1 = (uint32_t*)(w + L.v1);
    void* tmp = w + L.tmp;
    hipError_t e;
    hipLaunchKernelGGL(dev::query_terms, dim3(gq), dim3(256), 0, stream, n, grad_dist, positions, face_id, uvw,
                       vertices, faces, n_vertices, n_faces, gn, k0, v0);
    size_t tb = L.tmp_bytes;
    e = radix_sort_pairs_u32(tmp, tb, k0, k1, v0, v1, n, key_bits(n_faces), stream);
    if (e != hipSuccess) return e;
    uint32_t* fbeg = (uint32_t*)(w + L.fbeg);
    uint32_t* fend = (uint32_t*)(w + L.fend);
    uint32_t* nchunk = (uint32_t*)(w + L.nchunk);
    uint32_t* cbase = (uint32_t*)(w + L.cbase);
    uint32_t* cface = (uint32_t*)(w + L.cface);
    float* part = (float*)(w + L.part);
    const int64_t mc = max_chunks(n, n_faces);
    const unsigned gf = (unsigned)((n_faces + 1 + 255) / 256);
    if (n_faces > 0) {
        hipLaunchKernelGGL(dev::grad_fill_u32, dim3(gf), dim3(256), 0, stream, fbeg, n_faces, 0u);
        hipLaunchKernelGGL(dev::grad_fill_u32, dim3(gf), dim3(256), 0, stream, fend, n_faces, 0u);
    }
    hipLaunchKernelGGL(dev::run_bounds, dim3(gq), dim3(256), 0, stream, n, k1, n_faces, fbeg, fend);
    hipLaunchKernelGGL(dev::face_chunks, dim3(gf), dim3(256), 0, stream, n_faces, fbeg, fend, nchunk);
    tb = L.tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, nchunk, cbase, (int)(n_faces + 1), stream);
    if (e != hipSuccess) return e;
    uint32_t* mark = (uint32_t*)(w + L.mark);
    uint32_t* need = (uint32_t*)(w + L.need);
    hipLaunchKernelGGL(dev::grad_fill_u32, dim3((unsigned)((mc + 255) / 256)), dim3(256), 0, stream, mark, mc, 0u);
    hipLaunchKernelGGL(dev::grad_fill_u32, dim3(1), dim3(256), 0, stream, need, (int64_t)dev::kTreeLevels, 0u);
    hipLaunchKernelGGL(dev::chunk_marks, dim3(gf), dim3(256), 0, stream, n_faces, mc, nchunk, cbase, mark, need);
    tb = L.tmp_bytes;
    e = hipcub::DeviceScan::InclusiveScan(tmp, tb, mark, cface, hipcub::Max(), (int)mc, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::chunk_sums, dim3((unsigned)((mc + 255) / 256)), dim3(256), 0, stream, n_faces, mc, v1, fbeg,
                       fend, cbase, cface, gn, uvw, part);
    // the faces' sum trees: a face holds at most ceil(n / kChunk) chunks, so passes of
    // strides 1, kChunk, ... below that bound leave every face's total at its first chunk
    const int64_t face_chunks_max = (n + dev::kChunk - 1) / dev::kChunk;
    int64_t stride = 1;
    for (int j = 0; j < dev::kTreeLevels && stride < face_chunks_max; ++j, stride *= dev::kChunk)
        hipLaunchKernelGGL(dev::chunk_tree_level, dim3((unsigned)std::min<int64_t>((mc + 255) / 256, 2048)), dim3(256), 0,
                           stream, n_faces, mc, stride, need + j, cbase, cface, part);
    if (!have_tables) {
        corners = (const uint32_t*)(w + L.corners);
        corner_begin = (const uint32_t*)(w + L.vbeg);
        corner_end = (const uint32_t*)(w + L.vend);
        e = launch_vertex_corners(faces, n_faces, n_vertices, (uint32_t*)(w + L.corners), (uint32_t*)(w + L.vbeg),
                                  (uint32_t*)(w + L.vend), w + L.corner_ws, stream);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dev::vertex_sums, dim3((unsigned)((n_vertices + 255) / 256)), dim3(256), 0, stream, n_vertices,
                       3 * n_faces, mc, corners, corner_begin, corner_end, cbase, part, grad_vertices);
    return hipGetLastError();
}

}  // namespace cubvh_amd
