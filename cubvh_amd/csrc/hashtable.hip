// hashtable.hip -- static open-addressed hash table of int32 coordinate keys (sparse
// voxel lookup).
//
// Behaviour follows the reference's cuHashTable (include/gpu/hashtable.cuh:124-217):
// linear probing, one slot per inserted row (duplicate keys are not merged), a key
// that finds no free slot in a full table is dropped, and a search answers with the
// row index of a stored equal key or -1.  The layout is this library's own: the key is
// stored in the slot (words 0..D-1, row index in the last word of the power-of-two
// slot), so a probe is one aligned vector load, where the reference keeps (marker,
// row) pairs and compares each candidate through the caller's coordinate array.
// A search scans the probe run to its first empty slot and answers with the lowest
// matching row index: deterministic, and the answer of the reference's sequential
// table (include/cpu/hashtable.h:123-146), whose earlier rows take the earlier slots
// of a key's probe run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hashtable.hpp"

namespace cubvh_amd {
namespace {

// lowbias32 finalizer (C. Wellons, "hash-prospector", public domain constants).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Slot where the probe run of a key starts: a 32-bit hash of the words, scaled to
// [0, capacity) by a multiply-shift (no division).
template <int S>
__device__ __forceinline__ uint64_t home_slot(const int32_t (&key)[S - 1], int32_t d, int64_t capacity) {
    uint32_t h = 0x2545f491u;
#pragma unroll
    for (int k = 0; k < S - 1; ++k)
        if (k < d) h = mix32(h ^ (uint32_t)key[k]);
    return ((uint64_t)h * (uint64_t)capacity) >> 32;
}

template <int S>
__device__ __forceinline__ void load_key(const int32_t* __restrict__ row, int32_t d, int32_t (&key)[S - 1]) {
#pragma unroll
    for (int k = 0; k < S - 1; ++k) key[k] = k < d ? row[k] : 0;
}

__global__ __launch_bounds__(256) void ht_clear(int4* __restrict__ table, int64_t n_vec) {
    const int4 empty = make_int4(-1, -1, -1, -1);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_vec; i += (int64_t)gridDim.x * blockDim.x)
        table[i] = empty;
}

// One thread per row: claim the first free slot of the key's probe run with a
// compare-and-swap on its row word, then write the key words.  A slot goes from
// empty to claimed once, so a plain read that still shows it empty only costs a
// failed swap, and one that shows it claimed is final.
template <int S>
__global__ __launch_bounds__(256) void ht_insert(int32_t* __restrict__ table, int64_t capacity,
                                                 const int32_t* __restrict__ coords, int64_t n, int32_t d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t key[S - 1];
    load_key<S>(coords + i * d, d, key);
    uint64_t s = home_slot<S>(key, d, capacity);
    for (int64_t a = 0; a < capacity; ++a) {
        int32_t* w = table + s * S;
        if (w[S - 1] == -1 && atomicCAS(w + S - 1, -1, (int32_t)i) == -1) {
#pragma unroll
            for (int k = 0; k < S - 1; ++k)
                if (k < d) w[k] = key[k];
            return;
        }
        s = s + 1 == (uint64_t)capacity ? 0 : s + 1;
    }
}

// One thread per query: walk the key's probe run to its first empty slot (at most
// capacity slots) and keep the lowest row index whose key words are equal.
template <int S>
__global__ __launch_bounds__(256) void ht_search(const int32_t* __restrict__ table, int64_t capacity,
                                                 const int32_t* __restrict__ queries, int64_t m, int32_t d,
                                                 int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int32_t key[S - 1];
    load_key<S>(queries + i * d, d, key);
    uint64_t s = home_slot<S>(key, d, capacity);
    int32_t best = INT32_MAX;
    for (int64_t a = 0; a < capacity; ++a) {
        const int4* p = reinterpret_cast<const int4*>(table + s * S);
        int32_t w[S];
#pragma unroll
        for (int j = 0; j < S / 4; ++j) {
            const int4 v = p[j];
            w[4 * j] = v.x;
            w[4 * j + 1] = v.y;
            w[4 * j + 2] = v.z;
            w[4 * j + 3] = v.w;
        }
        const int32_t row = w[S - 1];
        if (row == -1) break;
        bool eq = true;
#pragma unroll
        for (int k = 0; k < S - 1; ++k)
            if (k < d) eq = eq && w[k] == key[k];
        if (eq && row < best) best = row;
        s = s + 1 == (uint64_t)capacity ? 0 : s + 1;
    }
    out[i] = best == INT32_MAX ? -1 : best;
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t hash_clear(int32_t* table, int64_t capacity, int32_t stride, hipStream_t s) {
    const int64_t n_vec = capacity * stride / 4;
    if (n_vec <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int64_t>((n_vec + 255) / 256, 8192);
    hipLaunchKernelGGL(ht_clear, dim3(blocks), dim3(256), 0, s, reinterpret_cast<int4*>(table), n_vec);
    return hipGetLastError();
}

hipError_t hash_insert(int32_t* table, int64_t capacity, int32_t stride, const int32_t* coords, int64_t n, int32_t d,
                       hipStream_t s) {
    if (n <= 0 || capacity <= 0) return hipSuccess;
    switch (stride) {
        case 4: hipLaunchKernelGGL(ht_insert<4>, dim3(blocks_of(n)), dim3(256), 0, s, table, capacity, coords, n, d); break;
        case 8: hipLaunchKernelGGL(ht_insert<8>, dim3(blocks_of(n)), dim3(256), 0, s, table, capacity, coords, n, d); break;
        case 16: hipLaunchKernelGGL(ht_insert<16>, dim3(blocks_of(n)), dim3(256), 0, s, table, capacity, coords, n, d); break;
        case 32: hipLaunchKernelGGL(ht_insert<32>, dim3(blocks_of(n)), dim3(256), 0, s, table, capacity, coords, n, d); break;
        case 64: hipLaunchKernelGGL(ht_insert<64>, dim3(blocks_of(n)), dim3(256), 0, s, table, capacity, coords, n, d); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t hash_search(const int32_t* table, int64_t capacity, int32_t stride, const int32_t* queries, int64_t m,
                       int32_t d, int32_t* out, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    switch (stride) {
        case 4: hipLaunchKernelGGL(ht_search<4>, dim3(blocks_of(m)), dim3(256), 0, s, table, capacity, queries, m, d, out); break;
        case 8: hipLaunchKernelGGL(ht_search<8>, dim3(blocks_of(m)), dim3(256), 0, s, table, capacity, queries, m, d, out); break;
        case 16: hipLaunchKernelGGL(ht_search<16>, dim3(blocks_of(m)), dim3(256), 0, s, table, capacity, queries, m, d, out); break;
        case 32: hipLaunchKernelGGL(ht_search<32>, dim3(blocks_of(m)), dim3(256), 0, s, table, capacity, queries, m, d, out); break;
        case 64: hipLaunchKernelGGL(ht_search<64>, dim3(blocks_of(m)), dim3(256), 0, s, table, capacity, queries, m, d, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cubvh_amd
