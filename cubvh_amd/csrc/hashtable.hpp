// hashtable.hpp -- launch interface of the sparse-voxel hash table (hashtable.hip).
//
// A static open-addressed table of int32 keys of D words (1 <= D <= 63) with linear
// probing.  Each slot is `stride` int32 words, stride = the power of two above D: the
// key in words 0..D-1 and the row index of the key in word stride-1 (-1 = empty), so
// a probe reads one aligned 16..256-byte record and needs no second load.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubvh_amd {

constexpr int32_t kHashMaxDims = 63;

// Words per slot for keys of d words.
inline int32_t hash_slot_words(int32_t d) {
    int32_t s = 4;
    while (s < d + 1) s <<= 1;
    return s;
}

// Marks every slot empty.
hipError_t hash_clear(int32_t* table, int64_t capacity, int32_t stride, hipStream_t s);
// Inserts rows 0..n-1 of coords [n, d]; a key finds no slot only when the table is full.
hipError_t hash_insert(int32_t* table, int64_t capacity, int32_t stride, const int32_t* coords, int64_t n, int32_t d,
                       hipStream_t s);
// out[i] = the lowest row index stored with key queries[i], or -1.
hipError_t hash_search(const int32_t* table, int64_t capacity, int32_t stride, const int32_t* queries, int64_t m,
                       int32_t d, int32_t* out, hipStream_t s);

}  // namespace cubvh_amd
