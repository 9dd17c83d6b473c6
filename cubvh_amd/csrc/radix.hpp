// radix.hpp -- stable radix sort of (uint32 key, uint32 value) pairs for the query and
// face orderings.
//
// rocPRIM's default dispatch sorts 4-byte keys of up to 2^20 items with a block sort
// and merge passes (about two dozen launches at 2^20); these orderings sort keys of at
// most 30 significant bits, which its onesweep radix sort does in ceil(end_bit / 8)
// passes.  A merge-sort limit of 0 keeps every call above one block on onesweep.  Both
// algorithms are stable, so the permutation is the same either way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/rocprim.hpp>

namespace cubvh_amd {

using RadixSortConfig =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>;

// Sorts n pairs by the key bits [0, end_bit); with tmp == nullptr only sets bytes.
inline hipError_t radix_sort_pairs_u32(void* tmp, size_t& bytes, const uint32_t* keys_in, uint32_t* keys_out,
                                       const uint32_t* vals_in, uint32_t* vals_out, int64_t n, int end_bit,
                                       hipStream_t stream) {
    return rocprim::radix_sort_pairs<RadixSortConfig>(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u,
                                                      (unsigned)end_bit, stream);
}

}  // namespace cubvh_amd
