// ref_hashtable.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own CPU hash table (cubvh::HashTableIntCPU,
// include/cpu/hashtable.h), compiled from the reference's header where it lies by
// `make -C oracle ref` into oracle/_ref/.  Used by the tests to pin the oracle's
// restatement slot for slot, and by the benchmark as the reference CPU baseline of the
// hash table.  Nothing in the product loads it.
#include <cstdint>
#include <cstring>

#include "cpu/hashtable.h"

extern "C" {

// Builds the reference table over coords [n, d] (capacity <= 0: the reference's
// build(), capacity max(16, 2n); else resize + prepare + insert), copies the table
// (2 * capacity int32) into table_out when non-null, and searches m queries.
// Returns the capacity used.
int ref_hash_build_search(const int32_t* coords, int n, int d, int capacity, const int32_t* queries, int m,
                          int32_t* out, int32_t* table_out) {
    cubvh::HashTableIntCPU ht;
    ht.set_num_dims(d);
    if (capacity <= 0) {
        ht.build(coords, n);
    } else {
        ht.resize(capacity);
        ht.prepare();
        ht.insert(coords, n);
    }
    if (table_out) std::memcpy(table_out, ht.table_kvs.data(), sizeof(int32_t) * ht.table_kvs.size());
    if (m > 0) ht.search(queries, m, out);
    return ht.capacity;
}

}  // extern "C"
