// ref_pcg32.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own random number generator (cubvh::pcg32,
// include/gpu/pcg32.h), compiled from the reference's header where it lies by
// `make -C oracle ref` into oracle/_ref/.  The header is plain C++ apart from its
// __host__ / __device__ qualifiers, which are defined empty here (host code only).
// Used by the tests to pin the oracle's PCG32 (the raystab offsets) against the
// reference's generator.  Nothing in the product loads it.
#include <cstdint>

#define __host__
#define __device__
#include "gpu/pcg32.h"

extern "C" {

// n outputs of next_uint() of a default-constructed generator advanced by `advance`.
void ref_pcg32_uints(int64_t advance, int64_t n, uint32_t* out) {
    cubvh::pcg32 r;
    r.advance(advance);
    for (int64_t i = 0; i < n; ++i) out[i] = r.next_uint();
}

// The raystab offsets of queries 0..n-1 (src/bvh.cu:663-668): a default generator
// advanced by 2i (the product i * 2 in uint32 arithmetic), then two next_float().
void ref_pcg32_offsets(int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        cubvh::pcg32 r;
        r.advance((int64_t)(uint32_t)((uint32_t)i * 2u));
        out[2 * i] = r.next_float();
        out[2 * i + 1] = r.next_float();
    }
}

// Internal state after `advance` steps, for checks of the jump-ahead alone.
uint64_t ref_pcg32_state(int64_t advance) {
    cubvh::pcg32 r;
    r.advance(advance);
    return r.state;
}

}  // extern "C"
