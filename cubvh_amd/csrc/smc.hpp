// smc.hpp -- sparse marching cubes (smc.hip): the surface through a set of cells of an
// integer lattice, each given by its lowest lattice point and its 8 corner values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace cubvh_amd {

// Scratch of one extraction between its two phases (device memory on the stream).
struct SmcState;

struct SmcStatus {
    int code = 0;  // 0 ok, -1 invalid argument, -2 HIP failure
    std::string message;
};

// Largest cell count and lattice span per axis one extraction accepts.
constexpr int64_t kSmcMaxCells = (int64_t)1 << 28;
constexpr int64_t kSmcMaxSpan = ((int64_t)1 << 20) - 2;

// Phase 1: classify the cells, key and order the edge vertices, count the distinct
// ones.  coords int32 [n, 3] and corners fp32 [n, 8] stay in use until smc_write.
// Synchronizes the stream two or three times (bounds, totals, distinct vertices).
SmcStatus smc_extract(const int32_t* coords, const float* corners, int64_t n, float iso, bool consistent,
                      hipStream_t s, SmcState** out, int64_t* n_verts, int64_t* n_tris);
// Phase 2: vertices fp32 [n_verts, 3] and triangles int32 [n_tris, 3] into the caller's
// buffers; releases the scratch (stream-ordered).
SmcStatus smc_write(SmcState* st, float* verts, int32_t* tris, hipStream_t s);
// Releases the scratch of an extraction that is not written (stream-ordered).
void smc_release(SmcState* st, hipStream_t s);

}  // namespace cubvh_amd
