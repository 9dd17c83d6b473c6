// ref_smc.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own CPU sparse marching cubes
// (cubvh::cpu::sparse_marching_cubes and its case tables, include/cpu/spmc.h), compiled
// from the reference's header where it lies by `make -C oracle ref` into oracle/_ref/.
// The tests use it to pin the oracle's restatement and the case table; nothing in the
// product loads it.
#include <cstdint>
#include <cstring>

#include "cpu/spmc.h"

extern "C" {

// The reference's 256 edge masks and 256 x 16 triangle lists.
void ref_mc_tables(int32_t* edge_masks, int32_t* triangles) {
    std::memcpy(edge_masks, cubvh::cpu::edgeTable, sizeof(int32_t) * 256);
    std::memcpy(triangles, cubvh::cpu::triTable, sizeof(int32_t) * 256 * 16);
}

// Runs the reference CPU extraction; copies at most max_v vertices and max_t triangles
// and returns their true counts.
void ref_smc(const int32_t* coords, const float* corners, int n, float iso, int consistent, float* verts,
             int64_t max_v, int32_t* tris, int64_t max_t, int64_t* n_verts, int64_t* n_tris) {
    auto mesh = cubvh::cpu::sparse_marching_cubes(coords, corners, n, iso, consistent != 0);
    *n_verts = (int64_t)mesh.first.size();
    *n_tris = (int64_t)mesh.second.size();
    const int64_t nv = *n_verts < max_v ? *n_verts : max_v, nt = *n_tris < max_t ? *n_tris : max_t;
    std::memcpy(verts, mesh.first.data(), sizeof(float) * 3 * nv);
    std::memcpy(tris, mesh.second.data(), sizeof(int32_t) * 3 * nt);
}

}  // extern "C"
