// grad.hip -- backward of the unsigned distance (no counterpart in the reference).
//
// d = |p - c| with c the closest point on the nearest face.  By the envelope
// theorem the derivative of the minimum over the face equals the derivative at the
// minimiser, so with n = (p - c) / d:
//   dd/dp   = n
//   dd/dv_k = -w_k n   for the face's vertex k with barycentric weight w_k
// c is rebuilt from the stored face id and barycentrics, against the current
// vertex values.  d = 0 gives the zero subgradient.  Vertex gradients are
// accumulated with fp32 atomics (order-dependent in the last bits).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "query.hpp"

namespace cubvh_amd {
namespace dev {

__global__ __launch_bounds__(256) void udf_backward_kernel(int64_t n, const float* __restrict__ g,
                                                           const float* __restrict__ p,
                                                           const int64_t* __restrict__ face_id,
                                                           const float* __restrict__ uvw,
                                                           const float* __restrict__ V,
                                                           const int32_t* __restrict__ F, int64_t n_vertices,
                                                           int64_t n_faces, float* __restrict__ gp,
                                                           float* __restrict__ gv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t f = face_id[i];
    // out-of-range face or vertex indices contribute nothing
    bool ok = f >= 0 && f < n_faces;
    const int32_t ia = ok ? F[3 * f] : 0, ib = ok ? F[3 * f + 1] : 0, ic = ok ? F[3 * f + 2] : 0;
    ok = ok && ia >= 0 && ia < n_vertices && ib >= 0 && ib < n_vertices && ic >= 0 && ic < n_vertices;
    if (!ok) {
        if (gp) {
            gp[3 * i] = 0.0f;
            gp[3 * i + 1] = 0.0f;
            gp[3 * i + 2] = 0.0f;
        }
        return;
    }
    const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
    float c[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = __builtin_fmaf(w, V[3 * ic + k], __builtin_fmaf(v, V[3 * ib + k], u * V[3 * ia + k]));
        d[k] = p[3 * i + k] - c[k];
    }
    const float len = sqrtf(__builtin_fmaf(d[2], d[2], __builtin_fmaf(d[1], d[1], d[0] * d[0])));
    const float s = len > 0.0f ? g[i] / len : 0.0f;
    const float nx = d[0] * s, ny = d[1] * s, nz = d[2] * s;  // g * n
    if (gp) {
        gp[3 * i] = nx;
        gp[3 * i + 1] = ny;
        gp[3 * i + 2] = nz;
    }
    if (gv && s != 0.0f) {
        atomicAdd(gv + 3 * ia, -u * nx);
        atomicAdd(gv + 3 * ia + 1, -u * ny);
        atomicAdd(gv + 3 * ia + 2, -u * nz);
        atomicAdd(gv + 3 * ib, -v * nx);
        atomicAdd(gv + 3 * ib + 1, -v * ny);
        atomicAdd(gv + 3 * ib + 2, -v * nz);
        atomicAdd(gv + 3 * ic, -w * nx);
        atomicAdd(gv + 3 * ic + 1, -w * ny);
        atomicAdd(gv + 3 * ic + 2, -w * nz);
    }
}

}  // namespace dev

hipError_t launch_udf_backward(int64_t n, const float* grad_dist, const float* positions, const int64_t* face_id,
                               const float* uvw, const float* vertices, int64_t n_vertices, const int32_t* faces,
                               int64_t n_faces, float* grad_points, float* grad_vertices, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const unsigned grid = (unsigned)((n + block - 1) / block);
    hipLaunchKernelGGL(dev::udf_backward_kernel, dim3(grid), dim3(block), 0, stream, n, grad_dist, positions, face_id,
                       uvw, vertices, faces, n_vertices, n_faces, grad_points, grad_vertices);
    return hipGetLastError();
}

}  // namespace cubvh_amd
