// order.hip -- Morton ordering of point queries.
//
// Consecutive lanes of a wave should hold queries that are close in space, so that
// they walk nearly the same part of the tree (the packet traversal in query.hip
// then visits close to one lane's worth of nodes per wave, and every node or face
// record is fetched once per wave).  The queries are not moved: a permutation is
// computed on the device and the kernels read query order[i] and write the outputs
// back to that position.
//
//   1. bounds : min/max of the query coordinates (order-preserving int atomics)
//   2. keys   : 30-bit Morton code (10 bits per axis) of each query in those bounds
//   3. sort   : rocPRIM radix sort (hipCUB interface) of (key, index) pairs
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <mutex>

#include "query.hpp"
#include "radix.hpp"

namespace cubvh_amd {
namespace dev {

// float <-> int mapping that preserves order under signed integer comparison
__device__ __forceinline__ int32_t ordered(float f) {
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float unordered(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void bounds_init(int32_t* b) {
    if (threadIdx.x < 3) b[threadIdx.x] = 0x7fffffff;
    else if (threadIdx.x < 6) b[threadIdx.x] = (int32_t)0x80000000;
}

// Bounds of every `stride`-th query (the ordering only needs approximate bounds:
// queries outside them are clamped to the boundary cells).  Waves reduce with
// shuffles, blocks through LDS, and one lane per block updates the six bounds: few
// blocks, so that the same-address atomics stay few.
__global__ __launch_bounds__(256) void bounds_kernel(int64_t n, int64_t stride, const float* __restrict__ p,
                                                     int32_t* b) {
    __shared__ float sh[6][4];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll 4
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * stride; i < n;
         i += (int64_t)gridDim.x * blockDim.x * stride) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float x = p[3 * i + k];
            lo[k] = fminf(lo[k], x);
            hi[k] = fmaxf(hi[k], x);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off, 64));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off, 64));
        }
    }
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            sh[k][w] = lo[k];
            sh[3 + k][w] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            for (int j = 1; j < 4; ++j) {
                lo[k] = fminf(lo[k], sh[k][j]);
                hi[k] = fmaxf(hi[k], sh[3 + k][j]);
            }
            if (lo[k] <= hi[k]) {  // skips blocks that saw no (finite) query
                atomicMin(b + k, ordered(lo[k]));
                atomicMax(b + 3 + k, ordered(hi[k]));
            }
        }
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ __launch_bounds__(256) void morton_kernel(int64_t n, const float* __restrict__ p, const int32_t* __restrict__ b,
                                                     uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = unordered(b[k]), hi = unordered(b[3 + k]);
        float ext = hi - lo;
        float s = ext > 0.0f ? 1023.0f / ext : 0.0f;
        float t = (p[3 * i + k] - lo) * s;
        t = fminf(fmaxf(t, 0.0f), 1023.0f);  // NaN coordinates map to 0
        q[k] = (uint32_t)t;
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    idx[i] = (uint32_t)i;
}

// Two 9-bit values interleaved (x in the even bits).
__device__ __forceinline__ uint32_t spread9x2(uint32_t x) {
    x &= 0x1ffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Ray sort key: the Morton code of the origin's cell in the origins' bounds (4 bits per
// axis) above the Morton code of the direction's cell on the octahedral map of the unit
// sphere (9 bits per axis).  Rays of one origin cell and similar directions -- e.g. a
// small square of pixels of one camera -- end up next to each other.
__global__ __launch_bounds__(256) void ray_key_kernel(int64_t n, const float* __restrict__ o,
                                                      const float* __restrict__ d, const int32_t* __restrict__ b,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = unordered(b[k]), hi = unordered(b[3 + k]);
        float ext = hi - lo;
        float sc = ext > 0.0f ? 15.0f / ext : 0.0f;
        float t = fminf(fmaxf((o[3 * i + k] - lo) * sc, 0.0f), 15.0f);  // NaN maps to 0
        c[k] = (uint32_t)t;
    }
    const uint32_t ko = (spread10(c[0]) << 2) | (spread10(c[1]) << 1) | spread10(c[2]);
    const float dx = d[3 * i], dy = d[3 * i + 1], dz = d[3 * i + 2];
    const float sum = fabsf(dx) + fabsf(dy) + fabsf(dz);
    float px = 0.0f, py = 0.0f;
    if (sum > 0.0f && sum < INFINITY) {
        px = dx / sum;
        py = dy / sum;
        if (dz < 0.0f) {
            const float tx = (1.0f - fabsf(py)) * copysignf(1.0f, px);
            const float ty = (1.0f - fabsf(px)) * copysignf(1.0f, py);
            px = tx;
            py = ty;
        }
    }
    const uint32_t u = (uint32_t)fminf(fmaxf((px + 1.0f) * 256.0f, 0.0f), 511.0f);
    const uint32_t v = (uint32_t)fminf(fmaxf((py + 1.0f) * 256.0f, 0.0f), 511.0f);
    keys[i] = (ko << 18) | (spread9x2(u) << 1) | spread9x2(v);
    idx[i] = (uint32_t)i;
}

}  // namespace dev

hipError_t launch_bounds(int64_t n, const float* positions, int32_t* bounds, hipStream_t s) {
    hipLaunchKernelGGL(dev::bounds_init, dim3(1), dim3(64), 0, s, bounds);
    if (n > 0) {
        const int64_t stride = std::max<int64_t>(1, n >> 20);
        const int64_t m = (n + stride - 1) / stride;
        const unsigned nb = (unsigned)std::min<int64_t>((m + 1023) / 1024, 256);  // >= 4 samples per thread
        hipLaunchKernelGGL(dev::bounds_kernel, dim3(nb), dim3(256), 0, s, n, stride, positions, bounds);
    }
    return hipGetLastError();
}

// Bytes of device workspace morton_order needs for n queries.
size_t morton_workspace_bytes(int64_t n) {
    size_t tmp = 0;
    radix_sort_pairs_u32(nullptr, tmp, nullptr, nullptr, nullptr, nullptr, n, 30, 0);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(64) + 3 * al(4 * (size_t)n) + al(tmp);
}

// Writes the permutation of n rays sorted by ray_key_kernel's key into order_out, using
// `ws` (morton_workspace_bytes(n) bytes: the same layout as the point ordering).
// n must be < 2^31.
hipError_t ray_order(int64_t n, const float* rays_o, const float* rays_d, void* ws, uint32_t* order_out, hipStream_t s) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char* w = (char*)ws;
    int32_t* bounds = (int32_t*)w;
    w += al(64);
    uint32_t* k0 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    uint32_t* k1 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    uint32_t* v0 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    size_t tmp = 0;
    hipError_t e = radix_sort_pairs_u32(nullptr, tmp, k0, k1, v0, order_out, n, 30, s);
    if (e != hipSuccess) return e;
    e = launch_bounds(n, rays_o, bounds, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::ray_key_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, rays_o, rays_d, bounds,
                       k0, v0);
    e = radix_sort_pairs_u32((void*)w, tmp, k0, k1, v0, order_out, n, 30, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// Scratch of the allocation-free distance entry point (cubvh_unsigned_distance_noalloc):
// the Morton permutation of up to kNoAllocMaxQueries queries and the ordering
// workspace, in a static device array.  Every device that runs the library holds its
// own copy until the process ends.  Its host-side name is not a device address: the
// launches take the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, which must be the current device.  Looked
// up once per device and kept.
hipError_t noalloc_scratch(int device, void** ptr) {
    if (device < 0 || device >= kNoAllocDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

// Writes the Morton permutation of the n queries into order_out (n uint32) using
// `ws` (morton_workspace_bytes(n) bytes).  n must be < 2^31.
hipError_t morton_order(int64_t n, const float* positions, void* ws, uint32_t* order_out, hipStream_t s) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char* w = (char*)ws;
    int32_t* bounds = (int32_t*)w;
    w += al(64);
    uint32_t* k0 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    uint32_t* k1 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    uint32_t* v0 = (uint32_t*)w;
    w += al(4 * (size_t)n);
    size_t tmp = 0;
    hipError_t e = radix_sort_pairs_u32(nullptr, tmp, k0, k1, v0, order_out, n, 30, s);
    if (e != hipSuccess) return e;
    e = launch_bounds(n, positions, bounds, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::morton_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, positions, bounds, k0,
                       v0);
    e = radix_sort_pairs_u32((void*)w, tmp, k0, k1, v0, order_out, n, 30, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace cubvh_amd
