// probe.hip -- the VALU issue ceiling of the device, measured: every SIMD runs 8 waves of
// 16 independent fp32 fma chains.  The walk kernels are bounded by VALU issue rather
// than by HBM, and their issue rate (SQ_INSTS_VALU per second) is compared with this
// kernel's.  Each loop iteration is 16 fma instructions per wave plus scalar loop
// control, in one of two forms whose issue rates differ by about 2x on gfx950:
//   form 0: operands in VGPRs and literals (v_fmamk_f32), the fastest fp32 form;
//   form 1: one operand in an SGPR (v_fma_f32 v, s, v, v), the form of every
//           operation that combines a wave-uniform node or face record (scalar
//           loads) with a lane's own point.
#include <hip/hip_runtime.h>

#include "query.hpp"

namespace cubvh_amd {
namespace dev {

template <int FORM>
__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, int32_t iters, float mul) {
    float a[kValuProbeChains];
#pragma unroll
    for (int k = 0; k < kValuProbeChains; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
#pragma unroll 1
    for (int32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kValuProbeChains; ++k) a[k] = __builtin_fmaf(a[k], FORM == 0 ? 0.999f : mul, 0.001f);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kValuProbeChains; ++k) s += a[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace dev

hipError_t launch_valu_probe(int form, int blocks, int32_t iters, float* out, hipStream_t s) {
    if (form == 0)
        hipLaunchKernelGGL(dev::valu_probe_kernel<0>, dim3((unsigned)blocks), dim3(256), 0, s, out, iters, 0.999f);
    else if (form == 1)
        hipLaunchKernelGGL(dev::valu_probe_kernel<1>, dim3((unsigned)blocks), dim3(256), 0, s, out, iters, 0.999f);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace cubvh_amd
