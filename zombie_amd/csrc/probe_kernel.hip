// probe_kernel.hip -- read-bandwidth probe used to measure the cache and HBM ceilings that the
// walk kernels are compared against (a streaming float4 read of a device buffer, repeated).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace zg {

__global__ void __launch_bounds__(256) readProbeKernel(const float4* __restrict__ buf, int64_t n4, int reps,
                                                       float* __restrict__ sink) {
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int r = 0; r < reps; r++) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
            float4 v = buf[i];
            acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
        }
    }
    // keeps the loads alive; the sum is never 1e30 for the probe's finite data
    float s = acc.x + acc.y + acc.z + acc.w;
    if (s == 1e30f) sink[0] = s;
}

hipError_t launchReadProbe(const void* buf, int64_t bytes, int reps, int blocks, float* sink, hipStream_t s) {
    int64_t n4 = bytes / 16;
    hipLaunchKernelGGL(readProbeKernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const float4*>(buf), n4, reps,
                       sink);
    return hipGetLastError();
}

}  // namespace zg
