// netdev.hpp -- device helpers shared by the network kernels (net.hip) and the flexible-CG
// kernels that feed the network (learned.hip): BF16 conversion.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace caffe {

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

}  // namespace caffe
