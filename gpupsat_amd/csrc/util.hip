// util.hip -- a plain streaming copy, used to measure the practical HBM peak and to
// scale the FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic.
#include "common.hpp"
#include "gpsat_internal.hpp"

namespace gpsat {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st)
{
    const int64_t n16 = bytes / 16;
    if (n16 <= 0) return hipSuccess;
    int blocks = (int)std::min<int64_t>((n16 + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, st, (const u32x4*)src, (u32x4*)dst, n16);
    return hipGetLastError();
}

}  // namespace gpsat
