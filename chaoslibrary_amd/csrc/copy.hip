// copy.hip -- a plain streaming copy (16 B per lane per iteration).  Used as
// the known-size reference when HBM counters (FETCH_SIZE / WRITE_SIZE) are
// scaled to bytes, and as a device memcpy for callers without a framework.
#include "runtime.hpp"

namespace chaos {

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_device_copy(void* dst, const void* src, size_t bytes, int device)
{
    return guarded("chaos_device_copy", [&] {
        CHAOS_REQUIRE(bytes % 16 == 0, "bytes must be a multiple of 16");
        CHAOS_REQUIRE(bytes == 0 || (dst && src), "NULL pointer");
        chaos_options o = resolve_options(nullptr);
        o.device = device;
        CallCtx ctx(o);
        const size_t n = bytes / 16;
        if (n == 0) return;
        const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx.stream(), static_cast<const uint4*>(src),
                           static_cast<uint4*>(dst), n);
        CHAOS_HIP(hipGetLastError());
        ctx.sync();
    });
}
