// Plain streaming copy kernel: 16-byte loads and stores, grid-stride.
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace {
__global__ void __launch_bounds__(256) k_copy(double2 *__restrict__ dst, const double2 *__restrict__ src, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

extern "C" int cfdx_device_copy(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CFDX_TRY(cfdx::enter(ctx));
    if ((!dst || !src) && bytes) return cfdx::fail(CFDX_E_INVALID, "device_copy: null pointer");
    if (bytes % 16) return cfdx::fail(CFDX_E_INVALID, "device_copy: %zu bytes is not a multiple of 16", bytes);
    const long long n = (long long)(bytes / 16);
    if (n == 0) return CFDX_OK;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 256LL * 64);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, ctx->stream, (double2 *)dst, (const double2 *)src, n);
    return cfdx::check_launch("k_copy");
}
