// Context (device + stream) and HIP error plumbing for the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>

#include "common.hpp"

// A context is shared by the blocks and batches made on it: each holds a reference, and the
// context is freed when cfdx_ctx_destroy and the last holder have released it.
struct cfdx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    hipEvent_t ev[8] = {};
    // a second stream of the device for work of a stage that may overlap the main stream's
    // (forked from and joined back to the main stream by the two events), and the device's CUs
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int n_cu = 0;
    bool overlap = true;   // mixed stages (and FP64 ones in graphs) overlap the viscous kernel with the face sweeps
    // bumped by every change that a captured batch graph depends on (overlap switch, stream): a batch
    // re-captures its step when the generation it captured under is not the current one
    std::atomic<uint64_t> gen{1};
    std::atomic<int> refs{1};
};

namespace cfdx {
void ctx_retain(cfdx_ctx *ctx);
// Resolves the device address of the fixed RHS scratch array on `device`, which must be the
// current device (once per device; later calls return at once).
int scratch_resolve(int device);
int ctx_release(cfdx_ctx *ctx);

// Device memory for blocks: a cached allocation of at least `bytes` (at most a quarter more;
// *got = its size) or a new one; nullptr when the device has none. pool_give keeps it for
// reuse (up to a quarter of the device memory) or frees it; pool_trim frees what is kept.
void *pool_take(int device, size_t bytes, size_t *got);
void pool_give(int device, void *ptr, size_t bytes);
void pool_trim(int device);
size_t pool_cached(int device);
}  // namespace cfdx

namespace cfdx {

inline int hip_fail(hipError_t e, const char *what)
{
    return fail(CFDX_E_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

// Makes ctx's device current; returns 0 or a CFDX error.
inline int enter(cfdx_ctx *ctx)
{
    if (!ctx) return fail(CFDX_E_INVALID, "null context");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return CFDX_OK;
}

// Checks the launch that was just issued.
inline int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return CFDX_OK;
}

inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace cfdx

#define CFDX_TRY(expr)                 \
    do {                               \
        int _rc = (expr);              \
        if (_rc != CFDX_OK) return _rc; \
    } while (0)

#define CFDX_HIP(expr, what)                                  \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return cfdx::hip_fail(_e, what); \
    } while (0)
