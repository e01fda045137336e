// Context lifecycle: one device, one stream (caller's or owned); the per-device pool of block
// memory.
#include "ctx.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace cfdx {
namespace {
// Block allocations returned by cfdx_block_destroy are kept for reuse instead of freed:
// hipFree waits for the whole device, and a serving loop creates and retires blocks of a few
// recurring sizes while other work runs. Bounded by a quarter of the device memory.
struct Pool {
    std::multimap<size_t, void *> free;   // size -> allocation
    size_t cached = 0, cap = 0;
};
std::mutex pool_mutex;
std::vector<Pool> pools;

Pool &pool_of(int device)
{
    if ((int)pools.size() <= device) pools.resize(device + 1);
    Pool &p = pools[device];
    if (!p.cap) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) == hipSuccess) p.cap = t / 4;
    }
    return p;
}
}  // namespace

void *pool_take(int device, size_t bytes, size_t *got)
{
    {
        std::lock_guard<std::mutex> lock(pool_mutex);
        Pool &p = pool_of(device);
        auto it = p.free.lower_bound(bytes);
        if (it != p.free.end() && it->first <= bytes + bytes / 4) {
            void *ptr = it->second;
            *got = it->first;
            p.cached -= it->first;
            p.free.erase(it);
            return ptr;
        }
    }
    void *ptr = nullptr;
    if (hipMalloc(&ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(device);
        if (hipMalloc(&ptr, bytes) != hipSuccess) return nullptr;
    }
    *got = bytes;
    return ptr;
}

void pool_give(int device, void *ptr, size_t bytes)
{
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lock(pool_mutex);
        Pool &p = pool_of(device);
        if (p.cached + bytes <= p.cap) {
            p.free.emplace(bytes, ptr);
            p.cached += bytes;
            return;
        }
    }
    (void)hipFree(ptr);
}

void pool_trim(int device)
{
    std::multimap<size_t, void *> out;
    {
        std::lock_guard<std::mutex> lock(pool_mutex);
        Pool &p = pool_of(device);
        out.swap(p.free);
        p.cached = 0;
    }
    for (auto &kv : out) (void)hipFree(kv.second);
}

size_t pool_cached(int device)
{
    std::lock_guard<std::mutex> lock(pool_mutex);
    return pool_of(device).cached;
}
}  // namespace cfdx

extern "C" int cfdx_pool_trim(cfdx_ctx *ctx)
{
    CFDX_TRY(cfdx::enter(ctx));
    cfdx::pool_trim(ctx->device);
    return CFDX_OK;
}

extern "C" int cfdx_ctx_create(int device, void *stream, cfdx_ctx **out)
{
    if (!out) return cfdx::fail(CFDX_E_INVALID, "ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    CFDX_HIP(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return cfdx::fail(CFDX_E_INVALID, "ctx_create: device %d of %d", device, ndev);
    CFDX_HIP(hipSetDevice(device), "hipSetDevice");
    // the device address of the no-allocation RHS's fixed array; a failure here is reported by
    // that entry point (which resolves again), not by every context
    (void)cfdx::scratch_resolve(device);
    cfdx_ctx *c = new cfdx_ctx;
    c->device = device;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return cfdx::hip_fail(e, "hipStreamCreateWithFlags");
        }
        c->owns_stream = true;
    }
    for (auto &ev : c->ev) {
        hipError_t e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            cfdx::ctx_release(c);
            return cfdx::hip_fail(e, "hipEventCreate");
        }
    }
    {
        hipError_t e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess) {
            cfdx::ctx_release(c);
            return cfdx::hip_fail(e, "ctx_create second stream");
        }
    }
    if (const char *v = std::getenv("CFDX_OVERLAP")) c->overlap = std::atoi(v) != 0;
    *out = c;
    return CFDX_OK;
}

namespace cfdx {
void ctx_retain(cfdx_ctx *ctx)
{
    if (ctx) ctx->refs.fetch_add(1);
}

int ctx_release(cfdx_ctx *ctx)
{
    if (!ctx || ctx->refs.fetch_sub(1) > 1) return CFDX_OK;
    int rc = CFDX_OK;
    if (hipSetDevice(ctx->device) == hipSuccess && ctx->owns_stream) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "hipStreamSynchronize");
        (void)hipStreamDestroy(ctx->stream);
    }
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->aux) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    delete ctx;
    return rc;
}
}  // namespace cfdx

extern "C" int cfdx_ctx_destroy(cfdx_ctx *ctx) { return cfdx::ctx_release(ctx); }

extern "C" int cfdx_ctx_set_stream(cfdx_ctx *ctx, void *stream)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (ctx->owns_stream) {
        CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        (void)hipStreamDestroy(ctx->stream);
        ctx->owns_stream = false;
    }
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        CFDX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        ctx->owns_stream = true;
    }
    ctx->gen.fetch_add(1);
    return CFDX_OK;
}

extern "C" int cfdx_ctx_set_overlap(cfdx_ctx *ctx, int on)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (ctx->overlap != (on != 0)) {
        ctx->overlap = on != 0;
        ctx->gen.fetch_add(1);   // captured batch steps hold the old fork / join structure
    }
    return CFDX_OK;
}

extern "C" void *cfdx_ctx_stream(cfdx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int cfdx_sync(cfdx_ctx *ctx)
{
    CFDX_TRY(cfdx::enter(ctx));
    CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return CFDX_OK;
}

extern "C" int cfdx_device_count(int *count)
{
    if (!count) return cfdx::fail(CFDX_E_INVALID, "device_count: null");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return cfdx::hip_fail(e, "hipGetDeviceCount");
    }
    return CFDX_OK;
}

extern "C" int cfdx_mem_info(cfdx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return cfdx::fail(CFDX_E_INVALID, "mem_info: null argument");
    CFDX_TRY(cfdx::enter(ctx));
    size_t f = 0, t = 0;
    CFDX_HIP(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    *free_bytes = (int64_t)(f + cfdx::pool_cached(ctx->device));   // pooled block memory is available
    *total_bytes = (int64_t)t;
    return CFDX_OK;
}

extern "C" int cfdx_device_alloc(cfdx_ctx *ctx, size_t bytes, void **dptr)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!dptr) return cfdx::fail(CFDX_E_INVALID, "device_alloc: null out");
    *dptr = nullptr;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) return cfdx::fail(CFDX_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return CFDX_OK;
}

extern "C" int cfdx_device_free(cfdx_ctx *ctx, void *dptr)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!dptr) return CFDX_OK;
    CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    CFDX_HIP(hipFree(dptr), "hipFree");
    return CFDX_OK;
}

extern "C" int cfdx_copy_to_device(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CFDX_TRY(cfdx::enter(ctx));
    if ((!dst || !src) && bytes) return cfdx::fail(CFDX_E_INVALID, "copy_to_device: null pointer");
    if (!bytes) return CFDX_OK;
    CFDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
    CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return CFDX_OK;
}

extern "C" int cfdx_copy_to_host(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CFDX_TRY(cfdx::enter(ctx));
    if ((!dst || !src) && bytes) return cfdx::fail(CFDX_E_INVALID, "copy_to_host: null pointer");
    if (!bytes) return CFDX_OK;
    CFDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
    CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return CFDX_OK;
}

extern "C" int cfdx_event_record(cfdx_ctx *ctx, int slot)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (slot < 0 || slot >= 8) return cfdx::fail(CFDX_E_INVALID, "event slot %d not in 0..7", slot);
    CFDX_HIP(hipEventRecord(ctx->ev[slot], ctx->stream), "hipEventRecord");
    return CFDX_OK;
}

extern "C" int cfdx_event_elapsed(cfdx_ctx *ctx, int a, int b, double *ms)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (a < 0 || a >= 8 || b < 0 || b >= 8 || !ms) return cfdx::fail(CFDX_E_INVALID, "event_elapsed: bad slot");
    CFDX_HIP(hipEventSynchronize(ctx->ev[b]), "hipEventSynchronize");
    float f = 0;
    CFDX_HIP(hipEventElapsedTime(&f, ctx->ev[a], ctx->ev[b]), "hipEventElapsedTime");
    *ms = f;
    return CFDX_OK;
}
