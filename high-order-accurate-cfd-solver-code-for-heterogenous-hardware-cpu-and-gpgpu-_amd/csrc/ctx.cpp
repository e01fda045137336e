// Context lifecycle: one device, one stream (caller's or owned).
#include "ctx.hpp"

extern "C" int cfdx_ctx_create(int device, void *stream, cfdx_ctx **out)
{
    if (!out) return cfdx::fail(CFDX_E_INVALID, "ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    CFDX_HIP(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return cfdx::fail(CFDX_E_INVALID, "ctx_create: device %d of %d", device, ndev);
    CFDX_HIP(hipSetDevice(device), "hipSetDevice");
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
    *out = c;
    return CFDX_OK;
}

extern "C" int cfdx_ctx_destroy(cfdx_ctx *ctx)
{
    if (!ctx) return CFDX_OK;
    int rc = CFDX_OK;
    if (hipSetDevice(ctx->device) == hipSuccess && ctx->owns_stream) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "hipStreamSynchronize");
        hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return rc;
}

extern "C" int cfdx_ctx_set_stream(cfdx_ctx *ctx, void *stream)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (ctx->owns_stream) {
        CFDX_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        hipStreamDestroy(ctx->stream);
        ctx->owns_stream = false;
    }
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        CFDX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        ctx->owns_stream = true;
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
