// Context, errors and buffer management of libncg.
#include "ncg_internal.hpp"

#include <cstdlib>
#include <new>

ncg_status ncg_fail(ncg_ctx *ctx, ncg_status code, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

void *ncg_dbuf(ncg_ctx *ctx, const char *name, size_t bytes)
{
    DevBuf &b = ctx->bufs[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        // The old contents may still be read by queued work on the stream.
        hipStreamSynchronize(ctx->stream);
        hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 256;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        b.ptr = nullptr;
        ncg_fail(ctx, NCG_ERR_NOMEM, "hipMalloc of %zu bytes for '%s' failed", want, name);
        return nullptr;
    }
    b.bytes = want;
    ctx->buf_gen++;
    return b.ptr;
}

extern "C" {

int ncg_abi_version(void) { return NCG_ABI_VERSION; }

ncg_status ncg_create(ncg_ctx **out, int device)
{
    if (!out) return NCG_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NCG_ERR_DEVICE;
    if (device < 0 || device >= ndev) return NCG_ERR_INVALID;
    ncg_ctx *ctx = new (std::nothrow) ncg_ctx();
    if (!ctx) return NCG_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return NCG_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    *out = ctx;
    return NCG_OK;
}

void ncg_destroy(ncg_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ncg_rk4_release(ctx);
    for (auto &kv : ctx->prof)
        for (auto &ev : kv.second.pending) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
    for (auto &kv : ctx->bufs)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *ncg_last_error(const ncg_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

ncg_status ncg_synchronize(ncg_ctx *ctx)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_HIP(hipStreamSynchronize(ctx->stream));
    return NCG_OK;
}

void *ncg_stream(ncg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

ncg_status ncg_profile_enable(ncg_ctx *ctx, int enable)
{
    if (!ctx) return NCG_ERR_INVALID;
    ctx->profiling = enable != 0;
    return NCG_OK;
}

ncg_status ncg_profile_read(ncg_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches,
                            int reset)
{
    if (!ctx || !kernel) return NCG_ERR_INVALID;
    ProfRec &p = ctx->prof[kernel];
    for (auto &ev : p.pending) {
        NCG_HIP(hipEventSynchronize(ev.second));
        float ms = 0.f;
        NCG_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
        p.ms += ms;
        p.launches++;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    p.pending.clear();
    if (total_ms) *total_ms = p.ms;
    if (launches) *launches = p.launches;
    if (reset) {
        p.ms = 0.0;
        p.launches = 0;
    }
    return NCG_OK;
}

}  // extern "C"
