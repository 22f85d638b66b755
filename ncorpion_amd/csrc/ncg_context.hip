// Context, errors and buffer management of libncg.
#include "ncg_internal.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
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
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
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

void *ncg_pbuf(ncg_ctx *ctx, const char *name, size_t bytes)
{
    DevBuf &b = ctx->pbufs[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipHostFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 256;
    if (hipHostMalloc(&b.ptr, want, hipHostMallocDefault) != hipSuccess) {
        b.ptr = nullptr;
        ncg_fail(ctx, NCG_ERR_NOMEM, "hipHostMalloc of %zu bytes for '%s' failed", want, name);
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}

// Reads the tuning table: $NCG_TUNING, else ncg_tuning.txt beside libncg.so.
// Lines: <form> <n_min> <ipl> <nsplit> <theta> <src>; '#' comments.  Forms:
// fast64, mixed (single-system tiled kernel, n_min = bodies) and ens_fast64,
// ens_mixed (ensemble kernel, n_min = mean bodies per system).
static void load_tuning(ncg_ctx *ctx)
{
    std::string path;
    if (const char *e = getenv("NCG_TUNING")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr((void *)&load_tuning, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
                   "/ncg_tuning.txt";
        }
    }
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        char pname[32];
        long nmin;
        int ipl, nsplit, src;
        double theta;
        if (sscanf(line, "%31s %ld %d %d %lf %d", pname, &nmin, &ipl, &nsplit, &theta, &src) != 6)
            continue;
        TuneRow r;
        r.prec = !strcmp(pname, "mixed")        ? NCG_MIXED
                 : !strcmp(pname, "fast64")     ? NCG_F64_FAST
                 : !strcmp(pname, "ens_mixed")  ? 100 + NCG_MIXED
                 : !strcmp(pname, "ens_fast64") ? 100 + NCG_F64_FAST
                                                : -1;
        if (r.prec < 0 || (ipl != 1 && ipl != 2 && ipl != 4) || nsplit < 0 || theta <= 0.0 ||
            (src != 0 && src != 1))
            continue;
        r.n_min = nmin;
        r.cfg.ipl = ipl;
        r.cfg.nsplit = nsplit;
        r.cfg.theta = theta;
        r.cfg.src = src;
        ctx->tune.push_back(r);
    }
    fclose(f);
    std::sort(ctx->tune.begin(), ctx->tune.end(),
              [](const TuneRow &a, const TuneRow &b) { return a.n_min < b.n_min; });
}

TiledCfg ncg_tiled_cfg(const ncg_ctx *ctx, int prec, long n)
{
    if (ctx->tiled_explicit) return ctx->tiled;
    TiledCfg c = ctx->tiled;
    for (const TuneRow &r : ctx->tune)
        if (r.prec == prec && r.n_min <= n) c = r.cfg;
    return c;
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
    load_tuning(ctx);
    if (const char *rc = getenv("NCG_REPLAY_CAP")) {
        const long v = atol(rc);
        if (v > 0 && v < (1L << 30)) ctx->replay_cap = (unsigned)v;
    }
    if (const char *cd = getenv("NCG_REPLAY_CLUSTER_DIST")) {
        const long v = atol(cd);
        if (v >= 0 && v <= 64) ctx->replay_cluster_dist = (int)v;
    }
    if (ncg_static_scratch_init(ctx) != NCG_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return NCG_ERR_DEVICE;
    }
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
    for (auto &kv : ctx->pbufs)
        if (kv.second.ptr) (void)hipHostFree(kv.second.ptr);
    if (ctx->aux) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
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
