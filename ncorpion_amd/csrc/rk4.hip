// Device-resident classic RK4 integrator over the tiled gravity forms.
//
// NcorpiON itself integrates with a leapfrog (rk4.c:68-127 kick/drift); this
// driver is a classic fourth-order Runge-Kutta added beside it, on which the
// pair-interaction throughput and the energy drift are measured.  State: the body array
// (x, v in the struct moonlet layout) and the central body
// {x, y, z, vx, vy, vz, M, present} in device memory; the central body moves
// under CM_acc exactly as kick() applies it (rk4.c:103-105).
//
// One step = 4 force evaluations + 4 fused stage updates.  The spatial order
// (sort) is refreshed eagerly at the start of each step; the 4 evaluations and
// stage updates are captured once into a hipGraph and replayed every step.
#include "ncg_internal.hpp"

struct Rk4Graph {
    bool valid = false;
    const void *key_b = nullptr;
    const void *key_c = nullptr;
    const void *key_ex = nullptr;
    int32_t n = 0;
    uint64_t buf_gen = 0;
    int prec = -1;
    bool central_present = false;  // the captured stages include the central body or not
    double G = 0, soft = 0, dt = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
};

void ncg_rk4_release(ncg_ctx *ctx)
{
    if (ctx->rk4) {
        ctx->rk4->release();
        delete ctx->rk4;
        ctx->rk4 = nullptr;
    }
}

namespace {

constexpr int NT = 256;

// Stage s of RK4 (s = 0..3).  Y holds the stage state used by evaluation s;
// acc its acceleration.  After the update Y holds the state for stage s+1, or
// (s == 3) y0 holds the new state.
__global__ __launch_bounds__(NT) void k_rk4_stage(int s, double dt, int n,
                                                  const int32_t *__restrict__ ex,
                                                  DBody *__restrict__ y0, DBody *__restrict__ Y,
                                                  const double *__restrict__ acc,
                                                  double *__restrict__ sx, double *__restrict__ sv,
                                                  double *__restrict__ c0, double *__restrict__ Yc,
                                                  const double *__restrict__ cm_acc,
                                                  double *__restrict__ sc)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    const double w = (s == 0 || s == 3) ? 1.0 : 2.0;
    const double h = s < 2 ? 0.5 * dt : dt;
    if (i < n && (ex == nullptr || ex[i])) {
        const DBody b0 = y0[i];
        const DBody bs = s == 0 ? b0 : Y[i];
        const double v[3] = {bs.vx, bs.vy, bs.vz};
        const double a[3] = {acc[3 * (size_t)i], acc[3 * (size_t)i + 1], acc[3 * (size_t)i + 2]};
        double kx[3], kv[3];
        for (int d = 0; d < 3; ++d) {
            kx[d] = s == 0 ? v[d] : sx[3 * (size_t)i + d] + w * v[d];
            kv[d] = s == 0 ? a[d] : sv[3 * (size_t)i + d] + w * a[d];
        }
        if (s < 3) {
            for (int d = 0; d < 3; ++d) {
                sx[3 * (size_t)i + d] = kx[d];
                sv[3 * (size_t)i + d] = kv[d];
            }
            DBody ny = b0;
            ny.x = b0.x + h * v[0];
            ny.y = b0.y + h * v[1];
            ny.z = b0.z + h * v[2];
            ny.vx = b0.vx + h * a[0];
            ny.vy = b0.vy + h * a[1];
            ny.vz = b0.vz + h * a[2];
            Y[i] = ny;
        } else {
            DBody nb = b0;
            const double f = dt / 6.0;
            nb.x = b0.x + f * kx[0];
            nb.y = b0.y + f * kx[1];
            nb.z = b0.z + f * kx[2];
            nb.vx = b0.vx + f * kv[0];
            nb.vy = b0.vy + f * kv[1];
            nb.vz = b0.vz + f * kv[2];
            y0[i] = nb;
        }
    }
    if (i == 0 && c0[7] != 0.0) {
        const double *cs = s == 0 ? c0 : Yc;
        double kx[3], kv[3];
        for (int d = 0; d < 3; ++d) {
            kx[d] = s == 0 ? cs[3 + d] : sc[d] + w * cs[3 + d];
            kv[d] = s == 0 ? cm_acc[d] : sc[3 + d] + w * cm_acc[d];
        }
        if (s < 3) {
            for (int d = 0; d < 3; ++d) {
                sc[d] = kx[d];
                sc[3 + d] = kv[d];
                Yc[d] = c0[d] + h * cs[3 + d];
                Yc[3 + d] = c0[3 + d] + h * cm_acc[d];
            }
            Yc[6] = c0[6];
            Yc[7] = c0[7];
        } else {
            const double f = dt / 6.0;
            for (int d = 0; d < 3; ++d) {
                c0[d] += f * kx[d];
                c0[3 + d] += f * kv[d];
            }
        }
    }
}

}  // namespace

ncg_status ncg_rk4_stage_launch(ncg_ctx *ctx, int s, double dt, int32_t n, const int32_t *ex,
                                DBody *y0, DBody *Y, const double *acc, double *sx, double *sv,
                                double *c0, double *Yc, const double *cm, double *sc,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_rk4_stage, dim3(cdiv(n, NT)), dim3(NT), 0, st, s, dt, (int)n, ex, y0, Y, acc,
                       sx, sv, c0, Yc, cm, sc);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// The 4 evaluations + stage updates of one step (no allocation, no sync).
static ncg_status rk4_body(ncg_ctx *ctx, DBody *y0, int32_t n, const int32_t *ex, double *c0,
                           double G, double soft, ncg_precision prec, double dt, hipStream_t st,
                           bool central_present)
{
    DBody *Y = dbuf<DBody>(ctx, "rk_Y", n);
    double *acc = dbuf<double>(ctx, "rk_acc", 3 * (size_t)n);
    double *sx = dbuf<double>(ctx, "rk_sx", 3 * (size_t)n);
    double *sv = dbuf<double>(ctx, "rk_sv", 3 * (size_t)n);
    double *Yc = dbuf<double>(ctx, "rk_Yc", 8);
    double *sc = dbuf<double>(ctx, "rk_sc", 8);
    double *cm = dbuf<double>(ctx, "rk_cm", 4);
    if (!Y || !acc || !sx || !sv || !Yc || !sc || !cm) return NCG_ERR_NOMEM;
    ncg_central cflag = {};
    cflag.present = central_present ? 1 : 0;
    for (int s = 0; s < 4; ++s) {
        const DBody *src = s == 0 ? y0 : Y;
        const double *dc = s == 0 ? c0 : Yc;
        ncg_status r = ncg_tiled_eval(ctx, src, n, ex, central_present ? &cflag : nullptr,
                                      central_present ? dc : nullptr, G, soft, prec, acc, nullptr,
                                      central_present ? cm : nullptr, st);
        if (r != NCG_OK) return r;
        hipLaunchKernelGGL(k_rk4_stage, dim3(cdiv(n, NT)), dim3(NT), 0, st, s, dt, (int)n, ex, y0, Y,
                           acc, sx, sv, c0, Yc, cm, sc);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}

extern "C" ncg_status ncg_rk4_steps_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                        const int32_t *d_exists, double *d_central,
                                        double grav_const, double softening, ncg_precision prec,
                                        double dt, int32_t nsteps, int32_t flags, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_central && nsteps >= 0, "invalid arguments");
    if (prec != NCG_F64_FAST && prec != NCG_MIXED)
        return ncg_fail(ctx, NCG_ERR_UNSUPPORTED,
                        "the RK4 driver runs the tiled forms (fast64, mixed); exact is for the "
                        "reference's own integrator");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    DBody *y0 = reinterpret_cast<DBody *>(d_bodies);
    double present = 0.0;
    NCG_HIP(hipMemcpyAsync(&present, d_central + 7, sizeof(double), hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    const bool cp = present != 0.0;
    const bool use_graph = (flags & 1) != 0;
    for (int32_t step = 0; step < nsteps; ++step) {
        ncg_status r = ncg_tiled_order(ctx, y0, n, d_exists, prec, st);
        if (r != NCG_OK) return r;
        if (!use_graph) {
            r = rk4_body(ctx, y0, n, d_exists, d_central, grav_const, softening, prec, dt, st, cp);
            if (r != NCG_OK) return r;
            continue;
        }
        Rk4Graph *g = ctx->rk4;
        const bool same = g && g->valid && g->key_b == y0 && g->key_c == d_central &&
                          g->key_ex == d_exists && g->n == n && g->prec == (int)prec &&
                          g->central_present == cp && g->buf_gen == ctx->buf_gen &&
                          g->G == grav_const && g->soft == softening && g->dt == dt &&
                          g->stream == st;
        if (!same) {
            // First step eagerly: allocates every buffer the graph will use.
            r = rk4_body(ctx, y0, n, d_exists, d_central, grav_const, softening, prec, dt, st, cp);
            if (r != NCG_OK) return r;
            NCG_HIP(hipStreamSynchronize(st));
            if (!ctx->rk4) ctx->rk4 = new Rk4Graph();
            g = ctx->rk4;
            g->release();
            ctx->capturing = true;
            NCG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            r = rk4_body(ctx, y0, n, d_exists, d_central, grav_const, softening, prec, dt, st, cp);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(st, &graph);
            ctx->capturing = false;
            if (r != NCG_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return r;
            }
            if (e != hipSuccess)
                return ncg_fail(ctx, NCG_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(e));
            g->graph = graph;
            NCG_HIP(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
            g->valid = true;
            g->key_b = y0;
            g->key_c = d_central;
            g->key_ex = d_exists;
            g->buf_gen = ctx->buf_gen;
            g->n = n;
            g->prec = (int)prec;
            g->central_present = cp;
            g->G = grav_const;
            g->soft = softening;
            g->dt = dt;
            g->stream = st;
            continue;  // this step was done eagerly
        }
        NCG_HIP(hipGraphLaunch(g->exec, st));
    }
    return NCG_OK;
}
