// GPU-resident leapfrog of NcorpiON's brute-force integrator
// (integration_brute_force_SABA1, rk4.c:840-954, with kick and drift of
// rk4.c:68-127): the bodies and the central body stay on the device and every
// step is
//   kick(dt)   v += dt * a(x) for every body, V += dt * CM_acc (rk4.c:68-107)
//   brute_force with dt (collision.c:425-456): candidate pass + device replay
//   drift(dt)  x += dt * v, X += dt * V (rk4.c:110-127)
//   did_collide cleared when one_collision_only (end_of_timestep, rk4.c:348-353)
// optionally preceded by the half drift of rk4.c:871-879 (brute_force and
// drift with dt/2, did_collide kept into the first step as there) and
// followed by the backward half drift of rk4.c:938-942.  The field is the
// exact form (bit for bit vector_field's, CM_acc summed in body order) or a
// tiled form; kick and drift are compiled without FMA contraction, so with
// the exact form a run reproduces the reference's arithmetic.
#include "ncg_internal.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int LT = 256;

// Bodies i < n; thread n: the central body.
__global__ __launch_bounds__(LT) void k_lf_kick(int n, const int32_t *__restrict__ ex, double h,
                                                DBody *__restrict__ b,
                                                const double *__restrict__ acc,
                                                double *__restrict__ cen,
                                                const double *__restrict__ cm)
{
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i < n) {
        if (ex && !ex[i]) return;
        b[i].vx += h * acc[3 * (size_t)i];
        b[i].vy += h * acc[3 * (size_t)i + 1];
        b[i].vz += h * acc[3 * (size_t)i + 2];
    } else if (i == n && cen[7] != 0.0) {
        cen[3] += h * cm[0];
        cen[4] += h * cm[1];
        cen[5] += h * cm[2];
    }
}

__global__ __launch_bounds__(LT) void k_lf_drift(int n, const int32_t *__restrict__ ex, double h,
                                                 DBody *__restrict__ b, double *__restrict__ cen)
{
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i < n) {
        if (ex && !ex[i]) return;
        b[i].x += h * b[i].vx;
        b[i].y += h * b[i].vy;
        b[i].z += h * b[i].vz;
    } else if (i == n && cen[7] != 0.0) {
        cen[0] += h * cen[3];
        cen[1] += h * cen[4];
        cen[2] += h * cen[5];
    }
}

__global__ void k_lf_zero_i32(int32_t *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_lf_zero_u64(unsigned long long *__restrict__ p) { p[threadIdx.x] = 0ull; }

__global__ void k_lf_off(int *__restrict__ off, int n)
{
    off[0] = 0;
    off[1] = n;
}

}  // namespace

extern "C" ncg_status ncg_leapfrog_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                       const int32_t *d_exists, int32_t *d_did_collide,
                                       double *d_central, double grav_const, double softening,
                                       ncg_precision prec, double dt, int32_t nsteps,
                                       double collision_f, int32_t one_collision_only,
                                       int32_t flags, int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_central && nsteps >= 0, "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    NCG_REQUIRE(!(collision_f > 0.0 && one_collision_only && !d_did_collide),
                "one_collision_only needs a did_collide array");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    if (n_collisions) *n_collisions = 0;
    DBody *b = reinterpret_cast<DBody *>(d_bodies);
    const int32_t *ex = d_exists;
    const bool coll = collision_f > 0.0;
    const bool use_graph = (flags & 1) != 0;
    const unsigned cap = 1u << 20;
    double *acc = dbuf<double>(ctx, "lf_acc", 3 * (size_t)n);
    double *cterm = dbuf<double>(ctx, "lf_cterm", 3 * (size_t)n);
    double *cm = dbuf<double>(ctx, "lf_cm", 4);
    int2 *cand = dbuf<int2>(ctx, "lf_cand", cap);
    unsigned long long *cnt = dbuf<unsigned long long>(ctx, "lf_cnt", 1);
    unsigned char *dirty = dbuf<unsigned char>(ctx, "lf_dirty", (size_t)n + 4);
    unsigned *stat = dbuf<unsigned>(ctx, "lf_stat", 2);
    int *off = dbuf<int>(ctx, "lf_off", 2);
    if (!acc || !cterm || !cm || !cand || !cnt || !dirty || !stat || !off) return NCG_ERR_NOMEM;
    double present = 0.0;
    NCG_HIP(hipMemcpyAsync(&present, d_central + 7, sizeof(double), hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    const bool cp = present != 0.0;
    hipLaunchKernelGGL(k_lf_off, dim3(1), dim3(1), 0, st, off, (int)n);
    hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(n / 4 + 2, 256)), dim3(256), 0, st,
                       reinterpret_cast<int32_t *>(dirty), (int)(n / 4 + 1));
    hipLaunchKernelGGL(k_lf_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t *>(stat), 2);
    NCG_CHECK_LAUNCH();
    const unsigned g = (unsigned)cdiv(n + 1, LT);
    const double G = grav_const, soft = softening;

    auto collide = [&](double h) -> ncg_status {
        hipLaunchKernelGGL(k_lf_zero_u64, dim3(1), dim3(1), 0, st, cnt);
        ncg_status r = ncg_candidates_launch(ctx, b, n, ex, h, cand, cap, cnt, st);
        if (r != NCG_OK) return r;
        return ncg_bf_replay_launch(ctx, b, ex, d_did_collide, one_collision_only, off, 1, cand,
                                    nullptr, cnt, cap, h, collision_f, dirty, stat, st);
    };
    // One step after the field's preparation (the tiled forms sort first).
    ncg_central cflag = {};
    cflag.present = cp ? 1 : 0;
    auto step = [&]() -> ncg_status {
        ncg_status r;
        if (prec == NCG_F64_EXACT)
            r = ncg_grav_exact_dev(ctx, b, n, ex, d_central, G, soft, acc, cterm, cp ? cm : nullptr, st);
        else
            r = ncg_tiled_eval(ctx, b, n, ex, cp ? &cflag : nullptr, cp ? d_central : nullptr, G,
                               soft, prec, acc, nullptr, cp ? cm : nullptr, st);
        if (r != NCG_OK) return r;
        hipLaunchKernelGGL(k_lf_kick, dim3(g), dim3(LT), 0, st, (int)n, ex, dt, b, acc, d_central, cm);
        NCG_CHECK_LAUNCH();
        if (coll) {
            r = collide(dt);
            if (r != NCG_OK) return r;
        }
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, dt, b, d_central);
        if (coll && one_collision_only)
            hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(n, 256)), dim3(256), 0, st, d_did_collide, (int)n);
        NCG_CHECK_LAUNCH();
        return NCG_OK;
    };

    if (flags & 2) {
        // rk4.c:871-879: brute_force and drift with dt / 2.
        if (coll) {
            ncg_status r = collide(0.5 * dt);
            if (r != NCG_OK) return r;
        }
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, 0.5 * dt, b, d_central);
        NCG_CHECK_LAUNCH();
    }
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    ncg_status res = NCG_OK;
    for (int32_t s = 0; s < nsteps && res == NCG_OK; ++s) {
        if (prec != NCG_F64_EXACT) {
            res = ncg_tiled_order(ctx, b, n, ex, st);  // eager: the sort stays out of graphs
            if (res != NCG_OK) break;
        }
        if (!use_graph || s == 0) {
            res = step();  // the first step also allocates what a graph uses
            continue;
        }
        if (!exec) {
            NCG_HIP(hipStreamSynchronize(st));
            ctx->capturing = true;
            NCG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            res = step();
            hipError_t e = hipStreamEndCapture(st, &graph);
            ctx->capturing = false;
            if (res != NCG_OK) break;
            if (e != hipSuccess) {
                res = ncg_fail(ctx, NCG_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(e));
                break;
            }
            if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                res = ncg_fail(ctx, NCG_ERR_DEVICE, "graph instantiation failed");
                break;
            }
        }
        if (hipGraphLaunch(exec, st) != hipSuccess)
            res = ncg_fail(ctx, NCG_ERR_DEVICE, "graph launch failed");
    }
    if (res == NCG_OK && (flags & 4)) {
        // rk4.c:938-942: the backward half drift.
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, -0.5 * dt, b, d_central);
        NCG_CHECK_LAUNCH();
    }
    unsigned hs[2] = {0, 0};
    (void)hipMemcpyAsync(hs, stat, sizeof hs, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (res != NCG_OK) return res;
    if (n_collisions) *n_collisions = hs[0];
    if (hs[1])
        return ncg_fail(ctx, NCG_ERR_CAPACITY,
                        "collision replay capacity exceeded (flags %u: 1 = more than %u candidates, "
                        "2 = more than %d to decide, 4 = more than %d changed bodies in one pass); "
                        "results are not valid",
                        hs[1], cap, NCG_REPLAY_RCAP, NCG_REPLAY_DCAP);
    return NCG_OK;
}
