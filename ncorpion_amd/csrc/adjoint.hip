// Hand-written adjoints: vector-Jacobian products of the gravity field and
// the discrete adjoint of the RK4 driver with checkpointing.
//
// Field (brute force, physics.c:120-202, plus the central body of
// physics.c:127-146 and its reaction CM_acc):
//   a_i   = sum_{j != i} G m_j d_ij / D_ij^3 - G M e_i / r_i^3
//   a_C   = sum_i G m_i e_i / r_i^3
// with d_ij = x_j - x_i, D^2 = |d|^2 + (s (R_i + R_j))^2, e_i = x_i - C.
// With T(d, D) = I / D^3 - 3 d d^T / D^5, for cotangents lam_i (of a_i) and
// lam_C (of a_C):
//   g_x_i += sum_j G T_ij (m_i lam_j - m_j lam_i)            (pair part)
//   g_m_i += -sum_j G lam_j . d_ij / D_ij^3                  (pair part)
//   g_x_i += -G M T_i lam_i + G m_i T_i lam_C                 (central part)
//   g_C   += sum_i (G M T_i lam_i - G m_i T_i lam_C)
//   g_M   += -sum_i G lam_i . e_i / r_i^3
//   g_m_i += G lam_C . e_i / r_i^3
// All FP64.  The RK4 adjoint differentiates the FP64-fast forward of
// rk4.hip step by step: checkpoints of the state every K steps on the
// forward pass; backward, each segment is recomputed from its checkpoint with
// the stage states stored, then the steps are reversed:
//   Kbar_s = h w_s / 6 * ybar',  for s = 4..1:
//     Ybar_s = (VJP_x(a at X_s; Kbar_s.v), Kbar_s.x);  ybar += Ybar_s;
//     Kbar_{s-1} += c_s h Ybar_s                     (c = 1/2, 1/2, 1)
#include "ncg_internal.hpp"

namespace {

constexpr int AT = 256;  // threads per workgroup
constexpr int VC = 64;   // j chunk: one wavefront's worth of bodies

// j records for the pair VJP: {x, y, z, m, lam x, lam y, lam z, s R}.
__global__ __launch_bounds__(AT) void k_vjp_pack(const DBody *__restrict__ b, int n,
                                                 const double *__restrict__ lam, double soft,
                                                 double *__restrict__ jv)
{
    const int i = blockIdx.x * AT + threadIdx.x;
    if (i >= n) return;
    const DBody q = b[i];
    double *o = jv + 8 * (size_t)i;
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
    o[3] = q.mass;
    o[4] = lam[3 * (size_t)i];
    o[5] = lam[3 * (size_t)i + 1];
    o[6] = lam[3 * (size_t)i + 2];
    o[7] = soft * q.radius;
}

struct VjpI {
    double x, y, z, m, lx, ly, lz, s;
};

// One ordered pair (i, j) of the VJP.  1/D from v_rsq_f64 plus one Newton
// step (FP64 rounding).  SELF: the chunk holds this lane's own body; its
// term is masked to 0 (D may be 0 there).
template <bool SELF>
__device__ inline void vjp_term(const double *__restrict__ q, int j, int i, const VjpI &a,
                                double &gxx, double &gxy, double &gxz, double &gmi)
{
    const double dx = q[0] - a.x, dy = q[1] - a.y, dz = q[2] - a.z;
    const double e = a.s + q[7];
    const double D2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, e * e)));
    double y = __builtin_amdgcn_rsq(D2);
    y = y * fma(-0.5 * D2, y * y, 1.5);
    double i3 = y * y * y;
    double i5 = i3 * y * y;
    if (SELF) {
        i3 = j == i ? 0.0 : i3;
        i5 = j == i ? 0.0 : i5;
    }
    // u = m_i lam_j - m_j lam_i ; T u = u / D^3 - 3 d (d . u) / D^5
    const double ux = a.m * q[4] - q[3] * a.lx;
    const double uy = a.m * q[5] - q[3] * a.ly;
    const double uz = a.m * q[6] - q[3] * a.lz;
    const double du3 = 3.0 * (dx * ux + dy * uy + dz * uz) * i5;
    gxx = fma(ux, i3, fma(-dx, du3, gxx));
    gxy = fma(uy, i3, fma(-dy, du3, gxy));
    gxz = fma(uz, i3, fma(-dz, du3, gxz));
    gmi = fma(-(q[4] * dx + q[5] * dy + q[6] * dz), i3, gmi);
}

// Pair part of the VJP.  Grid (bodies / AT, nsplit): lane = body i, the
// block's j range is [blockIdx.y * per, ...) with per a multiple of VC.  j
// records are read with wave-uniform addresses (scalar loads).  Output: one
// partial {g_x, g_m} per (split, body), summed in split order by k_vjp_reduce.
__global__ __launch_bounds__(AT) void k_vjp_pairs(const double *__restrict__ jv, int n, int per,
                                                  double4 *__restrict__ part)
{
    const int i = blockIdx.x * AT + threadIdx.x;
    const int ii = min(i, n - 1);
    VjpI a;
    {
        const double *p = jv + 8 * (size_t)ii;
        a = VjpI{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    }
    const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * AT + ((int)threadIdx.x & ~(VC - 1)));
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    double gxx = 0., gxy = 0., gxz = 0., gmi = 0.;
    for (int jc = j0; jc < j1; jc += VC) {
        const int je = min(j1, jc + VC);
        if (jc == wave0) {
            for (int j = jc; j < je; ++j) vjp_term<true>(jv + 8 * (size_t)j, j, i, a, gxx, gxy, gxz, gmi);
        } else if (je - jc == VC) {
#pragma unroll 4
            for (int j = jc; j < jc + VC; ++j)
                vjp_term<false>(jv + 8 * (size_t)j, j, i, a, gxx, gxy, gxz, gmi);
        } else {
            for (int j = jc; j < je; ++j) vjp_term<false>(jv + 8 * (size_t)j, j, i, a, gxx, gxy, gxz, gmi);
        }
    }
    if (i < n) part[(size_t)blockIdx.y * n + i] = make_double4(gxx, gxy, gxz, gmi);
}

// g_x (=) and g_m (=) from the split partials, in split order.
__global__ __launch_bounds__(AT) void k_vjp_reduce(const double4 *__restrict__ part, int nsplit,
                                                   int n, double G, double *__restrict__ gx,
                                                   double *__restrict__ gm)
{
    const int i = blockIdx.x * AT + threadIdx.x;
    if (i >= n) return;
    double4 t = part[i];
    for (int s = 1; s < nsplit; ++s) {
        const double4 u = part[(size_t)s * n + i];
        t.x += u.x;
        t.y += u.y;
        t.z += u.z;
        t.w += u.w;
    }
    gx[3 * (size_t)i + 0] = G * t.x;
    gx[3 * (size_t)i + 1] = G * t.y;
    gx[3 * (size_t)i + 2] = G * t.z;
    gm[i] = G * t.w;
}

// Central part, added to g_x / g_m; per-block partials of g_C (3) and g_M.
__global__ __launch_bounds__(AT) void k_vjp_central(const DBody *__restrict__ b, int n,
                                                    const double *__restrict__ cen,
                                                    const double *__restrict__ lam,
                                                    const double *__restrict__ lamc, double G,
                                                    double *__restrict__ gx,
                                                    double *__restrict__ gm,
                                                    double *__restrict__ part)
{
    __shared__ double red[4][AT];
    const int i = blockIdx.x * AT + threadIdx.x;
    double pc[4] = {0., 0., 0., 0.};
    if (i < n && cen[7] != 0.0) {
        const DBody q = b[i];
        const double ex = q.x - cen[0], ey = q.y - cen[1], ez = q.z - cen[2];
        const double r2 = ex * ex + ey * ey + ez * ez;
        const double r = sqrt(r2);
        const double i3 = 1.0 / (r2 * r), i5 = i3 / r2;
        const double M = cen[6];
        const double lx = lam[3 * (size_t)i], ly = lam[3 * (size_t)i + 1], lz = lam[3 * (size_t)i + 2];
        const double cx = lamc ? lamc[0] : 0., cy = lamc ? lamc[1] : 0., cz = lamc ? lamc[2] : 0.;
        // v = -M lam_i + m_i lam_C ; T v
        const double vx = -M * lx + q.mass * cx, vy = -M * ly + q.mass * cy, vz = -M * lz + q.mass * cz;
        const double ev = ex * vx + ey * vy + ez * vz;
        const double tx = G * (vx * i3 - 3.0 * ex * ev * i5);
        const double ty = G * (vy * i3 - 3.0 * ey * ev * i5);
        const double tz = G * (vz * i3 - 3.0 * ez * ev * i5);
        gx[3 * (size_t)i + 0] += tx;
        gx[3 * (size_t)i + 1] += ty;
        gx[3 * (size_t)i + 2] += tz;
        gm[i] += G * (cx * ex + cy * ey + cz * ez) * i3;
        pc[0] = -tx;
        pc[1] = -ty;
        pc[2] = -tz;
        pc[3] = -G * (lx * ex + ly * ey + lz * ez) * i3;
    }
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = pc[c];
    __syncthreads();
    for (int w = AT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) part[4 * (size_t)blockIdx.x + c] = red[c][0];
}

__global__ __launch_bounds__(AT) void k_sum4(const double *__restrict__ part, int np,
                                             double *__restrict__ out, int accumulate)
{
    __shared__ double red[4][AT];
    double a[4] = {0., 0., 0., 0.};
    for (int q = threadIdx.x; q < np; q += AT)
        for (int c = 0; c < 4; ++c) a[c] += part[4 * (size_t)q + c];
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = AT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) out[c] = (accumulate ? out[c] : 0.0) + red[c][0];
}

// Backward of RK4 stage s (1..4) for bodies (i < n) and the central body
// (i == n).  State adjoint ybar (A*), step cotangent ybar' (L*), stage
// cotangent Kbar_s (K*), VJP output g (gX for bodies, gC for the central
// body): A += (g, Kx); then Kbar_{s-1} = w_{s-1} h/6 L + c_s h (g, Kx).
__global__ void k_adj_stage(int s, double h, int n, const double *__restrict__ Lx,
                            const double *__restrict__ Lv, double *__restrict__ Kx,
                            double *__restrict__ Kv, const double *__restrict__ gX,
                            double *__restrict__ Ax, double *__restrict__ Av,
                            const double *__restrict__ cL, double *__restrict__ cK,
                            const double *__restrict__ cg, double *__restrict__ cA)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double ws[5] = {0., 1., 2., 2., 1.};
    const double cs[5] = {0., 0., 0.5, 0.5, 1.0};
    if (i < n) {
        for (int d = 0; d < 3; ++d) {
            const size_t k = 3 * (size_t)i + d;
            const double g = gX[k], kx = Kx[k];
            Ax[k] += g;
            Av[k] += kx;
            if (s > 1) {
                Kx[k] = ws[s - 1] * h / 6.0 * Lx[k] + cs[s] * h * g;
                Kv[k] = ws[s - 1] * h / 6.0 * Lv[k] + cs[s] * h * kx;
            }
        }
    } else if (i == n && cK) {
        // central: cL = ybar' (6), cK = Kbar (6), cg = g_C (3), cA = ybar (6)
        for (int d = 0; d < 3; ++d) {
            const double g = cg[d], kx = cK[d];
            cA[d] += g;
            cA[3 + d] += kx;
            if (s > 1) {
                cK[d] = ws[s - 1] * h / 6.0 * cL[d] + cs[s] * h * g;
                cK[3 + d] = ws[s - 1] * h / 6.0 * cL[3 + d] + cs[s] * h * kx;
            }
        }
    }
}

// Kbar_4 = h/6 L ; A = L (start of a step's backward).
__global__ void k_adj_begin(double h, int n, const double *__restrict__ Lx,
                            const double *__restrict__ Lv, double *__restrict__ Kx,
                            double *__restrict__ Kv, double *__restrict__ Ax,
                            double *__restrict__ Av, const double *__restrict__ cL,
                            double *__restrict__ cK, double *__restrict__ cA)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n) {
        Kx[i] = h / 6.0 * Lx[i];
        Kv[i] = h / 6.0 * Lv[i];
        Ax[i] = Lx[i];
        Av[i] = Lv[i];
    } else if (i == 3 * n && cK) {
        for (int d = 0; d < 6; ++d) {
            cK[d] = h / 6.0 * cL[d];
            cA[d] = cL[d];
        }
    }
}

// Final cotangent (x, v interleaved per body, and the central body's) into
// the adjoint's separate x and v arrays; thread n takes the central one.
__global__ void k_adj_split(const double *__restrict__ g, const double *__restrict__ gc, int n,
                            double *__restrict__ Lx, double *__restrict__ Lv,
                            double *__restrict__ cL)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        for (int d = 0; d < 3; ++d) {
            Lx[3 * (size_t)i + d] = g[6 * (size_t)i + d];
            Lv[3 * (size_t)i + d] = g[6 * (size_t)i + 3 + d];
        }
    } else if (i == n) {
        for (int d = 0; d < 6; ++d) cL[d] = gc ? gc[d] : 0.0;
    }
}

// The adjoint at t = 0 back to the interleaved layout, the central body's
// (x, v) and g_M; thread n takes the central one.
__global__ void k_adj_merge(const double *__restrict__ Lx, const double *__restrict__ Lv,
                            const double *__restrict__ cL, const double *__restrict__ gcm, int n,
                            double *__restrict__ g0, double *__restrict__ gc0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        for (int d = 0; d < 3; ++d) {
            g0[6 * (size_t)i + d] = Lx[3 * (size_t)i + d];
            g0[6 * (size_t)i + 3 + d] = Lv[3 * (size_t)i + d];
        }
    } else if (i == n) {
        for (int d = 0; d < 6; ++d) gc0[d] = cL[d];
        gc0[6] = gcm[3];
    }
}

__global__ void k_adj_zero(double *__restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ void k_add(double *__restrict__ a, const double *__restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

}  // namespace

// VJP at body state b (positions and masses) with central state cen (8):
// gx (3n) =; gm (n) += when acc_gm else =; gcm (4: g_C, g_M) =.
static ncg_status vjp_impl(ncg_ctx *ctx, const DBody *b, int32_t n, const double *cen, double G,
                           double soft, const double *lam, const double *lamc, double *gx,
                           double *gm, double *gcm, bool acc_gm, hipStream_t st)
{
    const bool accumulate = acc_gm;
    double *gm_tmp = accumulate ? dbuf<double>(ctx, "v_gm", n) : gm;
    const int nb = (int)cdiv(n, AT);
    // j splits: about 8 workgroups per CU, each split a multiple of VC bodies.
    const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
    const int nchunk = (int)cdiv(n, VC);
    int nsplit = std::max(1, std::min(nchunk, (int)cdiv(8 * cus, nb)));
    const int per = (int)cdiv(nchunk, nsplit) * VC;
    nsplit = (int)cdiv(n, per);
    double *part = dbuf<double>(ctx, "v_part", 4 * (size_t)nb);
    double *jv = dbuf<double>(ctx, "v_jv", 8 * (size_t)n);
    double4 *pp = dbuf<double4>(ctx, "v_pp", (size_t)nsplit * n);
    if (!gm_tmp || !part || !jv || !pp) return NCG_ERR_NOMEM;
    hipLaunchKernelGGL(k_vjp_pack, dim3(nb), dim3(AT), 0, st, b, (int)n, lam, soft, jv);
    {
        ProfScope ps(ctx, st, "vjp");
        hipLaunchKernelGGL(k_vjp_pairs, dim3(nb, nsplit), dim3(AT), 0, st, jv, (int)n, per, pp);
        NCG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_vjp_reduce, dim3(nb), dim3(AT), 0, st, pp, nsplit, (int)n, G, gx, gm_tmp);
    hipLaunchKernelGGL(k_vjp_central, dim3(nb), dim3(AT), 0, st, b, (int)n, cen, lam, lamc, G, gx,
                       gm_tmp, part);
    hipLaunchKernelGGL(k_sum4, dim3(1), dim3(AT), 0, st, part, nb, gcm, 0);
    if (accumulate) hipLaunchKernelGGL(k_add, dim3(cdiv(n, 256)), dim3(256), 0, st, gm, gm_tmp, (int)n);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

extern "C" ncg_status ncg_vjp_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                          const double *d_central, double grav_const,
                                          double softening, const double *d_lam,
                                          const double *d_lam_central, double *d_gx,
                                          double *d_gm, double *d_gcentral, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_central && d_lam && d_gx && d_gm && d_gcentral,
                "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    return vjp_impl(ctx, reinterpret_cast<const DBody *>(d_bodies), n, d_central, grav_const,
                    softening, d_lam, d_lam_central, d_gx, d_gm, d_gcentral, false,
                    pick_stream(ctx, stream));
}

// One forward RK4 step (FP64-fast) from y/c0, storing the stage states
// Ys[s] (s = 0..3) and central stage states Ycs[8 s].
static ncg_status fwd_step_store(ncg_ctx *ctx, DBody *y, int32_t n, double *c0, double G,
                                 double soft, double dt, DBody *Ys, double *Ycs, hipStream_t st)
{
    DBody *Y = dbuf<DBody>(ctx, "a_Y", n);
    double *acc = dbuf<double>(ctx, "a_acc", 3 * (size_t)n);
    double *sx = dbuf<double>(ctx, "a_sx", 3 * (size_t)n);
    double *sv = dbuf<double>(ctx, "a_sv", 3 * (size_t)n);
    double *Yc = dbuf<double>(ctx, "a_Yc", 8);
    double *sc = dbuf<double>(ctx, "a_sc", 8);
    double *cm = dbuf<double>(ctx, "a_cm", 4);
    if (!Y || !acc || !sx || !sv || !Yc || !sc || !cm) return NCG_ERR_NOMEM;
    ncg_status r = ncg_tiled_order(ctx, y, n, nullptr, NCG_F64_FAST, st);
    if (r != NCG_OK) return r;
    ncg_central cf = {};
    cf.present = 1;  // the device central record's own flag decides (k_finish reads it)
    for (int s = 0; s < 4; ++s) {
        const DBody *src = s == 0 ? y : Y;
        const double *dc = s == 0 ? c0 : Yc;
        if (Ys) {
            NCG_HIP(hipMemcpyAsync(Ys + (size_t)s * n, src, sizeof(DBody) * n, hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemcpyAsync(Ycs + 8 * s, dc, sizeof(double) * 8, hipMemcpyDeviceToDevice, st));
        }
        r = ncg_tiled_eval(ctx, src, n, nullptr, &cf, dc, G, soft, NCG_F64_FAST, acc, nullptr, cm, st);
        if (r != NCG_OK) return r;
        r = ncg_rk4_stage_launch(ctx, s, dt, n, nullptr, y, Y, acc, sx, sv, c0, Yc, cm, sc, st);
        if (r != NCG_OK) return r;
    }
    return NCG_OK;
}

// Backward pass from checkpoints ck (state before step q K, q = 0..) and
// ckc (central records): recompute each segment storing its stage states,
// then reverse it stage by stage.
static ncg_status adjoint_backward(ncg_ctx *ctx, const DBody *ck, const double *ckc, int32_t n,
                                   double G, double soft, double h, int32_t nsteps, int K,
                                   const double *d_gfinal, const double *d_gcentral_final,
                                   double *d_g0, double *d_gm, double *d_gcentral0, hipStream_t st)
{
    const double dt = h;
    DBody *y = dbuf<DBody>(ctx, "a_y", n);
    double *c = dbuf<double>(ctx, "a_c", 8);
    // Segment store: stage states of K steps.
    DBody *Ys = dbuf<DBody>(ctx, "a_Ys", (size_t)4 * K * n);
    double *Ycs = dbuf<double>(ctx, "a_Ycs", (size_t)32 * K);
    double *Lx = dbuf<double>(ctx, "a_Lx", 3 * (size_t)n), *Lv = dbuf<double>(ctx, "a_Lv", 3 * (size_t)n);
    double *Ax = dbuf<double>(ctx, "a_Ax", 3 * (size_t)n), *Av = dbuf<double>(ctx, "a_Av", 3 * (size_t)n);
    double *Kx = dbuf<double>(ctx, "a_Kx", 3 * (size_t)n), *Kv = dbuf<double>(ctx, "a_Kv", 3 * (size_t)n);
    double *gX = dbuf<double>(ctx, "a_gX", 3 * (size_t)n);
    double *cst = dbuf<double>(ctx, "a_cst", 32);  // cL[6] cK[6] cA[6] gcm[4] lamc[3]
    if (!y || !c || !Ys || !Ycs || !Lx || !Lv || !Ax || !Av || !Kx || !Kv || !gX || !cst)
        return NCG_ERR_NOMEM;
    double *cL = cst, *cK = cst + 6, *cA = cst + 12, *gcm = cst + 18, *lamc = cst + 22;
    // Final cotangent: d_gfinal is (x, v) per body, 6n, interleaved x,y,z,vx,vy,vz.
    hipLaunchKernelGGL(k_adj_split, dim3(cdiv(n + 1, 256)), dim3(256), 0, st, d_gfinal,
                       d_gcentral_final, (int)n, Lx, Lv, cL);
    hipLaunchKernelGGL(k_adj_zero, dim3(cdiv(n, 256) + 1), dim3(256), 0, st, d_gm, (int)n);
    hipLaunchKernelGGL(k_adj_zero, dim3(1), dim3(64), 0, st, gcm, 4);
    NCG_CHECK_LAUNCH();
    for (int seg = (nsteps - 1) / K; seg >= 0 && nsteps > 0; --seg) {
        const int s0 = seg * K, s1 = std::min(nsteps, s0 + K);
        NCG_HIP(hipMemcpyAsync(y, ck + (size_t)seg * n, sizeof(DBody) * n, hipMemcpyDeviceToDevice, st));
        NCG_HIP(hipMemcpyAsync(c, ckc + 8 * seg, sizeof(double) * 8, hipMemcpyDeviceToDevice, st));
        for (int step = s0; step < s1; ++step) {
            ncg_status r = fwd_step_store(ctx, y, n, c, G, soft, dt, Ys + (size_t)4 * (step - s0) * n,
                                          Ycs + 32 * (step - s0), st);
            if (r != NCG_OK) return r;
        }
        for (int step = s1 - 1; step >= s0; --step) {
            const DBody *Yst = Ys + (size_t)4 * (step - s0) * n;
            const double *Yct = Ycs + 32 * (step - s0);
            hipLaunchKernelGGL(k_adj_begin, dim3(cdiv(3 * n + 1, 256)), dim3(256), 0, st, h, (int)n, Lx,
                               Lv, Kx, Kv, Ax, Av, cL, cK, cA);
            for (int s = 4; s >= 1; --s) {
                // VJP at stage s with cotangent Kbar_s.v (bodies) and Kbar_s.v (central).
                NCG_HIP(hipMemcpyAsync(lamc, cK + 3, sizeof(double) * 3, hipMemcpyDeviceToDevice, st));
                double *gC = cst + 26;  // g_C (3) + g_M (1) of this stage
                ncg_status r = vjp_impl(ctx, Yst + (size_t)(s - 1) * n, n, Yct + 8 * (s - 1), G, soft,
                                        Kv, lamc, gX, d_gm, gC, true, st);
                if (r != NCG_OK) return r;
                hipLaunchKernelGGL(k_adj_stage, dim3(cdiv(n + 1, 256)), dim3(256), 0, st, s, h, (int)n,
                                   Lx, Lv, Kx, Kv, gX, Ax, Av, cL, cK, gC, cA);
                hipLaunchKernelGGL(k_add, dim3(1), dim3(64), 0, st, gcm + 3, gC + 3, 1);
                NCG_CHECK_LAUNCH();
            }
            NCG_HIP(hipMemcpyAsync(Lx, Ax, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemcpyAsync(Lv, Av, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemcpyAsync(cL, cA, sizeof(double) * 6, hipMemcpyDeviceToDevice, st));
        }
    }
    // Outputs: g0 interleaved (x, v) per body; central (x, v) and g_M.
    hipLaunchKernelGGL(k_adj_merge, dim3(cdiv(n + 1, 256)), dim3(256), 0, st, Lx, Lv, cL, gcm,
                       (int)n, d_g0, d_gcentral0);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

extern "C" ncg_status ncg_rk4_adjoint_dev(ncg_ctx *ctx, const ncg_body *d_bodies0, int32_t n,
                                          const double *d_central0, double grav_const,
                                          double softening, double dt, int32_t nsteps,
                                          int32_t checkpoint_every, const double *d_gfinal,
                                          const double *d_gcentral_final, double *d_g0,
                                          double *d_gm, double *d_gcentral0, ncg_body *d_final,
                                          void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies0 && d_central0 && nsteps >= 0 && checkpoint_every >= 1 &&
                    d_gfinal && d_g0 && d_gm && d_gcentral0,
                "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    const double G = grav_const, soft = softening;
    const int K = checkpoint_every;
    const int nck = nsteps / K + 1;
    // Checkpoints (bodies + central) every K steps.
    DBody *ck = dbuf<DBody>(ctx, "a_ck", (size_t)nck * n);
    double *ckc = dbuf<double>(ctx, "a_ckc", 8 * (size_t)nck);
    DBody *y = dbuf<DBody>(ctx, "a_fy", n);
    double *c = dbuf<double>(ctx, "a_fc", 8);
    if (!ck || !ckc || !y || !c) return NCG_ERR_NOMEM;
    // Forward with checkpoints.
    NCG_HIP(hipMemcpyAsync(y, d_bodies0, sizeof(DBody) * n, hipMemcpyDeviceToDevice, st));
    NCG_HIP(hipMemcpyAsync(c, d_central0, sizeof(double) * 8, hipMemcpyDeviceToDevice, st));
    for (int step = 0; step < nsteps; ++step) {
        if (step % K == 0) {
            NCG_HIP(hipMemcpyAsync(ck + (size_t)(step / K) * n, y, sizeof(DBody) * n, hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemcpyAsync(ckc + 8 * (step / K), c, sizeof(double) * 8, hipMemcpyDeviceToDevice, st));
        }
        ncg_status r = fwd_step_store(ctx, y, n, c, G, soft, dt, nullptr, nullptr, st);
        if (r != NCG_OK) return r;
    }
    if (d_final) NCG_HIP(hipMemcpyAsync(d_final, y, sizeof(DBody) * n, hipMemcpyDeviceToDevice, st));
    return adjoint_backward(ctx, ck, ckc, n, G, soft, dt, nsteps, K, d_gfinal, d_gcentral_final,
                            d_g0, d_gm, d_gcentral0, st);
}

extern "C" ncg_status ncg_rk4_forward_ckpt_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                               double *d_central, double grav_const,
                                               double softening, double dt, int32_t nsteps,
                                               int32_t checkpoint_every, ncg_body *d_ckpt,
                                               double *d_ckpt_central, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_central && nsteps >= 0 && checkpoint_every >= 1 &&
                    d_ckpt && d_ckpt_central,
                "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    const int K = checkpoint_every;
    for (int32_t s0 = 0; s0 < nsteps; s0 += K) {
        const int q = s0 / K;
        NCG_HIP(hipMemcpyAsync(d_ckpt + (size_t)q * n, d_bodies, sizeof(DBody) * n,
                               hipMemcpyDeviceToDevice, st));
        NCG_HIP(hipMemcpyAsync(d_ckpt_central + 8 * (size_t)q, d_central, sizeof(double) * 8,
                               hipMemcpyDeviceToDevice, st));
        // FP64-fast steps, replayed from the RK4 driver's graph.
        ncg_status r = ncg_rk4_steps_dev(ctx, d_bodies, n, nullptr, d_central, grav_const, softening,
                                         NCG_F64_FAST, dt, std::min(K, nsteps - s0), 1, st);
        if (r != NCG_OK) return r;
    }
    return NCG_OK;
}

extern "C" ncg_status ncg_rk4_adjoint_ckpt_dev(ncg_ctx *ctx, const ncg_body *d_ckpt, int32_t n,
                                               const double *d_ckpt_central, double grav_const,
                                               double softening, double dt, int32_t nsteps,
                                               int32_t checkpoint_every, const double *d_gfinal,
                                               const double *d_gcentral_final, double *d_g0,
                                               double *d_gm, double *d_gcentral0, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_ckpt && d_ckpt_central && nsteps >= 0 && checkpoint_every >= 1 &&
                    d_gfinal && d_g0 && d_gm && d_gcentral0,
                "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    return adjoint_backward(ctx, reinterpret_cast<const DBody *>(d_ckpt), d_ckpt_central, n,
                            grav_const, softening, dt, nsteps, checkpoint_every, d_gfinal,
                            d_gcentral_final, d_g0, d_gm, d_gcentral0, pick_stream(ctx, stream));
}
