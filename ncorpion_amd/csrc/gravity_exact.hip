// Brute-force gravity in the reference's FP64 expression tree and order.
//
// Restates vector_field's brute-force and central branches (NcorpiON
// physics.c:63-82, 120-202, 263-266) so that each acceleration is bit-identical.
// In the reference, body k starts from the central term, subtracts
// G*m_p*(X_k-X_p)/D^3 for p < k while k is the outer body, and later receives
// += G*m_q*(X_q-X_k)/D^3 from every q > k as they become the outer body.  For
// body k that is one sequential sum over j ascending, j != k.  One lane owns
// one body and walks j in that order through LDS tiles; the p < k terms are
// formed exactly as the reference forms them (X_k - X_p, then subtracted) and
// the q > k terms likewise (X_q - X_k, then added), so signed zeros match too.
// This translation unit is compiled with -ffp-contract=off: no FMA fusion.
#include "ncg_internal.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int EX_T = 256;

// Central body from the host (cp, cx, cy, cz, mu) or, when dcen is given,
// from a device record {x, y, z, vx, vy, vz, M, present}.
__global__ __launch_bounds__(256) void k_grav_exact(const DBody *__restrict__ b, int n,
                                                    const int32_t *__restrict__ ex, int cp,
                                                    double cx, double cy, double cz, double mu,
                                                    const double *__restrict__ dcen,
                                                    double G, double s,
                                                    double *__restrict__ acc,
                                                    double *__restrict__ cterm)
{
    __shared__ double tx[EX_T], ty[EX_T], tz[EX_T], tgm[EX_T], tr[EX_T];
    __shared__ int te[EX_T];
    if (dcen) {
        cp = dcen[7] != 0.0;
        cx = dcen[0];
        cy = dcen[1];
        cz = dcen[2];
        mu = G * dcen[6];  // physics.c:131
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = k < n && (ex == nullptr || ex[k] != 0);
    double X = 0., Y = 0., Z = 0., Rk = 0.;
    double aX = 0., aY = 0., aZ = 0.;
    if (mine) {
        X = b[k].x;
        Y = b[k].y;
        Z = b[k].z;
        Rk = b[k].radius;
        const double mk = b[k].mass;
        if (cp) {
            // physics.c:121-146
            const double DX = X - cx, DY = Y - cy, DZ = Z - cz;
            const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
            const double r3 = rk * rk * rk;
            aX = -mu * DX / r3;
            aY = -mu * DY / r3;
            aZ = -mu * DZ / r3;
            cterm[3 * (size_t)k + 0] = G * mk * DX / r3;
            cterm[3 * (size_t)k + 1] = G * mk * DY / r3;
            cterm[3 * (size_t)k + 2] = G * mk * DZ / r3;
        }
    }
    for (int base = 0; base < n; base += EX_T) {
        const int jl = base + (int)threadIdx.x;
        __syncthreads();
        if (jl < n) {
            const DBody q = b[jl];
            tx[threadIdx.x] = q.x;
            ty[threadIdx.x] = q.y;
            tz[threadIdx.x] = q.z;
            tgm[threadIdx.x] = G * q.mass;  // the reference's G*mp (physics.c:192)
            tr[threadIdx.x] = q.radius;
            te[threadIdx.x] = ex == nullptr ? 1 : (ex[jl] != 0);
        } else {
            te[threadIdx.x] = 0;
        }
        __syncthreads();
        if (!mine) continue;
        const int cnt = min(EX_T, n - base);
        for (int jj = 0; jj < cnt; ++jj) {
            const int j = base + jj;
            if (!te[jj] || j == k) continue;
            const bool lower = j < k;
            // physics.c:186-201: D = sqrt(DX*DX+DY*DY+DZ*DZ+soft*soft), D3 = D*D*D
            const double dx = lower ? X - tx[jj] : tx[jj] - X;
            const double dy = lower ? Y - ty[jj] : ty[jj] - Y;
            const double dz = lower ? Z - tz[jj] : tz[jj] - Z;
            const double soft = s * (lower ? (Rk + tr[jj]) : (tr[jj] + Rk));
            const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
            const double D3 = D * D * D;
            const double tX = tgm[jj] * dx / D3;
            const double tY = tgm[jj] * dy / D3;
            const double tZ = tgm[jj] * dz / D3;
            if (lower) {
                aX -= tX;
                aY -= tY;
                aZ -= tZ;
            } else {
                aX += tX;
                aY += tY;
                aZ += tZ;
            }
        }
    }
    if (mine) {
        acc[3 * (size_t)k + 0] = aX;
        acc[3 * (size_t)k + 1] = aY;
        acc[3 * (size_t)k + 2] = aZ;
    }
}

// CM_acc as vector_field forms it: the per-body reaction terms added in body
// order, one thread (physics.c:141-146 accumulates CM_acc body by body).
__global__ void k_cm_seq(const double *__restrict__ cterm, int n, const int32_t *__restrict__ ex,
                         const double *__restrict__ dcen, double *__restrict__ cm)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double a = 0.0, c = 0.0, d = 0.0;
    if (dcen[7] != 0.0)
        for (int k = 0; k < n; ++k) {
            if (ex && !ex[k]) continue;
            a += cterm[3 * (size_t)k];
            c += cterm[3 * (size_t)k + 1];
            d += cterm[3 * (size_t)k + 2];
        }
    cm[0] = a;
    cm[1] = c;
    cm[2] = d;
}

}  // namespace

// Launches the exact kernel; cterm (3n) receives the per-body central
// reaction terms when central->present.
ncg_status ncg_grav_exact_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                 const ncg_central *c, double G, double soft, double *d_acc,
                                 double *d_cterm, hipStream_t st)
{
    const int cp = (c && c->present) ? 1 : 0;
    const double mu = cp ? G * c->mass : 0.0;  // physics.c:131
    hipLaunchKernelGGL(k_grav_exact, dim3(cdiv(n, 256)), dim3(256), 0, st, d_b, (int)n, d_ex, cp,
                       cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0, mu,
                       (const double *)nullptr, G, soft, d_acc, d_cterm);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// The same with the central body in a device record (8 doubles), and
// CM_acc summed on the device in the reference's order.
ncg_status ncg_grav_exact_dev(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                              const double *d_cen, double G, double soft, double *d_acc,
                              double *d_cterm, double *d_cm_acc, hipStream_t st)
{
    {
        ProfScope ps(ctx, st, "exact");
        hipLaunchKernelGGL(k_grav_exact, dim3(cdiv(n, 256)), dim3(256), 0, st, d_b, (int)n, d_ex,
                           0, 0.0, 0.0, 0.0, 0.0, d_cen, G, soft, d_acc, d_cterm);
        NCG_CHECK_LAUNCH();
    }
    if (d_cm_acc) {
        hipLaunchKernelGGL(k_cm_seq, dim3(1), dim3(64), 0, st, d_cterm, (int)n, d_ex, d_cen,
                           d_cm_acc);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}
