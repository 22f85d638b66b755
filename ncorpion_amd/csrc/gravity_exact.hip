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
#include "exact_div.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int EX_T = 256;

// physics.c:186-201 for body k at (X, Y, Z) and body j: D = sqrt(DX*DX +
// DY*DY + DZ*DZ + soft*soft), D3 = D*D*D, term G m_j DX / D3, with DX formed
// as the reference forms it for j < k (X_k - X_j) and j > k (X_j - X_k).
__device__ inline void exact_term(double X, double Y, double Z, double Rk, double xj, double yj,
                                  double zj, double Rj, double gmj, double s, bool lower,
                                  double &tX, double &tY, double &tZ)
{
    const double dx = lower ? X - xj : xj - X;
    const double dy = lower ? Y - yj : yj - Y;
    const double dz = lower ? Z - zj : zj - Z;
    const double soft = s * (lower ? (Rk + Rj) : (Rj + Rk));
    const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
    const double D3 = D * D * D;
    {
        double q0, q1, q2;
        div3(gmj * dx, gmj * dy, gmj * dz, D3, q0, q1, q2);
        tX = q0;
        tY = q1;
        tZ = q2;
    }
}

// Central body from the host (cp, cx, cy, cz, mu) or, when dcen is given,
// from a device record {x, y, z, vx, vy, vz, M, present}.
__global__ __launch_bounds__(256) void k_grav_exact(const DBody *__restrict__ b, int n,
                                                    const int32_t *__restrict__ ex, int cp,
                                                    double cx, double cy, double cz, double mu,
                                                    const double *__restrict__ dcen,
                                                    double G, double s,
                                                    const double *__restrict__ init,
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
        if (init) {  // the caller's start value (physics.c:121-172 beyond the point mass)
            aX = init[3 * (size_t)k];
            aY = init[3 * (size_t)k + 1];
            aZ = init[3 * (size_t)k + 2];
        } else if (cp) {
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
        // Terms of four j at a time are formed independently (their square
        // roots and divisions overlap), then added in j order: the sum is the
        // reference's, skipped terms leave the accumulators untouched.
        int jj = 0;
        for (; jj + 4 <= cnt; jj += 4) {
            double tX[4], tY[4], tZ[4];
            bool use[4], low[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = base + jj + u;
                use[u] = te[jj + u] && j != k;
                low[u] = j < k;
                exact_term(X, Y, Z, Rk, tx[jj + u], ty[jj + u], tz[jj + u], tr[jj + u],
                           tgm[jj + u], s, low[u], tX[u], tY[u], tZ[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!use[u]) continue;
                if (low[u]) {
                    aX -= tX[u];
                    aY -= tY[u];
                    aZ -= tZ[u];
                } else {
                    aX += tX[u];
                    aY += tY[u];
                    aZ += tZ[u];
                }
            }
        }
        for (; jj < cnt; ++jj) {
            const int j = base + jj;
            if (!te[jj] || j == k) continue;
            const bool lower = j < k;
            double tX, tY, tZ;
            exact_term(X, Y, Z, Rk, tx[jj], ty[jj], tz[jj], tr[jj], tgm[jj], s, lower, tX, tY, tZ);
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
// order (physics.c:141-146 accumulates CM_acc body by body).  Waves 1..3
// stage chunks of CMS_CH bodies' terms in LDS, a term of an absent body (or
// past n) stored as -0.0, while lanes 0, 1, 2 of wave 0 add the x, y, z terms
// of the previous chunk: a run of additions with no branch (a + (-0.0) is a
// for every a, so a skipped body leaves the sum as it was).
constexpr int CMS_CH = 1024;

// The central body's presence comes from the device record dcen (its
// present field) or, when dcen is null, from cp.
__global__ __launch_bounds__(256) void k_cm_seq(const double *__restrict__ cterm, int n,
                                                const int32_t *__restrict__ ex,
                                                const double *__restrict__ dcen, int cp,
                                                double *__restrict__ cm)
{
    // Rows padded by two doubles so that the three summing lanes read
    // different banks.
    __shared__ double st[2][3][CMS_CH + 2];
    const int tid = threadIdx.x, w = tid >> 6;
    if (dcen ? dcen[7] == 0.0 : cp == 0) {
        if (tid < 3) cm[tid] = 0.0;
        return;
    }
    const int nch = (n + CMS_CH - 1) / CMS_CH;
    auto stage = [&](int buf, int base) {
        for (int q = tid - 64; q < CMS_CH; q += 192) {
            const int k = base + q;
            const bool on = k < n && (!ex || ex[k] != 0);
            st[buf][0][q] = on ? cterm[3 * (size_t)k] : -0.0;
            st[buf][1][q] = on ? cterm[3 * (size_t)k + 1] : -0.0;
            st[buf][2][q] = on ? cterm[3 * (size_t)k + 2] : -0.0;
        }
    };
    if (w > 0 && nch > 0) stage(0, 0);
    __syncthreads();
    double a = 0.0;
    for (int c = 0; c < nch; ++c) {
        if (w > 0) {
            if (c + 1 < nch) stage((c + 1) & 1, (c + 1) * CMS_CH);
        } else if (tid < 3) {
            // Whole groups of 16 (the padding past n is -0.0).
            const int len = (min(CMS_CH, n - c * CMS_CH) + 15) & ~15;
            const double *row = st[c & 1][tid];
            for (int q = 0; q < len; q += 16) {
#pragma unroll
                for (int u = 0; u < 16; ++u) a += row[q + u];
            }
        }
        __syncthreads();
    }
    if (tid < 3) cm[tid] = a;
}

// Small n: one wavefront per body k.  The lanes form the terms of 64 bodies j
// at a time (exact_term, as k_grav_exact does), and lanes 0, 1, 2 add the x,
// y, z terms in j order: the same sums, with 64 lanes forming the terms of one
// body instead of one lane forming them all.  A term to subtract is stored
// negated and a skipped one as -0.0, so the sum is a run of additions:
// a - t and a + (-t) are one operation, and a + (-0.0) is a for every a.
constexpr int EXW_MAX = 32768;  // bodies up to which the wavefront-per-body form is used

__global__ __launch_bounds__(256) void k_grav_exact_w(const DBody *__restrict__ b, int n,
                                                      const int32_t *__restrict__ ex, int cp,
                                                      double cx, double cy, double cz, double mu,
                                                      const double *__restrict__ dcen,
                                                      double G, double s,
                                                      const double *__restrict__ init,
                                                      double *__restrict__ acc,
                                                      double *__restrict__ cterm)
{
    __shared__ double st[4][3][64];
    if (dcen) {
        cp = dcen[7] != 0.0;
        cx = dcen[0];
        cy = dcen[1];
        cz = dcen[2];
        mu = G * dcen[6];  // physics.c:131
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + w;
    if (k >= n || (ex != nullptr && ex[k] == 0)) return;  // the whole wavefront
    const DBody me = b[k];
    const double X = me.x, Y = me.y, Z = me.z, Rk = me.radius;
    double a = 0.;  // component `lane` (lanes 0..2)
    if (init) {
        if (lane < 3) a = init[3 * (size_t)k + lane];
    } else if (cp) {
        // physics.c:121-146
        const double DX = X - cx, DY = Y - cy, DZ = Z - cz;
        const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
        const double r3 = rk * rk * rk;
        const double aX = -mu * DX / r3, aY = -mu * DY / r3, aZ = -mu * DZ / r3;
        a = lane == 0 ? aX : lane == 1 ? aY : aZ;
        if (lane == 0) {
            cterm[3 * (size_t)k + 0] = G * me.mass * DX / r3;
            cterm[3 * (size_t)k + 1] = G * me.mass * DY / r3;
            cterm[3 * (size_t)k + 2] = G * me.mass * DZ / r3;
        }
    }
    // The next group's records are loaded before the current group is
    // summed, so their latency overlaps the ordered additions.
    DBody q = {};
    int e = 0;
    if (lane < n) {
        q = b[lane];
        e = ex == nullptr ? 1 : ex[lane];
    }
    for (int base = 0; base < n; base += 64) {
        const int j = base + lane;
        double tX = -0.0, tY = -0.0, tZ = -0.0;
        if (j < n && j != k && e != 0) {
            const bool lower = j < k;
            exact_term(X, Y, Z, Rk, q.x, q.y, q.z, q.radius, G * q.mass, s, lower, tX, tY, tZ);
            if (lower) {
                tX = -tX;
                tY = -tY;
                tZ = -tZ;
            }
        }
        const int jn = j + 64;
        if (jn < n) {
            q = b[jn];
            e = ex == nullptr ? 1 : ex[jn];
        }
        st[w][0][lane] = tX;
        st[w][1][lane] = tY;
        st[w][2][lane] = tZ;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < 3) {
            const double *row = st[w][lane];
            for (int g = 0; g < 64; g += 16) {
                double v[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) v[i] = row[g + i];
#pragma unroll
                for (int i = 0; i < 16; ++i) a += v[i];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    if (lane < 3) acc[3 * (size_t)k + lane] = a;
}

}  // namespace

// Launches the exact kernel; cterm (3n) receives the per-body central
// reaction terms when central->present.
ncg_status ncg_grav_exact_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                 const ncg_central *c, double G, double soft, double *d_acc,
                                 double *d_cterm, hipStream_t st, const double *d_init)
{
    const int cp = (c && c->present) ? 1 : 0;
    const double mu = cp ? G * c->mass : 0.0;  // physics.c:131
    if (n <= EXW_MAX)
        hipLaunchKernelGGL(k_grav_exact_w, dim3(cdiv(n, 4)), dim3(256), 0, st, d_b, (int)n, d_ex, cp,
                           cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0, mu,
                           (const double *)nullptr, G, soft, d_init, d_acc, d_cterm);
    else
        hipLaunchKernelGGL(k_grav_exact, dim3(cdiv(n, 256)), dim3(256), 0, st, d_b, (int)n, d_ex, cp,
                           cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0, mu,
                           (const double *)nullptr, G, soft, d_init, d_acc, d_cterm);
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
        if (n <= EXW_MAX)
            hipLaunchKernelGGL(k_grav_exact_w, dim3(cdiv(n, 4)), dim3(256), 0, st, d_b, (int)n, d_ex,
                               0, 0.0, 0.0, 0.0, 0.0, d_cen, G, soft, (const double *)nullptr, d_acc,
                               d_cterm);
        else
            hipLaunchKernelGGL(k_grav_exact, dim3(cdiv(n, 256)), dim3(256), 0, st, d_b, (int)n, d_ex,
                               0, 0.0, 0.0, 0.0, 0.0, d_cen, G, soft, (const double *)nullptr, d_acc,
                               d_cterm);
        NCG_CHECK_LAUNCH();
    }
    if (d_cm_acc) {
        hipLaunchKernelGGL(k_cm_seq, dim3(1), dim3(256), 0, st, d_cterm, (int)n, d_ex, d_cen, 0,
                           d_cm_acc);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}

// CM_acc from the per-body terms d_cterm of an exact evaluation with the
// central body given on the host (cp: present), summed in body order.
ncg_status ncg_cm_seq_launch(ncg_ctx *ctx, const double *d_cterm, int32_t n, const int32_t *d_ex,
                             int cp, double *d_cm_acc, hipStream_t st)
{
    hipLaunchKernelGGL(k_cm_seq, dim3(1), dim3(256), 0, st, d_cterm, (int)n, d_ex,
                       (const double *)nullptr, cp, d_cm_acc);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}
