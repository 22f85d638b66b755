// Batched ensembles: many independent systems of uneven N in one launch.
//
// Every system is a body array of its own (concatenated, offsets[S+1]) with
// its own central body {x, y, z, vx, vy, vz, M, present}.  A force
// evaluation is ONE persistent launch: workgroups take work units from a
// queue (an atomic counter reset by a fill kernel, so the launch can be
// captured in a graph); a unit is (system, block of bodies) and computes those
// bodies' accelerations against every body of the system (brute-force, the
// field of vector_field physics.c:120-202 plus the central term), or -- for
// collision units -- runs the closest_approach predicate (collision.c:63-172)
// on the block's pairs (i, j > i) and appends candidates.  Units are ordered
// by decreasing cost so the persistent workgroups finish together.
//
// Forms: NCG_F64_FAST (FP64, rsqrt + one Newton step) and NCG_MIXED (FP64
// separations rounded to FP32, FP32 1/r^3 and pair force, FP32 partials over
// 64-body sub-tiles added into FP64 accumulators).
//
// RK4 (classic, as in rk4.hip) runs every stage of all systems at once; with
// collisions on, each step starts with a collision pass of every system and a
// replay kernel (one wavefront per system) that decides the candidates in the
// reference's lexicographic order with the elastic / inelastic outcome, so a
// whole step -- pass, replay, four stages -- is one captured graph.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "ncg_internal.hpp"

namespace {

constexpr int ENT = 256;  // threads per workgroup
constexpr unsigned ENS_CAND_CAP = 1u << 20;  // candidates of one pass, all systems
constexpr int ETJ = 256;  // j tile (LDS)

struct EnsUnit {
    int sys;   // system
    int i0;    // first body (local index)
    int cnt;   // bodies in the unit
    int kind;  // 0 gravity, 1 collision candidates
};

__device__ inline int next_unit(unsigned *counter)
{
    __shared__ int u;
    if (threadIdx.x == 0) u = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int r = u;
    __syncthreads();
    return r;
}

__global__ void k_zero_u32(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// One FP32 interaction of the mixed form: separation rounded from FP64,
// softening added only when SOFT.  With SELF the body at local index js is
// compared with the lane's own (li): a body against itself has r^2 = 1e-20
// without softening, so its mass is zeroed there (G m y^3 would overflow for
// G m > ~3e8 and inf * 0 poison the sum).
template <bool SOFT, bool SELF, int IPL>
__device__ inline void ens_mixed_term(const double4 p, const float2 g, int js, const int *li,
                                      const double *xi, const double *yi, const double *zi,
                                      const float *sif, float *fx, float *fy, float *fz)
{
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const float gm = (SELF && js == li[q]) ? 0.f : g.x;
        const float dx = (float)(p.x - xi[q]);
        const float dy = (float)(p.y - yi[q]);
        const float dz = (float)(p.z - zi[q]);
        // 1e-20 in the chain keeps a self pair's y finite; it is below half
        // an ulp of any r^2 above ~2e-13.
        float r2;
        if (SOFT) {
            const float e = sif[q] + g.y;
            r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, fmaf(e, e, 1e-20f))));
        } else {
            r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, 1e-20f)));
        }
        const float y = __builtin_amdgcn_rsqf(r2);
        const float s3 = (gm * y) * (y * y);
        fx[q] = fmaf(s3, dx, fx[q]);
        fy[q] = fmaf(s3, dy, fy[q]);
        fz[q] = fmaf(s3, dz, fz[q]);
    }
}

// Bodies j0 .. j1-1 (at most 64) of the staged tile (first local index jb)
// into FP32 partials.
template <bool SOFT, bool SELF, int IPL>
__device__ inline void ens_mixed_run(const double4 *tp, const float2 *tf, int j0, int j1, int jb,
                                     const int *li, const double *xi, const double *yi,
                                     const double *zi, const float *sif, float *fx, float *fy,
                                     float *fz)
{
    if (!SELF && j1 - j0 == 64) {
#pragma unroll 4
        for (int jj = j0; jj < j0 + 64; ++jj)
            ens_mixed_term<SOFT, false, IPL>(tp[jj], tf[jj], jb + jj, li, xi, yi, zi, sif, fx, fy,
                                             fz);
    } else {
        for (int jj = j0; jj < j1; ++jj)
            ens_mixed_term<SOFT, SELF, IPL>(tp[jj], tf[jj], jb + jj, li, xi, yi, zi, sif, fx, fy,
                                            fz);
    }
}

template <int MODE, int IPL>
__global__ __launch_bounds__(ENT) void k_ens_pass(const EnsUnit *__restrict__ units, int nunits,
                                                  unsigned *__restrict__ counter,
                                                  const DBody *__restrict__ b,
                                                  const int *__restrict__ off,
                                                  const double *__restrict__ cen, double G,
                                                  double soft, double dt,
                                                  double *__restrict__ acc,
                                                  double *__restrict__ cpart,
                                                  int2 *__restrict__ cand, unsigned cand_cap,
                                                  unsigned *__restrict__ cand_cnt)
{
    __shared__ double4 tp[ETJ];  // x, y, z, G m
    __shared__ double ts[ETJ];   // s R
    __shared__ float2 tf[ETJ];   // G m, s R in FP32 (mixed form)
    __shared__ double4 tv[ETJ];  // vx, vy, vz, R (collision units)
    __shared__ double red[3][ENT];
    while (true) {
        const int u = next_unit(counter);
        if (u >= nunits) break;
        const EnsUnit U = units[u];
        const int base = off[U.sys];
        const int ns = off[U.sys + 1] - base;
        if (U.kind == 0) {
            double xi[IPL], yi[IPL], zi[IPL], si[IPL], ax[IPL], ay[IPL], az[IPL];
            int li[IPL];
#pragma unroll
            for (int q = 0; q < IPL; ++q) {
                li[q] = U.i0 + (int)threadIdx.x + q * ENT;
                const int i = base + min(li[q], ns - 1);
                xi[q] = b[i].x;
                yi[q] = b[i].y;
                zi[q] = b[i].z;
                si[q] = soft * b[i].radius;
                ax[q] = ay[q] = az[q] = 0.0;
            }
            for (int jb = 0; jb < ns; jb += ETJ) {
                const int cnt = min(ETJ, ns - jb);
                __syncthreads();
                if ((int)threadIdx.x < cnt) {
                    const DBody q = b[base + jb + threadIdx.x];
                    tp[threadIdx.x] = make_double4(q.x, q.y, q.z, G * q.mass);
                    ts[threadIdx.x] = soft * q.radius;
                    tf[threadIdx.x] = make_float2((float)(G * q.mass), (float)(soft * q.radius));
                }
                __syncthreads();
                // The tile holds some of the unit's own bodies: their self
                // terms are zeroed (r^2 = 1e-200 without softening gives
                // y^3 = 1e300, which G m > ~1.8e8 would overflow).
                const bool diag = jb < U.i0 + ENT * IPL && jb + cnt > U.i0;
                if (MODE == 0) {
                    for (int jj = 0; jj < cnt; ++jj) {
                        const double4 p = tp[jj];
                        const double sj = ts[jj];
#pragma unroll
                        for (int q = 0; q < IPL; ++q) {
                            const double dx = p.x - xi[q], dy = p.y - yi[q], dz = p.z - zi[q];
                            const double e = si[q] + sj;
                            const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, fma(e, e, 1e-200))));
                            double y = __builtin_amdgcn_rsq(r2);
                            y = y * fma(-0.5 * r2, y * y, 1.5);
                            const double w = (diag && jb + jj == li[q]) ? 0.0 : p.w;
                            const double s3 = w * y * (y * y);
                            ax[q] = fma(s3, dx, ax[q]);
                            ay[q] = fma(s3, dy, ay[q]);
                            az[q] = fma(s3, dz, az[q]);
                        }
                    }
                } else {
                    float sif[IPL];
#pragma unroll
                    for (int q = 0; q < IPL; ++q) sif[q] = (float)si[q];
                    for (int j0 = 0; j0 < cnt; j0 += 64) {
                        const int j1 = min(cnt, j0 + 64);
                        float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
                        for (int q = 0; q < IPL; ++q) fx[q] = fy[q] = fz[q] = 0.f;
                        if (diag) {
                            if (soft != 0.0)
                                ens_mixed_run<true, true, IPL>(tp, tf, j0, j1, jb, li, xi, yi, zi,
                                                               sif, fx, fy, fz);
                            else
                                ens_mixed_run<false, true, IPL>(tp, tf, j0, j1, jb, li, xi, yi, zi,
                                                                sif, fx, fy, fz);
                        } else if (soft != 0.0) {
                            ens_mixed_run<true, false, IPL>(tp, tf, j0, j1, jb, li, xi, yi, zi,
                                                            sif, fx, fy, fz);
                        } else {
                            ens_mixed_run<false, false, IPL>(tp, tf, j0, j1, jb, li, xi, yi, zi,
                                                             sif, fx, fy, fz);
                        }
#pragma unroll
                        for (int q = 0; q < IPL; ++q) {
                            ax[q] += (double)fx[q];
                            ay[q] += (double)fy[q];
                            az[q] += (double)fz[q];
                        }
                    }
                }
            }
            // Central term (physics.c:127-146) and this unit's CM_acc partial.
            const double *c = cen + 8 * (size_t)U.sys;
            const bool cp = c[7] != 0.0;
            double cx = 0., cy = 0., cz = 0.;
#pragma unroll
            for (int q = 0; q < IPL; ++q) {
                if (li[q] >= ns || li[q] >= U.i0 + U.cnt) continue;
                const int i = base + li[q];
                if (cp) {
                    const double DX = xi[q] - c[0], DY = yi[q] - c[1], DZ = zi[q] - c[2];
                    const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
                    const double r3 = rk * rk * rk;
                    const double mu = G * c[6];
                    ax[q] += -mu * DX / r3;
                    ay[q] += -mu * DY / r3;
                    az[q] += -mu * DZ / r3;
                    const double gm = G * b[i].mass;
                    cx += gm * DX / r3;
                    cy += gm * DY / r3;
                    cz += gm * DZ / r3;
                }
                acc[3 * (size_t)i + 0] = ax[q];
                acc[3 * (size_t)i + 1] = ay[q];
                acc[3 * (size_t)i + 2] = az[q];
            }
            red[0][threadIdx.x] = cx;
            red[1][threadIdx.x] = cy;
            red[2][threadIdx.x] = cz;
            __syncthreads();
            for (int w = ENT / 2; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w)
                    for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0)
                for (int d = 0; d < 3; ++d) cpart[3 * (size_t)u + d] = red[d][0];
        } else {
            // Collision candidates among pairs (i, j > i), i in this unit.
            const int li = U.i0 + (int)threadIdx.x;
            const bool mine = (int)threadIdx.x < U.cnt && li < ns;
            DBody a;
            if (mine) a = b[base + li];
            for (int jb = U.i0; jb < ns; jb += ETJ) {
                const int cnt = min(ETJ, ns - jb);
                __syncthreads();
                if ((int)threadIdx.x < cnt) {
                    const DBody q = b[base + jb + threadIdx.x];
                    tp[threadIdx.x] = make_double4(q.x, q.y, q.z, 0.0);
                    tv[threadIdx.x] = make_double4(q.vx, q.vy, q.vz, q.radius);
                }
                __syncthreads();
                if (!mine) continue;
                for (int jj = 0; jj < cnt; ++jj) {
                    const int lj = jb + jj;
                    if (lj <= li) continue;
                    const double4 p = tp[jj], v = tv[jj];
                    // closest_approach's predicate, collision.c:112-137 (this file is
                    // compiled without FMA contraction).
                    const double dx = a.x - p.x, dy = a.y - p.y, dz = a.z - p.z;
                    const double dvx = a.vx - v.x, dvy = a.vy - v.y, dvz = a.vz - v.z;
                    const double bb = dx * dvx + dy * dvy + dz * dvz;
                    if (!(bb < 0.0)) continue;
                    const double dr2 = dx * dx + dy * dy + dz * dz;
                    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
                    const double R = a.radius + v.w;
                    const double disc = bb * bb + dv2 * (R * R - dr2);
                    if (!(disc > 0.0)) continue;
                    const double t = -(bb + sqrt(disc)) / dv2;
                    if (t <= dt + fabs(dt) * 9.094947017729282e-13) {
                        const unsigned slot = atomicAdd(cand_cnt, 1u);
                        if (slot < cand_cap) cand[slot] = make_int2(base + li, base + lj);
                    }
                }
            }
        }
    }
}

// Per-system CM_acc: the partials of the system's gravity units in unit order.
__global__ void k_ens_cm(const int *__restrict__ ufirst, const int *__restrict__ ulist,
                         const double *__restrict__ cpart, int nsys, double *__restrict__ cm)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    double a[3] = {0., 0., 0.};
    for (int q = ufirst[s]; q < ufirst[s + 1]; ++q)
        for (int d = 0; d < 3; ++d) a[d] += cpart[3 * (size_t)ulist[q] + d];
    for (int d = 0; d < 3; ++d) cm[3 * (size_t)s + d] = a[d];
}

// RK4 stage update of every body (as rk4.hip) and, one thread per system, of
// the central bodies.
__global__ void k_ens_stage(int s, double dt, int ntot, DBody *__restrict__ y0,
                            DBody *__restrict__ Y, const double *__restrict__ acc,
                            double *__restrict__ sx, double *__restrict__ sv, int nsys,
                            double *__restrict__ c0, double *__restrict__ Yc,
                            const double *__restrict__ cm, double *__restrict__ sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double w = (s == 0 || s == 3) ? 1.0 : 2.0;
    const double h = s < 2 ? 0.5 * dt : dt;
    if (i < ntot) {
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
    } else if (i < ntot + nsys) {
        const int q = i - ntot;
        double *c = c0 + 8 * (size_t)q;
        if (c[7] == 0.0) return;
        double *yc = Yc + 8 * (size_t)q;
        double *scq = sc + 6 * (size_t)q;
        const double *cs = s == 0 ? c : yc;
        const double *ca = cm + 3 * (size_t)q;
        double kx[3], kv[3];
        for (int d = 0; d < 3; ++d) {
            kx[d] = s == 0 ? cs[3 + d] : scq[d] + w * cs[3 + d];
            kv[d] = s == 0 ? ca[d] : scq[3 + d] + w * ca[d];
        }
        if (s < 3) {
            for (int d = 0; d < 3; ++d) {
                scq[d] = kx[d];
                scq[3 + d] = kv[d];
                yc[d] = c[d] + h * cs[3 + d];
                yc[3 + d] = c[3 + d] + h * ca[d];
            }
            yc[6] = c[6];
            yc[7] = c[7];
        } else {
            const double f = dt / 6.0;
            for (int d = 0; d < 3; ++d) {
                c[d] += f * kx[d];
                c[3 + d] += f * kv[d];
            }
        }
    }
}

}  // namespace

// ---- host side ------------------------------------------------------------

struct EnsPlan {
    int nsys = 0, ntot = 0, ipl = 1;
    int wg_per_cu = 8;  // persistent workgroups per CU
    std::vector<int> off;
    std::vector<EnsUnit> grav, coll;  // unit lists (cost-ordered)
};

static void make_units(EnsPlan &P, int ipl)
{
    const int per = ENT * ipl;
    P.grav.clear();
    P.coll.clear();
    for (int s = 0; s < P.nsys; ++s) {
        const int ns = P.off[s + 1] - P.off[s];
        for (int i0 = 0; i0 < ns; i0 += per) P.grav.push_back({s, i0, std::min(per, ns - i0), 0});
        for (int i0 = 0; i0 < ns; i0 += ENT) P.coll.push_back({s, i0, std::min(ENT, ns - i0), 1});
    }
    auto cost = [&](const EnsUnit &u) {
        const double ns = P.off[u.sys + 1] - P.off[u.sys];
        return u.kind == 0 ? ns * u.cnt * 1.0 : 0.35 * u.cnt * (ns - u.i0 - 0.5 * u.cnt);
    };
    auto by_cost = [&](const EnsUnit &a, const EnsUnit &b) { return cost(a) > cost(b); };
    std::stable_sort(P.grav.begin(), P.grav.end(), by_cost);
    std::stable_sort(P.coll.begin(), P.coll.end(), by_cost);
}

// Gravity-unit index lists per system for a unit array `units`.
static void unit_index(const EnsPlan &P, const std::vector<EnsUnit> &units, std::vector<int> &first,
                       std::vector<int> &list)
{
    std::vector<std::vector<int>> per(P.nsys);
    for (int q = 0; q < (int)units.size(); ++q)
        if (units[q].kind == 0) per[units[q].sys].push_back(q);
    first.assign(P.nsys + 1, 0);
    list.clear();
    for (int s = 0; s < P.nsys; ++s) {
        std::sort(per[s].begin(), per[s].end(), [&](int a, int b) { return units[a].i0 < units[b].i0; });
        first[s + 1] = first[s] + (int)per[s].size();
        list.insert(list.end(), per[s].begin(), per[s].end());
    }
}

struct EnsDev {
    EnsUnit *grav = nullptr, *coll = nullptr;
    int *gfirst = nullptr, *glist = nullptr;
    int *off = nullptr;
};

template <int MODE>
static void launch_pass(int ipl, int nblk, hipStream_t st, const EnsUnit *units, int nunits,
                        unsigned *counter, const DBody *b, const int *off, const double *cen,
                        double G, double soft, double dt, double *acc, double *cpart, int2 *cand,
                        unsigned cap, unsigned *ccnt)
{
    if (ipl == 2)
        hipLaunchKernelGGL((k_ens_pass<MODE, 2>), dim3(nblk), dim3(ENT), 0, st, units, nunits,
                           counter, b, off, cen, G, soft, dt, acc, cpart, cand, cap, ccnt);
    else
        hipLaunchKernelGGL((k_ens_pass<MODE, 1>), dim3(nblk), dim3(ENT), 0, st, units, nunits,
                           counter, b, off, cen, G, soft, dt, acc, cpart, cand, cap, ccnt);
}

// One pass over the unit list du[nunits] (gravity and/or collision units);
// with gravity units, the per-system CM_acc from the index lists.
static ncg_status ens_pass(ncg_ctx *ctx, const EnsPlan &P, const int *doff, const EnsUnit *du,
                           int nunits, const int *dfirst, const int *dlist, const DBody *b,
                           const double *cen, double G, double soft, ncg_precision prec,
                           double dt, double *acc, double *cm, hipStream_t st)
{
    unsigned *counter = dbuf<unsigned>(ctx, "e_counter", 2);
    double *cpart = dbuf<double>(ctx, "e_cpart", 3 * (size_t)std::max(nunits, 1));
    const unsigned cap = ncg_replay_cap(ctx, ENS_CAND_CAP);
    int2 *cand = dbuf<int2>(ctx, "e_cand", cap);
    if (!counter || !cpart || !cand) return NCG_ERR_NOMEM;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, counter, 2);
    if (nunits == 0) return NCG_OK;
    const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
    const int nblk = std::min(nunits, P.wg_per_cu * cus);
    {
        ProfScope ps(ctx, st, prec == NCG_MIXED ? "ens_mixed" : "ens_fast64");
        if (prec == NCG_MIXED)
            launch_pass<1>(P.ipl, nblk, st, du, nunits, counter, b, doff, cen, G, soft, dt, acc,
                           cpart, cand, cap, counter + 1);
        else
            launch_pass<0>(P.ipl, nblk, st, du, nunits, counter, b, doff, cen, G, soft, dt, acc,
                           cpart, cand, cap, counter + 1);
        NCG_CHECK_LAUNCH();
    }
    if (dfirst) {
        hipLaunchKernelGGL(k_ens_cm, dim3(cdiv(P.nsys, 128)), dim3(128), 0, st, dfirst, dlist,
                           cpart, P.nsys, cm);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}

// One evaluation of all systems.
static ncg_status ens_eval(ncg_ctx *ctx, const EnsPlan &P, const EnsDev &D, const DBody *b,
                           const double *cen, double G, double soft, ncg_precision prec,
                           double *acc, double *cm, hipStream_t st)
{
    return ens_pass(ctx, P, D.off, D.grav, (int)P.grav.size(), D.gfirst, D.glist, b, cen, G, soft,
                    prec, 0.0, acc, cm, st);
}

extern "C" ncg_status ncg_ensemble_rk4_dev(ncg_ctx *ctx, ncg_body *d_bodies,
                                           const int32_t *offsets, int32_t nsys,
                                           double *d_centrals, double grav_const,
                                           double softening, ncg_precision prec, double dt,
                                           int32_t nsteps, int32_t flags, double collision_f,
                                           int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(nsys >= 1 && offsets && d_bodies && d_centrals && nsteps >= 0, "invalid arguments");
    NCG_REQUIRE(prec == NCG_MIXED || prec == NCG_F64_FAST,
                "the ensemble runs the fast64 or mixed form");
    for (int s = 0; s < nsys; ++s) NCG_REQUIRE(offsets[s + 1] >= offsets[s], "offsets not monotone");
    NCG_REQUIRE(offsets[0] == 0, "offsets[0] must be 0");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    const bool use_graph = (flags & 1) != 0;
    const bool collide = (flags & 2) != 0;
    if (n_collisions) *n_collisions = 0;

    EnsPlan P;
    P.nsys = nsys;
    P.off.assign(offsets, offsets + nsys + 1);
    P.ntot = P.off[nsys];
    const TiledCfg tc = ncg_tiled_cfg(ctx, prec == NCG_MIXED ? 100 + NCG_MIXED : 100 + NCG_F64_FAST,
                                      P.ntot / std::max(nsys, 1));
    P.ipl = tc.ipl == 2 ? 2 : 1;
    // For the ensemble forms the table's split column is workgroups per CU.
    P.wg_per_cu = tc.nsplit > 0 ? std::min(tc.nsplit, 32) : 8;
    make_units(P, P.ipl);
    std::vector<int> gfirst, glist;
    unit_index(P, P.grav, gfirst, glist);

    const int ntot = P.ntot;
    EnsDev D;
    D.grav = dbuf<EnsUnit>(ctx, "e_grav", P.grav.size());
    D.coll = dbuf<EnsUnit>(ctx, "e_coll", P.coll.size());
    D.gfirst = dbuf<int>(ctx, "e_gfirst", nsys + 1);
    D.glist = dbuf<int>(ctx, "e_glist", std::max<size_t>(glist.size(), 1));
    D.off = dbuf<int>(ctx, "e_off", nsys + 1);
    DBody *Y = dbuf<DBody>(ctx, "e_Y", ntot);
    double *acc = dbuf<double>(ctx, "e_acc", 3 * (size_t)ntot);
    double *sx = dbuf<double>(ctx, "e_sx", 3 * (size_t)ntot);
    double *sv = dbuf<double>(ctx, "e_sv", 3 * (size_t)ntot);
    double *Yc = dbuf<double>(ctx, "e_Yc", 8 * (size_t)nsys);
    double *scen = dbuf<double>(ctx, "e_sc", 6 * (size_t)nsys);
    double *cm = dbuf<double>(ctx, "e_cm", 3 * (size_t)nsys);
    // Every buffer a captured step uses exists before the capture.
    unsigned *counter = dbuf<unsigned>(ctx, "e_counter", 2);
    double *cpart = dbuf<double>(ctx, "e_cpart", 3 * std::max(P.grav.size(), P.coll.size()) + 3);
    int2 *cand = dbuf<int2>(ctx, "e_cand", ENS_CAND_CAP);
    unsigned char *dirty = dbuf<unsigned char>(ctx, "e_dirty", (size_t)ntot + 4);
    unsigned *stat = dbuf<unsigned>(ctx, "e_stat", 2);  // collisions, error bits
    if (!D.grav || !D.coll || !D.gfirst || !D.glist || !D.off || !Y || !acc || !sx || !sv ||
        !Yc || !scen || !cm || !counter || !cpart || !cand || !dirty || !stat)
        return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(D.grav, P.grav.data(), sizeof(EnsUnit) * P.grav.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.coll, P.coll.data(), sizeof(EnsUnit) * P.coll.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.gfirst, gfirst.data(), sizeof(int) * gfirst.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.glist, glist.data(), sizeof(int) * glist.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.off, P.off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, stat, 2);
    hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(ntot / 4 + 1, 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned *>(dirty), ntot / 4 + 1);
    NCG_CHECK_LAUNCH();
    DBody *y0 = reinterpret_cast<DBody *>(d_bodies);
    const unsigned sgrid = cdiv(ntot + nsys, 256);
    const double G = grav_const, soft = softening;

    // The replay's capacity (candidates of a pass, per system: candidates and
    // changed bodies in LDS) is checked every ENS_CHECK steps.  A window in
    // which a pass exceeded it runs again from a snapshot taken at its start,
    // step by step with each collision pass checked at once: a pass that
    // exceeds the capacity is undone and every system's pass is replayed on
    // the host with the same decision and bounce (ncg_brute_force_pass with
    // the built-in ops), so results stay those of an unbounded replay.
    constexpr int ENS_CHECK = 16;
    DBody *snap_b = collide ? dbuf<DBody>(ctx, "e_snap_b", ntot) : nullptr;
    DBody *pass_b = collide ? dbuf<DBody>(ctx, "e_pass_b", ntot) : nullptr;
    double *snap_c = collide ? dbuf<double>(ctx, "e_snap_c", 8 * (size_t)nsys) : nullptr;
    unsigned *hst = pbuf<unsigned>(ctx, "e_hstat", 4);
    if ((collide && (!snap_b || !pass_b || !snap_c)) || !hst) return NCG_ERR_NOMEM;
    int64_t host_coll = 0;
    unsigned snap_coll = 0;
    auto read_stat = [&](unsigned *hs) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(hst, stat, sizeof(unsigned) * 2, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        hs[0] = hst[0];
        hs[1] = hst[1];
        return NCG_OK;
    };
    auto write_stat = [&](unsigned coll) -> ncg_status {
        hst[2] = coll;
        hst[3] = 0u;
        NCG_HIP(hipMemcpyAsync(stat, hst + 2, sizeof(unsigned) * 2, hipMemcpyHostToDevice, st));
        // A replay that ran out of room for changed bodies leaves marks set.
        hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(ntot / 4 + 1, 256)), dim3(256), 0, st,
                           reinterpret_cast<unsigned *>(dirty), ntot / 4 + 1);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    auto device_pass = [&]() -> ncg_status {
        ncg_status r = ens_pass(ctx, P, D.off, D.coll, (int)P.coll.size(), nullptr, nullptr, y0,
                                d_centrals, G, soft, prec, dt, acc, cm, st);
        if (r != NCG_OK) return r;
        return ncg_bf_replay_launch(ctx, y0, nullptr, nullptr, 0, D.off, nsys, cand, counter + 1,
                                    nullptr, ncg_replay_cap(ctx, ENS_CAND_CAP), dt, collision_f,
                                    dirty, stat, st);
    };
    auto checked_pass = [&]() -> ncg_status {
        unsigned hs[2];
        ncg_status r = read_stat(hs);
        if (r != NCG_OK) return r;
        const unsigned before = hs[0];
        NCG_HIP(hipMemcpyAsync(pass_b, y0, sizeof(DBody) * (size_t)ntot, hipMemcpyDeviceToDevice, st));
        if ((r = device_pass()) != NCG_OK) return r;
        if ((r = read_stat(hs)) != NCG_OK) return r;
        if (hs[1] == 0u) return NCG_OK;
        NCG_HIP(hipMemcpyAsync(y0, pass_b, sizeof(DBody) * (size_t)ntot, hipMemcpyDeviceToDevice, st));
        if ((r = write_stat(before)) != NCG_OK) return r;
        std::vector<DBody> hb((size_t)ntot);
        NCG_HIP(hipMemcpyAsync(hb.data(), y0, sizeof(DBody) * (size_t)ntot, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        for (int sy = 0; sy < nsys; ++sy) {
            const int ns = P.off[sy + 1] - P.off[sy];
            if (ns < 2) continue;
            ncg_builtin_collider bc = {};
            bc.bodies = reinterpret_cast<ncg_body *>(hb.data() + P.off[sy]);
            bc.timestep = dt;
            bc.collision_f = collision_f;
            ncg_replay_ops ops;
            if ((r = ncg_builtin_replay_ops(&bc, &ops)) != NCG_OK) return r;
            if ((r = ncg_brute_force_pass(ctx, bc.bodies, ns, nullptr, dt, &ops, nullptr)) != NCG_OK)
                return r;
            host_coll += bc.collisions;
        }
        NCG_HIP(hipMemcpyAsync(y0, hb.data(), sizeof(DBody) * (size_t)ntot, hipMemcpyHostToDevice, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    // One step: with collisions, the brute-force pass of every system and its
    // replay on the device, then the four RK4 stages.
    auto step_body = [&](bool checked) -> ncg_status {
        if (collide) {
            ncg_status r = checked ? checked_pass() : device_pass();
            if (r != NCG_OK) return r;
        }
        for (int s = 0; s < 4; ++s) {
            ncg_status r = ens_eval(ctx, P, D, s == 0 ? y0 : Y, s == 0 ? d_centrals : Yc, G, soft,
                                    prec, acc, cm, st);
            if (r != NCG_OK) return r;
            hipLaunchKernelGGL(k_ens_stage, dim3(sgrid), dim3(256), 0, st, s, dt, ntot, y0, Y, acc,
                               sx, sv, nsys, d_centrals, Yc, cm, scen);
            NCG_CHECK_LAUNCH();
        }
        return NCG_OK;
    };
    auto snapshot = [&]() -> ncg_status {
        unsigned hs[2];
        ncg_status r = read_stat(hs);
        if (r != NCG_OK) return r;
        snap_coll = hs[0];
        NCG_HIP(hipMemcpyAsync(snap_b, y0, sizeof(DBody) * (size_t)ntot, hipMemcpyDeviceToDevice, st));
        NCG_HIP(hipMemcpyAsync(snap_c, d_centrals, sizeof(double) * 8 * (size_t)nsys,
                               hipMemcpyDeviceToDevice, st));
        return NCG_OK;
    };

    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    struct GraphGuard {
        hipGraphExec_t &e;
        hipGraph_t &g;
        ~GraphGuard()
        {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } guard{exec, graph};
    if (collide) {
        ncg_status r = snapshot();
        if (r != NCG_OK) return r;
    }
    int32_t win0 = 0;
    for (int32_t step = 0; step < nsteps; ++step) {
        if (!use_graph || step == 0) {
            // Eager (the first step also allocates everything the graph uses).
            ncg_status r = step_body(false);
            if (r != NCG_OK) return r;
        } else {
            if (!exec) {
                NCG_HIP(hipStreamSynchronize(st));
                ctx->capturing = true;
                NCG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                ncg_status r = step_body(false);
                hipError_t e = hipStreamEndCapture(st, &graph);
                ctx->capturing = false;
                if (r != NCG_OK) return r;
                if (e != hipSuccess)
                    return ncg_fail(ctx, NCG_ERR_DEVICE, "stream capture failed: %s",
                                    hipGetErrorString(e));
                if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
                    return ncg_fail(ctx, NCG_ERR_DEVICE, "graph instantiation failed");
            }
            NCG_HIP(hipGraphLaunch(exec, st));
        }
        if (!collide || ((step + 1) % ENS_CHECK != 0 && step + 1 != nsteps)) continue;
        unsigned hs[2];
        ncg_status r = read_stat(hs);
        if (r != NCG_OK) return r;
        if (hs[1] != 0u) {
            // The window [win0, step] again, from its snapshot, checked.
            NCG_HIP(hipMemcpyAsync(y0, snap_b, sizeof(DBody) * (size_t)ntot,
                                   hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemcpyAsync(d_centrals, snap_c, sizeof(double) * 8 * (size_t)nsys,
                                   hipMemcpyDeviceToDevice, st));
            if ((r = write_stat(snap_coll)) != NCG_OK) return r;
            for (int32_t q = win0; q <= step; ++q)
                if ((r = step_body(true)) != NCG_OK) return r;
        }
        if ((r = snapshot()) != NCG_OK) return r;
        win0 = step + 1;
    }
    unsigned hs[2] = {0, 0};
    NCG_HIP(hipMemcpyAsync(hs, stat, sizeof hs, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    if (n_collisions) *n_collisions = hs[0] + host_coll;
    if (hs[1])
        return ncg_fail(ctx, NCG_ERR_CAPACITY,
                        "collision replay capacity exceeded (flags %u: 1 = more than %u candidates "
                        "in one pass, 2 = more than %d in one system, 4 = more than %d changed "
                        "bodies in one system); results are not valid",
                        hs[1], ncg_replay_cap(ctx, ENS_CAND_CAP), NCG_REPLAY_RCAP,
                        NCG_REPLAY_DCAP);
    return NCG_OK;
}

// One evaluation of every system (accelerations and per-system CM_acc), for
// checking the ensemble pass against single-system runs.
extern "C" ncg_status ncg_ensemble_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies,
                                               const int32_t *offsets, int32_t nsys,
                                               const double *d_centrals, double grav_const,
                                               double softening, ncg_precision prec,
                                               double *d_acc, double *d_cm_acc, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(nsys >= 1 && offsets && d_bodies && d_centrals && d_acc && d_cm_acc,
                "invalid arguments");
    NCG_REQUIRE(prec == NCG_MIXED || prec == NCG_F64_FAST,
                "the ensemble runs the fast64 or mixed form");
    NCG_REQUIRE(offsets[0] == 0, "offsets[0] must be 0");
    for (int s = 0; s < nsys; ++s) NCG_REQUIRE(offsets[s + 1] >= offsets[s], "offsets not monotone");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    EnsPlan P;
    P.nsys = nsys;
    P.off.assign(offsets, offsets + nsys + 1);
    P.ntot = P.off[nsys];
    const TiledCfg tc = ncg_tiled_cfg(ctx, prec == NCG_MIXED ? 100 + NCG_MIXED : 100 + NCG_F64_FAST,
                                      P.ntot / std::max(nsys, 1));
    P.ipl = tc.ipl == 2 ? 2 : 1;
    // For the ensemble forms the table's split column is workgroups per CU.
    P.wg_per_cu = tc.nsplit > 0 ? std::min(tc.nsplit, 32) : 8;
    make_units(P, P.ipl);
    std::vector<int> gfirst, glist;
    unit_index(P, P.grav, gfirst, glist);
    EnsDev D;
    D.grav = dbuf<EnsUnit>(ctx, "e_grav", P.grav.size());
    D.gfirst = dbuf<int>(ctx, "e_gfirst", nsys + 1);
    D.glist = dbuf<int>(ctx, "e_glist", std::max<size_t>(glist.size(), 1));
    D.off = dbuf<int>(ctx, "e_off", nsys + 1);
    if (!D.grav || !D.gfirst || !D.glist || !D.off) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(D.grav, P.grav.data(), sizeof(EnsUnit) * P.grav.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.gfirst, gfirst.data(), sizeof(int) * gfirst.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.glist, glist.data(), sizeof(int) * glist.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.off, P.off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, st));
    return ens_eval(ctx, P, D, reinterpret_cast<const DBody *>(d_bodies), d_centrals, grav_const,
                    softening, prec, d_acc, d_cm_acc, st);
}
