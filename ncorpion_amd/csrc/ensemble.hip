// Batched ensembles: many independent systems of uneven N in one launch.
//
// Every system is a body array of its own (concatenated, offsets[S+1]) with
// its own central body {x, y, z, vx, vy, vz, M, present}.  A force
// evaluation is ONE launch of one workgroup per work unit: a unit is (system,
// block of 256 bodies) and computes those bodies' accelerations against every
// body of the system (brute-force, the field of vector_field
// physics.c:120-202 plus the central term), or -- in a collision pass -- runs
// the closest_approach predicate (collision.c:63-172) on the block's pairs
// and appends candidates.  Units are ordered by decreasing cost, so the
// dispatcher starts the large ones first and the workgroups finish together.
//
// Order (once per step, outside captured graphs): each system's bodies are
// put along a Hilbert curve over the system's bounding cube -- one radix sort
// of (system << 30 | key) for all systems -- and laid out in padded slots:
// system s owns slots [poff_s, poff_s + n_s rounded up to TJ), its j tiles are
// TJ slots and its wavefront groups 64; slots past n_s are padding.
//
// Forms:
//   NCG_F64_FAST  body order, FP64 (rsqrt + one Newton step), no tiles, so
//                 the sums do not depend on where the bodies are;
//   NCG_MIXED     the single-system mixed form per system (tiled_dev.hpp):
//                 for each (wavefront group, j tile) pair a uniform test
//                 (precomputed by k_ens_mask, one bit per pair) picks FP64 near
//                 terms or scaled FP32 far terms t = s_j (w_i + u_j) with FP32
//                 tile partials added into FP64.
// Collision pass: the same order; a wavefront group tests the pairs (p < q
// in slot order) of the tiles its position and velocity boxes can reach
// within dt.  The test is conservative (below), so the candidates are those
// of the all-pairs predicate.
//
// RK4 (classic, as in rk4.hip) runs every stage of all systems at once; with
// collisions on, each step starts with a collision pass of every system and a
// replay kernel (one wavefront per system) that decides the candidates in the
// reference's lexicographic order with the elastic / inelastic outcome, so a
// whole step -- pass, replay, four stages -- is one captured graph.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <numeric>

#include "ncg_internal.hpp"
#include "tiled_dev.hpp"

namespace {

constexpr int ENT = 256;  // threads per workgroup
constexpr unsigned ENS_CAND_CAP = 1u << 20;  // candidates of one pass, all systems
constexpr int ETJ = 256;  // j tile (LDS) of the FP64-fast form

struct EnsUnit {
    int sys;   // system
    int i0;    // first body (local index; in slot order for tiled passes)
    int cnt;   // bodies in the unit
    int kind;  // 0 gravity, 1 collision candidates
};

struct EnsSys {
    int off;   // first body
    int poff;  // first padded slot (a multiple of TJ)
    int n;     // bodies
    int pad;
};

// Position and velocity boxes of a tile or wavefront group (collision pass).
struct CollHdr {
    double px, py, pz;     // position box centre
    double vx, vy, vz;     // velocity box centre
    float phx, phy, phz;   // half extents, rounded up
    float vhx, vhy, vhz;
    float rmax;            // largest radius, rounded up
    int count;
};

__global__ void k_zero_u32(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Centre and half extent (rounded up to FP32) of [lo, hi].
__device__ inline void box_axis(double lo, double hi, double &c, float &h)
{
    c = 0.5 * (lo + hi);
    const double e = fmax(hi - c, c - lo);
    float f = (float)e;
    if ((double)f < e) f = nextafterf(f, 3.0e38f);
    h = f;
}

// ---- order ------------------------------------------------------------------
// Index along the Hilbert curve of a cell with 10-bit coordinates, after
// J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004):
// the coordinates are turned in place into the transposed index, whose bits
// interleaved (x most significant) are the position along the curve.  Unlike
// a Morton curve it has no jumps, so runs of keys are compact in space also
// on a thin ring, where Morton runs straddle quadrant boundaries.
__device__ inline unsigned hilbert10(unsigned x, unsigned y, unsigned z)
{
    unsigned X[3] = {x, y, z};
    for (unsigned Q = 1u << 9; Q > 1; Q >>= 1) {
        const unsigned P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;  // invert
            } else {        // exchange
                const unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];  // Gray encode
    X[2] ^= X[1];
    unsigned t = 0;
    for (unsigned Q = 1u << 9; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    return (expand10(X[0]) << 2) | (expand10(X[1]) << 1) | expand10(X[2]);
}

// One workgroup per system: bounding cube, largest radius, and the keys
// (system << 30) | Hilbert index over the cube.
__global__ __launch_bounds__(ENT) void k_ens_keys(const DBody *__restrict__ b,
                                                  const EnsSys *__restrict__ sys,
                                                  unsigned long long *__restrict__ key,
                                                  int *__restrict__ idx,
                                                  double *__restrict__ srmax)
{
    __shared__ double sh[ENT];
    const EnsSys S = sys[blockIdx.x];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, rm = 0.0;
    for (int k = threadIdx.x; k < S.n; k += ENT) {
        const DBody q = b[S.off + k];
        const double p[3] = {q.x, q.y, q.z};
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], p[c]);
            mx[c] = fmax(mx[c], p[c]);
        }
        rm = fmax(rm, q.radius);
    }
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
        lo[c] = block_reduce(mn[c], sh, 0);
        hi[c] = block_reduce(mx[c], sh, 1);
    }
    rm = block_reduce(rm, sh, 1);
    if (threadIdx.x == 0) srmax[blockIdx.x] = rm;
    const double w = fmax(hi[0] - lo[0], fmax(hi[1] - lo[1], hi[2] - lo[2]));
    for (int k = threadIdx.x; k < S.n; k += ENT) {
        const DBody q = b[S.off + k];
        const double p[3] = {q.x, q.y, q.z};
        unsigned g[3];
        for (int c = 0; c < 3; ++c) {
            const double base = 0.5 * (lo[c] + hi[c]) - 0.5 * w;
            const double f = w > 0.0 ? (p[c] - base) / w : 0.0;
            int v = (int)(f * 1023.999);
            v = v < 0 ? 0 : (v > 1023 ? 1023 : v);
            g[c] = (unsigned)v;
        }
        key[S.off + k] = ((unsigned long long)blockIdx.x << 30) | hilbert10(g[0], g[1], g[2]);
        idx[S.off + k] = S.off + k;
    }
}

// Slot -> body (-1 for padding) from the sorted body indices.
__global__ __launch_bounds__(ENT) void k_ens_place(const EnsSys *__restrict__ sys,
                                                   const int *__restrict__ tsys,
                                                   const int *__restrict__ sidx, int nslots,
                                                   int *__restrict__ pidx)
{
    const int s = blockIdx.x * ENT + threadIdx.x;
    if (s >= nslots) return;
    const EnsSys S = sys[tsys[s / TJ]];
    const int l = s - S.poff;
    pidx[s] = l < S.n ? sidx[S.off + l] : -1;
}

// ---- mixed form: records and tiles --------------------------------------------
// One wavefront per j tile: the sorted FP64 records {x, y, z, G m} and s R of
// its slots (padding: far away and massless), the tile header with its far
// records, and the headers of its two wavefront groups.
__global__ __launch_bounds__(ENT) void k_ens_tiles(const DBody *__restrict__ b,
                                                   const int *__restrict__ pidx,
                                                   const int *__restrict__ tsys,
                                                   const EnsSys *__restrict__ sys, int ntiles,
                                                   double G, double soft,
                                                   double4 *__restrict__ jd,
                                                   double *__restrict__ jsig,
                                                   float4 *__restrict__ ju,
                                                   TileHdr *__restrict__ thdr,
                                                   TileHdr *__restrict__ ghdr)
{
    const int t = blockIdx.x * (ENT / WV) + (int)threadIdx.x / WV;
    const int lane = threadIdx.x % WV;
    if (t >= ntiles) return;
    const EnsSys S = sys[tsys[t]];
    const int cnt = min(TJ, S.n - (t * TJ - S.poff));
    double4 p[2];
    double mn[2][3], mx[2][3], gmin = 1e300, gmax = 0.0;
    bool bad = false;
    for (int r = 0; r < 2; ++r) {
        const int s = t * TJ + r * WV + lane;
        const int k = pidx[s];
        double sig = 0.0;
        p[r] = make_double4(1e30, 1e30, 1e30, 0.0);
        for (int c = 0; c < 3; ++c) {
            mn[r][c] = 1e300;
            mx[r][c] = -1e300;
        }
        if (k >= 0) {
            const DBody q = b[k];
            p[r] = make_double4(q.x, q.y, q.z, G * q.mass);
            sig = soft * q.radius;
            const double v[3] = {q.x, q.y, q.z};
            for (int c = 0; c < 3; ++c) mn[r][c] = mx[r][c] = v[c];
            if (p[r].w > 0.0 && p[r].w < 1e300) {
                gmin = fmin(gmin, p[r].w);
                gmax = fmax(gmax, p[r].w);
            } else if (!(p[r].w == 0.0)) {
                bad = true;
            }
        }
        jd[s] = p[r];
        jsig[s] = sig;
    }
    TileHdr th, gh[2];
    double A[3], Ag[2][3];
    float H[3], Hg[2][3];
    for (int c = 0; c < 3; ++c) {
        const double l0 = wave_min(mn[0][c]), h0 = wave_max(mx[0][c]);
        const double l1 = wave_min(mn[1][c]), h1 = wave_max(mx[1][c]);
        box_axis(l0, h0, Ag[0][c], Hg[0][c]);
        box_axis(l1, h1, Ag[1][c], Hg[1][c]);
        box_axis(fmin(l0, l1), fmax(h0, h1), A[c], H[c]);
    }
    gmin = wave_min(gmin);
    gmax = wave_max(gmax);
    const bool anybad = __any(bad);
    th.ax = A[0];
    th.ay = A[1];
    th.az = A[2];
    th.hx = H[0];
    th.hy = H[1];
    th.hz = H[2];
    th.count = cnt;
    tile_s_range(gmin, gmax, anybad, th);
    for (int r = 0; r < 2; ++r) {
        gh[r].ax = Ag[r][0];
        gh[r].ay = Ag[r][1];
        gh[r].az = Ag[r][2];
        gh[r].hx = Hg[r][0];
        gh[r].hy = Hg[r][1];
        gh[r].hz = Hg[r][2];
        gh[r].count = max(0, min(WV, cnt - r * WV));
        gh[r].smax = gh[r].smin = 0.f;
    }
    if (lane == 0) {
        thdr[t] = th;
        ghdr[2 * t] = gh[0];
        ghdr[2 * t + 1] = gh[1];
    }
    for (int r = 0; r < 2; ++r) ju[t * TJ + r * WV + lane] = far_record(p[r], A);
}

// ---- mixed form: near/far decisions ---------------------------------------------
// The tile_near decision of every (wavefront group, j tile of its system)
// pair, one bit each: group g0 / WV's row of `nw` words, bit t - t0 for the
// system's tile t.  Same unit and wavefront layout as k_ens_tpass, which then
// tests a bit with scalar instructions instead of ~65 vector instructions
// per group and tile.
__global__ __launch_bounds__(ENT) void k_ens_mask(const EnsUnit *__restrict__ units,
                                                  const EnsSys *__restrict__ sys,
                                                  const double *__restrict__ srmax,
                                                  const TileHdr *__restrict__ thdr,
                                                  const TileHdr *__restrict__ ghdr, double soft,
                                                  double theta, int nw,
                                                  unsigned long long *__restrict__ mask)
{
    const int w = (int)threadIdx.x / WV, lane = (int)threadIdx.x % WV;
    const EnsUnit U = units[blockIdx.x];
    const EnsSys S = sys[U.sys];
    const int l0 = U.i0 + w * WV;
    if (l0 >= S.n) return;  // the whole wavefront
    const int g0 = S.poff + l0;
    const TileHdr ih = ghdr[g0 / WV];
    const double eps = 2.0 * soft * srmax[U.sys];
    const double eps2x = eps * eps * 67108864.0;  // 2^26
    const double hi = sqrt((double)ih.hx * ih.hx + (double)ih.hy * ih.hy + (double)ih.hz * ih.hz);
    const int t0 = S.poff / TJ, nt = (S.n + TJ - 1) / TJ;
    unsigned long long *row = mask + (size_t)(g0 / WV) * nw;
    for (int word = 0; word < nw; ++word) {
        const int lt = word * WV + lane;
        const bool near = lt < nt && tile_near(ih, hi, thdr[t0 + lt], theta, eps2x);
        const unsigned long long bits = __ballot(near);
        if (lane == 0) row[word] = bits;
    }
}

// ---- mixed form: the pass -----------------------------------------------------
// One workgroup per gravity unit (system, 256 slots); wavefront w takes the
// unit's group w and walks every j tile of its system, near tiles
// with FP64 terms read with wave-uniform addresses, far tiles staged in the
// wavefront's own LDS slice.  Then the central term (physics.c:127-146), the
// accelerations scattered to body order and the unit's CM_acc partial.
// (At most 73 VGPRs: 7 wavefronts per SIMD.)  SOFT false: no softening (the
// near terms skip s_i + s_j, same bits).
template <bool SOFT>
__global__ __launch_bounds__(ENT, 7) void k_ens_tpass(const EnsUnit *__restrict__ units,
                                                   const EnsSys *__restrict__ sys,
                                                   const double *__restrict__ srmax,
                                                   const double4 *__restrict__ jd,
                                                   const double *__restrict__ jsig,
                                                   const float4 *__restrict__ ju,
                                                   const TileHdr *__restrict__ thdr,
                                                   const TileHdr *__restrict__ ghdr,
                                                   const int *__restrict__ pidx,
                                                   const double *__restrict__ cen, double G,
                                                   double soft,
                                                   const unsigned long long *__restrict__ mask,
                                                   int nw, double *__restrict__ acc,
                                                   double *__restrict__ cpart)
{
    __shared__ float4 lu[(ENT / WV) * TJ];
    __shared__ double red[3][ENT];
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WV);
    const int lane = (int)threadIdx.x % WV;
    float4 *tu = lu + w * TJ;
    const int u = blockIdx.x;
    {
        const EnsUnit U = units[u];
        const EnsSys S = sys[U.sys];
        const int l0 = U.i0 + w * WV;  // the wavefront group's first local slot
        const bool gvalid = l0 < S.n;
        const int g0 = S.poff + l0;
        const bool mine = l0 + lane < S.n;
        double xi = 0.0, yi = 0.0, zi = 0.0, si = 0.0, gmi = 0.0;
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (gvalid) {
            const double4 p = jd[g0 + lane];
            xi = p.x;
            yi = p.y;
            zi = p.z;
            gmi = p.w;
            si = jsig[g0 + lane];
            // The group's decisions (k_ens_mask).
            const unsigned long long *mrow = mask + (size_t)(g0 / WV) * nw;
            const int t0 = S.poff / TJ, t1 = t0 + (S.n + TJ - 1) / TJ;
            for (int t = t0; t < t1; ++t) {
                const TileHdr th = thdr[t];
                const int jb = t * TJ;
                const int cnt = th.count;
                if ((mrow[(t - t0) >> 6] >> ((t - t0) & 63)) & 1ull) {
                    if (jb < g0 + WV && jb + TJ > g0) {
                        for (int jj = 0; jj < cnt; ++jj)
                            near_term_self<1, SOFT>(jd[jb + jj], SOFT ? jsig[jb + jj] : 0.0,
                                                    jb + jj, g0 + lane, WV, &xi, &yi, &zi, &si,
                                                    &ax, &ay, &az);
                    } else if (cnt == TJ) {
#pragma unroll 4
                        for (int jj = 0; jj < TJ; ++jj)
                            near_term<1, SOFT>(jd[jb + jj], SOFT ? jsig[jb + jj] : 0.0, &xi, &yi,
                                               &zi, &si, &ax, &ay, &az);
                    } else {
                        for (int jj = 0; jj < cnt; ++jj)
                            near_term<1, SOFT>(jd[jb + jj], SOFT ? jsig[jb + jj] : 0.0, &xi, &yi,
                                               &zi, &si, &ax, &ay, &az);
                    }
                } else {
                    // The wavefront's own slice: ordered after its reads of
                    // the previous far tile and before the reads below.
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
#pragma unroll
                    for (int r = 0; r < TJ / WV; ++r) tu[r * WV + lane] = ju[jb + r * WV + lane];
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    // w_i = A_J - x_i, rounded once from FP64.
                    const float wx = (float)(th.ax - xi), wy = (float)(th.ay - yi),
                                wz = (float)(th.az - zi);
                    float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll 8
                    for (int jj = 0; jj < TJ; ++jj) far_term<1>(tu[jj], &wx, &wy, &wz, &fx, &fy, &fz);
                    ax += (double)fx;
                    ay += (double)fy;
                    az += (double)fz;
                }
            }
        }
        // Central term (physics.c:127-146) and this unit's CM_acc partial.
        const double *c = cen + 8 * (size_t)U.sys;
        double cx = 0., cy = 0., cz = 0.;
        if (mine) {
            if (c[7] != 0.0) {
                const double DX = xi - c[0], DY = yi - c[1], DZ = zi - c[2];
                const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
                const double r3 = rk * rk * rk;
                const double mu = G * c[6];
                ax += -mu * DX / r3;
                ay += -mu * DY / r3;
                az += -mu * DZ / r3;
                cx = gmi * DX / r3;
                cy = gmi * DY / r3;
                cz = gmi * DZ / r3;
            }
            const int k = pidx[g0 + lane];
            acc[3 * (size_t)k + 0] = ax;
            acc[3 * (size_t)k + 1] = ay;
            acc[3 * (size_t)k + 2] = az;
        }
        red[0][threadIdx.x] = cx;
        red[1][threadIdx.x] = cy;
        red[2][threadIdx.x] = cz;
        __syncthreads();
        for (int h = ENT / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h)
                for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int d = 0; d < 3; ++d) cpart[3 * (size_t)u + d] = red[d][0];
    }
}

// ---- FP64-fast form (body order) ------------------------------------------------
// One workgroup per gravity unit (system, 256 IPL bodies in body order), j
// records staged in LDS ETJ at a time.
template <int IPL>
__global__ __launch_bounds__(ENT) void k_ens_pass(const EnsUnit *__restrict__ units,
                                                  const DBody *__restrict__ b,
                                                  const int *__restrict__ off,
                                                  const double *__restrict__ cen, double G,
                                                  double soft, double *__restrict__ acc,
                                                  double *__restrict__ cpart)
{
    __shared__ double4 tp[ETJ];  // x, y, z, G m
    __shared__ double ts[ETJ];   // s R
    __shared__ double red[3][ENT];
    const int u = blockIdx.x;
    {
        const EnsUnit U = units[u];
        const int base = off[U.sys];
        const int ns = off[U.sys + 1] - base;
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
            }
            __syncthreads();
            // The tile holds some of the unit's own bodies: their self terms
            // are zeroed (r^2 = 1e-200 without softening gives y^3 = 1e300,
            // which G m > ~1.8e8 would overflow).
            const bool diag = jb < U.i0 + ENT * IPL && jb + cnt > U.i0;
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
        for (int h = ENT / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h)
                for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int d = 0; d < 3; ++d) cpart[3 * (size_t)u + d] = red[d][0];
    }
}

// ---- collision pass -------------------------------------------------------------
// One wavefront per tile: the sorted copy of its bodies and the position /
// velocity boxes and largest radius of the tile and of its two groups.
__global__ __launch_bounds__(ENT) void k_ens_ctiles(const DBody *__restrict__ b,
                                                    const int *__restrict__ pidx,
                                                    const int *__restrict__ tsys,
                                                    const EnsSys *__restrict__ sys, int ntiles,
                                                    DBody *__restrict__ jc,
                                                    CollHdr *__restrict__ tch,
                                                    CollHdr *__restrict__ gch)
{
    const int t = blockIdx.x * (ENT / WV) + (int)threadIdx.x / WV;
    const int lane = threadIdx.x % WV;
    if (t >= ntiles) return;
    const EnsSys S = sys[tsys[t]];
    const int cnt = min(TJ, S.n - (t * TJ - S.poff));
    // Per slot r: 0-2 position, 3-5 velocity.
    double mn[2][6], mx[2][6], rm[2];
    for (int r = 0; r < 2; ++r) {
        const int s = t * TJ + r * WV + lane;
        const int k = pidx[s];
        for (int c = 0; c < 6; ++c) {
            mn[r][c] = 1e300;
            mx[r][c] = -1e300;
        }
        rm[r] = 0.0;
        if (k >= 0) {
            const DBody q = b[k];
            jc[s] = q;
            const double v[6] = {q.x, q.y, q.z, q.vx, q.vy, q.vz};
            for (int c = 0; c < 6; ++c) mn[r][c] = mx[r][c] = v[c];
            rm[r] = q.radius;
        }
    }
    double cg[2][6], ct[6];
    float hg[2][6], ht[6];
    for (int c = 0; c < 6; ++c) {
        const double l0 = wave_min(mn[0][c]), h0 = wave_max(mx[0][c]);
        const double l1 = wave_min(mn[1][c]), h1 = wave_max(mx[1][c]);
        box_axis(l0, h0, cg[0][c], hg[0][c]);
        box_axis(l1, h1, cg[1][c], hg[1][c]);
        box_axis(fmin(l0, l1), fmax(h0, h1), ct[c], ht[c]);
    }
    const double r0 = wave_max(rm[0]), r1 = wave_max(rm[1]);
    if (lane != 0) return;
    auto up = [](double v) {
        float f = (float)v;
        if ((double)f < v) f = nextafterf(f, 3.0e38f);
        return f;
    };
    auto put = [&](CollHdr &h, const double *cc, const float *hh, double r, int count) {
        h.px = cc[0];
        h.py = cc[1];
        h.pz = cc[2];
        h.vx = cc[3];
        h.vy = cc[4];
        h.vz = cc[5];
        h.phx = hh[0];
        h.phy = hh[1];
        h.phz = hh[2];
        h.vhx = hh[3];
        h.vhy = hh[4];
        h.vhz = hh[5];
        h.rmax = up(r);
        h.count = count;
    };
    CollHdr h;
    put(h, ct, ht, fmax(r0, r1), cnt);
    tch[t] = h;
    put(h, cg[0], hg[0], r0, max(0, min(WV, cnt)));
    gch[2 * t] = h;
    put(h, cg[1], hg[1], r1, max(0, min(WV, cnt - WV)));
    gch[2 * t + 1] = h;
}

// Whether some pair of group g and tile t may meet within |dt|: the pair's
// closest_approach (below) lists it only if its first contact time
// t_c = (|d| - R) / |dv| or less lies within dt (up to the predicate's
// tolerance of 9.1e-13 |dt| and its rounding, a relative 4e-8 of |d|), so
// |d| <= (R + |dv| |dt|) (1 + 1e-7).  The boxes bound |d| from below and
// R, |dv| from above; the margin 1e-6 and the absolute term (a few ulps of
// the coordinates, for the rounding of the box centres) keep the test
// conservative.  NaN anywhere keeps the pair.
__device__ inline bool may_meet(const CollHdr &g, const CollHdr &t, double adt)
{
    const double gx = fmax(0.0, fabs(t.px - g.px) - ((double)g.phx + t.phx));
    const double gy = fmax(0.0, fabs(t.py - g.py) - ((double)g.phy + t.phy));
    const double gz = fmax(0.0, fabs(t.pz - g.pz) - ((double)g.phz + t.phz));
    const double ux = fabs(t.vx - g.vx) + ((double)g.vhx + t.vhx);
    const double uy = fabs(t.vy - g.vy) + ((double)g.vhy + t.vhy);
    const double uz = fabs(t.vz - g.vz) + ((double)g.vhz + t.vhz);
    const double mag = fmax(fmax(fabs(g.px), fabs(t.px)), fmax(fmax(fabs(g.py), fabs(t.py)),
                                                               fmax(fabs(g.pz), fabs(t.pz))));
    const double reach = ((double)g.rmax + t.rmax + sqrt(ux * ux + uy * uy + uz * uz) * adt) *
                             1.000001 +
                         1e-14 * mag;
    return !(gx * gx + gy * gy + gz * gz > reach * reach);
}

// One workgroup per collision unit (system, 256 slots); wavefront w takes
// the unit's group w and tests its pairs (p, q > p in slot order) with every
// tile from its own on that it may meet, staging each such tile in its own
// LDS slice.  The predicate is
// closest_approach's (collision.c:112-137; this file is compiled without FMA
// contraction), symmetric in the pair bit for bit, with a tolerance of
// 9.1e-13 |dt| on the contact time: candidates are a superset of the
// reference's hits, decided exactly by the replay.
__global__ __launch_bounds__(ENT) void k_ens_cpass(const EnsUnit *__restrict__ units,
                                                   const EnsSys *__restrict__ sys,
                                                   const DBody *__restrict__ jc,
                                                   const int *__restrict__ pidx,
                                                   const CollHdr *__restrict__ tch,
                                                   const CollHdr *__restrict__ gch, double dt,
                                                   int2 *__restrict__ cand, unsigned cand_cap,
                                                   unsigned *__restrict__ cand_cnt)
{
    __shared__ double4 lp[(ENT / WV) * TJ];  // x, y, z, R
    __shared__ double4 lv[(ENT / WV) * TJ];  // vx, vy, vz
    __shared__ int lk[(ENT / WV) * TJ];      // body index
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WV);
    const int lane = (int)threadIdx.x % WV;
    const double adt = fabs(dt);
    double4 *tp = lp + w * TJ, *tv = lv + w * TJ;
    int *tk = lk + w * TJ;
    {
        const EnsUnit U = units[blockIdx.x];
        const EnsSys S = sys[U.sys];
        const int l0 = U.i0 + w * WV;
        if (l0 >= S.n) return;  // no synchronisation in this kernel
        const int g0 = S.poff + l0;
        const int p = g0 + lane;
        const bool mine = l0 + lane < S.n;
        DBody a = {};
        int ka = -1;
        if (mine) {
            a = jc[p];
            ka = pidx[p];
        }
        const CollHdr gh = gch[g0 / WV];
        const int t1 = S.poff / TJ + (S.n + TJ - 1) / TJ;
        for (int t = g0 / TJ; t < t1; ++t) {
            const CollHdr th = tch[t];
            if (!may_meet(gh, th, adt)) continue;
            const int jb = t * TJ;
            // The wavefront's own slice: ordered after its reads of the
            // previous tile and before the reads below.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < TJ / WV; ++r) {
                const int l = r * WV + lane;
                if (l < th.count) {
                    const DBody c = jc[jb + l];
                    tp[l] = make_double4(c.x, c.y, c.z, c.radius);
                    tv[l] = make_double4(c.vx, c.vy, c.vz, 0.0);
                    tk[l] = pidx[jb + l];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int jj = 0; jj < th.count; ++jj) {
                const int q = jb + jj;
                if (!mine || q <= p) continue;
                const double4 cp = tp[jj], cv = tv[jj];
                const double dx = a.x - cp.x, dy = a.y - cp.y, dz = a.z - cp.z;
                const double dvx = a.vx - cv.x, dvy = a.vy - cv.y, dvz = a.vz - cv.z;
                const double bb = dx * dvx + dy * dvy + dz * dvz;
                if (!(bb < 0.0)) continue;
                const double dr2 = dx * dx + dy * dy + dz * dz;
                const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
                const double R = a.radius + cp.w;
                const double disc = bb * bb + dv2 * (R * R - dr2);
                if (!(disc > 0.0)) continue;
                const double tc = -(bb + sqrt(disc)) / dv2;
                if (tc <= dt + adt * 9.094947017729282e-13) {
                    const int kc = tk[jj];
                    const unsigned slot = atomicAdd(cand_cnt, 1u);
                    if (slot < cand_cap) cand[slot] = make_int2(min(ka, kc), max(ka, kc));
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
    double theta = 0.35;
    int nslots = 0, ntiles = 0, key_bits = 30;
    int maxt = 0;                     // tiles of the largest system
    std::vector<int> off;
    std::vector<EnsSys> sys;
    std::vector<int> tsys;            // system of each tile
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
        return u.kind == 0 ? ns * u.cnt * 1.0 : ns;
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

// Systems, slot layout and unit lists from the offsets; bodies per lane (the
// FP64-fast form) and theta (the mixed form) from the tuning table.
static void make_plan(ncg_ctx *ctx, EnsPlan &P, const int32_t *offsets, int32_t nsys,
                      ncg_precision prec)
{
    P.nsys = nsys;
    P.off.assign(offsets, offsets + nsys + 1);
    P.ntot = P.off[nsys];
    P.sys.resize(nsys);
    P.tsys.clear();
    int poff = 0;
    for (int s = 0; s < nsys; ++s) {
        const int n = P.off[s + 1] - P.off[s];
        P.sys[s] = {P.off[s], poff, n, 0};
        const int nt = (n + TJ - 1) / TJ;
        for (int t = 0; t < nt; ++t) P.tsys.push_back(s);
        poff += nt * TJ;
        P.maxt = std::max(P.maxt, nt);
    }
    P.nslots = poff;
    P.ntiles = poff / TJ;
    P.key_bits = 30;
    while ((1LL << (P.key_bits - 30)) < nsys) ++P.key_bits;
    const TiledCfg tc = ncg_tiled_cfg(ctx, prec == NCG_MIXED ? 100 + NCG_MIXED : 100 + NCG_F64_FAST,
                                      P.ntot / std::max(nsys, 1));
    // The mixed form runs one body per lane.
    P.ipl = prec == NCG_MIXED ? 1 : (tc.ipl == 2 ? 2 : 1);
    P.theta = tc.theta;
    make_units(P, P.ipl);
}

struct EnsDev {
    EnsUnit *grav = nullptr, *coll = nullptr;
    int *gfirst = nullptr, *glist = nullptr;
    int *off = nullptr;
    EnsSys *sys = nullptr;
    int *tsys = nullptr;
    // Order (mixed form and collision passes).
    unsigned long long *key = nullptr, *key2 = nullptr;
    int *idx = nullptr, *idx2 = nullptr, *pidx = nullptr;
    double *srmax = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
    // Mixed-form records and tiles.
    double4 *jd = nullptr;
    double *jsig = nullptr;
    float4 *ju = nullptr;
    TileHdr *thdr = nullptr, *ghdr = nullptr;
    unsigned long long *mask = nullptr;  // near/far decisions, 2 nt rows of nw words
    int nw = 1;
    // Collision pass.
    DBody *jc = nullptr;
    CollHdr *tch = nullptr, *gch = nullptr;
};

// Device buffers and uploads of a plan: the unit lists always, the order when
// `order`, the mixed form's tiles when `tiles`, the collision pass's when
// `coll`.
static ncg_status ens_setup(ncg_ctx *ctx, const EnsPlan &P, EnsDev &D, bool order, bool tiles,
                            bool coll, hipStream_t st)
{
    const int nsys = P.nsys;
    std::vector<int> gfirst, glist;
    unit_index(P, P.grav, gfirst, glist);
    D.grav = dbuf<EnsUnit>(ctx, "e_grav", P.grav.size());
    D.coll = dbuf<EnsUnit>(ctx, "e_coll", P.coll.size());
    D.gfirst = dbuf<int>(ctx, "e_gfirst", nsys + 1);
    D.glist = dbuf<int>(ctx, "e_glist", std::max<size_t>(glist.size(), 1));
    D.off = dbuf<int>(ctx, "e_off", nsys + 1);
    D.sys = dbuf<EnsSys>(ctx, "e_sys", nsys);
    D.tsys = dbuf<int>(ctx, "e_tsys", std::max(P.ntiles, 1));
    if (!D.grav || !D.coll || !D.gfirst || !D.glist || !D.off || !D.sys || !D.tsys)
        return NCG_ERR_NOMEM;
    const size_t ns = std::max(P.nslots, 1), nt = std::max(P.ntiles, 1);
    if (order) {
        D.key = dbuf<unsigned long long>(ctx, "e_key", std::max(P.ntot, 1));
        D.key2 = dbuf<unsigned long long>(ctx, "e_key2", std::max(P.ntot, 1));
        D.idx = dbuf<int>(ctx, "e_idx", std::max(P.ntot, 1));
        D.idx2 = dbuf<int>(ctx, "e_idx2", std::max(P.ntot, 1));
        D.pidx = dbuf<int>(ctx, "e_pidx", ns);
        D.srmax = dbuf<double>(ctx, "e_srmax", nsys);
        NCG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, D.sort_bytes,
                                                   (unsigned long long *)nullptr,
                                                   (unsigned long long *)nullptr, (int *)nullptr,
                                                   (int *)nullptr, std::max(P.ntot, 1), 0,
                                                   P.key_bits, st));
        D.sort_tmp = dbuf<unsigned char>(ctx, "e_sort", D.sort_bytes);
        if (!D.key || !D.key2 || !D.idx || !D.idx2 || !D.pidx || !D.srmax || !D.sort_tmp)
            return NCG_ERR_NOMEM;
    }
    if (tiles) {
        D.jd = dbuf<double4>(ctx, "e_jd", ns);
        D.jsig = dbuf<double>(ctx, "e_jsig", ns);
        D.ju = dbuf<float4>(ctx, "e_ju", ns);
        D.thdr = dbuf<TileHdr>(ctx, "e_thdr", nt);
        D.ghdr = dbuf<TileHdr>(ctx, "e_ghdr", 2 * nt);
        D.nw = std::max(1, (int)cdiv(P.maxt, WV));
        D.mask = dbuf<unsigned long long>(ctx, "e_mask", 2 * (size_t)nt * D.nw);
        if (!D.jd || !D.jsig || !D.ju || !D.thdr || !D.ghdr || !D.mask) return NCG_ERR_NOMEM;
    }
    if (coll) {
        D.jc = dbuf<DBody>(ctx, "e_jc", ns);
        D.tch = dbuf<CollHdr>(ctx, "e_tch", nt);
        D.gch = dbuf<CollHdr>(ctx, "e_gch", 2 * nt);
        if (!D.jc || !D.tch || !D.gch) return NCG_ERR_NOMEM;
    }
    NCG_HIP(hipMemcpyAsync(D.grav, P.grav.data(), sizeof(EnsUnit) * P.grav.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.coll, P.coll.data(), sizeof(EnsUnit) * P.coll.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.gfirst, gfirst.data(), sizeof(int) * gfirst.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.glist, glist.data(), sizeof(int) * glist.size(), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.off, P.off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, st));
    NCG_HIP(hipMemcpyAsync(D.sys, P.sys.data(), sizeof(EnsSys) * nsys, hipMemcpyHostToDevice, st));
    if (P.ntiles)
        NCG_HIP(hipMemcpyAsync(D.tsys, P.tsys.data(), sizeof(int) * P.ntiles, hipMemcpyHostToDevice, st));
    // The host vectors are temporaries: the copies finish before they go.
    NCG_HIP(hipStreamSynchronize(st));
    return NCG_OK;
}

// The step's order from the positions in b: keys, one sort, slots.
static ncg_status ens_order(ncg_ctx *ctx, const EnsPlan &P, const EnsDev &D, const DBody *b,
                            hipStream_t st)
{
    if (P.ntot == 0) return NCG_OK;
    size_t bytes = D.sort_bytes;
    hipLaunchKernelGGL(k_ens_keys, dim3(P.nsys), dim3(ENT), 0, st, b, D.sys, D.key, D.idx, D.srmax);
    NCG_CHECK_LAUNCH();
    NCG_HIP(hipcub::DeviceRadixSort::SortPairs(D.sort_tmp, bytes, D.key, D.key2, D.idx, D.idx2,
                                               P.ntot, 0, P.key_bits, st));
    hipLaunchKernelGGL(k_ens_place, dim3(cdiv(P.nslots, ENT)), dim3(ENT), 0, st, D.sys, D.tsys,
                       D.idx2, P.nslots, D.pidx);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// One evaluation of all systems (the mixed form with the order of the last
// ens_order): accelerations and per-system CM_acc.
static ncg_status ens_eval(ncg_ctx *ctx, const EnsPlan &P, const EnsDev &D, const DBody *b,
                           const double *cen, double G, double soft, ncg_precision prec,
                           double *acc, double *cm, hipStream_t st)
{
    const int nunits = (int)P.grav.size();
    double *cpart = dbuf<double>(ctx, "e_cpart", 3 * (size_t)std::max(nunits, 1));
    if (!cpart) return NCG_ERR_NOMEM;
    if (nunits == 0) return NCG_OK;
    if (prec == NCG_MIXED) {
        hipLaunchKernelGGL(k_ens_tiles, dim3(cdiv(P.ntiles, ENT / WV)), dim3(ENT), 0, st, b, D.pidx,
                           D.tsys, D.sys, P.ntiles, G, soft, D.jd, D.jsig, D.ju, D.thdr, D.ghdr);
        NCG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ens_mask, dim3(nunits), dim3(ENT), 0, st, D.grav, D.sys, D.srmax,
                           D.thdr, D.ghdr, soft, P.theta, D.nw, D.mask);
        NCG_CHECK_LAUNCH();
        ProfScope ps(ctx, st, "ens_mixed");
        if (soft != 0.0)
            hipLaunchKernelGGL(k_ens_tpass<true>, dim3(nunits), dim3(ENT), 0, st, D.grav, D.sys,
                               D.srmax, D.jd, D.jsig, D.ju, D.thdr, D.ghdr, D.pidx, cen, G, soft,
                               D.mask, D.nw, acc, cpart);
        else
            hipLaunchKernelGGL(k_ens_tpass<false>, dim3(nunits), dim3(ENT), 0, st, D.grav, D.sys,
                               D.srmax, D.jd, D.jsig, D.ju, D.thdr, D.ghdr, D.pidx, cen, G, soft,
                               D.mask, D.nw, acc, cpart);
        NCG_CHECK_LAUNCH();
    } else {
        ProfScope ps(ctx, st, "ens_fast64");
        if (P.ipl == 2)
            hipLaunchKernelGGL(k_ens_pass<2>, dim3(nunits), dim3(ENT), 0, st, D.grav, b, D.off,
                               cen, G, soft, acc, cpart);
        else
            hipLaunchKernelGGL(k_ens_pass<1>, dim3(nunits), dim3(ENT), 0, st, D.grav, b, D.off,
                               cen, G, soft, acc, cpart);
        NCG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ens_cm, dim3(cdiv(P.nsys, 128)), dim3(128), 0, st, D.gfirst, D.glist,
                       cpart, P.nsys, cm);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// The collision pass of all systems (order of the last ens_order): candidates
// in cand, their count in counter[1] (counter[0] unused).
static ncg_status ens_collide(ncg_ctx *ctx, const EnsPlan &P, const EnsDev &D, const DBody *b,
                              double dt, int2 *cand, unsigned cap, unsigned *counter,
                              hipStream_t st)
{
    const int nunits = (int)P.coll.size();
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, counter, 2);
    NCG_CHECK_LAUNCH();
    if (nunits == 0) return NCG_OK;
    hipLaunchKernelGGL(k_ens_ctiles, dim3(cdiv(P.ntiles, ENT / WV)), dim3(ENT), 0, st, b, D.pidx,
                       D.tsys, D.sys, P.ntiles, D.jc, D.tch, D.gch);
    NCG_CHECK_LAUNCH();
    ProfScope ps(ctx, st, "ens_collide");
    hipLaunchKernelGGL(k_ens_cpass, dim3(nunits), dim3(ENT), 0, st, D.coll, D.sys, D.jc, D.pidx,
                       D.tch, D.gch, dt, cand, cap, counter + 1);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
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
    const bool tiled = prec == NCG_MIXED;
    if (n_collisions) *n_collisions = 0;

    EnsPlan P;
    make_plan(ctx, P, offsets, nsys, prec);
    EnsDev D;
    ncg_status rs = ens_setup(ctx, P, D, tiled || collide, tiled, collide, st);
    if (rs != NCG_OK) return rs;

    const int ntot = P.ntot;
    DBody *Y = dbuf<DBody>(ctx, "e_Y", ntot);
    double *acc = dbuf<double>(ctx, "e_acc", 3 * (size_t)ntot);
    double *sx = dbuf<double>(ctx, "e_sx", 3 * (size_t)ntot);
    double *sv = dbuf<double>(ctx, "e_sv", 3 * (size_t)ntot);
    double *Yc = dbuf<double>(ctx, "e_Yc", 8 * (size_t)nsys);
    double *scen = dbuf<double>(ctx, "e_sc", 6 * (size_t)nsys);
    double *cm = dbuf<double>(ctx, "e_cm", 3 * (size_t)nsys);
    // Every buffer a captured step uses exists before the capture.
    unsigned *counter = dbuf<unsigned>(ctx, "e_counter", 2);
    double *cpart = dbuf<double>(ctx, "e_cpart", 3 * P.grav.size() + 3);
    const unsigned cap = ncg_replay_cap(ctx, ENS_CAND_CAP);
    int2 *cand = dbuf<int2>(ctx, "e_cand", cap);
    unsigned char *dirty = dbuf<unsigned char>(ctx, "e_dirty", (size_t)ntot + 4);
    unsigned *stat = dbuf<unsigned>(ctx, "e_stat", 2);  // collisions, error bits
    if (!Y || !acc || !sx || !sv || !Yc || !scen || !cm || !counter || !cpart || !cand || !dirty ||
        !stat)
        return NCG_ERR_NOMEM;
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
        ncg_status r = ens_collide(ctx, P, D, y0, dt, cand, cap, counter, st);
        if (r != NCG_OK) return r;
        return ncg_bf_replay_launch(ctx, y0, nullptr, nullptr, 0, D.off, nsys, cand, counter + 1,
                                    nullptr, cap, dt, collision_f, dirty, stat, st);
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
    // The step's order (mixed form, collision pass) is sorted eagerly, outside
    // the captured graph.
    auto order = [&]() -> ncg_status {
        return tiled || collide ? ens_order(ctx, P, D, y0, st) : NCG_OK;
    };
    int32_t win0 = 0;
    for (int32_t step = 0; step < nsteps; ++step) {
        ncg_status ro = order();
        if (ro != NCG_OK) return ro;
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
                if ((r = order()) != NCG_OK || (r = step_body(true)) != NCG_OK) return r;
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
                        hs[1], cap, NCG_REPLAY_RCAP,
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
    const bool tiled = prec == NCG_MIXED;
    EnsPlan P;
    make_plan(ctx, P, offsets, nsys, prec);
    EnsDev D;
    ncg_status r = ens_setup(ctx, P, D, tiled, tiled, false, st);
    if (r != NCG_OK) return r;
    const DBody *b = reinterpret_cast<const DBody *>(d_bodies);
    if (tiled && (r = ens_order(ctx, P, D, b, st)) != NCG_OK) return r;
    return ens_eval(ctx, P, D, b, d_centrals, grav_const, softening, prec, d_acc, d_cm_acc, st);
}
