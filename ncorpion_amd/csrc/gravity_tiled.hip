// All-pairs gravity, FP64-fast and mixed FP32/FP64 forms.
//
// Computes the same field as the brute-force branch of NcorpiON's
// vector_field (physics.c:120-202): a_i = sum_{j != i} G m_j (x_j - x_i) / D^3
// with D^2 = |x_j - x_i|^2 + (s (R_i + R_j))^2, plus the central term
// (physics.c:127-146), but in tile order on the vector ALUs.
//
// Pipeline of one evaluation (all on one stream, no host synchronisation):
//   1. k_bbox_*      bounding box, largest radius and count of existing bodies
//   2. k_morton      mixed form: 30-bit Morton key per body over the bounding
//                    CUBE (the same cell size on every axis, so that runs of
//                    keys are compact in space also for a thin disk); FP64-fast
//                    form (no near/far split): existence only, so that its
//                    sums keep body order whatever the positions; holes last
//   3. radix sort    (key, index) pairs -> spatially coherent (mixed) or body
//                    (FP64-fast) order
//   4. k_gather      sorted FP64 records {x, y, z, G m} and s R
//   5. k_tiles       per tile of TJ bodies: FP64 anchor (box centre), FP32
//                    half extent, the range of (G m)^(-1/2) over the tile and
//                    the scaled far-form records (below); the anchor and
//                    extent also per wavefront group of 64*IPL i-bodies
//   5b. k_near_mask  mixed form: the near/far decision (below) of every
//                    (wavefront group, tile) pair, one bit each, made once
//                    instead of once per j split by vector instructions
//   6. k_pairs<MODE> one workgroup per (i-block, j-split); every lane owns IPL
//                    bodies, every wavefront a contiguous group of 64*IPL of
//                    them, and walks the j tiles of its split.  Tile data is
//                    read with wave-uniform addresses (scalar loads).  For each
//                    (wavefront group, tile) pair a uniform bit picks the form:
//                      near: FP64 separations, FP64 rsqrt refined by one Newton
//                            step, FP64 accumulation (also every tile in
//                            NCG_F64_FAST);
//                      far:  scaled FP32 separations t = s_j (w_i + u_j)
//                            with s_j ~ (G m_j)^(-1/2) (the FP32 value whose
//                            s^-2 is nearest G m_j), w_i = A_J - x_i rounded
//                            once from FP64 per tile and u_j the offset from
//                            the tile's anchor A_J; then G m_j d / r^3 =
//                            t / |t|^3, so a term is 12 FP32 operations; FP32
//                            tile partial added into FP64.
//                    A pair is far when its box distance is at least theta
//                    times (anchor distance + half extents), which bounds the
//                    FP32 separation error, when the largest softening length
//                    is below FP32 resolution of that distance, so the far
//                    form can omit it without changing an FP32 result, and
//                    when every scaled distance |t|^2 lies in [1e-24, 1e24],
//                    so t / |t|^3 neither overflows nor underflows in FP32.
//   7. k_finish      sums the j-split partials in fixed order, adds the FP64
//                    central term and scatters back to body order.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "ncg_internal.hpp"
#include "tiled_dev.hpp"

namespace {

constexpr int NT = 256;     // threads per workgroup

struct Scal {           // per-evaluation scalars produced on the device
    double bmin[3], bmax[3];
    double rmax;
    int nvalid;
    int pad;
};

// ---- 1. bounding box / count ---------------------------------------------
__global__ __launch_bounds__(NT) void k_bbox_partial(const DBody *__restrict__ b, int n,
                                                     const int32_t *__restrict__ ex,
                                                     double *__restrict__ part)
{
    __shared__ double sh[NT];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, rm = 0.0, cnt = 0.0;
    for (int k = blockIdx.x * NT + threadIdx.x; k < n; k += gridDim.x * NT) {
        if (ex && !ex[k]) continue;
        const double p[3] = {b[k].x, b[k].y, b[k].z};
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], p[c]);
            mx[c] = fmax(mx[c], p[c]);
        }
        rm = fmax(rm, b[k].radius);
        cnt += 1.0;
    }
    double *o = part + 8 * blockIdx.x;
    for (int c = 0; c < 3; ++c) {
        double v = block_reduce(mn[c], sh, 0);
        if (threadIdx.x == 0) o[c] = v;
        v = block_reduce(mx[c], sh, 1);
        if (threadIdx.x == 0) o[3 + c] = v;
    }
    double v = block_reduce(rm, sh, 1);
    if (threadIdx.x == 0) o[6] = v;
    v = block_reduce(cnt, sh, 2);
    if (threadIdx.x == 0) o[7] = v;
}

__global__ __launch_bounds__(NT) void k_bbox_final(const double *__restrict__ part, int nparts,
                                                   Scal *__restrict__ sc)
{
    __shared__ double sh[NT];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, rm = 0.0, cnt = 0.0;
    for (int q = threadIdx.x; q < nparts; q += NT) {
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], part[8 * q + c]);
            mx[c] = fmax(mx[c], part[8 * q + 3 + c]);
        }
        rm = fmax(rm, part[8 * q + 6]);
        cnt += part[8 * q + 7];
    }
    for (int c = 0; c < 3; ++c) {
        double v = block_reduce(mn[c], sh, 0);
        if (threadIdx.x == 0) sc->bmin[c] = v;
        v = block_reduce(mx[c], sh, 1);
        if (threadIdx.x == 0) sc->bmax[c] = v;
    }
    double v = block_reduce(rm, sh, 1);
    if (threadIdx.x == 0) sc->rmax = v;
    v = block_reduce(cnt, sh, 2);
    if (threadIdx.x == 0) sc->nvalid = (int)v;
}

// ---- 2. Morton keys -------------------------------------------------------
// spatial = 0 (the FP64-fast form, which has no near/far split): key 0 for
// existing bodies and 1 for holes, so the stable sort keeps body order and
// the field's sums do not depend on where the bodies are.
__global__ __launch_bounds__(NT) void k_morton(const DBody *__restrict__ b, int n,
                                               const int32_t *__restrict__ ex,
                                               const Scal *__restrict__ sc, int spatial,
                                               unsigned *__restrict__ key, int *__restrict__ idx)
{
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    idx[k] = k;
    if (ex && !ex[k]) {
        key[k] = spatial ? 0xffffffffu : 1u;
        return;
    }
    if (!spatial) {
        key[k] = 0u;
        return;
    }
    unsigned q[3];
    const double p[3] = {b[k].x, b[k].y, b[k].z};
    // Cells of one size on all axes: the bounding cube of edge w around the
    // box centre.
    const double w = fmax(sc->bmax[0] - sc->bmin[0],
                          fmax(sc->bmax[1] - sc->bmin[1], sc->bmax[2] - sc->bmin[2]));
    for (int c = 0; c < 3; ++c) {
        const double lo = 0.5 * (sc->bmin[c] + sc->bmax[c]) - 0.5 * w;
        double f = w > 0.0 ? (p[c] - lo) / w : 0.0;
        int v = (int)(f * 1023.999);
        v = v < 0 ? 0 : (v > 1023 ? 1023 : v);
        q[c] = (unsigned)v;
    }
    key[k] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
}

// ---- 4. sorted records ----------------------------------------------------
__global__ __launch_bounds__(NT) void k_gather(const DBody *__restrict__ b, int n,
                                               const int *__restrict__ perm,
                                               const Scal *__restrict__ sc, double G, double soft,
                                               double4 *__restrict__ jd, double *__restrict__ jsig)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= n) return;
    const int nv = sc->nvalid;
    if (nv <= 0) return;
    // Slots past the last valid body repeat it with zero mass, so tile boxes
    // are unchanged and their terms vanish.
    const int k = perm[s < nv ? s : nv - 1];
    const DBody q = b[k];
    jd[s] = make_double4(q.x, q.y, q.z, s < nv ? G * q.mass : 0.0);
    jsig[s] = soft * q.radius;
}

// ---- 5. tile anchors and scaled FP32 offsets --------------------------------
// One workgroup per tile of tsize bodies: header (anchor, half extents, count,
// range of s) and, for j tiles, the far records (far_record; slots past the
// count get the massless record).
__global__ __launch_bounds__(NT) void k_tiles(const double4 *__restrict__ jd,
                                              const Scal *__restrict__ sc, int tsize,
                                              TileHdr *__restrict__ hdr,
                                              float4 *__restrict__ ju /* null: header only */)
{
    __shared__ double sh[NT];
    const int nv = sc->nvalid;
    const int t = blockIdx.x;
    const int b0 = t * tsize;
    if (b0 >= nv) return;
    const int cnt = min(tsize, nv - b0);
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    double gmin = 1e300, gmax = 0.0;  // over G m > 0
    bool bad = false;                 // a negative or non-finite G m
    for (int q = threadIdx.x; q < cnt; q += NT) {
        const double4 p = jd[b0 + q];
        mn[0] = fmin(mn[0], p.x);
        mn[1] = fmin(mn[1], p.y);
        mn[2] = fmin(mn[2], p.z);
        mx[0] = fmax(mx[0], p.x);
        mx[1] = fmax(mx[1], p.y);
        mx[2] = fmax(mx[2], p.z);
        if (p.w > 0.0 && p.w < 1e300) {
            gmin = fmin(gmin, p.w);
            gmax = fmax(gmax, p.w);
        } else if (!(p.w == 0.0)) {
            bad = true;
        }
    }
    double A[3];
    float H[3];
    for (int c = 0; c < 3; ++c) {
        const double lo = block_reduce(mn[c], sh, 0);
        const double hi = block_reduce(mx[c], sh, 1);
        A[c] = 0.5 * (lo + hi);
        const double h = fmax(hi - A[c], A[c] - lo);
        float hf = (float)h;
        if ((double)hf < h) hf = nextafterf(hf, 3.0e38f);
        H[c] = hf;
    }
    const double gl = block_reduce(gmin, sh, 0);
    const double gh = block_reduce(gmax, sh, 1);
    const double nbad = block_reduce(bad ? 1.0 : 0.0, sh, 2);
    if (threadIdx.x == 0) {
        TileHdr th;
        th.ax = A[0];
        th.ay = A[1];
        th.az = A[2];
        th.hx = H[0];
        th.hy = H[1];
        th.hz = H[2];
        th.count = cnt;
        tile_s_range(gl, gh, nbad != 0.0, th);
        hdr[t] = th;
    }
    if (ju) {
        for (int q = threadIdx.x; q < tsize; q += NT) {
            const int s = b0 + q;
            ju[s] = q < cnt ? far_record(jd[s], A) : make_float4(1e30f, 1e30f, 1e30f, 0.f);
        }
    }
}

// ---- 5b. near/far decisions of the mixed form ------------------------------
// One lane per (wavefront group g, j tile t): the tile_near decision of the
// pair kernel, made once for all of its j splits, as one bit of
// mask[g nw + t / 64].  The pair kernel then tests a bit with scalar
// instructions instead of spending ~60 vector instructions (two FP64 square
// roots) per group and tile.
__global__ __launch_bounds__(WV) void k_near_mask(const TileHdr *__restrict__ thdr,
                                                  const TileHdr *__restrict__ ihdr,
                                                  const Scal *__restrict__ sc, int gw, int nw,
                                                  double theta, double soft,
                                                  unsigned long long *__restrict__ mask)
{
    const int nv = sc->nvalid;
    const int g = blockIdx.x;  // groups on x: more than 65,535 above 4M bodies
    const int t = blockIdx.y * WV + (int)threadIdx.x;
    bool near = false;
    if (g * gw < nv && t * TJ < nv) {
        const TileHdr ih = ihdr[g];
        const double eps = 2.0 * soft * sc->rmax;
        const double eps2x = eps * eps * 67108864.0;  // 2^26
        const double hi = sqrt((double)ih.hx * ih.hx + (double)ih.hy * ih.hy + (double)ih.hz * ih.hz);
        near = tile_near(ih, hi, thdr[t], theta, eps2x);
    }
    const unsigned long long bits = __ballot(near);
    if (threadIdx.x == 0) mask[(size_t)g * nw + blockIdx.y] = bits;
}

// ---- 6. pair kernel -------------------------------------------------------
// MODE 0: FP64-fast on every tile; MODE 1: mixed (near FP64, far FP32).
// SRC 0: the form picked per wavefront; near tiles read with wave-uniform
//        addresses (scalar loads), far tiles staged by the wavefront in its
//        own LDS slice and read as broadcasts (a far term takes three per-j
//        operands in one instruction, more than the one scalar operand a
//        VALU instruction can read);
// SRC 1: tile data staged in LDS by the workgroup, read as broadcasts; the
//        form picked per workgroup (near when any of its wavefronts is near).
// Lane `lane` of wavefront w owns bodies i0 + w 64 IPL + q 64 + lane.
template <int MODE, int IPL, int SRC>
__global__ __launch_bounds__(NT) void k_pairs(const double4 *__restrict__ jd,
                                              const double *__restrict__ jsig,
                                              const float4 *__restrict__ ju,
                                              const TileHdr *__restrict__ thdr,
                                              const TileHdr *__restrict__ ihdr,
                                              const Scal *__restrict__ sc, int nsplit,
                                              double theta, double soft, int nslots,
                                              double *__restrict__ part,
                                              const unsigned long long *__restrict__ mask, int nw)
{
    constexpr bool LDS = SRC == 1;
    constexpr int GW = WV * IPL;  // bodies of one wavefront group
    __shared__ float4 lu[LDS ? TJ : (NT / WV) * TJ];
    __shared__ double4 ld[LDS ? TJ : 1];
    __shared__ double ls[LDS ? TJ : 1];
    const int nv = sc->nvalid;
    const int iblk = blockIdx.x;
    const int i0 = iblk * NT * IPL;
    if (i0 >= nv) return;
    const int ntiles = (nv + TJ - 1) / TJ;
    const int per = (ntiles + nsplit - 1) / nsplit;
    const int t0 = blockIdx.y * per;
    const int t1 = min(ntiles, t0 + per);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WV);
    const int lane = (int)threadIdx.x % WV;
    const int g0 = i0 + w * GW;  // the wavefront group's first body
    const bool gvalid = g0 < nv;
    if (!LDS && !gvalid) return;

    const int gidx = gvalid ? iblk * (NT / WV) + w : 0;  // the group's decision row
    (void)ihdr;  // the group headers enter through k_near_mask
    const unsigned long long *mrow = MODE == 1 ? mask + (size_t)gidx * nw : nullptr;
    double xi[IPL], yi[IPL], zi[IPL], si[IPL];
    double ax[IPL], ay[IPL], az[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = min(g0 + q * WV + lane, nv - 1);
        const double4 p = jd[i];
        xi[q] = p.x;
        yi[q] = p.y;
        zi[q] = p.z;
        si[q] = jsig[i];
        ax[q] = ay[q] = az[q] = 0.0;
    }
    for (int t = t0; t < t1; ++t) {
        const TileHdr th = thdr[t];
        const int jb = t * TJ;
        const int cnt = th.count;
        bool near = true;
        if (MODE == 1) {
            // The decision of k_near_mask for this group and tile.
            near = !gvalid || ((mrow[t >> 6] >> (t & 63)) & 1ull) != 0;
            if (LDS) near = __syncthreads_or(near && gvalid);
        }
        if (near) {
            if (LDS) {
                __syncthreads();
                if ((int)threadIdx.x < cnt) {
                    ld[threadIdx.x] = jd[jb + threadIdx.x];
                    ls[threadIdx.x] = jsig[jb + threadIdx.x];
                }
                __syncthreads();
            }
            if (jb < g0 + GW && jb + TJ > g0) {
                // The tile holds some of the wavefront's own bodies.
                const int is0 = g0 + lane;
                for (int jj = 0; jj < cnt; ++jj)
                    near_term_self<IPL>(LDS ? ld[jj] : jd[jb + jj], LDS ? ls[jj] : jsig[jb + jj],
                                        jb + jj, is0, WV, xi, yi, zi, si, ax, ay, az);
            } else if (cnt == TJ) {
#pragma unroll 4
                for (int jj = 0; jj < TJ; ++jj)
                    near_term<IPL>(LDS ? ld[jj] : jd[jb + jj], LDS ? ls[jj] : jsig[jb + jj], xi,
                                   yi, zi, si, ax, ay, az);
            } else {
                for (int jj = 0; jj < cnt; ++jj)
                    near_term<IPL>(LDS ? ld[jj] : jd[jb + jj], LDS ? ls[jj] : jsig[jb + jj], xi,
                                   yi, zi, si, ax, ay, az);
            }
        } else {
            float4 *tu = LDS ? lu : lu + w * TJ;
            if (LDS) {
                __syncthreads();
                if ((int)threadIdx.x < TJ) lu[threadIdx.x] = ju[jb + threadIdx.x];
                __syncthreads();
            } else {
                // The wavefront's own slice: ordered after its reads of the
                // previous far tile and before the reads below.
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int r = 0; r < TJ / WV; ++r) tu[r * WV + lane] = ju[jb + r * WV + lane];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            // w_i = A_J - x_i, rounded once from FP64.
            float wx[IPL], wy[IPL], wz[IPL], fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
            for (int q = 0; q < IPL; ++q) {
                wx[q] = (float)(th.ax - xi[q]);
                wy[q] = (float)(th.ay - yi[q]);
                wz[q] = (float)(th.az - zi[q]);
                fx[q] = fy[q] = fz[q] = 0.f;
            }
            // Slots past cnt hold massless records (exact zeros), so whole
            // tiles are walked.
#pragma unroll 8
            for (int jj = 0; jj < TJ; ++jj) far_term<IPL>(tu[jj], wx, wy, wz, fx, fy, fz);
#pragma unroll
            for (int q = 0; q < IPL; ++q) {
                ax[q] += (double)fx[q];
                ay[q] += (double)fy[q];
                az[q] += (double)fz[q];
            }
        }
    }
    double *o = part + (size_t)blockIdx.y * 3 * (size_t)nslots;
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = g0 + q * WV + lane;
        if (i < nv) {
            o[3 * (size_t)i + 0] = ax[q];
            o[3 * (size_t)i + 1] = ay[q];
            o[3 * (size_t)i + 2] = az[q];
        }
    }
}

// ---- 6b. small n: direct form ------------------------------------------------
// Up to SMALL_N bodies both fast forms skip the spatial order and the tiles:
// one wavefront per body, the workgroup's four bodies sharing the j records
// staged in LDS NT at a time.  Each lane takes the j of its columns in the
// FP64 form (near_term); at this size that costs less than the tiled
// pipeline's set-up launches.  An absent j (or one past n) is staged with
// zero mass at the origin and softening length 1, so its term is exactly 0.
// The lanes' partials are added by a fixed butterfly, then the central term
// as k_finish adds it.
constexpr int SMALL_N = 4096;

// SOFT false: no softening, the near terms skip s_k + s_j (same bits).
template <bool SOFT>
__global__ __launch_bounds__(NT) void k_small(const DBody *__restrict__ b, int n,
                                              const int32_t *__restrict__ ex, int cp, double cx,
                                              double cy, double cz, double mu,
                                              const double *__restrict__ dcen, double G,
                                              double soft, double *__restrict__ acc,
                                              double *__restrict__ cterm)
{
    __shared__ double4 sp[NT];
    __shared__ double ss[NT];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * (NT / 64) + w;
    const bool mine = k < n && (!ex || ex[k] != 0);
    double X = 0.0, Y = 0.0, Z = 0.0, Sk = 0.0;
    if (mine) {
        const DBody q = b[k];
        X = q.x;
        Y = q.y;
        Z = q.z;
        Sk = soft * q.radius;
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int base = 0; base < n; base += NT) {
        const int j = base + (int)threadIdx.x;
        __syncthreads();
        if (j < n && (!ex || ex[j] != 0)) {
            const DBody q = b[j];
            sp[threadIdx.x] = make_double4(q.x, q.y, q.z, G * q.mass);
            ss[threadIdx.x] = soft * q.radius;
        } else {
            sp[threadIdx.x] = make_double4(0.0, 0.0, 0.0, 0.0);
            ss[threadIdx.x] = 1.0;
        }
        __syncthreads();
        if (k >= base && k < base + NT) {
            // This chunk holds body k itself: its own term is zeroed
            // (near_term_self).
#pragma unroll
            for (int u = 0; u < NT / 64; ++u)
                near_term_self<1, SOFT>(sp[lane + 64 * u], ss[lane + 64 * u],
                                        base + lane + 64 * u, k, 0, &X, &Y, &Z, &Sk, &ax, &ay,
                                        &az);
        } else {
#pragma unroll
            for (int u = 0; u < NT / 64; ++u)
                near_term<1, SOFT>(sp[lane + 64 * u], ss[lane + 64 * u], &X, &Y, &Z, &Sk, &ax, &ay,
                                   &az);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_xor(ax, off);
        ay += __shfl_xor(ay, off);
        az += __shfl_xor(az, off);
    }
    if (!mine || lane != 0) return;
    if (dcen) {
        cx = dcen[0];
        cy = dcen[1];
        cz = dcen[2];
        mu = G * dcen[6];
        cp = dcen[7] != 0.0;
    }
    if (cp) {
        const DBody q = b[k];
        const double DX = q.x - cx, DY = q.y - cy, DZ = q.z - cz;
        const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
        const double r3 = rk * rk * rk;
        ax += -mu * DX / r3;
        ay += -mu * DY / r3;
        az += -mu * DZ / r3;
        if (cterm) {
            cterm[3 * (size_t)k + 0] = G * q.mass * DX / r3;
            cterm[3 * (size_t)k + 1] = G * q.mass * DY / r3;
            cterm[3 * (size_t)k + 2] = G * q.mass * DZ / r3;
        }
    }
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

// ---- 7. finish ------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_finish(const DBody *__restrict__ b,
                                               const int *__restrict__ perm,
                                               const Scal *__restrict__ sc,
                                               const double *__restrict__ part, int nsplit,
                                               int nslots, int cp, double cx, double cy, double cz,
                                               double mu, const double *__restrict__ dcen,
                                               double G, double *__restrict__ acc,
                                               double *__restrict__ cterm)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= sc->nvalid) return;
    if (dcen) {  // central body state {x, y, z, vx, vy, vz, M, present} on the device
        cx = dcen[0];
        cy = dcen[1];
        cz = dcen[2];
        mu = G * dcen[6];
        cp = dcen[7] != 0.0;
    }
    const int k = perm[s];
    double a[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < nsplit; ++p) {
        const double *o = part + (size_t)p * 3 * (size_t)nslots + 3 * (size_t)s;
        a[0] += o[0];
        a[1] += o[1];
        a[2] += o[2];
    }
    if (cp) {
        const DBody q = b[k];
        const double DX = q.x - cx, DY = q.y - cy, DZ = q.z - cz;
        const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
        const double r3 = rk * rk * rk;
        a[0] += -mu * DX / r3;
        a[1] += -mu * DY / r3;
        a[2] += -mu * DZ / r3;
        if (cterm) {
            cterm[3 * (size_t)k + 0] = G * q.mass * DX / r3;
            cterm[3 * (size_t)k + 1] = G * q.mass * DY / r3;
            cterm[3 * (size_t)k + 2] = G * q.mass * DZ / r3;
        }
    }
    acc[3 * (size_t)k + 0] = a[0];
    acc[3 * (size_t)k + 1] = a[1];
    acc[3 * (size_t)k + 2] = a[2];
}

// Deterministic sum of the central reaction terms of existing bodies.
__global__ __launch_bounds__(NT) void k_sum3(const double *__restrict__ v, int n,
                                             const int32_t *__restrict__ ex,
                                             double *__restrict__ part)
{
    __shared__ double sh[NT];
    double a[3] = {0.0, 0.0, 0.0};
    for (int k = blockIdx.x * NT + threadIdx.x; k < n; k += gridDim.x * NT) {
        if (ex && !ex[k]) continue;
        a[0] += v[3 * (size_t)k];
        a[1] += v[3 * (size_t)k + 1];
        a[2] += v[3 * (size_t)k + 2];
    }
    for (int c = 0; c < 3; ++c) {
        const double r = block_reduce(a[c], sh, 2);
        if (threadIdx.x == 0) part[3 * blockIdx.x + c] = r;
    }
}

__global__ __launch_bounds__(NT) void k_sum3_final(const double *__restrict__ part, int np,
                                                   double *__restrict__ out)
{
    __shared__ double sh[NT];
    double a[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < np; q += NT) {
        a[0] += part[3 * q];
        a[1] += part[3 * q + 1];
        a[2] += part[3 * q + 2];
    }
    for (int c = 0; c < 3; ++c) {
        const double r = block_reduce(a[c], sh, 2);
        if (threadIdx.x == 0) out[c] = r;
    }
}

}  // namespace

// Tuning knobs of the tiled pipeline (chosen by the autotuner; defaults here).

extern "C" ncg_status ncg_set_tiling(ncg_ctx *ctx, int32_t ipl, int32_t nsplit, double theta,
                                     int32_t source)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(ipl == 0 || ipl == 1 || ipl == 2 || ipl == 4, "ipl must be 1, 2 or 4 (0 keeps)");
    NCG_REQUIRE(nsplit >= -1, "nsplit must be >= 0 (0 = automatic) or -1 (keep)");
    NCG_REQUIRE(source >= -1 && source <= 1, "source must be 0 (scalar), 1 (LDS) or -1 (keep)");
    TiledCfg &c = ctx->tiled;
    ctx->tiled_explicit = true;
    if (ipl > 0) c.ipl = ipl;
    if (nsplit >= 0) c.nsplit = nsplit;
    if (theta > 0.0) c.theta = theta;
    if (source >= 0) c.src = source;
    if (ctx->rk4) ctx->buf_gen++;  // captured graphs hold the old launch shape
    return NCG_OK;
}


template <int MODE, int IPL>
static void launch_pairs(int src, dim3 grid, hipStream_t st, const double4 *jd,
                         const double *jsig, const float4 *ju, const TileHdr *th,
                         const TileHdr *ih, const Scal *sc, int nsplit, double theta, double soft,
                         int n, double *part, const unsigned long long *mask, int nw)
{
    if (src == 1)
        hipLaunchKernelGGL((k_pairs<MODE, IPL, 1>), grid, dim3(NT), 0, st, jd, jsig, ju, th, ih, sc,
                           nsplit, theta, soft, n, part, mask, nw);
    else
        hipLaunchKernelGGL((k_pairs<MODE, IPL, 0>), grid, dim3(NT), 0, st, jd, jsig, ju, th, ih, sc,
                           nsplit, theta, soft, n, part, mask, nw);
}

// Grid shape of the tiled forms for n slots (same in order and eval).
struct TiledShape {
    int ipl, iblock, nib, ngroups, ntiles, nsplit, nbb, src;
    double theta;
};

static TiledShape tiled_shape(ncg_ctx *ctx, int32_t n, int prec)
{
    const TiledCfg cfg = ncg_tiled_cfg(ctx, prec, n);
    TiledShape t;
    t.ipl = cfg.ipl;
    t.theta = cfg.theta;
    t.src = cfg.src;
    t.iblock = NT * t.ipl;
    t.nib = (int)cdiv(n, t.iblock);
    t.ngroups = t.nib * (NT / WV);  // wavefront groups of 64 ipl bodies
    t.ntiles = (int)cdiv(n, TJ);
    int nsplit = cfg.nsplit;
    if (nsplit <= 0) {
        // Enough workgroups for 64 per CU: the near tiles make the work per
        // workgroup uneven, and small workgroups balance it.
        const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
        nsplit = (int)cdiv(256L * cus, t.nib);
    }
    t.nsplit = std::max(1, std::min(nsplit, t.ntiles));
    t.nbb = std::min(1024, (int)cdiv(n, NT));
    return t;
}

// Device buffers of the tiled pipeline.  The order part (bounding box, keys,
// sort) and the evaluation part are used by separate calls: the RK4 and
// leapfrog drivers sort eagerly and capture only the evaluation in graphs.
struct TiledWork {
    Scal *sc = nullptr;
    double *bpart = nullptr;
    unsigned *key = nullptr, *key2 = nullptr;
    int *idx = nullptr, *perm = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
    double4 *jd = nullptr;
    double *jsig = nullptr;
    float4 *ju = nullptr;
    TileHdr *th = nullptr, *ih = nullptr;
    double *part = nullptr;
    double *cterm = nullptr, *p3 = nullptr;
    unsigned long long *mask = nullptr;  // near/far decisions (mixed form)
};

// 64-bit words of one group's row of decisions.
static int mask_words(const TiledShape &t) { return (int)cdiv(t.ntiles, 64); }

// Temporary storage of the order's sort, for both key widths (32-bit Morton
// keys, 1-bit existence keys).
static ncg_status sort_bytes(ncg_ctx *ctx, int32_t n, hipStream_t st, size_t *bytes)
{
    size_t b32 = 0, b1 = 0;
    NCG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b32, (unsigned *)nullptr,
                                               (unsigned *)nullptr, (int *)nullptr, (int *)nullptr,
                                               n, 0, 32, st));
    NCG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, (unsigned *)nullptr,
                                               (unsigned *)nullptr, (int *)nullptr, (int *)nullptr,
                                               n, 0, 1, st));
    *bytes = std::max(b32, b1);
    return NCG_OK;
}

// Buffers of the context (grown on demand): order part, evaluation part.
static ncg_status work_order_ctx(ncg_ctx *ctx, const TiledShape &t, int32_t n, hipStream_t st,
                                 TiledWork &w)
{
    w.sc = dbuf<Scal>(ctx, "t_scal", 1);
    w.bpart = dbuf<double>(ctx, "t_bpart", 8 * (size_t)t.nbb);
    w.key = dbuf<unsigned>(ctx, "t_key", n);
    w.key2 = dbuf<unsigned>(ctx, "t_key2", n);
    w.idx = dbuf<int>(ctx, "t_idx", n);
    w.perm = dbuf<int>(ctx, "t_perm", n);
    if (!w.sc || !w.bpart || !w.key || !w.key2 || !w.idx || !w.perm) return NCG_ERR_NOMEM;
    ncg_status r = sort_bytes(ctx, n, st, &w.sort_bytes);
    if (r != NCG_OK) return r;
    w.sort_tmp = ncg_dbuf(ctx, "t_sorttmp", w.sort_bytes);
    return w.sort_tmp ? NCG_OK : NCG_ERR_NOMEM;
}

static ncg_status work_eval_ctx(ncg_ctx *ctx, const TiledShape &t, int32_t n, bool want_cm,
                                double *d_cterm, TiledWork &w)
{
    w.sc = dbuf<Scal>(ctx, "t_scal", 1);
    w.perm = dbuf<int>(ctx, "t_perm", n);
    w.jd = dbuf<double4>(ctx, "t_jd", (size_t)t.ntiles * TJ + t.iblock);
    w.jsig = dbuf<double>(ctx, "t_jsig", (size_t)t.ntiles * TJ + t.iblock);
    w.ju = dbuf<float4>(ctx, "t_ju", (size_t)t.ntiles * TJ);
    w.th = dbuf<TileHdr>(ctx, "t_th", t.ntiles);
    w.ih = dbuf<TileHdr>(ctx, "t_ih", t.ngroups);
    w.part = dbuf<double>(ctx, "t_part", (size_t)t.nsplit * 3 * n);
    w.cterm = d_cterm ? d_cterm : (want_cm ? dbuf<double>(ctx, "cterm", 3 * (size_t)n) : nullptr);
    w.p3 = want_cm ? dbuf<double>(ctx, "sum3", 3 * (size_t)t.nbb) : nullptr;
    w.mask = dbuf<unsigned long long>(ctx, "t_mask", (size_t)t.ngroups * mask_words(t));
    if (!w.sc || !w.perm || !w.jd || !w.jsig || !w.ju || !w.th || !w.ih || !w.part ||
        (want_cm && (!w.cterm || !w.p3)) || !w.mask)
        return NCG_ERR_NOMEM;
    return NCG_OK;
}

// The order of the evaluations that follow: spatial (Morton keys over the
// bounding cube) for the mixed form, whose near/far test needs compact tiles;
// body order (existing bodies first) for the FP64-fast form.
static ncg_status tiled_order_w(ncg_ctx *ctx, const TiledShape &t, const TiledWork &w,
                                const DBody *d_b, int32_t n, const int32_t *d_ex,
                                ncg_precision prec, hipStream_t st)
{
    const int spatial = prec == NCG_MIXED ? 1 : 0;
    hipLaunchKernelGGL(k_bbox_partial, dim3(t.nbb), dim3(NT), 0, st, d_b, n, d_ex, w.bpart);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(NT), 0, st, w.bpart, t.nbb, w.sc);
    hipLaunchKernelGGL(k_morton, dim3(cdiv(n, NT)), dim3(NT), 0, st, d_b, n, d_ex, w.sc, spatial,
                       w.key, w.idx);
    NCG_CHECK_LAUNCH();
    size_t bytes = w.sort_bytes;
    NCG_HIP(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, bytes, w.key, w.key2, w.idx, w.perm, n,
                                               0, spatial ? 32 : 1, st));
    return NCG_OK;
}

ncg_status ncg_tiled_order(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                           ncg_precision prec, hipStream_t st)
{
    if (n <= SMALL_N) return NCG_OK;  // the direct form needs no order
    const TiledShape t = tiled_shape(ctx, n, NCG_MIXED);
    TiledWork w;
    ncg_status r = work_order_ctx(ctx, t, n, st, w);
    if (r != NCG_OK) return r;
    return tiled_order_w(ctx, t, w, d_b, n, d_ex, prec, st);
}

int ncg_small_partials(int32_t n) { return std::max(1, std::min(1024, (int)cdiv(n, NT))); }

ncg_status ncg_small_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                            const ncg_central *c, const double *d_central, double G, double soft,
                            double *d_acc, double *d_cterm, double *d_p3, double *d_cm_acc,
                            hipStream_t st)
{
    NCG_REQUIRE(n >= 1 && n <= SMALL_N, "the direct form takes 1..%d bodies", SMALL_N);
    const int cp = (c && c->present) ? 1 : 0;
    const dim3 grid(std::max(1, (int)cdiv(n, NT / 64)));
    if (soft != 0.0)
        hipLaunchKernelGGL(k_small<true>, grid, dim3(NT), 0, st, d_b, n, d_ex, cp,
                           cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0,
                           cp ? G * c->mass : 0.0, d_central, G, soft, d_acc, d_cterm);
    else
        hipLaunchKernelGGL(k_small<false>, grid, dim3(NT), 0, st, d_b, n, d_ex, cp,
                           cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0,
                           cp ? G * c->mass : 0.0, d_central, G, soft, d_acc, d_cterm);
    NCG_CHECK_LAUNCH();
    if (d_cm_acc) {
        const int nb = ncg_small_partials(n);
        hipLaunchKernelGGL(k_sum3, dim3(nb), dim3(NT), 0, st, d_cterm, n, d_ex, d_p3);
        hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(NT), 0, st, d_p3, nb, d_cm_acc);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}

static ncg_status tiled_eval_w(ncg_ctx *ctx, const TiledShape &t, const TiledWork &w,
                               const DBody *d_b, int32_t n, const int32_t *d_ex,
                               const ncg_central *c, const double *d_central, double G,
                               double soft, ncg_precision prec, double *d_acc, double *d_cm_acc,
                               hipStream_t st)
{
    const bool want_cm = d_cm_acc != nullptr;
    hipLaunchKernelGGL(k_gather, dim3(cdiv(n, NT)), dim3(NT), 0, st, d_b, n, w.perm, w.sc, G, soft,
                       w.jd, w.jsig);
    hipLaunchKernelGGL(k_tiles, dim3(t.ntiles), dim3(NT), 0, st, w.jd, w.sc, TJ, w.th, w.ju);
    hipLaunchKernelGGL(k_tiles, dim3(t.ngroups), dim3(NT), 0, st, w.jd, w.sc, WV * t.ipl, w.ih,
                       (float4 *)nullptr);
    NCG_CHECK_LAUNCH();
    const dim3 grid(t.nib, t.nsplit);
    const double theta = t.theta;
    const int src = t.src;
    const double4 *jd = w.jd;
    const double *jsig = w.jsig;
    const float4 *ju = w.ju;
    const TileHdr *th = w.th, *ih = w.ih;
    const Scal *sc = w.sc;
    double *part = w.part;
    const int nw = mask_words(t);
    const unsigned long long *mask = w.mask;
    if (prec != NCG_F64_FAST) {
        hipLaunchKernelGGL(k_near_mask, dim3(t.ngroups, nw), dim3(WV), 0, st, th, ih, sc,
                           WV * t.ipl, nw, theta, soft, w.mask);
        NCG_CHECK_LAUNCH();
    }
    {
        ProfScope ps(ctx, st, prec == NCG_F64_FAST ? "pairs_fast64" : "pairs_mixed");
        if (prec == NCG_F64_FAST) {
            if (t.ipl == 1) launch_pairs<0, 1>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
            else if (t.ipl == 2) launch_pairs<0, 2>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
            else launch_pairs<0, 4>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
        } else {
            if (t.ipl == 1) launch_pairs<1, 1>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
            else if (t.ipl == 2) launch_pairs<1, 2>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
            else launch_pairs<1, 4>(src, grid, st, jd, jsig, ju, th, ih, sc, t.nsplit, theta, soft, n, part, mask, nw);
        }
        NCG_CHECK_LAUNCH();
    }
    const int cp = (c && c->present) ? 1 : 0;
    const double mu = cp ? G * c->mass : 0.0;
    hipLaunchKernelGGL(k_finish, dim3(cdiv(n, NT)), dim3(NT), 0, st, d_b, w.perm, w.sc, w.part,
                       t.nsplit, n, cp, cp ? c->x : 0.0, cp ? c->y : 0.0, cp ? c->z : 0.0, mu,
                       d_central, G, d_acc, w.cterm);
    NCG_CHECK_LAUNCH();
    if (want_cm) {
        hipLaunchKernelGGL(k_sum3, dim3(t.nbb), dim3(NT), 0, st, w.cterm, n, d_ex, w.p3);
        hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(NT), 0, st, w.p3, t.nbb, d_cm_acc);
        NCG_CHECK_LAUNCH();
    }
    return NCG_OK;
}

ncg_status ncg_tiled_eval(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                          const ncg_central *c, const double *d_central, double G, double soft,
                          ncg_precision prec, double *d_acc, double *d_cterm, double *d_cm_acc,
                          hipStream_t st)
{
    const bool want_cm = d_cm_acc != nullptr;
    if (n <= SMALL_N) {
        double *cterm = d_cterm ? d_cterm : (want_cm ? dbuf<double>(ctx, "cterm", 3 * (size_t)n) : nullptr);
        double *p3 = want_cm ? dbuf<double>(ctx, "sum3", 3 * (size_t)ncg_small_partials(n)) : nullptr;
        if (want_cm && (!cterm || !p3)) return NCG_ERR_NOMEM;
        ProfScope ps(ctx, st, prec == NCG_F64_FAST ? "pairs_fast64" : "pairs_mixed");
        return ncg_small_launch(ctx, d_b, n, d_ex, c, d_central, G, soft, d_acc, cterm, p3,
                                d_cm_acc, st);
    }
    const TiledShape t = tiled_shape(ctx, n, prec);
    TiledWork w;
    ncg_status r = work_eval_ctx(ctx, t, n, want_cm, d_cterm, w);
    if (r != NCG_OK) return r;
    return tiled_eval_w(ctx, t, w, d_b, n, d_ex, c, d_central, G, soft, prec, d_acc, d_cm_acc, st);
}

// acc += start value for every existing body (the fast forms with start values).
__global__ __launch_bounds__(NT) void k_add_start3(int n, const int32_t *__restrict__ ex,
                                                   const double *__restrict__ init,
                                                   double *__restrict__ acc)
{
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= n || (ex && !ex[k])) return;
    for (int c = 0; c < 3; ++c) acc[3 * (size_t)k + c] += init[3 * (size_t)k + c];
}

ncg_status ncg_gravity_dev_impl(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                const ncg_central *c, double G, double soft, ncg_precision prec,
                                double *d_acc, double *d_cterm, double *d_cm_acc, hipStream_t st,
                                const double *d_init)
{
    NCG_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return NCG_OK;
    const int cp = (c && c->present) ? 1 : 0;
    if (prec == NCG_F64_EXACT) {
        double *cterm = d_cterm;
        if (cp && !cterm) {
            cterm = dbuf<double>(ctx, "cterm", 3 * (size_t)n);
            if (!cterm) return NCG_ERR_NOMEM;
        }
        ncg_status r;
        {
            ProfScope ps(ctx, st, "exact");
            r = ncg_grav_exact_launch(ctx, d_b, n, d_ex, c, G, soft, d_acc, cterm, st, d_init);
        }
        if (r != NCG_OK) return r;
        // CM_acc in the reference's order: the terms added body by body.
        if (cp && d_cm_acc) return ncg_cm_seq_launch(ctx, cterm, n, d_ex, 1, d_cm_acc, st);
        return NCG_OK;
    }
    NCG_REQUIRE(prec == NCG_F64_FAST || prec == NCG_MIXED, "unknown precision %d", (int)prec);
    ncg_status r = ncg_tiled_order(ctx, d_b, n, d_ex, prec, st);
    if (r != NCG_OK) return r;
    r = ncg_tiled_eval(ctx, d_b, n, d_ex, cp ? c : nullptr, nullptr, G, soft, prec, d_acc, d_cterm,
                       cp ? d_cm_acc : nullptr, st);
    if (r != NCG_OK || !d_init) return r;
    hipLaunchKernelGGL(k_add_start3, dim3(cdiv(n, NT)), dim3(NT), 0, st, (int)n, d_ex, d_init, d_acc);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// ---- allocation-free evaluation ----------------------------------------------
// Scratch of ncg_gravity_allpairs_static_dev: one array per GPU, in the code
// object's static device memory (each GPU the library runs on holds it until
// the process ends).  Host code never passes `ncg_scratch` itself to a
// kernel or a copy (in host code the name is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress returns for the current
// device, looked up once per device at context creation.
static __device__ __align__(256) unsigned char ncg_scratch[NCG_STATIC_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
}  // namespace

ncg_status ncg_static_scratch_init(ncg_ctx *ctx)
{
    const int dev = ctx->device;
    NCG_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return NCG_OK;
    void *p = nullptr;
    NCG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(ncg_scratch)));
    NCG_REQUIRE(p != nullptr, "no address for the static scratch array");
    g_scratch_ptr[dev] = p;
    return NCG_OK;
}

// Carves `w` (and the exact form's per-body terms) out of [base, base + cap);
// lowers t.nsplit when the j-split partials would not fit.  Returns false when
// even one split does not fit.
static bool work_carve(unsigned char *base, size_t cap, TiledShape &t, int32_t n, bool tiled,
                       bool want_cm, size_t sort_tmp_bytes, TiledWork &w)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        const size_t at = (off + 255) & ~(size_t)255;
        if (at + bytes > cap) return nullptr;
        off = at + bytes;
        return base + at;
    };
    w = TiledWork();
    w.cterm = (double *)take(sizeof(double) * 3 * (size_t)n);
    w.p3 = (double *)take(sizeof(double) * 3 * (size_t)std::max(t.nbb, ncg_small_partials(n)));
    if (!w.cterm || !w.p3) return false;
    if (!tiled) return true;
    w.sc = (Scal *)take(sizeof(Scal));
    w.bpart = (double *)take(sizeof(double) * 8 * (size_t)t.nbb);
    w.key = (unsigned *)take(sizeof(unsigned) * (size_t)n);
    w.key2 = (unsigned *)take(sizeof(unsigned) * (size_t)n);
    w.idx = (int *)take(sizeof(int) * (size_t)n);
    w.perm = (int *)take(sizeof(int) * (size_t)n);
    w.sort_bytes = sort_tmp_bytes;
    w.sort_tmp = take(sort_tmp_bytes);
    const size_t slots = (size_t)t.ntiles * TJ + t.iblock;
    w.jd = (double4 *)take(sizeof(double4) * slots);
    w.jsig = (double *)take(sizeof(double) * slots);
    w.ju = (float4 *)take(sizeof(float4) * (size_t)t.ntiles * TJ);
    w.th = (TileHdr *)take(sizeof(TileHdr) * (size_t)t.ntiles);
    w.ih = (TileHdr *)take(sizeof(TileHdr) * (size_t)t.ngroups);
    w.mask = (unsigned long long *)take(sizeof(unsigned long long) * (size_t)t.ngroups *
                                        mask_words(t));
    if (!w.sc || !w.bpart || !w.key || !w.key2 || !w.idx || !w.perm || !w.sort_tmp || !w.jd ||
        !w.jsig || !w.ju || !w.th || !w.ih || !w.mask)
        return false;
    const size_t at = (off + 255) & ~(size_t)255;
    const size_t per_split = sizeof(double) * 3 * (size_t)n;
    if (at + per_split > cap) return false;
    t.nsplit = (int)std::min<size_t>((size_t)t.nsplit, (cap - at) / per_split);
    w.part = (double *)take(per_split * (size_t)t.nsplit);
    return w.part != nullptr;
}

// ---- C ABI ------------------------------------------------------------------
extern "C" ncg_status ncg_gravity_allpairs_static_dev(ncg_ctx *ctx, const ncg_body *d_bodies,
                                                      int32_t n, const int32_t *d_exists,
                                                      const ncg_central *central, double G,
                                                      double softening_parameter,
                                                      ncg_precision prec, double *d_acc,
                                                      double *d_cm_acc, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (d_bodies && d_acc)), "invalid arguments");
    NCG_REQUIRE(n <= NCG_STATIC_MAX_N, "n = %d exceeds the static scratch's limit of %d bodies",
                (int)n, NCG_STATIC_MAX_N);
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    if (n == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < SCRATCH_MAX_DEVICES)
            base = static_cast<unsigned char *>(g_scratch_ptr[ctx->device]);
    }
    NCG_REQUIRE(base != nullptr, "static scratch not set up for device %d", ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    const DBody *db = reinterpret_cast<const DBody *>(d_bodies);
    const int cp = (central && central->present) ? 1 : 0;
    const bool want_cm = cp && d_cm_acc;
    const bool tiled = prec != NCG_F64_EXACT && n > SMALL_N;
    TiledShape t = tiled_shape(ctx, n, prec == NCG_F64_EXACT ? NCG_MIXED : prec);
    size_t sbytes = 0;
    if (tiled) {
        ncg_status r = sort_bytes(ctx, n, st, &sbytes);
        if (r != NCG_OK) return r;
    }
    TiledWork w;
    if (!work_carve(base, NCG_STATIC_SCRATCH_BYTES, t, n, tiled, want_cm, sbytes, w))
        return ncg_fail(ctx, NCG_ERR_CAPACITY, "n = %d does not fit the static scratch (%d bytes)",
                        (int)n, NCG_STATIC_SCRATCH_BYTES);
    // Every scratch word the kernels read is written first in the same call:
    // nothing needs clearing between calls.
    if (prec == NCG_F64_EXACT) {
        ncg_status r = ncg_grav_exact_launch(ctx, db, n, d_exists, cp ? central : nullptr, G,
                                             softening_parameter, d_acc, w.cterm, st, nullptr);
        if (r != NCG_OK) return r;
        if (want_cm) return ncg_cm_seq_launch(ctx, w.cterm, n, d_exists, 1, d_cm_acc, st);
        return NCG_OK;
    }
    if (!tiled)
        return ncg_small_launch(ctx, db, n, d_exists, cp ? central : nullptr, nullptr, G,
                                softening_parameter, d_acc, w.cterm, w.p3,
                                want_cm ? d_cm_acc : nullptr, st);
    // The order's launch shape (bounding-box partials) is the mixed form's,
    // as in ncg_tiled_order.
    const TiledShape to = tiled_shape(ctx, n, NCG_MIXED);
    ncg_status r = tiled_order_w(ctx, to, w, db, n, d_exists, prec, st);
    if (r != NCG_OK) return r;
    return tiled_eval_w(ctx, t, w, db, n, d_exists, cp ? central : nullptr, nullptr, G,
                        softening_parameter, prec, d_acc, want_cm ? d_cm_acc : nullptr, st);
}

// ---- C ABI ------------------------------------------------------------------
extern "C" ncg_status ncg_gravity_allpairs_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                               const int32_t *d_exists, const ncg_central *central,
                                               double G, double softening_parameter,
                                               ncg_precision prec, double *d_acc,
                                               double *d_cm_acc, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (d_bodies && d_acc)), "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    return ncg_gravity_dev_impl(ctx, reinterpret_cast<const DBody *>(d_bodies), n, d_exists,
                                central, G, softening_parameter, prec, d_acc, nullptr, d_cm_acc,
                                pick_stream(ctx, stream));
}

// ncg_gravity_allpairs and ncg_gravity_allpairs_from (start: 3n start values, or null).
static ncg_status gravity_host(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                               const int32_t *exists, const ncg_central *central,
                               const double *start, double G, double softening_parameter,
                               ncg_precision prec, double *acc, double *cm_acc)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (bodies && acc)), "invalid arguments");
    if (n == 0) {
        if (cm_acc && central && central->present) cm_acc[0] = cm_acc[1] = cm_acc[2] = 0.0;
        return NCG_OK;
    }
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int cp = (central && central->present) ? 1 : 0;
    DBody *db = dbuf<DBody>(ctx, "h_bodies", n);
    int32_t *dex = exists ? dbuf<int32_t>(ctx, "h_exists", n) : nullptr;
    double *dacc = dbuf<double>(ctx, "h_acc", 3 * (size_t)n);
    double *dct = cp ? dbuf<double>(ctx, "h_cterm", 3 * (size_t)n) : nullptr;
    double *dcm = dbuf<double>(ctx, "h_cm", 4);
    double *dinit = start ? dbuf<double>(ctx, "h_start", 3 * (size_t)n) : nullptr;
    if (!db || (exists && !dex) || !dacc || (cp && !dct) || !dcm || (start && !dinit))
        return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(db, bodies, sizeof(DBody) * (size_t)n, hipMemcpyHostToDevice, st));
    if (start)
        NCG_HIP(hipMemcpyAsync(dinit, start, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice,
                               st));
    if (exists)
        NCG_HIP(hipMemcpyAsync(dex, exists, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    const bool exact = prec == NCG_F64_EXACT;
    ncg_status r = ncg_gravity_dev_impl(ctx, db, n, dex, central, G, softening_parameter, prec,
                                        dacc, dct, (cp && !exact) ? dcm : nullptr, st, dinit);
    if (r != NCG_OK) return r;
    // Holes keep the caller's values: copy back per existing body.
    double *hacc = hbuf<double>(ctx, "acc", 3 * (size_t)n);
    NCG_HIP(hipMemcpyAsync(hacc, dacc, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, st));
    double *hct = nullptr;
    if (cp && exact) {
        hct = hbuf<double>(ctx, "cterm", 3 * (size_t)n);
        NCG_HIP(hipMemcpyAsync(hct, dct, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, st));
    }
    double hcm[4] = {0, 0, 0, 0};
    if (cp && !exact) NCG_HIP(hipMemcpyAsync(hcm, dcm, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    for (int32_t k = 0; k < n; ++k) {
        if (exists && !exists[k]) continue;
        acc[3 * (size_t)k] = hacc[3 * (size_t)k];
        acc[3 * (size_t)k + 1] = hacc[3 * (size_t)k + 1];
        acc[3 * (size_t)k + 2] = hacc[3 * (size_t)k + 2];
    }
    if (cp && cm_acc) {
        if (exact) {
            // CM_acc accumulates body by body in index order (physics.c:80, 140-142).
            double s0 = 0., s1 = 0., s2 = 0.;
            for (int32_t k = 0; k < n; ++k) {
                if (exists && !exists[k]) continue;
                s0 += hct[3 * (size_t)k];
                s1 += hct[3 * (size_t)k + 1];
                s2 += hct[3 * (size_t)k + 2];
            }
            cm_acc[0] = s0;
            cm_acc[1] = s1;
            cm_acc[2] = s2;
        } else {
            cm_acc[0] = hcm[0];
            cm_acc[1] = hcm[1];
            cm_acc[2] = hcm[2];
        }
    }
    return NCG_OK;
}

extern "C" ncg_status ncg_gravity_allpairs(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                           const int32_t *exists, const ncg_central *central,
                                           double G, double softening_parameter,
                                           ncg_precision prec, double *acc, double *cm_acc)
{
    return gravity_host(ctx, bodies, n, exists, central, nullptr, G, softening_parameter, prec, acc,
                        cm_acc);
}

extern "C" ncg_status ncg_gravity_allpairs_from(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                                const int32_t *exists, const double *start,
                                                double G, double softening_parameter,
                                                ncg_precision prec, double *acc)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n == 0 || start, "invalid arguments");
    return gravity_host(ctx, bodies, n, exists, nullptr, start, G, softening_parameter, prec, acc,
                        nullptr);
}
