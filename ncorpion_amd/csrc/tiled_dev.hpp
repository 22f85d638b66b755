// Device pieces shared by the tiled all-pairs forms (gravity_tiled.hip: one
// system; ensemble.hip: many systems in one launch): tile headers, the FP64
// near term, the scaled FP32 far term, the far-record construction and the
// near/far decision of a (wavefront group, j tile) pair.
#pragma once
#include "ncg_internal.hpp"

namespace {

constexpr int TJ = 128;     // bodies per j tile
constexpr double TINY2 = 1e-200;  // floor of r^2 in the FP64 form (keeps a self pair's y finite)

constexpr int WV = 64;      // lanes per wavefront

struct TileHdr {
    double ax, ay, az;  // anchor (box centre)
    float hx, hy, hz;   // half extents, rounded up
    int count;          // valid bodies in the tile
    // Largest and smallest s = (G m)^(-1/2) over the tile's bodies with
    // G m > 0 (j tiles only); smax is +inf when a mass is negative or not
    // finite, which keeps the tile in the FP64 form.
    float smax, smin;
};

__device__ inline unsigned expand10(unsigned v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

template <class T>
__device__ inline T block_reduce(T v, T *sh, int op /*0 min 1 max 2 sum*/)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            T a = sh[t], c = sh[t + w];
            sh[t] = op == 0 ? (a < c ? a : c) : op == 1 ? (a > c ? a : c) : a + c;
        }
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// One interaction of the FP64 form: a += G m_j (x_j - x_i) / (r^2 + eps^2)^(3/2),
// hardware 1/sqrt refined by one Newton step (v_rsq_f64 is good to ~2^-29;
// one step brings it to FP64 rounding).  TINY2 enters the FMA chain with
// eps^2, so that a body against itself (r = 0, no softening) has a finite
// y = 1e100 (its mass is zeroed by near_term_self, so it adds exactly 0);
// above r^2 ~ 1e-184 it is below half an ulp and leaves r^2 as it is (one
// FMA instead of a multiply and a floor).
// SOFT false: no softening (s_i = s_j = 0), two instructions fewer and the
// same bits (fma(0, 0, TINY2) is TINY2).
template <int IPL, bool SOFT = true>
__device__ inline void near_term(const double4 p, double sj, const double *xi, const double *yi,
                                 const double *zi, const double *si, double *ax, double *ay,
                                 double *az)
{
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const double dx = p.x - xi[q], dy = p.y - yi[q], dz = p.z - zi[q];
        double seed = TINY2;
        if (SOFT) {
            const double e = si[q] + sj;
            seed = fma(e, e, TINY2);
        }
        const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, seed)));
        double y = __builtin_amdgcn_rsq(r2);
        const double h = 0.5 * r2;
        y = y * fma(-h, y * y, 1.5);
        const double sc3 = p.w * y * (y * y);
        ax[q] = fma(sc3, dx, ax[q]);
        ay[q] = fma(sc3, dy, ay[q]);
        az[q] = fma(sc3, dz, az[q]);
    }
}

// near_term for a j slot that may be one of the lane's own bodies (i slot q
// is is0 + q stride): that pair's mass is zeroed.  Without softening its
// r^2 is TINY2 and y^3 = 1e300, so G m y^3 would overflow to inf for
// G m > ~1.8e8 and inf * 0 poison the sum; with the mass zeroed the term is
// 0 * 1e300 * 0 = 0 whatever G m is.  Used only where the j tile overlaps
// the lane's i slots.
template <int IPL, bool SOFT = true>
__device__ inline void near_term_self(const double4 p, double sj, int js, int is0, int stride,
                                      const double *xi, const double *yi, const double *zi,
                                      const double *si, double *ax, double *ay, double *az)
{
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const double4 pq = make_double4(p.x, p.y, p.z, js == is0 + q * stride ? 0.0 : p.w);
        near_term<1, SOFT>(pq, sj, xi + q, yi + q, zi + q, si + q, ax + q, ay + q, az + q);
    }
}

// One interaction of the FP32 far form from the scaled record u = {s u_j, s}
// (k_tiles) and the lane's w_i = C - u_i: t = s w_i + s u_j, y = 1/|t|, and
// the term G m_j d / r^3 = y^3 t.
template <int IPL>
__device__ inline void far_term(const float4 u, const float *wx, const float *wy, const float *wz,
                                float *fx, float *fy, float *fz)
{
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const float tx = fmaf(u.w, wx[q], u.x);
        const float ty = fmaf(u.w, wy[q], u.y);
        const float tz = fmaf(u.w, wz[q], u.z);
        const float r2 = fmaf(tz, tz, fmaf(ty, ty, tx * tx));
        const float y = __builtin_amdgcn_rsqf(r2);
        const float y3 = (y * y) * y;
        fx[q] = fmaf(y3, tx, fx[q]);
        fy[q] = fmaf(y3, ty, fy[q]);
        fz[q] = fmaf(y3, tz, fz[q]);
    }
}


// Far-form record of body p = {x, y, z, G m} in a tile with anchor A:
// {s (x - A), s} with s = (G m)^(-1/2), each product formed in FP64 and
// rounded once.  A body without mass (and a padding slot) gets
// {1e30, 1e30, 1e30, 0}: its scaled separation is 1e30 whatever the i-body,
// |t|^2 overflows to +inf, 1/|t| = 0 and the term t / |t|^3 is exactly 0.
__device__ inline float4 far_record(const double4 p, const double *A)
{
    if (!(p.w > 0.0)) return make_float4(1e30f, 1e30f, 1e30f, 0.f);
    // The term is G m_j d / r^3 only as far as s^-2 = G m_j: of the FP32
    // values next to (G m)^-1/2, the one whose s^-2 is nearest G m (within
    // 2^-24 relative, as G m rounded to FP32 would be), and the offsets
    // scaled by that same s.
    const double sd = 1.0 / sqrt(p.w);
    float sf = (float)sd, best = sf;
    double berr = fabs(1.0 / ((double)sf * (double)sf) - p.w);
    const float cand[2] = {nextafterf(sf, 0.f), nextafterf(sf, 3.0e38f)};
    for (int c = 0; c < 2; ++c) {
        const double e = fabs(1.0 / ((double)cand[c] * (double)cand[c]) - p.w);
        if (e < berr) {
            berr = e;
            best = cand[c];
        }
    }
    const double sb = (double)best;
    return make_float4((float)(sb * (p.x - A[0])), (float)(sb * (p.y - A[1])),
                       (float)(sb * (p.z - A[2])), best);
}

// Bounds of s = (G m)^(-1/2) over a tile from the extremes of its positive
// G m (gl, gh) and whether any G m is negative or not finite (bad): smax is
// +inf for a bad tile (kept in the FP64 form), both 0 for a massless one.
// Rounded outwards, so the bounds hold for the FP32 values.
__device__ inline void tile_s_range(double gl, double gh, bool bad, TileHdr &th)
{
    if (bad) {
        th.smax = __int_as_float(0x7f800000);
        th.smin = 0.f;
    } else if (gh == 0.0) {
        th.smax = th.smin = 0.f;
    } else {
        th.smax = (float)(1.0 / sqrt(gl)) * 1.0000002f;
        th.smin = (float)(1.0 / sqrt(gh)) * 0.9999998f;
    }
}

// Near/far decision of wavefront group ih (hi = |its half extents|) against
// j tile th.  Far means: box distance dmin >= theta (anchor distance + half
// extents), which bounds the FP32 separation error; the largest softening
// length below FP32 resolution of dmin (eps2x = 2^26 eps_max^2), so the far
// form can omit it without changing an FP32 result; and every scaled
// distance s^2 r^2 of the pair of boxes within [1e-24, 1e24], so t / |t|^3
// neither overflows nor underflows in FP32.  Any NaN makes the pair near.
__device__ inline bool tile_near(const TileHdr &ih, double hi, const TileHdr &th, double theta,
                                 double eps2x)
{
    const double Cx = th.ax - ih.ax, Cy = th.ay - ih.ay, Cz = th.az - ih.az;
    const double gx = fmax(0.0, fabs(Cx) - ((double)ih.hx + th.hx));
    const double gy = fmax(0.0, fabs(Cy) - ((double)ih.hy + th.hy));
    const double gz = fmax(0.0, fabs(Cz) - ((double)ih.hz + th.hz));
    const double dmin2 = gx * gx + gy * gy + gz * gz;
    const double hj = sqrt((double)th.hx * th.hx + (double)th.hy * th.hy + (double)th.hz * th.hz);
    const double size = sqrt(Cx * Cx + Cy * Cy + Cz * Cz) + hi + hj;
    const double smax = th.smax, smin = th.smin;
    const bool range = size * size * (smax * smax) <= 1e24 &&
                       (smin == 0.0 || dmin2 * (smin * smin) >= 1e-24);
    return (dmin2 < theta * theta * size * size) || (dmin2 <= eps2x) || !range;
}

}  // namespace
