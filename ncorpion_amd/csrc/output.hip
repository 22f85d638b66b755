// output.hip -- NcorpiON's output path (display(), display.c:239-451) on the
// GPU: the decimal text that fprintf(file, "%.14lf ", v) writes for every
// body value, and the elliptic elements of cart2ell (structure.c:172-284).
//
// Text: the exact binary value m 2^e of each double times 10^D is an
// integer multiple of a power of two, m 5^D 2^(e+D); it is held in 32-bit
// limbs, shifted (rounding to nearest, ties to even, as glibc's printf does
// for every finite value), and converted to decimal by division by 10^9.
// Two passes: lengths, then text at the scanned offsets.
//
// cart2ell: FP64 in the reference's expression order (this file is built
// without contraction); the mean longitude's atan2 / atanh / sinh are taken
// on the host from libm, with arguments the kernel computed bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "ncg_internal.hpp"

namespace {

constexpr int OT = 256;
constexpr int FMT_MAXD = 17;
// m 5^D < 2^93, shifted left by at most 971 + 17 bits: 1081 bits, 34 limbs.
constexpr int LIMBS = 35;

// Length of printf("%.<D>f ", v) as glibc writes it; the text too when o
// is not null.  A '-' for every set sign bit (-0.0 and negatives that round
// to zero included), "nan" / "inf" for the non-finite values.
__device__ int fmt_fixed(double v, int D, char *o)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const int neg = (int)(bits >> 63);
    const int bexp = (int)((bits >> 52) & 0x7ff);
    unsigned long long m = bits & 0xfffffffffffffull;
    int len = 0;
    if (neg) {
        if (o) o[len] = '-';
        ++len;
    }
    if (bexp == 0x7ff) {
        const char *w = m ? "nan " : "inf ";
        for (int i = 0; i < 4; ++i) {
            if (o) o[len] = w[i];
            ++len;
        }
        return len;
    }
    int e = -1074;
    if (bexp) {
        m |= 1ull << 52;
        e = bexp - 1075;
    }
    unsigned long long p5 = 1;
    for (int i = 0; i < D; ++i) p5 *= 5;
    // P = m 5^D as (hi, lo).
    const unsigned long long lo = m * p5, hi = __umul64hi(m, p5);
    unsigned L[LIMBS];
    int top = 0;  // limbs in use
    const int sh = e + D;
    if (sh < 0) {
        const int s = -sh;
        unsigned long long nlo = 0, nhi = 0;
        if (s < 128) {
            // N = P >> s, rounded: half bit and sticky bits below it.
            bool half, sticky;
            if (s < 64) {
                nlo = s ? (lo >> s) | (hi << (64 - s)) : lo;
                nhi = s ? hi >> s : hi;
                half = s ? (lo >> (s - 1)) & 1 : false;
                sticky = s > 1 ? (lo & ((1ull << (s - 1)) - 1)) != 0 : false;
            } else {
                const int t = s - 64;
                nlo = t ? hi >> t : hi;
                nhi = 0;
                half = t ? (hi >> (t - 1)) & 1 : (lo >> 63) & 1;
                sticky = t ? ((hi & ((1ull << (t - 1)) - 1)) != 0 || lo != 0)
                           : (lo & ((1ull << 63) - 1)) != 0;
            }
            if (half && (sticky || (nlo & 1))) {
                ++nlo;
                if (nlo == 0) ++nhi;
            }
        }
        L[0] = (unsigned)nlo;
        L[1] = (unsigned)(nlo >> 32);
        L[2] = (unsigned)nhi;
        L[3] = (unsigned)(nhi >> 32);
        top = 4;
    } else {
        // N = P << sh (exact).
        for (int i = 0; i < LIMBS; ++i) L[i] = 0;
        const unsigned P[4] = {(unsigned)lo, (unsigned)(lo >> 32), (unsigned)hi, (unsigned)(hi >> 32)};
        const int ls = sh >> 5, bs = sh & 31;
        for (int i = 0; i < 4; ++i) {
            const unsigned long long w = (unsigned long long)P[i] << bs;
            L[i + ls] |= (unsigned)w;
            if (i + ls + 1 < LIMBS) L[i + ls + 1] |= (unsigned)(w >> 32);
        }
        top = ls + 5 < LIMBS ? ls + 5 : LIMBS;
    }
    while (top > 0 && L[top - 1] == 0) --top;
    // Decimal digits of N, least significant first, by division by 10^9.
    unsigned ch[40];
    int nch = 0;
    while (top > 0) {
        unsigned long long r = 0;
        for (int i = top - 1; i >= 0; --i) {
            const unsigned long long cur = (r << 32) | L[i];
            L[i] = (unsigned)(cur / 1000000000ull);
            r = cur % 1000000000ull;
        }
        ch[nch++] = (unsigned)r;
        while (top > 0 && L[top - 1] == 0) --top;
    }
    int nd = 0;  // digits of N (0 for N = 0)
    if (nch) {
        unsigned t = ch[nch - 1];
        int d = 0;
        while (t) {
            ++d;
            t /= 10;
        }
        nd = 9 * (nch - 1) + d;
    }
    const int T = nd > D + 1 ? nd : D + 1;  // digits printed, at least "0." + D
    const int total = len + T + (D > 0 ? 1 : 0) + 1;
    if (o) {
        // Digit i from the right goes before or after the point.
        int ci = 0, cd = 0;
        unsigned cur = nch ? ch[0] : 0;
        for (int i = 0; i < T; ++i) {
            char c = '0';
            if (i < nd) {
                c = (char)('0' + cur % 10);
                cur /= 10;
                if (++cd == 9 && ++ci < nch) {
                    cur = ch[ci];
                    cd = 0;
                }
            }
            const int pos = len + (T - 1 - i) + (D > 0 && i < D ? 1 : 0);
            o[pos] = c;
        }
        if (D > 0) o[len + T - D] = '.';
        o[total - 1] = ' ';
    }
    return total;
}

__global__ __launch_bounds__(OT) void k_fmt_len(const double *__restrict__ v, long long n, int D,
                                                unsigned long long *__restrict__ len)
{
    const long long i = (long long)blockIdx.x * OT + threadIdx.x;
    if (i < n) len[i] = (unsigned long long)fmt_fixed(v[i], D, nullptr);
}

__global__ __launch_bounds__(OT) void k_fmt_write(const double *__restrict__ v, long long n, int D,
                                                  const unsigned long long *__restrict__ off,
                                                  char *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * OT + threadIdx.x;
    if (i < n) fmt_fixed(v[i], D, out + off[i]);
}

// cart2ell (structure.c:172-284, J2 off) of every existing body k with
// mu_k = mu + G m_k (display.c:340) relative to the central state cen
// (x, y, z, vx, vy, vz, or null: origin at rest).  el: a, -, k, h, q, p;
// the mean longitude from the host: mode 0: 0; 1: atan2(x0, x1) - x2;
// 2: x1 sinh(E) - E + atan2(h, k) with E = atanh(x0).  warn: the (k, h)
// consistency test of structure.c:245.
__global__ __launch_bounds__(OT) void k_cart2ell(const DBody *__restrict__ b, int n,
                                                 const int32_t *__restrict__ ex,
                                                 const double *__restrict__ cen, double mu0,
                                                 double G, double *__restrict__ el,
                                                 double *__restrict__ xl, int *__restrict__ mode,
                                                 int *__restrict__ warn)
{
    const int k = blockIdx.x * OT + threadIdx.x;
    if (k >= n || (ex && !ex[k])) return;
    const DBody q = b[k];
    const double mu = mu0 + G * q.mass;
    double X = q.x, Y = q.y, Z = q.z, vX = q.vx, vY = q.vy, vZ = q.vz;
    if (cen) {
        X -= cen[0];
        Y -= cen[1];
        Z -= cen[2];
        vX -= cen[3];
        vY -= cen[4];
        vZ -= cen[5];
    } else {
        X -= 0.;
        Y -= 0.;
        Z -= 0.;
        vX -= 0.;
        vY -= 0.;
        vZ -= 0.;
    }
    const double R = sqrt(X * X + Y * Y + Z * Z);
    double V2 = vX * vX + vY * vY + vZ * vZ;
    const double AA = R * mu / (2. * mu - R * V2);
    const double SMU = sqrt(mu);
    vX /= SMU;
    vY /= SMU;
    vZ /= SMU;
    V2 = vX * vX + vY * vY + vZ * vZ;
    const double RV = X * vX + Y * vY + Z * vZ;
    const double C1 = Y * vZ - Z * vY, C2 = Z * vX - X * vZ, C3 = X * vY - Y * vX;
    const double CC = C1 * C1 + C2 * C2 + C3 * C3;
    const double DC = sqrt(CC);
    const double a0 = sqrt(2. * (CC + DC * C3));
    double *o = el + 6 * (size_t)k;
    o[0] = AA;
    o[4] = a0 == 0. ? 0. : -C2 / a0;
    o[5] = a0 == 0. ? 0. : C1 / a0;
    double K = 0., H = 0., FAC1 = 0., FAC2 = 0.;
    int w = 0;
    const bool flat = R == 0. || V2 == 0.;
    if (!flat) {
        const double a11 = V2 - 1. / R, a12 = RV / (R * DC), a21 = -RV, a22 = DC - R / DC;
        const double c11 = X * a11 + vX * a21, c12 = X * a12 + vX * a22;
        const double c21 = Y * a11 + vY * a21, c22 = Y * a12 + vY * a22;
        FAC1 = C1 / (DC + C3);
        FAC2 = C2 / (DC + C3);
        const double K1 = c11 - FAC1 * (Z * a11 + vZ * a21);
        const double H1 = -c12 + FAC1 * (Z * a12 + vZ * a22);
        const double H2 = c21 - FAC2 * (Z * a11 + vZ * a21);
        const double K2 = c22 - FAC2 * (Z * a12 + vZ * a22);
        w = fabs(H1 - H2) + fabs(K1 - K2) > 1.e-6;
        K = 0.5 * (K1 + K2);
        H = 0.5 * (H1 + H2);
    }
    o[2] = K;
    o[3] = H;
    o[1] = 0.;
    double *x = xl + 3 * (size_t)k;
    x[0] = x[1] = x[2] = 0.;
    int md = 0;
    if (!flat) {
        double USQA = sqrt(2. / R - V2);
        if (USQA >= 0.) {
            const double b12 = vX - FAC1 * vZ, b22 = vY - FAC2 * vZ;
            const double aux1 = (R * V2 - 1.) / (1. + DC * USQA);
            x[0] = -b12 * R * USQA + H * aux1;
            x[1] = b22 * R * USQA + K * aux1;
            x[2] = RV * USQA;
            md = 1;
        } else {
            USQA = sqrt(-(2. / R - V2));
            x[0] = RV * USQA / (R * V2 - 1.);
            x[1] = sqrt(K * K + H * H);
            md = 2;
        }
    }
    mode[k] = md;
    warn[k] = w;
}

}  // namespace

extern "C" ncg_status ncg_format_fixed(ncg_ctx *ctx, const double *values, int64_t count,
                                       int32_t decimals, char *text, int64_t capacity,
                                       int64_t *length)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(length && count >= 0 && (count == 0 || values) && capacity >= 0 &&
                    (capacity == 0 || text),
                "invalid arguments");
    NCG_REQUIRE(decimals >= 0 && decimals <= FMT_MAXD, "decimals %d outside 0..%d", (int)decimals,
                FMT_MAXD);
    NCG_REQUIRE(count < (1LL << 30), "at most 2^30 values per call");
    *length = 0;
    if (count == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    double *dv = dbuf<double>(ctx, "o_val", (size_t)count);
    unsigned long long *dl = dbuf<unsigned long long>(ctx, "o_len", (size_t)count + 1);
    unsigned long long *doff = dbuf<unsigned long long>(ctx, "o_off", (size_t)count + 1);
    unsigned long long *htot = pbuf<unsigned long long>(ctx, "o_htot", 2);
    if (!dv || !dl || !doff || !htot) return NCG_ERR_NOMEM;
    size_t tb = 0;
    NCG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dl, doff, (int)count + 1, st));
    void *tmp = ncg_dbuf(ctx, "o_scantmp", tb);
    if (!tmp) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(dv, values, sizeof(double) * (size_t)count, hipMemcpyHostToDevice, st));
    const unsigned g = cdiv(count, OT);
    hipLaunchKernelGGL(k_fmt_len, dim3(g), dim3(OT), 0, st, (const double *)dv, (long long)count,
                       (int)decimals, dl);
    NCG_CHECK_LAUNCH();
    // One more (zero) entry makes the scan's last output the total.
    NCG_HIP(hipMemsetAsync(dl + count, 0, sizeof(unsigned long long), st));
    NCG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, dl, doff, (int)count + 1, st));
    NCG_HIP(hipMemcpyAsync(htot, doff + count, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    const long long total = (long long)htot[0];
    *length = total;
    if (total > capacity)
        return ncg_fail(ctx, NCG_ERR_CAPACITY, "%lld bytes of text exceed capacity %lld", total,
                        (long long)capacity);
    char *dt = dbuf<char>(ctx, "o_text", (size_t)total);
    if (!dt) return NCG_ERR_NOMEM;
    hipLaunchKernelGGL(k_fmt_write, dim3(g), dim3(OT), 0, st, (const double *)dv, (long long)count,
                       (int)decimals, (const unsigned long long *)doff, dt);
    NCG_CHECK_LAUNCH();
    NCG_HIP(hipMemcpyAsync(text, dt, (size_t)total, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    return NCG_OK;
}

extern "C" ncg_status ncg_cart2ell(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                   const int32_t *exists, const double *central, double mu,
                                   double grav_const, double *alkhqp, int32_t *warn)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (bodies && alkhqp)), "invalid arguments");
    if (n == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DBody *db = dbuf<DBody>(ctx, "e_b", (size_t)n);
    int32_t *dex = exists ? dbuf<int32_t>(ctx, "e_ex", (size_t)n) : nullptr;
    double *dcen = dbuf<double>(ctx, "e_cen", 6);
    double *del = dbuf<double>(ctx, "e_el", 6 * (size_t)n);
    double *dxl = dbuf<double>(ctx, "e_xl", 3 * (size_t)n);
    int *dmode = dbuf<int>(ctx, "e_mode", (size_t)n);
    int *dwarn = dbuf<int>(ctx, "e_warn", (size_t)n);
    if (!db || (exists && !dex) || !dcen || !del || !dxl || !dmode || !dwarn) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(db, bodies, sizeof(DBody) * (size_t)n, hipMemcpyHostToDevice, st));
    if (exists)
        NCG_HIP(hipMemcpyAsync(dex, exists, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    if (central)
        NCG_HIP(hipMemcpyAsync(dcen, central, sizeof(double) * 6, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cart2ell, dim3(cdiv(n, OT)), dim3(OT), 0, st, (const DBody *)db, (int)n,
                       (const int32_t *)dex, central ? (const double *)dcen : nullptr, mu, grav_const,
                       del, dxl, dmode, dwarn);
    NCG_CHECK_LAUNCH();
    std::vector<double> xl(3 * (size_t)n);
    std::vector<int> md((size_t)n), wn((size_t)n);
    NCG_HIP(hipMemcpyAsync(alkhqp, del, sizeof(double) * 6 * (size_t)n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(xl.data(), dxl, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(md.data(), dmode, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(wn.data(), dwarn, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    // The mean longitude's transcendental part, with the C library's
    // functions the reference calls (structure.c:266, 278-281).
    for (int32_t k = 0; k < n; ++k) {
        if (exists && !exists[k]) continue;
        double *o = alkhqp + 6 * (size_t)k;
        const double *x = xl.data() + 3 * (size_t)k;
        if (md[k] == 1) {
            const double F = std::atan2(x[0], x[1]);
            o[1] = F - x[2];
        } else if (md[k] == 2) {
            const double E = std::atanh(x[0]);
            const double M = x[1] * std::sinh(E) - E;
            o[1] = M + std::atan2(o[3], o[2]);
        }
        if (warn) warn[k] = wn[k];
    }
    return NCG_OK;
}
