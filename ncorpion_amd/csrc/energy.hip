// Total energy of a brute-force system, FP64, for energy-drift measurements.
// NcorpiON has no energy diagnostic on this path; the definition is the one
// whose gradient is the field of vector_field (physics.c:127-146, 175-201):
//   E = sum_i m_i |v_i|^2 / 2 + M |V|^2 / 2
//     - sum_i G M m_i / |x_i - C|
//     - sum_{i<j} G m_i m_j / sqrt(|x_i - x_j|^2 + (s (R_i + R_j))^2)
// The mutual sum is evaluated as -1/2 sum_i m_i phi_i with
// phi_i = sum_{j != i} G m_j / D_ij (one lane per body i, j split over
// workgroups, partials summed in a fixed order: results do not depend on
// scheduling).
#include "ncg_internal.hpp"

namespace {

constexpr int ET = 256;  // threads per workgroup
constexpr int EC = 64;   // j chunk: one wavefront's worth of bodies

// j records {x, y, z, m, s R, 0, 0, 0}; missing bodies get m = 0 at the origin.
__global__ __launch_bounds__(ET) void k_en_pack(const DBody *__restrict__ b, int n,
                                                const int *__restrict__ ex, double soft,
                                                double *__restrict__ jv)
{
    const int i = blockIdx.x * ET + threadIdx.x;
    if (i >= n) return;
    const DBody q = b[i];
    const bool live = !ex || ex[i];
    double *o = jv + 8 * (size_t)i;
    o[0] = live ? q.x : 0.0;
    o[1] = live ? q.y : 0.0;
    o[2] = live ? q.z : 0.0;
    o[3] = live ? q.mass : 0.0;
    o[4] = live ? soft * q.radius : 0.0;
    o[5] = o[6] = o[7] = 0.0;
}

// phi partials.  r^2 is floored at 1e-200 so that two massless records at one
// point give 0 * finite; the lane's own record is masked in its own chunk.
template <bool SELF>
__device__ inline void en_term(const double *__restrict__ q, int j, int i, double xi, double yi,
                               double zi, double si, double &phi)
{
    const double dx = q[0] - xi, dy = q[1] - yi, dz = q[2] - zi;
    const double e = si + q[4];
    double D2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, e * e)));
    D2 = fmax(D2, 1e-200);
    double y = __builtin_amdgcn_rsq(D2);
    y = y * fma(-0.5 * D2, y * y, 1.5);
    if (SELF) y = j == i ? 0.0 : y;
    phi = fma(q[3], y, phi);
}

__global__ __launch_bounds__(ET) void k_en_pairs(const double *__restrict__ jv, int n, int per,
                                                 double *__restrict__ part)
{
    const int i = blockIdx.x * ET + threadIdx.x;
    const int ii = min(i, n - 1);
    const double *p = jv + 8 * (size_t)ii;
    const double xi = p[0], yi = p[1], zi = p[2], si = p[4];
    const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * ET + ((int)threadIdx.x & ~(EC - 1)));
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    double phi = 0.0;
    for (int jc = j0; jc < j1; jc += EC) {
        const int je = min(j1, jc + EC);
        if (jc == wave0) {
            for (int j = jc; j < je; ++j) en_term<true>(jv + 8 * (size_t)j, j, i, xi, yi, zi, si, phi);
        } else if (je - jc == EC) {
#pragma unroll 8
            for (int j = jc; j < jc + EC; ++j)
                en_term<false>(jv + 8 * (size_t)j, j, i, xi, yi, zi, si, phi);
        } else {
            for (int j = jc; j < je; ++j) en_term<false>(jv + 8 * (size_t)j, j, i, xi, yi, zi, si, phi);
        }
    }
    if (i < n) part[(size_t)blockIdx.y * n + i] = phi;
}

// Per body: kinetic, mutual and central potential; per-workgroup sums.
__global__ __launch_bounds__(ET) void k_en_body(const DBody *__restrict__ b, int n,
                                                const int *__restrict__ ex,
                                                const double *__restrict__ part, int nsplit,
                                                const double *__restrict__ cen, double G,
                                                double *__restrict__ blk)
{
    __shared__ double red[3][ET];
    const int i = blockIdx.x * ET + threadIdx.x;
    double kin = 0.0, pm = 0.0, pc = 0.0;
    if (i < n && (!ex || ex[i])) {
        const DBody q = b[i];
        double phi = part[i];
        for (int s = 1; s < nsplit; ++s) phi += part[(size_t)s * n + i];
        kin = 0.5 * q.mass * (q.vx * q.vx + q.vy * q.vy + q.vz * q.vz);
        pm = -0.5 * G * q.mass * phi;
        if (cen && cen[7] != 0.0) {
            const double dx = q.x - cen[0], dy = q.y - cen[1], dz = q.z - cen[2];
            pc = -G * cen[6] * q.mass / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    red[0][threadIdx.x] = kin;
    red[1][threadIdx.x] = pm;
    red[2][threadIdx.x] = pc;
    __syncthreads();
    for (int w = ET / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) blk[3 * (size_t)blockIdx.x + c] = red[c][0];
}

// out = {kinetic, mutual potential, central potential, total}.
__global__ __launch_bounds__(ET) void k_en_final(const double *__restrict__ blk, int nb,
                                                 const double *__restrict__ cen,
                                                 double *__restrict__ out)
{
    __shared__ double red[3][ET];
    double a[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nb; q += ET)
        for (int c = 0; c < 3; ++c) a[c] += blk[3 * (size_t)q + c];
    for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = ET / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double kin = red[0][0];
        if (cen && cen[7] != 0.0) kin += 0.5 * cen[6] * (cen[3] * cen[3] + cen[4] * cen[4] + cen[5] * cen[5]);
        out[0] = kin;
        out[1] = red[1][0];
        out[2] = red[2][0];
        out[3] = kin + red[1][0] + red[2][0];
    }
}

}  // namespace

extern "C" ncg_status ncg_energy_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                     const int32_t *d_exists, const double *d_central,
                                     double grav_const, double softening, double *d_out,
                                     void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_out, "invalid arguments");
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    const DBody *b = reinterpret_cast<const DBody *>(d_bodies);
    const int nb = (int)cdiv(n, ET);
    const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
    const int nchunk = (int)cdiv(n, EC);
    int nsplit = std::max(1, std::min(nchunk, (int)cdiv(8 * cus, nb)));
    const int per = (int)cdiv(nchunk, nsplit) * EC;
    nsplit = (int)cdiv(n, per);
    double *jv = dbuf<double>(ctx, "e_jv", 8 * (size_t)n);
    double *part = dbuf<double>(ctx, "e_part", (size_t)nsplit * n);
    double *blk = dbuf<double>(ctx, "e_blk", 3 * (size_t)nb);
    if (!jv || !part || !blk) return NCG_ERR_NOMEM;
    hipLaunchKernelGGL(k_en_pack, dim3(nb), dim3(ET), 0, st, b, (int)n, d_exists, softening, jv);
    {
        ProfScope ps(ctx, st, "energy");
        hipLaunchKernelGGL(k_en_pairs, dim3(nb, nsplit), dim3(ET), 0, st, jv, (int)n, per, part);
        NCG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_en_body, dim3(nb), dim3(ET), 0, st, b, (int)n, d_exists, part, nsplit,
                       d_central, grav_const, blk);
    hipLaunchKernelGGL(k_en_final, dim3(1), dim3(ET), 0, st, blk, nb, d_central, d_out);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}
