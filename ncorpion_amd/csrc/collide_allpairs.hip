// Brute-force collision search: GPU candidate filter + host replay.
//
// The reference (NcorpiON collision.c:425-456) visits every pair i < j in
// lexicographic order, calls closest_approach (collision.c:63-172) and resolves
// a hit at once, so later pairs see the updated bodies.  Here the GPU
// evaluates the closest_approach predicate for all pairs at the start-of-pass
// state and the host replays the lexicographic order: a pair whose bodies are
// untouched so far is decided by the GPU result (confirmed by the caller's own
// test), and every pair that involves a body modified earlier in the pass is
// re-tested on the host against the current state.
#include <algorithm>
#include <set>

#include "ncg_internal.hpp"

#pragma clang fp contract(off)

namespace {

// Tile edge (bodies) of the upper-triangular tiling: 128, or 64 up to
// CT_SMALL_N bodies, where 128-body tiles leave most SIMDs idle (n = 2000:
// 136 workgroups of 128 lanes each walking 128 partners; with 64, 528
// wavefronts walking 64).
constexpr int CT_SMALL_N = 16384;
__host__ __device__ constexpr int cand_tile(int n) { return n <= CT_SMALL_N ? 64 : 128; }

// closest_approach's predicate (collision.c:112-137), same operations in the
// same order.  The final time test keeps a relative slack of 2^-40 so the set
// is a superset of the exact one whatever the last bit of sqrt/div; the host
// re-evaluates every candidate exactly.
__device__ inline bool approach_hit(double xa, double ya, double za, double vxa, double vya,
                                    double vza, double ra, double xb, double yb, double zb,
                                    double vxb, double vyb, double vzb, double rb, double dt)
{
    const double dx = xa - xb, dy = ya - yb, dz = za - zb;
    const double dvx = vxa - vxb, dvy = vya - vyb, dvz = vza - vzb;
    const double b = dx * dvx + dy * dvy + dz * dvz;
    if (!(b < 0.0)) return false;
    const double dr2 = dx * dx + dy * dy + dz * dz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double R = ra + rb;
    const double disc = b * b + dv2 * (R * R - dr2);
    if (!(disc > 0.0)) return false;
    const double t = -(b + sqrt(disc)) / dv2;
    return t <= dt + fabs(dt) * 9.094947017729282e-13;
}

template <int CT>
__global__ __launch_bounds__(CT) void k_candidates(const DBody *__restrict__ b, int n,
                                                   const int32_t *__restrict__ ex, double dt,
                                                   int ntiles, int2 *__restrict__ out,
                                                   unsigned long long cap,
                                                   unsigned long long *__restrict__ count)
{
    __shared__ double sx[CT], sy[CT], sz[CT], svx[CT], svy[CT], svz[CT], sr[CT];
    __shared__ int se[CT];
    // Map blockIdx.x to an upper-triangular tile pair (I <= J).
    const long t = blockIdx.x;
    int I = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)I * (I + 1) / 2 > t) --I;
    while ((long)(I + 1) * (I + 2) / 2 <= t) ++I;
    // t = I*(I+1)/2 + r, r in [0, I]; pair (tile r, tile I) with r <= I.
    const int J = I;
    const int Itile = (int)(t - (long)I * (I + 1) / 2);
    const int i = Itile * CT + (int)threadIdx.x;
    const int jb = J * CT;
    {
        const int j = jb + (int)threadIdx.x;
        if (j < n) {
            const DBody q = b[j];
            sx[threadIdx.x] = q.x;
            sy[threadIdx.x] = q.y;
            sz[threadIdx.x] = q.z;
            svx[threadIdx.x] = q.vx;
            svy[threadIdx.x] = q.vy;
            svz[threadIdx.x] = q.vz;
            sr[threadIdx.x] = q.radius;
            se[threadIdx.x] = ex ? (ex[j] != 0) : 1;
        } else {
            se[threadIdx.x] = 0;
        }
    }
    __syncthreads();
    if (i >= n || (ex && !ex[i])) return;
    const DBody a = b[i];
    const int cnt = min(CT, n - jb);
    for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        if (j <= i || !se[jj]) continue;
        if (approach_hit(a.x, a.y, a.z, a.vx, a.vy, a.vz, a.radius, sx[jj], sy[jj], sz[jj],
                         svx[jj], svy[jj], svz[jj], sr[jj], dt)) {
            const unsigned long long slot = atomicAdd(count, 1ull);
            if (slot < cap) out[slot] = make_int2(i, j);
        }
    }
    (void)ntiles;
}

}  // namespace

ncg_status ncg_candidates_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                 double dt, int2 *d_out, unsigned long long cap,
                                 unsigned long long *d_count, hipStream_t st)
{
    const int ct = cand_tile(n);
    const int ntiles = (int)cdiv(n, ct);
    const long nblk = (long)ntiles * (ntiles + 1) / 2;
    if (nblk == 0) return NCG_OK;
    ProfScope ps(ctx, st, "candidates");
    if (ct == 64)
        hipLaunchKernelGGL(k_candidates<64>, dim3((unsigned)nblk), dim3(64), 0, st, d_b, (int)n,
                           d_ex, dt, ntiles, d_out, cap, d_count);
    else
        hipLaunchKernelGGL(k_candidates<128>, dim3((unsigned)nblk), dim3(128), 0, st, d_b, (int)n,
                           d_ex, dt, ntiles, d_out, cap, d_count);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

// Runs the candidate kernel on device buffers; returns the sorted list in `out`.
static ncg_status candidates_dev(ncg_ctx *ctx, const DBody *db, int32_t n, const int32_t *dex,
                                 double dt, std::vector<ncg_pair> &out)
{
    hipStream_t st = ctx->stream;
    const int ct = cand_tile(n);
    const int ntiles = (int)cdiv(n, ct);
    const long nblk = (long)ntiles * (ntiles + 1) / 2;
    unsigned long long cap = 1ull << 16;
    for (int attempt = 0; attempt < 2; ++attempt) {
        int2 *dout = dbuf<int2>(ctx, "cand", cap);
        unsigned long long *dcnt = dbuf<unsigned long long>(ctx, "cand_cnt", 1);
        if (!dout || !dcnt) return NCG_ERR_NOMEM;
        NCG_HIP(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), st));
        {
            ProfScope ps(ctx, st, "candidates");
            if (ct == 64)
                hipLaunchKernelGGL(k_candidates<64>, dim3((unsigned)nblk), dim3(64), 0, st, db,
                                   (int)n, dex, dt, ntiles, dout, cap, dcnt);
            else
                hipLaunchKernelGGL(k_candidates<128>, dim3((unsigned)nblk), dim3(128), 0, st, db,
                                   (int)n, dex, dt, ntiles, dout, cap, dcnt);
            NCG_CHECK_LAUNCH();
        }
        unsigned long long cnt = 0;
        NCG_HIP(hipMemcpyAsync(&cnt, dcnt, sizeof cnt, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        if (cnt <= cap) {
            out.resize(cnt);
            if (cnt) {
                NCG_HIP(hipMemcpyAsync(out.data(), dout, sizeof(int2) * cnt, hipMemcpyDeviceToHost, st));
                NCG_HIP(hipStreamSynchronize(st));
            }
            std::sort(out.begin(), out.end(), [](const ncg_pair &p, const ncg_pair &q) {
                return p.fst != q.fst ? p.fst < q.fst : p.snd < q.snd;
            });
            return NCG_OK;
        }
        cap = cnt;
    }
    return ncg_fail(ctx, NCG_ERR_DEVICE, "candidate count changed between launches");
}

static ncg_status upload_state(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                               const int32_t *exists, DBody **db, int32_t **dex)
{
    hipStream_t st = ctx->stream;
    *db = dbuf<DBody>(ctx, "c_bodies", n);
    *dex = exists ? dbuf<int32_t>(ctx, "c_exists", n) : nullptr;
    if (!*db || (exists && !*dex)) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(*db, bodies, sizeof(DBody) * (size_t)n, hipMemcpyHostToDevice, st));
    if (exists)
        NCG_HIP(hipMemcpyAsync(*dex, exists, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    return NCG_OK;
}

extern "C" ncg_status ncg_collide_candidates_allpairs(ncg_ctx *ctx, const ncg_body *bodies,
                                                      int32_t n, const int32_t *exists,
                                                      double timestep, ncg_pair *out,
                                                      int64_t capacity, int64_t *count)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || bodies) && count && capacity >= 0 && (capacity == 0 || out),
                "invalid arguments");
    *count = 0;
    if (n < 2) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    DBody *db;
    int32_t *dex;
    ncg_status r = upload_state(ctx, bodies, n, exists, &db, &dex);
    if (r != NCG_OK) return r;
    std::vector<ncg_pair> c;
    r = candidates_dev(ctx, db, n, dex, timestep, c);
    if (r != NCG_OK) return r;
    *count = (int64_t)c.size();
    if ((int64_t)c.size() > capacity)
        return ncg_fail(ctx, NCG_ERR_CAPACITY, "%zu candidates exceed capacity %lld", c.size(),
                        (long long)capacity);
    std::copy(c.begin(), c.end(), out);
    return NCG_OK;
}

extern "C" ncg_status ncg_brute_force_pass(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                           const int32_t *exists, double timestep,
                                           const ncg_replay_ops *ops, int64_t *n_collisions)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || bodies) && ops && ops->test && ops->resolve,
                "invalid arguments");
    if (n_collisions) *n_collisions = 0;
    if (n < 2) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    DBody *db;
    int32_t *dex;
    ncg_status r = upload_state(ctx, bodies, n, exists, &db, &dex);
    if (r != NCG_OK) return r;
    std::vector<ncg_pair> cand;
    r = candidates_dev(ctx, db, n, dex, timestep, cand);
    if (r != NCG_OK) return r;

    // Replay of collision.c:433-455.  `dirty` holds every body modified or
    // created so far in this pass; pairs touching it are decided on the host.
    // Rows stop at the n slots of the call (current_largest_id); columns run
    // to the live end, which fragmentation can extend (largest_id).
    std::vector<char> dirty((size_t)n, 0);
    std::set<int32_t> dirty_set;
    int32_t ncol = n;
    int64_t ncoll = 0;
    size_t ci = 0;
    std::vector<int32_t> born(64);
    auto live = [&](int32_t k) { return !exists || exists[k] != 0; };
    auto mark = [&](int32_t k) {
        if ((size_t)k >= dirty.size()) dirty.resize((size_t)k + 1, 0);
        if (!dirty[k]) {
            dirty[k] = 1;
            dirty_set.insert(k);
        }
    };
    auto collide = [&](int32_t i, int32_t j) -> ncg_status {
        ops->resolve(ops->user, i, j);
        ++ncoll;
        mark(i);
        mark(j);
        if (ops->created) {
            const int32_t nb = ops->created(ops->user, born.data(), (int32_t)born.size());
            NCG_REQUIRE(nb >= 0 && nb <= (int32_t)born.size(),
                        "created callback reported %d bodies (at most %d)", nb, (int)born.size());
            for (int32_t q = 0; q < nb; ++q) {
                NCG_REQUIRE(born[q] >= 0, "created callback reported slot %d", born[q]);
                mark(born[q]);
                ncol = std::max(ncol, born[q] + 1);
            }
        }
        return NCG_OK;
    };
    for (int32_t i = 0; i < n; ++i) {
        // Candidates of earlier rows are consumed; position the cursor on row i.
        while (ci < cand.size() && cand[ci].fst < i) ++ci;
        if (!live(i)) continue;
        int32_t jc = i + 1;  // next j to consider in this row
        // A row body removed by an outcome takes part in no further collision
        // (collision_treatment returns at physics.c:1059-1061).
        while (jc < ncol && live(i)) {
            if (dirty[i]) {
                // Row of a modified body: every remaining pair on the host.
                for (int32_t j = jc; j < ncol && live(i); ++j)
                    if (live(j) && ops->test(ops->user, i, j)) {
                        r = collide(i, j);
                        if (r != NCG_OK) return r;
                    }
                break;
            }
            while (ci < cand.size() && cand[ci].fst == i && cand[ci].snd < jc) ++ci;
            const int32_t j_cand = (ci < cand.size() && cand[ci].fst == i) ? cand[ci].snd : ncol;
            auto it = dirty_set.lower_bound(jc);
            const int32_t j_dirty = it == dirty_set.end() ? ncol : *it;
            const int32_t j = std::min(j_cand, j_dirty);
            if (j >= ncol) break;
            if (live(j) && ops->test(ops->user, i, j)) {
                r = collide(i, j);
                if (r != NCG_OK) return r;
            }
            jc = j + 1;
        }
    }
    if (n_collisions) *n_collisions = ncoll;
    return NCG_OK;
}
