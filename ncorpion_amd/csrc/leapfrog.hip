// GPU-resident leapfrog of NcorpiON's brute-force integrator
// (integration_brute_force_SABA1, rk4.c:840-954, with kick and drift of
// rk4.c:68-127): the bodies and the central body stay on the device and every
// step is
//   kick(dt)   v += dt * a(x) for every body, V += dt * CM_acc (rk4.c:68-107)
//   brute_force with dt (collision.c:425-456): candidate pass + device replay
//   drift(dt)  x += dt * v, X += dt * V (rk4.c:110-127)
//   did_collide cleared when one_collision_only (end_of_timestep, rk4.c:348-353)
// optionally preceded by the half drift of rk4.c:871-879 (brute_force and
// drift with dt/2, did_collide kept into the first step as there) and
// followed by the backward half drift of rk4.c:938-942.  The field is the
// exact form (bit for bit vector_field's, CM_acc summed in body order) or a
// tiled form; kick and drift are compiled without FMA contraction, so with
// the exact form a run reproduces the reference's arithmetic.
#include "ncg_internal.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int LT = 256;

// Bodies i < n; thread n: the central body.
__global__ __launch_bounds__(LT) void k_lf_kick(int n, const int32_t *__restrict__ ex, double h,
                                                DBody *__restrict__ b,
                                                const double *__restrict__ acc,
                                                double *__restrict__ cen,
                                                const double *__restrict__ cm)
{
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i < n) {
        if (ex && !ex[i]) return;
        b[i].vx += h * acc[3 * (size_t)i];
        b[i].vy += h * acc[3 * (size_t)i + 1];
        b[i].vz += h * acc[3 * (size_t)i + 2];
    } else if (i == n && cen[7] != 0.0) {
        cen[3] += h * cm[0];
        cen[4] += h * cm[1];
        cen[5] += h * cm[2];
    }
}

__global__ __launch_bounds__(LT) void k_lf_drift(int n, const int32_t *__restrict__ ex, double h,
                                                 DBody *__restrict__ b, double *__restrict__ cen)
{
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i < n) {
        if (ex && !ex[i]) return;
        b[i].x += h * b[i].vx;
        b[i].y += h * b[i].vy;
        b[i].z += h * b[i].vz;
    } else if (i == n && cen[7] != 0.0) {
        cen[0] += h * cen[3];
        cen[1] += h * cen[4];
        cen[2] += h * cen[5];
    }
}

__global__ void k_lf_zero_i32(int32_t *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_lf_zero_u64(unsigned long long *__restrict__ p) { p[threadIdx.x] = 0ull; }

__global__ void k_lf_off(int *__restrict__ off, int n)
{
    off[0] = 0;
    off[1] = n;
}

// three_largest_moonlets (structure.c:914-998) then three_largest_three_first
// (structure.c:1001-1064), one block.  The sequential scan keeps, after its
// first three existing bodies (ordered by its own rules, under which equal
// radii put the second body first), a list sorted by decreasing radius in
// which a later body displaces an earlier one only when strictly larger; so
// the result is those three merged with the three largest of the rest by
// (radius descending, index ascending), the first three winning ties.  Then
// the swaps into slots 0, 1, 2 (moving a body into an empty slot frees its
// old one, structure.c:1019-1060).  err |= 1 if fewer than three bodies exist.
constexpr int T3 = 1024;

struct Top3 {
    double r[3];
    int i[3];
};

__device__ inline bool t3_better(double ra, int ia, double rb, int ib)
{
    return ra > rb || (ra == rb && ia < ib);
}

__device__ inline void t3_insert(Top3 &t, double r, int i)
{
    if (i < 0) return;
    if (t.i[0] < 0 || t3_better(r, i, t.r[0], t.i[0])) {
        t.r[2] = t.r[1]; t.i[2] = t.i[1];
        t.r[1] = t.r[0]; t.i[1] = t.i[0];
        t.r[0] = r; t.i[0] = i;
    } else if (t.i[1] < 0 || t3_better(r, i, t.r[1], t.i[1])) {
        t.r[2] = t.r[1]; t.i[2] = t.i[1];
        t.r[1] = r; t.i[1] = i;
    } else if (t.i[2] < 0 || t3_better(r, i, t.r[2], t.i[2])) {
        t.r[2] = r; t.i[2] = i;
    }
}

// Per block of k_top3_part: the three largest (radius descending, index
// ascending) existing bodies after the first three existing ones, over a
// grid-stride share of the slots.
constexpr int T3B = 256;
constexpr int T3_PARTS = 1024;

__global__ __launch_bounds__(T3B) void k_top3_part(const DBody *__restrict__ b, int n,
                                                   const int32_t *__restrict__ ex,
                                                   Top3 *__restrict__ part)
{
    __shared__ Top3 sh[T3B];
    __shared__ int s_start;
    const int t = threadIdx.x;
    if (t == 0) {
        int c = 0, j = 0;
        for (; j < n && c < 3; ++j)
            if (!ex || ex[j]) ++c;
        s_start = c == 3 ? j : n;
    }
    __syncthreads();
    Top3 mine;
    for (int q = 0; q < 3; ++q) {
        mine.r[q] = 0.0;
        mine.i[q] = -1;
    }
    for (int j = s_start + blockIdx.x * T3B + t; j < n; j += gridDim.x * T3B)
        if (!ex || ex[j]) t3_insert(mine, b[j].radius, j);
    sh[t] = mine;
    __syncthreads();
    for (int w = T3B / 2; w > 0; w >>= 1) {
        if (t < w) {
            Top3 a = sh[t];
            const Top3 &o = sh[t + w];
            for (int q = 0; q < 3; ++q) t3_insert(a, o.r[q], o.i[q]);
            sh[t] = a;
        }
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = sh[0];
}

// The first three existing bodies, the partials of k_top3_part merged, then
// the swaps.  One block.
__global__ __launch_bounds__(T3) void k_top3_first(DBody *__restrict__ b, int n,
                                                   int32_t *__restrict__ ex,
                                                   const Top3 *__restrict__ part, int nparts,
                                                   unsigned *__restrict__ err)
{
    __shared__ Top3 sh[T3];
    __shared__ int s_j[3];
    const int t = threadIdx.x;
    if (t == 0) {
        int c = 0;
        for (int j = 0; j < n && c < 3; ++j)
            if (!ex || ex[j]) s_j[c++] = j;
        for (; c < 3; ++c) s_j[c] = -1;
    }
    __syncthreads();
    if (s_j[2] < 0) {
        if (t == 0) atomicOr(err, 1u);
        return;
    }
    Top3 mine;
    for (int q = 0; q < 3; ++q) {
        mine.r[q] = 0.0;
        mine.i[q] = -1;
    }
    for (int q = t; q < nparts; q += T3)
        for (int c = 0; c < 3; ++c) t3_insert(mine, part[q].r[c], part[q].i[c]);
    sh[t] = mine;
    __syncthreads();
    for (int w = T3 / 2; w > 0; w >>= 1) {
        if (t < w) {
            Top3 a = sh[t];
            const Top3 &o = sh[t + w];
            for (int q = 0; q < 3; ++q) t3_insert(a, o.r[q], o.i[q]);
            sh[t] = a;
        }
        __syncthreads();
    }
    if (t != 0) return;
    // The first three existing bodies by the scan's own rules (structure.c:932-970).
    int L[3];
    double R[3];
    const int j0 = s_j[0], j1 = s_j[1], j2 = s_j[2];
    const double r0 = b[j0].radius, r1 = b[j1].radius, r2 = b[j2].radius;
    if (r0 > r1) {
        L[0] = j0; R[0] = r0; L[1] = j1; R[1] = r1;
    } else {
        L[0] = j1; R[0] = r1; L[1] = j0; R[1] = r0;
    }
    if (R[1] >= r2) {
        L[2] = j2; R[2] = r2;
    } else if (r2 > R[1] && r2 <= R[0]) {
        L[2] = L[1]; R[2] = R[1]; L[1] = j2; R[1] = r2;
    } else {
        L[2] = L[1]; R[2] = R[1]; L[1] = L[0]; R[1] = R[0]; L[0] = j2; R[0] = r2;
    }
    // Merge with the rest: a body of the rest enters only if strictly larger.
    const Top3 rest = sh[0];
    for (int q = 0; q < 3; ++q) {
        const int i = rest.i[q];
        if (i < 0) continue;
        const double r = rest.r[q];
        if (r > R[0]) {
            L[2] = L[1]; R[2] = R[1]; L[1] = L[0]; R[1] = R[0]; L[0] = i; R[0] = r;
        } else if (r > R[1]) {
            L[2] = L[1]; R[2] = R[1]; L[1] = i; R[1] = r;
        } else if (r > R[2]) {
            L[2] = i; R[2] = r;
        }
    }
    // three_largest_three_first.
    int first = L[0], second = L[1], third = L[2];
    auto place = [&](int slot, int from) {
        if (!ex || ex[slot]) {
            const DBody tmp = b[slot];
            b[slot] = b[from];
            b[from] = tmp;
        } else {
            b[slot] = b[from];
            ex[from] = 0;
            ex[slot] = 1;
        }
    };
    if (first != 0) {
        place(0, first);
        if (second == 0) second = first;
        else if (third == 0) third = first;
    }
    if (second != 1) {
        place(1, second);
        if (third == 1) third = second;
    }
    if (third != 2) place(2, third);
}

// Existing bodies (how_many_moonlets), and whether one is within
// disruption_threshold of the central body or beyond high_dumping_threshold
// (end_of_timestep would remove it, rk4.c:159-183).
__global__ __launch_bounds__(LT) void k_lf_census(const DBody *__restrict__ b, int n,
                                                  const int32_t *__restrict__ ex,
                                                  const double *__restrict__ cen, double rlo,
                                                  double rhi, int *__restrict__ count,
                                                  unsigned *__restrict__ err)
{
    // Grid-stride share of the slots; one atomic per block for the count.
    __shared__ int wc[LT / 64];
    const int cp = cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    int c = 0;
    bool out = false;
    for (int i = blockIdx.x * LT + threadIdx.x; i < n; i += gridDim.x * LT) {
        if (ex && !ex[i]) continue;
        ++c;
        const double DX = b[i].x - cx, DY = b[i].y - cy, DZ = b[i].z - cz;
        const double r2 = DX * DX + DY * DY + DZ * DZ;
        out |= (cp && r2 < rlo * rlo) || r2 > rhi * rhi;
    }
    if (out) atomicOr(err, 2u);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < LT / 64; ++w) tot += wc[w];
        if (tot) atomicAdd(count, tot);
    }
}

__global__ void k_lf_zero_i32_one(int *__restrict__ p) { *p = 0; }

}  // namespace

namespace {

// Slots, outcome model and did_collide of a model run.
ncg_status check_model_params(ncg_ctx *ctx, const ncg_leapfrog_model_params &P,
                              int32_t one_collision_only, const int32_t *d_did_collide)
{
    NCG_REQUIRE(P.n_max >= 1 && P.largest_id >= 0 && P.largest_id < P.n_max &&
                    P.how_many_free >= 0 && P.how_many_free <= P.n_max && P.free_indexes,
                "invalid slots (n_max %d, largest_id %d, how_many_free %d)", (int)P.n_max,
                (int)P.largest_id, (int)P.how_many_free);
    const ncg_outcome_model &M = P.model;
    NCG_REQUIRE(M.kind >= NCG_OUTCOME_BOUNCE && M.kind <= NCG_OUTCOME_FRAGMENTATION,
                "unknown outcome model %d", (int)M.kind);
    NCG_REQUIRE(M.kind != NCG_OUTCOME_FRAGMENTATION ||
                    (M.n_tilde >= 1 && M.n_tilde <= NCG_MAX_TILDE && M.pq[1] >= M.pq[0] &&
                     M.pq[3] >= M.pq[2] &&
                     (int64_t)(M.pq[1] - M.pq[0] + 1) * (M.pq[3] - M.pq[2] + 1) == M.n_tilde),
                "invalid fragmentation lattice (n_tilde %d, pq %d %d %d %d)", (int)M.n_tilde,
                (int)M.pq[0], (int)M.pq[1], (int)M.pq[2], (int)M.pq[3]);
    const bool coll = M.kind != NCG_OUTCOME_BOUNCE || M.collision_f > 0.0;
    NCG_REQUIRE(!(coll && one_collision_only && !d_did_collide),
                "one_collision_only needs a did_collide array");
    for (int32_t q = 0; q < P.how_many_free; ++q)
        NCG_REQUIRE(P.free_indexes[q] >= 0 && P.free_indexes[q] < P.n_max,
                    "free slot %d out of range", (int)P.free_indexes[q]);
    return NCG_OK;
}

// NcorpiON's mesh integrator; lp == nullptr: ncg_mesh_leapfrog_dev (bounce
// with collision_f, fixed slots, no removals); else ncg_mesh_leapfrog_model_dev.
ncg_status mesh_loop_run(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n_in, int32_t *d_exists,
                         int32_t *d_did_collide, double *d_central, double grav_const,
                         double softening, ncg_precision prec, double dt, int32_t nsteps,
                         ncg_mesh_loop_params *mp, double collision_f, int32_t one_collision_only,
                         int32_t flags, ncg_leapfrog_model_params *lp, int64_t *n_collisions,
                         void *stream)
{
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    if (n_collisions) *n_collisions = 0;
    DBody *b = reinterpret_cast<DBody *>(d_bodies);
    const bool model_host = lp && lp->model.kind != NCG_OUTCOME_BOUNCE;
    const bool frag = lp && lp->model.kind == NCG_OUTCOME_FRAGMENTATION;
    // The device replay applies the bounce; under the other models a pass in
    // which it resolved a pair is replayed on the host.
    const double f = model_host ? 2.0 : (lp ? lp->model.collision_f : collision_f);
    const bool coll = model_host || f > 0.0;
    const int32_t ncap = lp ? lp->n_max : n_in;
    int32_t n = n_in;
    // Removal distances: the model's, else the mesh parameters' (checked only).
    const double r_dis = lp ? (lp->r_disruption > 0.0 ? lp->r_disruption : 0.0) : mp->r_disruption;
    const double r_hd = lp ? (lp->r_high_dumping > 0.0 ? lp->r_high_dumping : INFINITY)
                           : mp->r_high_dumping;
    double *acc = dbuf<double>(ctx, "ml_acc", 3 * (size_t)ncap);
    double *cm = dbuf<double>(ctx, "ml_cm", 4);
    unsigned *err = dbuf<unsigned>(ctx, "ml_err", 2);
    int *count = dbuf<int>(ctx, "ml_count", 1);
    int *hpin = pbuf<int>(ctx, "ml_hpin", 4);
    if (!acc || !cm || !err || !count || !hpin) return NCG_ERR_NOMEM;
    unsigned g = (unsigned)cdiv(n + 1, LT);
    const double G = grav_const, soft = softening;
    int64_t ncoll = 0;
    // Host-side slot state of a model run.
    int32_t largest_id = lp ? lp->largest_id : n - 1;
    int32_t how_many_free = lp ? lp->how_many_free : 0;
    int32_t need = lp ? lp->need_to_reduce_com : 0;
    std::vector<int32_t> free_list;
    if (lp) free_list.assign(lp->free_indexes, lp->free_indexes + ncap);
    int32_t host_step_ends = 0, removed = 0, tidy_ups = 0, host_passes = 0;
    int64_t mergers = 0, super_catastrophic = 0, partial = 0, full = 0;
    // Status of the device-side checks (top-3 and census) read with the
    // counters of each step.  *crossed: a body crossed a removal distance
    // (an error without a model).
    auto census = [&](int &how_many, bool *crossed) -> ncg_status {
        hipLaunchKernelGGL(k_lf_zero_i32_one, dim3(1), dim3(1), 0, st, count);
        hipLaunchKernelGGL(k_lf_census, dim3(std::min<long>(1024, cdiv(n, LT))), dim3(LT), 0, st, b,
                           (int)n, d_exists, d_central, r_dis, r_hd, count, err);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipMemcpyAsync(hpin, count, sizeof(int), hipMemcpyDeviceToHost, st));
        NCG_HIP(hipMemcpyAsync(hpin + 1, err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        how_many = hpin[0];
        if (hpin[1] & 1u)
            return ncg_fail(ctx, NCG_ERR_INVALID, "fewer than three bodies: three_largest_moonlets "
                                                  "is undefined");
        if (hpin[1] & 2u) {
            if (!lp)
                return ncg_fail(ctx, NCG_ERR_INVALID,
                                "a body crossed disruption_threshold or high_dumping_threshold: "
                                "end_of_timestep would remove it, which this loop does not do");
            hipLaunchKernelGGL(k_lf_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t *>(err), 2);
            NCG_CHECK_LAUNCH();
        }
        if (crossed) *crossed = (hpin[1] & 2u) != 0;
        return NCG_OK;
    };
    const int nparts_max = (int)std::min<long>(T3_PARTS, cdiv(ncap, T3B));
    Top3 *parts = dbuf<Top3>(ctx, "ml_top3", (size_t)nparts_max);
    if (!parts) return NCG_ERR_NOMEM;
    auto largest3 = [&]() {
        const int nparts = (int)std::min<long>(T3_PARTS, cdiv(n, T3B));
        hipLaunchKernelGGL(k_top3_part, dim3(nparts), dim3(T3B), 0, st, (const DBody *)b, (int)n,
                           (const int32_t *)d_exists, parts);
        hipLaunchKernelGGL(k_top3_first, dim3(1), dim3(T3), 0, st, b, (int)n, d_exists,
                           (const Top3 *)parts, nparts, err);
    };
    // Mesh pass at the current gam (rk4.c:697-700, 765-767, 774-777).
    int64_t tn = 0;
    // With the exact field, each pass starts the field's neighbourhood lists
    // on the second stream during its candidate search and replay.
    struct ListsFlag {
        MeshState &m;
        ~ListsFlag() { m.want_lists = false; }
    } lists_flag{ctx->mesh};
    ctx->mesh.want_lists = prec == NCG_F64_EXACT;
    // A mesh() pass the device replay cannot take -- a body larger than the
    // mesh size (whose visit scans every body, collision.c:381-414), more
    // candidates or pending re-tests than it holds, or, under the merger and
    // fragmentation models, any resolved pair -- is undone (a snapshot of the
    // bodies and did_collide before it) and replayed on the host with the
    // same decision and outcome (ncg_mesh_pass with the built-in ops).
    DBody *snap_b = dbuf<DBody>(ctx, "ml_snap_b", ncap);
    int32_t *snap_dc = d_did_collide ? dbuf<int32_t>(ctx, "ml_snap_dc", ncap) : nullptr;
    if (!snap_b || (d_did_collide && !snap_dc)) return NCG_ERR_NOMEM;
    mp->host_passes = 0;
    mp->conflict_passes = 0;
    std::vector<DBody> hb;
    std::vector<int32_t> hex, hdc;
    double hcen[8];
    // Host copies of the state (slots 0 .. n - 1; zero beyond) and back.
    auto download = [&](const DBody *from_b, const int32_t *from_dc) -> ncg_status {
        hb.assign((size_t)ncap, DBody{});
        hex.assign(d_exists ? (size_t)ncap : 0, 0);
        hdc.assign(d_did_collide ? (size_t)ncap : 0, 0);
        NCG_HIP(hipMemcpyAsync(hb.data(), from_b, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToHost, st));
        if (d_exists)
            NCG_HIP(hipMemcpyAsync(hex.data(), d_exists, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToHost, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(hdc.data(), from_dc, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToHost, st));
        NCG_HIP(hipMemcpyAsync(hcen, d_central, sizeof hcen, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    auto upload = [&](int32_t slots) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(b, hb.data(), sizeof(DBody) * (size_t)slots, hipMemcpyHostToDevice, st));
        if (d_exists && lp)
            NCG_HIP(hipMemcpyAsync(d_exists, hex.data(), sizeof(int32_t) * (size_t)slots,
                                   hipMemcpyHostToDevice, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(d_did_collide, hdc.data(), sizeof(int32_t) * (size_t)slots,
                                   hipMemcpyHostToDevice, st));
        if (lp) NCG_HIP(hipMemcpyAsync(d_central, hcen, sizeof hcen, hipMemcpyHostToDevice, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    auto set_n = [&](int32_t nn) {
        n = nn;
        g = (unsigned)cdiv(n + 1, LT);
    };
    auto host_pass = [&](double h, double cube, ncg_mesh_stats &ms) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(b, snap_b, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToDevice, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(d_did_collide, snap_dc, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToDevice, st));
        ncg_status r = download(snap_b, snap_dc);
        if (r != NCG_OK) return r;
        ncg_replay_ops ops;
        if (!lp) {
            ncg_builtin_collider bc = {};
            bc.bodies = reinterpret_cast<ncg_body *>(hb.data());
            bc.exists = d_exists ? hex.data() : nullptr;
            bc.did_collide = d_did_collide ? hdc.data() : nullptr;
            bc.timestep = h;
            bc.collision_f = f;
            bc.one_collision_only = one_collision_only;
            if ((r = ncg_builtin_replay_ops(&bc, &ops)) != NCG_OK) return r;
            r = ncg_mesh_pass(ctx, bc.bodies, n, bc.exists, h, mp->mesh_size, cube, mp->cells, 0,
                              &ops, &ms);
            if (r != NCG_OK) return r;
            ms.collisions = bc.collisions;
            ++mp->host_passes;
            return upload(n);
        }
        ncg_outcome_state os = {};
        os.c.bodies = reinterpret_cast<ncg_body *>(hb.data());
        os.c.did_collide = d_did_collide ? hdc.data() : nullptr;
        os.c.timestep = h;
        os.c.one_collision_only = one_collision_only;
        os.model = lp->model;
        os.exists = hex.data();
        os.free_indexes = free_list.data();
        os.central = hcen;
        os.n_max = ncap;
        os.largest_id = largest_id;
        os.how_many_free = how_many_free;
        os.mesh = 1;
        os.need_to_reduce_com = need;
        if ((r = ncg_builtin_outcome_ops(&os, &ops)) != NCG_OK)
            return ncg_fail(ctx, r, "invalid outcome model");
        r = ncg_mesh_pass(ctx, os.c.bodies, n, os.exists, h, mp->mesh_size, cube, mp->cells, 0,
                          &ops, &ms);
        if (r != NCG_OK) return r;
        if (os.error)
            return ncg_fail(ctx, NCG_ERR_INVALID,
                            os.error == 1 ? "a fragmentation needs more than n_max = %d slots"
                                          : "an ejection speed fell below the escape speed "
                                            "(n_max %d); NcorpiON aborts here",
                            (int)ncap);
        ms.collisions = os.c.collisions;
        largest_id = os.largest_id;
        how_many_free = os.how_many_free;
        need = os.need_to_reduce_com;
        mergers += os.mergers;
        super_catastrophic += os.super_catastrophic;
        partial += os.partial;
        full += os.full;
        ++host_passes;
        ++mp->host_passes;
        if ((r = upload(largest_id + 1)) != NCG_OK) return r;
        set_n(largest_id + 1);
        return NCG_OK;
    };
    // Without an outcome model no body is removed or created, so the count
    // of a step's census is the first one's: the census only checks the
    // removal distances, and its flags are read at the next pass's
    // synchronisation instead of one of their own.
    bool census_pending = false;
    auto census_deferred = [&]() -> ncg_status {
        hipLaunchKernelGGL(k_lf_census, dim3(std::min<long>(1024, cdiv(n, LT))), dim3(LT), 0, st, b,
                           (int)n, d_exists, d_central, r_dis, r_hd, count, err);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipMemcpyAsync(hpin + 2, err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        census_pending = true;
        return NCG_OK;
    };
    auto census_read = [&]() -> ncg_status {  // after a synchronisation of st
        if (!census_pending) return NCG_OK;
        census_pending = false;
        if (hpin[2] & 1u)
            return ncg_fail(ctx, NCG_ERR_INVALID, "fewer than three bodies: three_largest_moonlets "
                                                  "is undefined");
        if (hpin[2] & 2u)
            return ncg_fail(ctx, NCG_ERR_INVALID,
                            "a body crossed disruption_threshold or high_dumping_threshold: "
                            "end_of_timestep would remove it, which this loop does not do");
        return NCG_OK;
    };
    auto pass = [&](int mode, double h) -> ncg_status {
        largest3();
        NCG_CHECK_LAUNCH();
        ncg_mesh_stats ms;
        const double cube = mp->mesh_size * (double)mp->cells;
        if (mode == 0) {
            NCG_HIP(hipMemcpyAsync(snap_b, b, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToDevice, st));
            if (d_did_collide)
                NCG_HIP(hipMemcpyAsync(snap_dc, d_did_collide, sizeof(int32_t) * (size_t)n,
                                       hipMemcpyDeviceToDevice, st));
        }
        unsigned pflags = 0u;
        ncg_status r = ncg_mesh_pass_dev_flags(ctx, d_bodies, n, d_exists, d_did_collide, h,
                                               mp->mesh_size, cube, mp->cells, mode, f,
                                               one_collision_only, &ms, &pflags, st);
        if (r != NCG_OK) return r;
        if ((r = census_read()) != NCG_OK) return r;  // the pass synchronised st
        mp->conflict_passes += (int32_t)ms.replay_conflicts;
        if (pflags != 0u || (model_host && mode == 0 && ms.collisions > 0)) {
            if (mode != 0)
                return ncg_fail(ctx, NCG_ERR_INVALID, "neighbour pass failed (flags %u)", pflags);
            if ((r = host_pass(h, cube, ms)) != NCG_OK) return r;
        }
        tn = ms.total_neighbours;
        ncoll += ms.collisions;
        return NCG_OK;
    };
    auto kick = [&](double h) -> ncg_status {
        ncg_status r = ncg_mesh_gravity_dev(ctx, d_bodies, n, d_exists, d_central, G, soft, prec, acc,
                                            cm, (void *)st);
        if (r != NCG_OK) return r;
        hipLaunchKernelGGL(k_lf_kick, dim3(g), dim3(LT), 0, st, (int)n, d_exists, h, b, acc,
                           d_central, cm);
        NCG_CHECK_LAUNCH();
        return NCG_OK;
    };
    // end_of_timestep's mesh-size update (rk4.c:370-377).
    auto update_gam = [&](int how_many) {
        const double avg = 2.0 * (double)tn / (double)how_many;
        double gam = mp->mesh_size * pow(mp->target_neighbours / avg, 1.0 / 3.0);
        if (gam < mp->mesh_size_min) gam = mp->mesh_size_min;
        mp->mesh_size = gam;
    };
    // end_of_timestep's host half (removals, reduction of the centre of
    // mass, tidy_up) when a body crossed a distance or the slot state asks.
    auto step_end = [&](int &how_many, bool crossed) -> ncg_status {
        if (!lp) return NCG_OK;
        if (!crossed && !(need && lp->reduce_to_com) && !(how_many < 9 * largest_id / 10))
            return NCG_OK;
        ncg_status r = download(b, d_did_collide);
        if (r != NCG_OK) return r;
        const int32_t n_before = n;
        how_many = ncg_host_step_end(reinterpret_cast<ncg_body *>(hb.data()), hex.data(),
                                     free_list.data(), largest_id, how_many_free, hcen, r_dis,
                                     r_hd, lp->reduce_to_com, need, removed, tidy_ups, 1);
        ++host_step_ends;
        if ((r = upload(n_before)) != NCG_OK) return r;
        set_n(largest_id + 1);
        return NCG_OK;
    };
    hipLaunchKernelGGL(k_lf_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t *>(err), 2);
    NCG_CHECK_LAUNCH();
    int how_many = 0;
    bool crossed = false;
    ncg_status r = census(how_many, nullptr);
    if (r != NCG_OK) return r;
    if (flags & 2) {
        // Half kick before the loop (rk4.c:697-717).
        if ((r = pass(1, 0.5 * dt)) != NCG_OK) return r;
        if ((r = kick(0.5 * dt)) != NCG_OK) return r;
        if ((r = census(how_many, nullptr)) != NCG_OK) return r;
        update_gam(how_many);
    }
    for (int32_t s = 0; s < nsteps; ++s) {
        // SBAB1 (rk4.c:761-780): mesh() or get_neighbours_mesh(), drift, kick.
        if (coll && (r = pass(0, dt)) != NCG_OK) return r;
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, d_exists, dt, b, d_central);
        NCG_CHECK_LAUNCH();
        if (!coll && (r = pass(1, dt)) != NCG_OK) return r;
        if ((r = kick(dt)) != NCG_OK) return r;
        // end_of_timestep (rk4.c:130-393): removals, centre of mass,
        // did_collide, census, mesh size, tidy_up.
        if (coll && d_did_collide && (one_collision_only || frag))
            hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(n, 256)), dim3(256), 0, st, d_did_collide,
                               (int)n);
        if (lp) {
            if ((r = census(how_many, &crossed)) != NCG_OK) return r;
            if ((r = step_end(how_many, crossed)) != NCG_OK) return r;
        } else if ((r = census_deferred()) != NCG_OK) {
            return r;
        }
        update_gam(how_many);
    }
    if (flags & 4) {
        // Backward half kick after the loop (rk4.c:807-812), the central
        // record taking CM_buffer's kick.
        if ((r = pass(1, -0.5 * dt)) != NCG_OK) return r;
        if ((r = kick(-0.5 * dt)) != NCG_OK) return r;
    }
    NCG_HIP(hipStreamSynchronize(st));
    if ((r = census_read()) != NCG_OK) return r;
    if (n_collisions) *n_collisions = ncoll;
    if (lp) {
        lp->largest_id = largest_id;
        lp->how_many_free = how_many_free;
        std::copy(free_list.begin(), free_list.end(), lp->free_indexes);
        lp->need_to_reduce_com = need;
        lp->host_passes = host_passes;
        lp->host_step_ends = host_step_ends;
        lp->removed = removed;
        lp->tidy_ups = tidy_ups;
        lp->mergers = mergers;
        lp->super_catastrophic = super_catastrophic;
        lp->partial = partial;
        lp->full = full;
    }
    return NCG_OK;
}

}  // namespace

extern "C" ncg_status ncg_mesh_leapfrog_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                            int32_t *d_exists, int32_t *d_did_collide,
                                            double *d_central, double grav_const,
                                            double softening, ncg_precision prec, double dt,
                                            int32_t nsteps, ncg_mesh_loop_params *mp,
                                            double collision_f, int32_t one_collision_only,
                                            int32_t flags, int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 3 && d_bodies && d_central && mp && nsteps >= 0, "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    NCG_REQUIRE(mp->mesh_size > 0.0 && mp->mesh_size_min > 0.0 && mp->cells >= 3 &&
                    mp->target_neighbours > 0.0,
                "invalid mesh parameters");
    NCG_REQUIRE(!(collision_f > 0.0 && one_collision_only && !d_did_collide),
                "one_collision_only needs a did_collide array");
    return mesh_loop_run(ctx, d_bodies, n, d_exists, d_did_collide, d_central, grav_const,
                         softening, prec, dt, nsteps, mp, collision_f, one_collision_only, flags,
                         nullptr, n_collisions, stream);
}

extern "C" ncg_status ncg_mesh_leapfrog_model_dev(ncg_ctx *ctx, ncg_body *d_bodies,
                                                  int32_t *d_exists, int32_t *d_did_collide,
                                                  double *d_central, double grav_const,
                                                  double softening, ncg_precision prec, double dt,
                                                  int32_t nsteps, ncg_mesh_loop_params *mp,
                                                  ncg_leapfrog_model_params *params,
                                                  int32_t one_collision_only, int32_t flags,
                                                  int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(d_bodies && d_exists && d_central && mp && params && nsteps >= 0,
                "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    NCG_REQUIRE(mp->mesh_size > 0.0 && mp->mesh_size_min > 0.0 && mp->cells >= 3 &&
                    mp->target_neighbours > 0.0,
                "invalid mesh parameters");
    ncg_status r = check_model_params(ctx, *params, one_collision_only, d_did_collide);
    if (r != NCG_OK) return r;
    NCG_REQUIRE(params->largest_id >= 2, "fewer than three slots");
    return mesh_loop_run(ctx, d_bodies, params->largest_id + 1, d_exists, d_did_collide, d_central,
                         grav_const, softening, prec, dt, nsteps, mp, 0.0, one_collision_only,
                         flags, params, n_collisions, stream);
}

namespace {

// end_of_timestep's removal tests (rk4.c:158-179): sets *flag when an existing
// body j < n is closer to the central body than r_dis (central body present)
// or farther than r_hd from it (from the origin without one).
__global__ __launch_bounds__(LT) void k_lf_thresholds(int n, const int32_t *__restrict__ ex,
                                                      const DBody *__restrict__ b,
                                                      const double *__restrict__ cen,
                                                      double r_dis, double r_hd,
                                                      unsigned *__restrict__ flag)
{
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i >= n || (ex && !ex[i])) return;
    const bool cm = cen[7] != 0.0;
    const double DX = b[i].x - (cm ? cen[0] : 0.);
    const double DY = b[i].y - (cm ? cen[1] : 0.);
    const double DZ = b[i].z - (cm ? cen[2] : 0.);
    const double d2 = DX * DX + DY * DY + DZ * DZ;
    if ((d2 < r_dis * r_dis && cm) || d2 > r_hd * r_hd) atomicOr(flag, 1u);
}

// Host-side state of a run with an outcome model (slots and flags of the
// reference's globals), kept in window snapshots.
struct ModelHost {
    int32_t largest_id = 0, how_many_free = 0, need = 0, live = 0;
    std::vector<int32_t> free;
    int32_t host_passes = 0, host_step_ends = 0, removed = 0, tidy_ups = 0;
    int64_t mergers = 0, super_catastrophic = 0, partial = 0, full = 0, coll = 0;
};

// The brute-force loop; mp == nullptr: ncg_leapfrog_dev (bounce with
// collision_f, fixed slots); else ncg_leapfrog_model_dev.
ncg_status leapfrog_run(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n_in, int32_t *d_exists_w,
                        const int32_t *d_exists, int32_t *d_did_collide, double *d_central,
                        double grav_const, double softening, ncg_precision prec, double dt,
                        int32_t nsteps, double collision_f, int32_t one_collision_only,
                        int32_t flags, ncg_leapfrog_model_params *mp, int64_t *n_collisions,
                        void *stream)
{
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    if (n_collisions) *n_collisions = 0;
    DBody *b = reinterpret_cast<DBody *>(d_bodies);
    const int32_t *ex = d_exists;
    const bool model_host = mp && mp->model.kind != NCG_OUTCOME_BOUNCE;
    const bool frag = mp && mp->model.kind == NCG_OUTCOME_FRAGMENTATION;
    // The device replay applies the bounce; under the other models any pass
    // it resolves a pair in is replayed on the host, so its f only needs to
    // be a valid restitution.
    const double f = model_host ? 2.0 : (mp ? mp->model.collision_f : collision_f);
    const bool coll = model_host || f > 0.0;
    const bool thr = mp && (mp->r_disruption > 0.0 || mp->r_high_dumping > 0.0);
    const double r_dis = mp && mp->r_disruption > 0.0 ? mp->r_disruption : 0.0;
    const double r_hd = mp && mp->r_high_dumping > 0.0 ? mp->r_high_dumping : INFINITY;
    const bool clear_dc = coll && d_did_collide && (one_collision_only || frag);
    const int32_t ncap = mp ? mp->n_max : n_in;
    int32_t n = n_in;
    const bool use_graph = (flags & 1) != 0;
    const unsigned cap = ncg_replay_cap(ctx, 1u << 20);
    double *acc = dbuf<double>(ctx, "lf_acc", 3 * (size_t)ncap);
    double *cterm = dbuf<double>(ctx, "lf_cterm", 3 * (size_t)ncap);
    double *cm = dbuf<double>(ctx, "lf_cm", 4);
    int2 *cand = dbuf<int2>(ctx, "lf_cand", cap);
    unsigned long long *cnt = dbuf<unsigned long long>(ctx, "lf_cnt", 1);
    unsigned char *dirty = dbuf<unsigned char>(ctx, "lf_dirty", (size_t)ncap + 4);
    unsigned *stat = dbuf<unsigned>(ctx, "lf_stat", 4);
    int *off = dbuf<int>(ctx, "lf_off", 2);
    if (!acc || !cterm || !cm || !cand || !cnt || !dirty || !stat || !off) return NCG_ERR_NOMEM;
    double present = 0.0;
    NCG_HIP(hipMemcpyAsync(&present, d_central + 7, sizeof(double), hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    const bool cp = present != 0.0;
    hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(ncap / 4 + 2, 256)), dim3(256), 0, st,
                       reinterpret_cast<int32_t *>(dirty), (int)(ncap / 4 + 1));
    hipLaunchKernelGGL(k_lf_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t *>(stat), 3);
    NCG_CHECK_LAUNCH();
    unsigned g = 0;
    auto set_n = [&](int32_t nn) {
        n = nn;
        g = (unsigned)cdiv(n + 1, LT);
        hipLaunchKernelGGL(k_lf_off, dim3(1), dim3(1), 0, st, off, (int)n);
    };
    set_n(n);
    NCG_CHECK_LAUNCH();
    const double G = grav_const, soft = softening;

    auto collide = [&](double h) -> ncg_status {
        hipLaunchKernelGGL(k_lf_zero_u64, dim3(1), dim3(1), 0, st, cnt);
        ncg_status r = ncg_candidates_launch(ctx, b, n, ex, h, cand, cap, cnt, st);
        if (r != NCG_OK) return r;
        return ncg_bf_replay_launch(ctx, b, ex, d_did_collide, one_collision_only, off, 1, cand,
                                    nullptr, cnt, cap, h, f, dirty, stat, st);
    };
    // The replay's capacity (candidates, changed bodies in LDS) is checked
    // every LF_CHECK steps.  A window in which a pass exceeded it -- or, with
    // an outcome model, in which a pass resolved a pair under the merger or
    // fragmentation model or a body crossed a removal threshold -- is run
    // again from a snapshot of the state taken at the window's start, step
    // by step, each collision pass and step end checked at once: such a pass
    // is undone (a snapshot before it) and replayed on the host with the
    // same decision and the outcome (ncg_brute_force_pass with the built-in
    // ops), such a step end runs end_of_timestep's host half, so results are
    // those of the reference's loop.
    constexpr int LF_CHECK = 64;
    DBody *snap_b = dbuf<DBody>(ctx, "lf_snap_b", ncap);
    DBody *pass_b = dbuf<DBody>(ctx, "lf_pass_b", ncap);
    double *snap_c = dbuf<double>(ctx, "lf_snap_c", 8);
    int32_t *snap_dc = d_did_collide ? dbuf<int32_t>(ctx, "lf_snap_dc", ncap) : nullptr;
    int32_t *pass_dc = d_did_collide ? dbuf<int32_t>(ctx, "lf_pass_dc", ncap) : nullptr;
    int32_t *snap_ex = mp ? dbuf<int32_t>(ctx, "lf_snap_ex", ncap) : nullptr;
    unsigned *hst = pbuf<unsigned>(ctx, "lf_hstat", 8);
    if (!snap_b || !pass_b || !snap_c || (d_did_collide && (!snap_dc || !pass_dc)) ||
        (mp && !snap_ex) || !hst)
        return NCG_ERR_NOMEM;
    int64_t host_coll = 0;      // collisions of passes replayed on the host (bounce only)
    unsigned snap_coll = 0;     // device collision count at the snapshot
    ModelHost mh, snap_mh;      // model runs: host state now and at the snapshot
    std::vector<DBody> hb;
    std::vector<int32_t> hex, hdc;
    double hcen[8];
    if (mp) {
        mh.largest_id = mp->largest_id;
        mh.how_many_free = mp->how_many_free;
        mh.need = mp->need_to_reduce_com;
        mh.free.assign(mp->free_indexes, mp->free_indexes + ncap);
        hex.assign((size_t)ncap, 0);
        NCG_HIP(hipMemcpyAsync(hex.data(), d_exists_w, sizeof(int32_t) * (size_t)n,
                               hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        for (int32_t j = 0; j < n; ++j) mh.live += hex[j] != 0;
    }
    auto read_stat = [&](unsigned *hs) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(hst, stat, sizeof(unsigned) * 3, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        hs[0] = hst[0];
        hs[1] = hst[1];
        hs[2] = hst[2];
        return NCG_OK;
    };
    auto write_stat = [&](unsigned coll_count) -> ncg_status {
        hst[4] = coll_count;
        hst[5] = 0u;
        hst[6] = 0u;
        NCG_HIP(hipMemcpyAsync(stat, hst + 4, sizeof(unsigned) * 3, hipMemcpyHostToDevice, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    auto copy_state = [&](DBody *to_b, int32_t *to_dc, double *to_c, const DBody *from_b,
                          const int32_t *from_dc, const double *from_c, int32_t slots) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(to_b, from_b, sizeof(DBody) * (size_t)slots, hipMemcpyDeviceToDevice, st));
        if (to_dc && from_dc)
            NCG_HIP(hipMemcpyAsync(to_dc, from_dc, sizeof(int32_t) * (size_t)slots,
                                   hipMemcpyDeviceToDevice, st));
        if (to_c && from_c)
            NCG_HIP(hipMemcpyAsync(to_c, from_c, sizeof(double) * 8, hipMemcpyDeviceToDevice, st));
        return NCG_OK;
    };
    // A replay that ran out of room for changed bodies leaves some of their
    // marks set.
    auto clear_dirty = [&]() -> ncg_status {
        hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(ncap / 4 + 2, 256)), dim3(256), 0, st,
                           reinterpret_cast<int32_t *>(dirty), (int)(ncap / 4 + 1));
        NCG_CHECK_LAUNCH();
        return NCG_OK;
    };
    // Host copies of the state (slots 0 .. n - 1; exists and did_collide
    // zero beyond), and back (`slots` slots of exists, which covers the slots
    // a compaction emptied).
    auto download = [&]() -> ncg_status {
        hb.assign((size_t)ncap, DBody{});
        if (!mp) hex.assign(ex ? (size_t)n : 0, 0);
        else std::fill(hex.begin(), hex.end(), 0);
        hdc.assign(d_did_collide ? (size_t)ncap : 0, 0);
        NCG_HIP(hipMemcpyAsync(hb.data(), b, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToHost, st));
        if (ex)
            NCG_HIP(hipMemcpyAsync(hex.data(), ex, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToHost, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(hdc.data(), d_did_collide, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToHost, st));
        NCG_HIP(hipMemcpyAsync(hcen, d_central, sizeof hcen, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    auto upload = [&](int32_t slots) -> ncg_status {
        NCG_HIP(hipMemcpyAsync(b, hb.data(), sizeof(DBody) * (size_t)slots, hipMemcpyHostToDevice, st));
        if (mp)
            NCG_HIP(hipMemcpyAsync(d_exists_w, hex.data(), sizeof(int32_t) * (size_t)slots,
                                   hipMemcpyHostToDevice, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(d_did_collide, hdc.data(), sizeof(int32_t) * (size_t)slots,
                                   hipMemcpyHostToDevice, st));
        NCG_HIP(hipMemcpyAsync(d_central, hcen, sizeof hcen, hipMemcpyHostToDevice, st));
        NCG_HIP(hipStreamSynchronize(st));
        return NCG_OK;
    };
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    bool need_eager = true;  // the next unchecked step runs eagerly (and allocates)
    auto drop_graph = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        need_eager = true;
    };
    // A collision pass with dt h, checked: on a capacity error (or, under the
    // merger and fragmentation models, a resolved pair) the pass is undone
    // and replayed on the host.
    auto collide_checked = [&](double h) -> ncg_status {
        unsigned hs[3];
        ncg_status r = read_stat(hs);
        if (r != NCG_OK) return r;
        const unsigned before = hs[0];
        if ((r = copy_state(pass_b, pass_dc, nullptr, b, d_did_collide, nullptr, n)) != NCG_OK) return r;
        if ((r = collide(h)) != NCG_OK) return r;
        if ((r = read_stat(hs)) != NCG_OK) return r;
        if (hs[1] == 0u && !(model_host && hs[0] != before)) return NCG_OK;
        if ((r = copy_state(b, d_did_collide, nullptr, pass_b, pass_dc, nullptr, n)) != NCG_OK) return r;
        if ((r = clear_dirty()) != NCG_OK) return r;
        if ((r = write_stat(before)) != NCG_OK) return r;
        if ((r = download()) != NCG_OK) return r;
        ncg_replay_ops ops;
        if (!mp) {
            ncg_builtin_collider bc = {};
            bc.bodies = reinterpret_cast<ncg_body *>(hb.data());
            bc.exists = ex ? hex.data() : nullptr;
            bc.did_collide = d_did_collide ? hdc.data() : nullptr;
            bc.timestep = h;
            bc.collision_f = f;
            bc.one_collision_only = one_collision_only;
            if ((r = ncg_builtin_replay_ops(&bc, &ops)) != NCG_OK) return r;
            if ((r = ncg_brute_force_pass(ctx, bc.bodies, n, bc.exists, h, &ops, nullptr)) != NCG_OK)
                return r;
            host_coll += bc.collisions;
            return upload(n);
        }
        ncg_outcome_state os = {};
        os.c.bodies = reinterpret_cast<ncg_body *>(hb.data());
        os.c.did_collide = d_did_collide ? hdc.data() : nullptr;
        os.c.timestep = h;
        os.c.one_collision_only = one_collision_only;
        os.model = mp->model;
        os.exists = hex.data();
        os.free_indexes = mh.free.data();
        os.central = hcen;
        os.n_max = ncap;
        os.largest_id = mh.largest_id;
        os.how_many_free = mh.how_many_free;
        os.need_to_reduce_com = mh.need;
        if ((r = ncg_builtin_outcome_ops(&os, &ops)) != NCG_OK)
            return ncg_fail(ctx, r, "invalid outcome model");
        if ((r = ncg_brute_force_pass(ctx, os.c.bodies, n, os.exists, h, &ops, nullptr)) != NCG_OK)
            return r;
        if (os.error)
            return ncg_fail(ctx, NCG_ERR_INVALID,
                            os.error == 1 ? "a fragmentation needs more than n_max = %d slots"
                                          : "an ejection speed fell below the escape speed "
                                            "(n_max %d); NcorpiON aborts here",
                            (int)ncap);
        mh.largest_id = os.largest_id;
        mh.how_many_free = os.how_many_free;
        mh.need = os.need_to_reduce_com;
        mh.coll += os.c.collisions;
        mh.mergers += os.mergers;
        mh.super_catastrophic += os.super_catastrophic;
        mh.partial += os.partial;
        mh.full += os.full;
        ++mh.host_passes;
        const int32_t nn = mh.largest_id + 1;
        mh.live = 0;
        for (int32_t j = 0; j < nn; ++j) mh.live += hex[j] != 0;
        if ((r = upload(std::max(n, nn))) != NCG_OK) return r;
        if (nn != n) {
            set_n(nn);
            drop_graph();
        }
        return NCG_OK;
    };
    // end_of_timestep's host half (removals, reduction of the centre of
    // mass, tidy_up) when a body crossed a threshold or the host state asks
    // for it.
    auto tidy_due = [&]() { return mh.live < 9 * mh.largest_id / 10; };
    auto end_checked = [&]() -> ncg_status {
        if (!mp) return NCG_OK;
        bool due = (mh.need && mp->reduce_to_com) || tidy_due();
        ncg_status r;
        if (thr) {
            unsigned hs[3];
            if ((r = read_stat(hs)) != NCG_OK) return r;
            due = due || hs[2] != 0u;
        }
        if (!due) return NCG_OK;
        if ((r = download()) != NCG_OK) return r;
        const int32_t n_before = n;
        int32_t lid = mh.largest_id, hmf = mh.how_many_free;
        mh.live = ncg_host_step_end(reinterpret_cast<ncg_body *>(hb.data()), hex.data(),
                                    mh.free.data(), lid, hmf, hcen, r_dis, r_hd,
                                    mp->reduce_to_com, mh.need, mh.removed, mh.tidy_ups, 0);
        mh.largest_id = lid;
        mh.how_many_free = hmf;
        ++mh.host_step_ends;
        if ((r = upload(n_before)) != NCG_OK) return r;
        hst[7] = 0u;
        NCG_HIP(hipMemcpyAsync(stat + 2, hst + 7, sizeof(unsigned), hipMemcpyHostToDevice, st));
        NCG_HIP(hipStreamSynchronize(st));
        if (lid + 1 != n) {
            set_n(lid + 1);
            drop_graph();
        }
        return NCG_OK;
    };
    // One step after the field's preparation (the tiled forms sort first).
    ncg_central cflag = {};
    cflag.present = cp ? 1 : 0;
    auto step = [&](bool checked) -> ncg_status {
        ncg_status r;
        if (prec == NCG_F64_EXACT)
            r = ncg_grav_exact_dev(ctx, b, n, ex, d_central, G, soft, acc, cterm, cp ? cm : nullptr, st);
        else
            r = ncg_tiled_eval(ctx, b, n, ex, cp ? &cflag : nullptr, cp ? d_central : nullptr, G,
                               soft, prec, acc, nullptr, cp ? cm : nullptr, st);
        if (r != NCG_OK) return r;
        hipLaunchKernelGGL(k_lf_kick, dim3(g), dim3(LT), 0, st, (int)n, ex, dt, b, acc, d_central, cm);
        NCG_CHECK_LAUNCH();
        if (coll) {
            r = checked ? collide_checked(dt) : collide(dt);
            if (r != NCG_OK) return r;
        }
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, dt, b, d_central);
        if (thr)
            hipLaunchKernelGGL(k_lf_thresholds, dim3(g), dim3(LT), 0, st, (int)n, ex, b, d_central,
                               r_dis, r_hd, stat + 2);
        if (clear_dc)
            hipLaunchKernelGGL(k_lf_zero_i32, dim3(cdiv(n, 256)), dim3(256), 0, st, d_did_collide, (int)n);
        NCG_CHECK_LAUNCH();
        return checked ? end_checked() : NCG_OK;
    };

    if (flags & 2) {
        // rk4.c:871-879: brute_force and drift with dt / 2.
        if (coll) {
            ncg_status r = collide_checked(0.5 * dt);
            if (r != NCG_OK) return r;
        }
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, 0.5 * dt, b, d_central);
        NCG_CHECK_LAUNCH();
    }
    ncg_status res = NCG_OK;
    int32_t snap_n = n;
    const bool windows = coll || thr;
    auto snapshot = [&]() -> ncg_status {
        unsigned hs[3];
        ncg_status r = read_stat(hs);
        if (r != NCG_OK) return r;
        snap_coll = hs[0];
        snap_mh = mh;
        snap_n = n;
        if (snap_ex)
            NCG_HIP(hipMemcpyAsync(snap_ex, d_exists_w, sizeof(int32_t) * (size_t)ncap,
                                   hipMemcpyDeviceToDevice, st));
        return copy_state(snap_b, snap_dc, snap_c, b, d_did_collide, d_central, mp ? ncap : n);
    };
    if (windows) res = snapshot();
    int32_t win0 = 0;  // first step of the current window
    for (int32_t s = 0; s < nsteps && res == NCG_OK; ++s) {
        // A step end the host state already asks for (only possible at the
        // call's start: afterwards such states arise in checked steps, which
        // settle them) runs checked.
        const bool host_due = mp && ((mh.need && mp->reduce_to_com) || tidy_due());
        if (prec != NCG_F64_EXACT) {
            res = ncg_tiled_order(ctx, b, n, ex, prec, st);  // eager: the sort stays out of graphs
            if (res != NCG_OK) break;
        }
        if (host_due) {
            res = step(true);
        } else if (!use_graph || need_eager) {
            res = step(false);  // the first step also allocates what a graph uses
            need_eager = false;
        } else {
            if (!exec) {
                NCG_HIP(hipStreamSynchronize(st));
                ctx->capturing = true;
                NCG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                res = step(false);
                hipError_t e = hipStreamEndCapture(st, &graph);
                ctx->capturing = false;
                if (res != NCG_OK) break;
                if (e != hipSuccess) {
                    res = ncg_fail(ctx, NCG_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(e));
                    break;
                }
                if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                    res = ncg_fail(ctx, NCG_ERR_DEVICE, "graph instantiation failed");
                    break;
                }
            }
            if (hipGraphLaunch(exec, st) != hipSuccess) {
                res = ncg_fail(ctx, NCG_ERR_DEVICE, "graph launch failed");
                break;
            }
        }
        if (res != NCG_OK || !windows) continue;
        if ((s + 1) % LF_CHECK != 0 && s + 1 != nsteps) continue;
        unsigned hs[3];
        if ((res = read_stat(hs)) != NCG_OK) break;
        if (hs[1] != 0u || (model_host && hs[0] != snap_coll) || (thr && hs[2] != 0u)) {
            // The window [win0, s] again, from its snapshot, checked.
            mh = snap_mh;
            if (n != snap_n) {
                set_n(snap_n);
                drop_graph();
            }
            if (snap_ex)
                NCG_HIP(hipMemcpyAsync(d_exists_w, snap_ex, sizeof(int32_t) * (size_t)ncap,
                                       hipMemcpyDeviceToDevice, st));
            if ((res = copy_state(b, d_did_collide, d_central, snap_b, snap_dc, snap_c,
                                  mp ? ncap : n)) != NCG_OK)
                break;
            if ((res = clear_dirty()) != NCG_OK) break;
            if ((res = write_stat(snap_coll)) != NCG_OK) break;
            for (int32_t q = win0; q <= s && res == NCG_OK; ++q) {
                if (prec != NCG_F64_EXACT && (res = ncg_tiled_order(ctx, b, n, ex, prec, st)) != NCG_OK)
                    break;
                res = step(true);
            }
            if (res != NCG_OK) break;
        }
        if ((res = snapshot()) != NCG_OK) break;
        win0 = s + 1;
    }
    if (res == NCG_OK && (flags & 4)) {
        // rk4.c:938-942: the backward half drift.
        hipLaunchKernelGGL(k_lf_drift, dim3(g), dim3(LT), 0, st, (int)n, ex, -0.5 * dt, b, d_central);
        NCG_CHECK_LAUNCH();
    }
    unsigned hs[2] = {0, 0};
    (void)hipMemcpyAsync(hs, stat, sizeof hs, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    drop_graph();
    if (res != NCG_OK) return res;
    if (n_collisions) *n_collisions = hs[0] + host_coll + mh.coll;
    if (mp) {
        mp->largest_id = mh.largest_id;
        mp->how_many_free = mh.how_many_free;
        std::copy(mh.free.begin(), mh.free.end(), mp->free_indexes);
        mp->need_to_reduce_com = mh.need;
        mp->host_passes = mh.host_passes;
        mp->host_step_ends = mh.host_step_ends;
        mp->removed = mh.removed;
        mp->tidy_ups = mh.tidy_ups;
        mp->mergers = mh.mergers;
        mp->super_catastrophic = mh.super_catastrophic;
        mp->partial = mh.partial;
        mp->full = mh.full;
    }
    if (hs[1])
        return ncg_fail(ctx, NCG_ERR_CAPACITY,
                        "collision replay capacity exceeded (flags %u: 1 = more than %u candidates, "
                        "2 = more than %d to decide, 4 = more than %d changed bodies in one pass); "
                        "results are not valid",
                        hs[1], cap, NCG_REPLAY_RCAP, NCG_REPLAY_DCAP);
    return NCG_OK;
}

}  // namespace

extern "C" ncg_status ncg_leapfrog_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                       const int32_t *d_exists, int32_t *d_did_collide,
                                       double *d_central, double grav_const, double softening,
                                       ncg_precision prec, double dt, int32_t nsteps,
                                       double collision_f, int32_t one_collision_only,
                                       int32_t flags, int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 1 && d_bodies && d_central && nsteps >= 0, "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    NCG_REQUIRE(!(collision_f > 0.0 && one_collision_only && !d_did_collide),
                "one_collision_only needs a did_collide array");
    return leapfrog_run(ctx, d_bodies, n, nullptr, d_exists, d_did_collide, d_central, grav_const,
                        softening, prec, dt, nsteps, collision_f, one_collision_only, flags,
                        nullptr, n_collisions, stream);
}

extern "C" ncg_status ncg_leapfrog_model_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t *d_exists,
                                             int32_t *d_did_collide, double *d_central,
                                             double grav_const, double softening,
                                             ncg_precision prec, double dt, int32_t nsteps,
                                             int32_t one_collision_only, int32_t flags,
                                             ncg_leapfrog_model_params *params,
                                             int64_t *n_collisions, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(params && d_bodies && d_exists && d_central && nsteps >= 0, "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    ncg_status r = check_model_params(ctx, *params, one_collision_only, d_did_collide);
    if (r != NCG_OK) return r;
    const ncg_leapfrog_model_params &P = *params;
    return leapfrog_run(ctx, d_bodies, P.largest_id + 1, d_exists, d_exists, d_did_collide,
                        d_central, grav_const, softening, prec, dt, nsteps, 0.0, one_collision_only,
                        flags, params, n_collisions, stream);
}
