// Mesh module: cell binning, neighbour counts, collision candidates, host
// replay and neighbour-cell gravity.
//
// Restates the cell semantics of NcorpiON's mesh algorithm
// (collision.c:175-417, physics.c:1094-1118):
//   - a body is inserted in the hash table when it exists, lies inside the
//     collision cube (|X|, |Y|, |Z| <= collision_cube/2) and, in mesh(), has
//     radius <= gam (get_neighbours_mesh() inserts every existing body);
//   - its cell is (i, j, k) = cells/2 + 1 + floor(X/gam) per axis and its key
//     i + j*d + k*d^2 with d = collision_cube_cells (collision.c:203-207), so
//     cells near the +x/+y faces alias into the next row exactly as in the
//     reference;
//   - body k's neighbours are the bodies inserted before it (p < k) whose key
//     is one of k's 27 keys, visited for cell l = 26..0 and ascending id within
//     a cell (the LIFO of chains, collision.c:285-302 and 349-377).
// The reference walks linked lists; here the bodies are sorted by (key, id)
// with a radix sort, a dense table gives each key its [start, end) range, and
// one GPU thread per inserted body counts its neighbours (total_neighbours)
// and runs the closest_approach predicate on each neighbour pair.  The host
// then replays the reference's visiting order over the candidates; every pair
// touching a body modified earlier in the pass is re-tested on the host.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <queue>
#include <set>
#include <tuple>

#include "ncg_internal.hpp"
#include "exact_div.hpp"
#include "collide_dev.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int MT = 256;
constexpr unsigned NOKEY = 0xffffffffu;

__device__ __host__ inline long stencil_offset(int l, long d)
{
    const int di = l % 3 - 1, dj = (l / 3) % 3 - 1, dk = l / 9 - 1;
    return di + dj * d + dk * d * d;
}

template <class T>
__device__ inline T bsum(T v, T *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = MT / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// Key of every body (NOKEY when not inserted) and the pass counters:
// cnt[0] inserted, cnt[1] small (R <= gam), cnt[2] big, cnt[3] big-list length.
__global__ __launch_bounds__(MT) void k_mkey(const DBody *__restrict__ b, int n,
                                             const int32_t *__restrict__ ex, double gam,
                                             double cube, int d, int mode,
                                             unsigned *__restrict__ key, int *__restrict__ idx,
                                             int *__restrict__ cnt, int *__restrict__ big,
                                             int big_cap)
{
    __shared__ int sh[MT];
    const int k = blockIdx.x * MT + threadIdx.x;
    int ins = 0, sm = 0, bg = 0;
    if (k < n) {
        idx[k] = k;
        unsigned kk = NOKEY;
        if (ex == nullptr || ex[k] != 0) {
            const DBody q = b[k];
            const bool small = q.radius <= gam;  // collision.c:344
            if (mode == 0) {
                sm = small;
                bg = !small;
                if (!small) {
                    const int s = atomicAdd(&cnt[3], 1);
                    if (s < big_cap) big[s] = k;
                }
            }
            if (mode == 1 || small) {
                const double half = cube / 2.0;  // collision.c:197
                if (!(fabs(q.x) > half || fabs(q.y) > half || fabs(q.z) > half)) {
                    const long i = d / 2 + 1 + (int)floor(q.x / gam);
                    const long j = d / 2 + 1 + (int)floor(q.y / gam);
                    const long l = d / 2 + 1 + (int)floor(q.z / gam);
                    kk = (unsigned)(i + j * (long)d + l * (long)d * d);
                    ins = 1;
                }
            }
        }
        key[k] = kk;
    }
    const int a = bsum(ins, sh), s2 = bsum(sm, sh), b2 = bsum(bg, sh);
    if (threadIdx.x == 0) {
        if (a) atomicAdd(&cnt[0], a);
        if (s2) atomicAdd(&cnt[1], s2);
        if (b2) atomicAdd(&cnt[2], b2);
    }
}

__global__ __launch_bounds__(MT) void k_clear_cells(const unsigned *__restrict__ keys, int cnt,
                                                    unsigned *__restrict__ cs,
                                                    unsigned *__restrict__ ce)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s < cnt) {
        cs[keys[s]] = 0u;
        ce[keys[s]] = 0u;
    }
}

__global__ __launch_bounds__(MT) void k_iota(int *__restrict__ v, int n)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s < n) v[s] = s;
}

// Inserted-body count: *dins on the device when given (launches that do not
// wait for the host), else ins.
__device__ inline int ins_count(int ins, const int *dins) { return dins ? *dins : ins; }

__global__ __launch_bounds__(MT) void k_bounds(const unsigned *__restrict__ skey, int ins_h,
                                               const int *__restrict__ dins,
                                               unsigned *__restrict__ cs, unsigned *__restrict__ ce)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    const int ins = ins_count(ins_h, dins);
    if (s >= ins) return;
    const unsigned k = skey[s];
    if (s == 0 || skey[s - 1] != k) cs[k] = (unsigned)s;
    if (s == ins - 1 || skey[s + 1] != k) ce[k] = (unsigned)(s + 1);
}

// Sorted copies {x, y, z, R} and {vx, vy, vz, m} of the inserted bodies.
__global__ __launch_bounds__(MT) void k_msorted(const DBody *__restrict__ b,
                                                const int *__restrict__ sid, int ins_h,
                                                const int *__restrict__ dins,
                                                double4 *__restrict__ pr, double4 *__restrict__ vm)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s >= ins_count(ins_h, dins)) return;
    const DBody q = b[sid[s]];
    pr[s] = make_double4(q.x, q.y, q.z, q.radius);
    vm[s] = make_double4(q.vx, q.vy, q.vz, q.mass);
}

// Keys of this pass, kept to clear their cells at the start of the next one.
__global__ __launch_bounds__(MT) void k_copy_keys(const unsigned *__restrict__ skey,
                                                  const int *__restrict__ dins,
                                                  unsigned *__restrict__ out)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s < *dins) out[s] = skey[s];
}

// closest_approach's predicate (collision.c:112-137) with a 2^-40 relative
// slack on the time test; the host decides every candidate exactly.
// Symmetric in the pair bit for bit: swapping a and b negates every
// difference exactly, which leaves the products, sums, R and t unchanged (this
// file is compiled without FMA contraction), so one order serves both ids.
__device__ inline bool approach_hit(double4 pa, double4 va, double4 pb, double4 vb, double dt)
{
    const double dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
    const double dvx = va.x - vb.x, dvy = va.y - vb.y, dvz = va.z - vb.z;
    const double bb = dx * dvx + dy * dvy + dz * dvz;
    if (!(bb < 0.0)) return false;
    const double dr2 = dx * dx + dy * dy + dz * dz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double R = pa.w + pb.w;
    const double disc = bb * bb + dv2 * (R * R - dr2);
    if (!(disc > 0.0)) return false;
    const double t = -(bb + sqrt(disc)) / dv2;
    return t <= dt + fabs(dt) * 9.094947017729282e-13;
}

// Device passes (ncg_mesh_pass_dev) need only the total of the counts, not
// each body's.  With at least 3 cells per axis each unordered neighbour pair
// is met exactly once from the side of its higher cell (the 13 cells below a
// body's own) or, within a cell, from its higher id; so the total is the sum
// over bodies of the members of the 13 lower cells and of the own cell's
// lower ids, which needs no search in the cells above.  The 12 cells of the
// 4 lower stencil rows are 4 runs of the sorted order (a row's three cells
// have consecutive keys), and cell 12 (key - 1) with the own cell's lower ids
// (positions before the body's own, l = 13) is a fifth.  Mode 0 tests every
// pair of the runs (the candidates of k_mcount's per-body form).
template <int MODE>
__global__ __launch_bounds__(MT) void k_mcount_total(const unsigned *__restrict__ skey,
                                                     const int *__restrict__ sid,
                                                     const double4 *__restrict__ pr,
                                                     const double4 *__restrict__ vm,
                                                     const unsigned *__restrict__ cs,
                                                     const unsigned *__restrict__ ce,
                                                     const int *__restrict__ dins, int d,
                                                     double dt, int4 *__restrict__ cand,
                                                     unsigned long long cap,
                                                     unsigned long long *__restrict__ ccnt,
                                                     unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long sh[MT];
    const int s = blockIdx.x * MT + threadIdx.x;
    const int ins = *dins;
    unsigned long long c = 0;
    if (s < ins) {
        const long key = skey[s];
        int k = 0;
        double4 pk = make_double4(0, 0, 0, 0), vk = pk;
        if (MODE == 0) {
            k = sid[s];
            pk = pr[s];
            vk = vm[s];
        }
        for (int r = 0; r < 5; ++r) {
            // Cells c0, c1, c2 of the run (c2 unused for the fifth run).
            const long k1 = r < 4 ? key + (long)(r % 3 - 1) * d + (long)(r / 3 - 1) * d * d : key;
            const unsigned s0 = cs[k1 - 1], e0 = ce[k1 - 1], s1 = cs[k1], e1 = ce[k1];
            unsigned a, e, s2 = 0, e2 = 0;
            if (r < 4) {
                s2 = cs[k1 + 1];
                e2 = ce[k1 + 1];
                a = e0 > s0 ? s0 : (e1 > s1 ? s1 : s2);
                e = e2 > s2 ? e2 : (e1 > s1 ? e1 : e0);
            } else {
                a = e0 > s0 ? s0 : s1;  // the own cell is not empty
                e = (unsigned)s;
            }
            if (a >= e) continue;
            c += e - a;
            if (MODE != 0) continue;
            int np = sid[a];
            double4 nr = pr[a], nv = vm[a];
            for (unsigned q = a; q < e; ++q) {
                const int p = np;
                const double4 rq = nr, vq = nv;
                if (q + 1 < e) {
                    np = sid[q + 1];
                    nr = pr[q + 1];
                    nv = vm[q + 1];
                }
                const bool lower = p < k;
                if (approach_hit(pk, vk, rq, vq, dt)) {
                    // Stencil index of q's cell.
                    const int ci = (e0 > s0 && q < e0) ? 0 : ((r == 4 || (e1 > s1 && q < e1)) ? 1 : 2);
                    const int l = r < 4 ? 3 * r + ci : 12 + ci;
                    const unsigned long long slot = atomicAdd(ccnt, 1ull);
                    if (slot < cap)
                        cand[slot] = lower ? make_int4(k, l, p, 0) : make_int4(p, 26 - l, k, 0);
                }
            }
        }
    }
    const unsigned long long t = bsum(c, sh);
    if (threadIdx.x == 0 && t) atomicAdd(total, t);
}

// The candidate pass of k_mcount_total (mode 0) with each stencil row staged
// once per workgroup: the 256 consecutive bodies of a workgroup lie in a few
// neighbouring cells, so the union of their runs of one row (cells i - 1 ..
// i + 3 for cells i .. i + 2) is a short stretch of the sorted order; it is
// copied to LDS (records, ids) and the lanes walk their own runs there
// instead of loading every neighbour record from memory.  A workgroup whose
// union is longer than MC_CAP (its bodies straddle distant cells) walks
// memory as before.  Same tests, same candidates, same total.
constexpr int MC_CAP = 512;

__global__ __launch_bounds__(MT) void k_mcand_staged(const unsigned *__restrict__ skey,
                                                     const int *__restrict__ sid,
                                                     const double4 *__restrict__ pr,
                                                     const double4 *__restrict__ vm,
                                                     const unsigned *__restrict__ cs,
                                                     const unsigned *__restrict__ ce,
                                                     const int *__restrict__ dins, int d,
                                                     double dt, int4 *__restrict__ cand,
                                                     unsigned long long cap,
                                                     unsigned long long *__restrict__ ccnt,
                                                     unsigned long long *__restrict__ total)
{
    __shared__ double4 sp[MC_CAP], sv[MC_CAP];
    __shared__ int sq[MC_CAP];
    __shared__ unsigned s_lo, s_hi;
    __shared__ unsigned long long sh[MT];
    const int s = blockIdx.x * MT + threadIdx.x;
    const int ins = *dins;
    const bool act = s < ins;
    unsigned long long c = 0;
    long key = 0;
    int k = 0;
    double4 pk = make_double4(0, 0, 0, 0), vk = pk;
    if (act) {
        key = skey[s];
        k = sid[s];
        pk = pr[s];
        vk = vm[s];
    }
    for (int r = 0; r < 5; ++r) {
        unsigned a = 0, e = 0, s0 = 0, e0 = 0, s1 = 0, e1 = 0;
        if (act) {
            // Cells c0, c1, c2 of the run (c2 unused for the fifth run).
            const long k1 = r < 4 ? key + (long)(r % 3 - 1) * d + (long)(r / 3 - 1) * d * d : key;
            s0 = cs[k1 - 1];
            e0 = ce[k1 - 1];
            s1 = cs[k1];
            e1 = ce[k1];
            if (r < 4) {
                const unsigned s2 = cs[k1 + 1], e2 = ce[k1 + 1];
                a = e0 > s0 ? s0 : (e1 > s1 ? s1 : s2);
                e = e2 > s2 ? e2 : (e1 > s1 ? e1 : e0);
            } else {
                a = e0 > s0 ? s0 : s1;  // the own cell is not empty
                e = (unsigned)s;
            }
            if (a < e) c += e - a;
            else a = e = 0;
        }
        if (threadIdx.x == 0) {
            s_lo = 0xffffffffu;
            s_hi = 0u;
        }
        __syncthreads();
        if (a < e) {
            atomicMin(&s_lo, a);
            atomicMax(&s_hi, e);
        }
        __syncthreads();
        const unsigned lo = s_lo, hi = s_hi;
        const bool staged = hi > lo && hi - lo <= (unsigned)MC_CAP;
        if (staged) {
            for (unsigned q = lo + threadIdx.x; q < hi; q += MT) {
                sp[q - lo] = pr[q];
                sv[q - lo] = vm[q];
                sq[q - lo] = sid[q];
            }
            __syncthreads();
        }
        auto test = [&](unsigned q, int p, double4 rq, double4 vq) {
            const bool lower = p < k;
            if (approach_hit(pk, vk, rq, vq, dt)) {
                // Stencil index of q's cell.
                const int ci = (e0 > s0 && q < e0) ? 0 : ((r == 4 || (e1 > s1 && q < e1)) ? 1 : 2);
                const int l = r < 4 ? 3 * r + ci : 12 + ci;
                const unsigned long long slot = atomicAdd(ccnt, 1ull);
                if (slot < cap)
                    cand[slot] = lower ? make_int4(k, l, p, 0) : make_int4(p, 26 - l, k, 0);
            }
        };
        if (staged) {
            for (unsigned q = a; q < e; ++q) test(q, sq[q - lo], sp[q - lo], sv[q - lo]);
        } else {
            for (unsigned q = a; q < e; ++q) test(q, sid[q], pr[q], vm[q]);
        }
        __syncthreads();  // the row's reads are done before the next row is staged
    }
    const unsigned long long t = bsum(c, sh);
    if (threadIdx.x == 0 && t) atomicAdd(total, t);
}

// One thread per inserted body (sorted position s): neighbour count (p < k in
// the 27 cells) and, for mesh(), the collision candidates (k, l, p): every
// neighbour pair is tested once.  With at least 3 cells per axis the 27
// offsets are distinct and the 13 cells l < 13 lie below the own one, so a
// body tests every member of those cells and the lower ids of its own cell;
// a pair found from its lower id is recorded from its higher id's side
// (stencil index 26 - l).  With 2 cells per axis (a cell can appear twice in
// the stencil, and the reference then gathers its members twice) a body
// tests the lower ids of all 27 cells.
template <int MODE>
__global__ __launch_bounds__(MT) void k_mcount(const unsigned *__restrict__ skey,
                                               const int *__restrict__ sid,
                                               const double4 *__restrict__ pr,
                                               const double4 *__restrict__ vm,
                                               const unsigned *__restrict__ cs,
                                               const unsigned *__restrict__ ce, int ins_h,
                                               const int *__restrict__ dins, int d,
                                               double dt, int *__restrict__ count_by_id,
                                               int4 *__restrict__ cand, unsigned long long cap,
                                               unsigned long long *__restrict__ ccnt,
                                               unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long sh[MT];
    const int s = blockIdx.x * MT + threadIdx.x;
    const int ins = ins_count(ins_h, dins);
    unsigned long long c = 0;
    if (s < ins) {
        const int k = sid[s];
        const long key = skey[s];
        double4 pk = make_double4(0, 0, 0, 0), vk = pk;
        if (MODE == 0) {
            pk = pr[s];
            vk = vm[s];
        }
        for (int l = 0; l < 27; ++l) {
            const unsigned ck = (unsigned)(key + stencil_offset(l, d));
            const unsigned a = cs[ck], e = ce[ck];
            if (a == e) continue;
            if (MODE == 0 && d >= 3 && l < 13) {
                // The whole cell, counting the lower ids on the way; the next
                // member's record is loaded while the current one is tested.
                int np = sid[a];
                double4 nr = pr[a], nv = vm[a];
                for (unsigned q = a; q < e; ++q) {
                    const int p = np;
                    const double4 rq = nr, vq = nv;
                    if (q + 1 < e) {
                        np = sid[q + 1];
                        nr = pr[q + 1];
                        nv = vm[q + 1];
                    }
                    const bool lower = p < k;
                    c += lower;
                    if (approach_hit(pk, vk, rq, vq, dt)) {
                        const unsigned long long slot = atomicAdd(ccnt, 1ull);
                        if (slot < cap)
                            cand[slot] = lower ? make_int4(k, l, p, 0) : make_int4(p, 26 - l, k, 0);
                    }
                }
                continue;
            }
            unsigned lo = a, hi = e;  // first position with id >= k
            while (lo < hi) {
                const unsigned mid = (lo + hi) >> 1;
                if (sid[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            c += lo - a;
            if (MODE == 0 && (d < 3 || l == 13)) {
                for (unsigned q = a; q < lo; ++q) {
                    if (approach_hit(pr[q], vm[q], pk, vk, dt)) {
                        const unsigned long long slot = atomicAdd(ccnt, 1ull);
                        if (slot < cap) cand[slot] = make_int4(k, l, sid[q], 0);
                    }
                }
            }
        }
        count_by_id[k] = (int)c;
    }
    const unsigned long long t = bsum(c, sh);
    if (threadIdx.x == 0 && t) atomicAdd(total, t);
}

// ---- device replay of mesh() ---------------------------------------------------
// For callers that keep the bodies on the GPU, with the library's own bounce
// (elastic or inelastic, collision.c:351-366 with physics.c:545-605) and no
// body larger than the mesh size.  Then a body changes only during its own
// visit or as the partner p < k of a later visit, so every body still ahead
// of the current visit holds its start-of-pass state, and since a bounce
// removes no body, every binned body (only existing ones are binned) exists
// throughout the pass: the replay reads no existence flags.  The replay walks
// the visits with something to decide, in ascending id:
//   - events of visit k: candidate pairs (k, p) of the pass and pending
//     re-tests, in traversal order (cell l = 26..0, ascending p), each tested
//     exactly with the current states; the first hit is resolved;
//   - after a hit, the rest of visit k is scanned in full (k changed), MR_T
//     pairs at a time, the first hit in traversal order winning;
//   - whenever a body changes, its later visitors q > k (the members of its
//     27 cells) are tested against its new state at once; the ones that may
//     hit become pending re-tests of visit q.  A visitor's own state is its
//     start-of-pass state (it is still ahead), and a re-test is exact when it
//     runs, so the pending list only has to be a superset.
// One block of MR_T threads; a scan round takes MR_R entries per thread, whose
// loads are issued together (a scan is bound by the latency of its rounds).
// stat[0] += collisions; stat[1] |= 1 (more candidates than cap), 2 (more
// than MR_PCAP pending re-tests), 4 (a body larger than the mesh size).
constexpr int MR_PCAP = 8192;
constexpr int MR_T = 512;  // threads of k_mesh_replay
constexpr int MR_R = 2;    // entries per thread in a scan round
constexpr int MR_W = MR_R * MR_T / 64;  // hit masks of a round
static_assert(MR_R == 2, "the round votes below test h[0] and h[1]");

__global__ void k_lf_fill_u32(unsigned *__restrict__ p, unsigned n)
{
    if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Candidates (k, l, p) as sort keys (k, 26 - l, p), for k, p < 2^29.
__global__ __launch_bounds__(MT) void k_mcand_keys(const int4 *__restrict__ cand,
                                                   const unsigned long long *__restrict__ ccnt,
                                                   unsigned cap, unsigned long long *__restrict__ keys,
                                                   unsigned *__restrict__ vals)
{
    const unsigned i = blockIdx.x * MT + threadIdx.x;
    if (i >= cap) return;
    const unsigned long long m = *ccnt;
    unsigned long long k = ~0ull;
    if (i < m) {
        const int4 c = cand[i];
        k = ((unsigned long long)c.x << 34) | ((unsigned long long)(26 - c.y) << 29) |
            (unsigned long long)c.z;
    }
    keys[i] = k;
    vals[i] = i;
}

// The candidates in visiting order, gathered once for the replay's serial walk.
__global__ __launch_bounds__(MT) void k_mcand_gather(const int4 *__restrict__ cand,
                                                     const unsigned *__restrict__ order,
                                                     const unsigned long long *__restrict__ ccnt,
                                                     unsigned cap, int4 *__restrict__ sorted)
{
    const unsigned i = blockIdx.x * MT + threadIdx.x;
    const unsigned long long m = *ccnt;
    if (i < cap && i < m) sorted[i] = cand[order[i]];
}

// ---- clusters of the replay --------------------------------------------------
// Visits far apart in the mesh do not interact: a visit reads and changes
// only bodies of its 27 cells, and a changed body's later visitors are in
// its own 27 cells.  The candidates are grouped into clusters whose visits'
// cells are within `dist` cells (4 by default) of another visit of the cluster
// (k_mcluster), and each cluster is replayed by its own workgroup in id
// order.  Every workgroup claims the cells it reads or changes in a hash
// table; when two clusters claim one cell the clustered results are thrown
// away (the bodies restored from a snapshot) and the pass is replayed by one
// workgroup over all candidates, so the result is always the one-workgroup
// replay's.
constexpr int MCL_MAXM = 1024;     // candidates clustered (more: one cluster)
constexpr int MCL_TAB = 1 << 16;   // claim table entries (cell << 32 | cluster + 1)

__device__ inline bool claim_cell(unsigned long long *__restrict__ tab, unsigned c, int cl)
{
    const unsigned long long mine = ((unsigned long long)c << 32) | (unsigned)(cl + 1);
    unsigned h = (c * 2654435761u) >> 16;
    for (int probe = 0; probe < MCL_TAB; ++probe, h = (h + 1) & (MCL_TAB - 1)) {
        unsigned long long cur = tab[h];
        if (cur == 0ull) {
            cur = atomicCAS(&tab[h], 0ull, mine);
            if (cur == 0ull) return true;
        }
        if ((unsigned)(cur >> 32) == c) return cur == mine;
    }
    return false;  // table full: treated as a conflict
}

__global__ __launch_bounds__(MT) void k_fill_u64(unsigned long long *__restrict__ p, int n)
{
    const int i = blockIdx.x * MT + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// One workgroup of MCL_T threads: clusters of the m <= MCL_MAXM candidates
// (connected components of their visits' cells under Chebyshev distance
// <= dist, by hooking roots with atomicMin and path halving), then the
// candidates regrouped by cluster (each cluster's in their visiting order),
// coff[0 .. ncl] their offsets.  More candidates: one cluster of all.
constexpr int MCL_T = 1024;

__global__ __launch_bounds__(MCL_T) void k_mcluster(const int4 *__restrict__ cs_in,
                                                    const unsigned long long *__restrict__ ccnt,
                                                    unsigned cap,
                                                    const unsigned *__restrict__ key, int d,
                                                    int dist, int4 *__restrict__ cs_out,
                                                    int *__restrict__ coff,
                                                    int *__restrict__ ncl)
{
    __shared__ int lab[MCL_MAXM], cx[MCL_MAXM], cy[MCL_MAXM], cz[MCL_MAXM];
    __shared__ int s_changed;
    const int t = threadIdx.x;
    const unsigned long long total = *ccnt;
    const int m = (int)(total < cap ? total : cap);
    if (m > MCL_MAXM || m <= 1) {
        for (int i = t; i < m; i += MCL_T) cs_out[i] = cs_in[i];
        if (t == 0) {
            *ncl = m > 0 ? 1 : 0;
            coff[0] = 0;
            coff[1] = m;
        }
        return;
    }
    for (int i = t; i < m; i += MCL_T) {
        const unsigned c = key[cs_in[i].x];
        cx[i] = (int)(c % (unsigned)d);
        cy[i] = (int)((c / (unsigned)d) % (unsigned)d);
        cz[i] = (int)(c / ((unsigned)d * (unsigned)d));
        lab[i] = i;
    }
    __syncthreads();
    auto root = [&](int x) {
        while (true) {
            const int p = lab[x];
            if (p == x) return x;
            const int g = lab[p];
            if (g != p) atomicMin(&lab[x], g);  // path halving
            x = p;
        }
    };
    while (true) {
        if (t == 0) s_changed = 0;
        __syncthreads();
        for (int i = t; i < m; i += MCL_T) {
            for (int j = i + 1; j < m; ++j) {
                if (abs(cx[i] - cx[j]) > dist || abs(cy[i] - cy[j]) > dist ||
                    abs(cz[i] - cz[j]) > dist)
                    continue;
                const int ri = root(i), rj = root(j);
                if (ri != rj) {
                    atomicMin(&lab[ri > rj ? ri : rj], ri < rj ? ri : rj);
                    s_changed = 1;
                }
            }
        }
        __syncthreads();
        const int ch = s_changed;
        __syncthreads();
        if (!ch) break;
    }
    for (int i = t; i < m; i += MCL_T) lab[i] = root(i);
    __syncthreads();
    __shared__ int s_off[MCL_MAXM + 1];
    if (t == 0) {
        // Cluster ids in order of their first candidate, and their offsets.
        int nc = 0;
        for (int i = 0; i < m; ++i)
            if (lab[i] == i) cx[i] = nc++;  // cx reused: cluster id of root i
        for (int c = 0; c <= nc; ++c) s_off[c] = 0;
        for (int i = 0; i < m; ++i) s_off[cx[lab[i]] + 1]++;
        for (int c = 0; c < nc; ++c) s_off[c + 1] += s_off[c];
        *ncl = nc;
        cy[0] = nc;  // cy reused: the count, for the other threads
    }
    __syncthreads();
    const int nc = cy[0];
    for (int c = t; c <= nc; c += MCL_T) coff[c] = s_off[c];
    // Stable regrouping: candidate i goes after the earlier ones of its cluster.
    for (int i = t; i < m; i += MCL_T) {
        const int r = lab[i];
        int pos = s_off[cx[r]];
        for (int j = 0; j < i; ++j) pos += lab[j] == r;
        cs_out[pos] = cs_in[i];
    }
}

__device__ inline bool mr_before(int4 a, int4 b)  // (q, 26 - l, d) order
{
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y > b.y;
    return a.z < b.z;
}

__global__ __launch_bounds__(MR_T) void k_mesh_replay(
    DBody *__restrict__ b, int32_t *__restrict__ dcf, int one_coll,
    double f, double dt, const unsigned *__restrict__ key, const int *__restrict__ sid,
    const unsigned *__restrict__ cs, const unsigned *__restrict__ ce,
    const double4 *__restrict__ pr, const double4 *__restrict__ vm, int d,
    const int4 *__restrict__ csorted, const unsigned long long *__restrict__ ccnt, unsigned cap,
    const int *__restrict__ cnt, unsigned *__restrict__ stat, const int *__restrict__ coff,
    const int *__restrict__ ncl, unsigned long long *__restrict__ ctab,
    unsigned *__restrict__ conflict)
{
    __shared__ int4 P[MR_PCAP];
    __shared__ int s_np, s_pi, s_ci, s_k, s_hl, s_hp, s_hit, s_stop;
    // Entry r MR_T + lane of a scan round: its visitor and stencil cell; hit
    // mask w covers entries 64 w .. 64 w + 63 of the round.
    __shared__ int s_q[MR_R * MR_T], s_l[MR_R * MR_T];
    __shared__ unsigned long long s_m[MR_W];
    // 27 position ranges [ra[i], ra[i] + rn[i]) walked as one index space,
    // pre[i] = positions before range i, rl[i] = its stencil cell: set 0 for
    // the later-visitor scans, set 1 for the rest of a visit (which calls
    // those scans).
    __shared__ unsigned s_ra[2][27], s_rn[2][27];
    __shared__ __align__(16) unsigned s_pre[2][28];
    __shared__ int s_rl[2][27];
    __shared__ double s_ap[6], s_tc;
    const int lane = threadIdx.x, wv = lane >> 6;
    if (cnt[3] > 0) {
        if (lane == 0) atomicOr(stat + 1, 4u);
        return;
    }
    const unsigned long long total = *ccnt;
    if (total > cap) {
        if (lane == 0) atomicOr(stat + 1, 1u);
        return;
    }
    // coff: this workgroup replays cluster blockIdx.x (candidates
    // coff[b] .. coff[b + 1] of csorted), claiming the cells it touches when
    // there is more than one cluster; else the whole list.
    int c0 = 0, m = (int)total;
    bool claims = false;
    if (coff) {
        const int nc = *ncl;
        if ((int)blockIdx.x >= nc) return;
        c0 = coff[blockIdx.x];
        m = coff[blockIdx.x + 1] - c0;
        claims = nc > 1;
    }
    if (lane == 0) {
        s_np = 0;
        s_pi = 0;
        s_ci = 0;
    }
    __syncthreads();
    unsigned hits = 0;
    auto cand_at = [&](int i) { return csorted[c0 + i]; };
    // The 27 cells around cell key kd, claimed for this cluster (every lane
    // calls it).
    auto claim = [&](unsigned kd) {
        if (claims && lane < 27 && kd != NOKEY &&
            !claim_cell(ctab, (unsigned)((long)kd + stencil_offset(lane, d)), (int)blockIdx.x))
            atomicOr(conflict, 1u);
    };
    // Prefix of the 27 ranges of set r (every lane calls it): a scan across
    // the first wavefront's lanes.
    auto flatten = [&](int r) {
        __syncthreads();
        if (lane < 64) {
            const unsigned v = lane < 27 ? s_rn[r][lane] : 0u;
            unsigned incl = v;
            for (int o = 1; o < 32; o <<= 1) {
                const unsigned t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            if (lane < 27) s_pre[r][lane] = incl - v;
            if (lane == 26) s_pre[r][27] = incl;
        }
        __syncthreads();
    };
    // Range of flat index x < s_pre[r][27] in set r: the last i with
    // s_pre[r][i] <= x (a non-empty range), counted over all 28 prefixes.
    auto range_of = [&](int r, unsigned x) {
        const uint4 *pp = reinterpret_cast<const uint4 *>(s_pre[r]);
        int c = -1;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const uint4 v = pp[j];
            c += (v.x <= x) + (v.y <= x) + (v.z <= x) + (v.w <= x);
        }
        return c;
    };
    // Ranges of set 1 (the rest of visit kk from cell l0 down), loaded by
    // the second wavefront: called before an eager scan, whose set-0 loads
    // then run beside these.
    auto rest_ranges = [&](unsigned kk, int l0) {
        if (lane >= 64 && lane < 64 + 27) {
            const int l = l0 - (lane - 64);  // range `lane - 64` is cell l0 - (lane - 64)
            unsigned lo = 0, e = 0;
            if (l >= 0) {
                const unsigned c = (unsigned)((long)kk + stencil_offset(l, d));
                lo = cs[c];
                e = ce[c];
            }
            s_ra[1][lane - 64] = lo;
            s_rn[1][lane - 64] = e - lo;
            s_rl[1][lane - 64] = l;
        }
    };
    // The round's later visitors that may hit changed body dd (h[r] for
    // entry r MR_T + lane, whose visitor and cell are in s_q, s_l) become
    // pending re-tests.  Every lane calls it; most rounds hit nothing, and
    // one barrier vote decides that.
    auto record = [&](const bool *h, int dd) {
        if (!__syncthreads_or(h[0] || h[1])) return;
#pragma unroll
        for (int r = 0; r < MR_R; ++r) {
            const unsigned long long wm = __ballot(h[r]);
            if ((lane & 63) == 0) s_m[r * (MR_T / 64) + wv] = wm;
        }
        __syncthreads();
        if (lane == 0) {
            for (int w = 0; w < MR_W; ++w) {
              unsigned long long mask = s_m[w];
              while (mask) {
                const int bit = w * 64 + __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                // In visit q, dd sits in cell 26 - l (the stencil is symmetric).
                const int4 ev = make_int4(s_q[bit], 26 - s_l[bit], dd, 0);
                if (s_np >= MR_PCAP) {
                    atomicOr(stat + 1, 2u);
                    continue;
                }
                int at = s_np++;
                while (at > s_pi && mr_before(ev, P[at - 1])) {
                    P[at] = P[at - 1];
                    --at;
                }
                P[at] = ev;
              }
            }
        }
        __syncthreads();
    };
    // Later visitors of body dd (just changed during visit kc) that may hit
    // it: the members above kc of its 27 cells, MR_T at a time across cells.
    auto eager = [&](int dd, int kc) {
        const unsigned kd = key[dd];
        if (kd == NOKEY) return;
        claim(kd);
        const DBody D = b[dd];
        const double4 pd = make_double4(D.x, D.y, D.z, D.radius);
        const double4 vd = make_double4(D.vx, D.vy, D.vz, D.mass);
        // Whole cells: the members at or below kc are skipped in the scan
        // (cheaper than locating them first with dependent loads).
        if (lane < 27) {
            const unsigned c = (unsigned)((long)kd + stencil_offset(lane, d));
            const unsigned lo = cs[c];
            s_ra[0][lane] = lo;
            s_rn[0][lane] = ce[c] - lo;
            s_rl[0][lane] = lane;
        }
        flatten(0);
        const unsigned total = s_pre[0][27];
        for (unsigned base = 0; base < total; base += MR_R * MR_T) {
            bool h[MR_R];
#pragma unroll
            for (int r = 0; r < MR_R; ++r) {
                const unsigned x = base + r * MR_T + lane;
                h[r] = false;
                if (x < total) {
                    const int i = range_of(0, x);
                    const unsigned u = s_ra[0][i] + (x - s_pre[0][i]);
                    const int q = sid[u];
                    const double4 pu = pr[u], vu = vm[u];  // loaded beside sid[u]
                    s_q[r * MR_T + lane] = q;
                    s_l[r * MR_T + lane] = s_rl[0][i];
                    if (q > kc) h[r] = approach_hit(pd, vd, pu, vu, dt);
                }
            }
            record(h, dd);
        }
    };
    auto resolve = [&](int p, int k) {  // lane 0
        bf_bounce(b, p, k, f, s_ap, s_tc);
        if (dcf) {
            dcf[p] = one_coll;
            dcf[k] = one_coll;
        }
        ++hits;
    };
    while (true) {
        if (lane == 0) {
            // Another cluster met this one: the results are thrown away.
            s_stop = claims && __hip_atomic_load(conflict, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) != 0u;
            if (s_pi >= MR_PCAP / 2) {  // drop the consumed re-tests
                for (int q = s_pi; q < s_np; ++q) P[q - s_pi] = P[q];
                s_np -= s_pi;
                s_pi = 0;
            }
            const int kc = s_ci < m ? cand_at(s_ci).x : INT_MAX;
            const int kp = s_pi < s_np ? P[s_pi].x : INT_MAX;
            const int k = min(kc, kp);
            s_k = k;
            s_hit = 0;
            if (k != INT_MAX) {
                // Events of visit k in traversal order until the first hit.
                int ci = s_ci, pi = s_pi;
                int4 prev = make_int4(-1, -1, -1, -1);
                while (true) {
                    const bool hc = ci < m && cand_at(ci).x == k;
                    const bool hp = pi < s_np && P[pi].x == k;
                    if (!hc && !hp) break;
                    int4 ev;
                    if (hc && (!hp || !mr_before(P[pi], make_int4(k, cand_at(ci).y, cand_at(ci).z, 0)))) {
                        const int4 c = cand_at(ci++);
                        ev = make_int4(k, c.y, c.z, 0);
                    } else {
                        ev = P[pi++];
                    }
                    if (ev.y == prev.y && ev.z == prev.z) continue;
                    prev = ev;
                    double ap[6], tc;
                    if (bf_hit(b, nullptr, dcf, ev.z, k, dt, ap, &tc)) {
                        for (int q = 0; q < 6; ++q) s_ap[q] = ap[q];
                        s_tc = tc;
                        s_hl = ev.y;
                        s_hp = ev.z;
                        s_hit = 1;
                        resolve(ev.z, k);
                        break;
                    }
                }
                // Every event of visit k is consumed (after a hit the full scan
                // below covers the rest of the visit).
                while (s_ci < m && cand_at(s_ci).x == k) ++s_ci;
                while (s_pi < s_np && P[s_pi].x == k) ++s_pi;
            }
        }
        __threadfence_block();
        __syncthreads();
        const int k = s_k;
        if (k == INT_MAX || s_stop) break;
        claim(key[k]);  // the cells visit k read (its events and, after a hit, the rest)
        if (!s_hit) continue;
        // The rest of visit k in traversal order -- cells s_hl down to 0,
        // ids below k, after s_hp in cell s_hl -- MR_R MR_T pairs at a time
        // across cells; the first hit wins and the walk resumes after it.
        // Whole cells, walked in traversal order; the members outside the
        // rest of the visit (ids at or above k, and up to p0 in cell l0) are
        // skipped in the scan.
        // One walk over all 27 cells of k (l = 26 .. 0) does two things: the
        // rest of visit k (ids below k in cells l0 .. 0, after p0 in cell l0)
        // and, while k keeps the state of its first hit, k's own
        // later-visitor scan (ids above k).  A further hit changes k: the
        // scan's pending re-tests so far stay (a superset does no harm) and
        // k's scan runs again at the end with its final state.
        const int l0 = s_hl, p0 = s_hp;
        rest_ranges(key[k], 26);
        eager(s_hp, k);
        flatten(1);
        const unsigned total = s_pre[1][27];
        const DBody K = b[k];
        const double4 pkk = make_double4(K.x, K.y, K.z, K.radius);
        const double4 vkk = make_double4(K.vx, K.vy, K.vz, K.mass);
        bool stale = false;  // k changed after its first hit (the same in every lane)
        unsigned base = 0;
        while (base < total) {
            int pv[MR_R];
            double ap[MR_R][6], tc[MR_R];
            bool h[MR_R], g[MR_R];
#pragma unroll
            for (int r = 0; r < MR_R; ++r) {
                const unsigned x = base + r * MR_T + lane;
                pv[r] = -1;
                h[r] = false;
                g[r] = false;
                if (x < total) {
                    const int i = range_of(1, x);
                    const unsigned u = s_ra[1][i] + (x - s_pre[1][i]);
                    const int p = sid[u];
                    const double4 pu = pr[u], vu = vm[u];  // loaded beside sid[u]
                    const int l = s_rl[1][i];
                    pv[r] = p;
                    s_q[r * MR_T + lane] = p;
                    s_l[r * MR_T + lane] = l;
                    if (p < k && (l < l0 || (l == l0 && p > p0)))
                        h[r] = bf_hit(b, nullptr, dcf, p, k, dt, ap[r], &tc[r]);
                    else if (!stale && p > k)
                        g[r] = approach_hit(pkk, vkk, pu, vu, dt);
                }
            }
            // Most rounds hit nothing: one barrier decides (and orders the
            // reads of the previous round's masks before the writes below).
            if (!__syncthreads_or(h[0] || h[1])) {
                if (!stale) record(g, k);
                base += MR_R * MR_T;
                continue;
            }
#pragma unroll
            for (int r = 0; r < MR_R; ++r) {
                const unsigned long long hm = __ballot(h[r]);
                if ((lane & 63) == 0) s_m[r * (MR_T / 64) + wv] = hm;
            }
            __syncthreads();
            int first = -1;  // first hit of the round in traversal order
            for (int w = 0; w < MR_W && first < 0; ++w)
                if (s_m[w]) first = w * 64 + __ffsll((long long)s_m[w]) - 1;
            if (first < 0) {
                base += MR_R * MR_T;
                continue;
            }
#pragma unroll
            for (int r = 0; r < MR_R; ++r) {
                if (r * MR_T + lane == first) {
                    for (int q = 0; q < 6; ++q) s_ap[q] = ap[r][q];
                    s_tc = tc[r];
                    s_hp = pv[r];
                }
            }
            __syncthreads();
            if (lane == 0) resolve(s_hp, k);
            __threadfence_block();
            __syncthreads();
            eager(s_hp, k);
            stale = true;
            base = base + first + 1;
        }
        if (stale) eager(k, k);
    }
    if (lane == 0 && hits) atomicAdd(stat, hits);
}

// ---- neighbour-cell gravity ---------------------------------------------------
// Term G m_p (X_p - X_k) / D^3 of physics.c:186-200 (reference expression).
__device__ inline void term_exact(double X, double Y, double Z, double Rk, double4 p, double Rp,
                                  double G, double s, double &ax, double &ay, double &az)
{
    const double dx = X - p.x, dy = Y - p.y, dz = Z - p.z;
    const double soft = s * (Rk + Rp);
    const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
    const double D3 = D * D * D;
    const double gm = G * p.w;
    {
        double q0, q1, q2;
        div3(gm * dx, gm * dy, gm * dz, D3, q0, q1, q2);
        ax -= q0;
        ay -= q1;
        az -= q2;
    }
}

__device__ inline void term_fast(double X, double Y, double Z, double Rk, double4 p, double Rp,
                                 double G, double s, double &ax, double &ay, double &az)
{
    const double dx = p.x - X, dy = p.y - Y, dz = p.z - Z;
    const double e = s * (Rk + Rp);
    // 1e-200 in the chain: a body against itself adds 0 * finite.
    const double r2 =
        __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, __builtin_fma(e, e, 1e-200))));
    double y = __builtin_amdgcn_rsq(r2);
    y = y * __builtin_fma(-0.5 * r2, y * y, 1.5);
    const double f = (G * p.w) * y * (y * y);
    ax = __builtin_fma(f, dx, ax);
    ay = __builtin_fma(f, dy, ay);
    az = __builtin_fma(f, dz, az);
}

// Sorted current positions {x, y, z, m}, radius and existence of the binned bodies.
__global__ __launch_bounds__(MT) void k_gsorted(const DBody *__restrict__ b,
                                                const int32_t *__restrict__ ex,
                                                const int *__restrict__ sid, int ins,
                                                double4 *__restrict__ pm, double *__restrict__ rr,
                                                int *__restrict__ sx)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s >= ins) return;
    const int p = sid[s];
    const DBody q = b[p];
    pm[s] = make_double4(q.x, q.y, q.z, q.mass);
    rr[s] = q.radius;
    sx[s] = (ex == nullptr || ex[p] != 0) && p > 2;
}

// Central term of body k (physics.c:127-146): acceleration and CM_acc term.
__device__ inline void central_term(double x, double y, double z, double m, double cx, double cy,
                                    double cz, double mu, double G, double &ax, double &ay,
                                    double &az, double *ct)
{
    const double DX = x - cx, DY = y - cy, DZ = z - cz;
    const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
    const double r3 = rk * rk * rk;
    ax = -mu * DX / r3;
    ay = -mu * DY / r3;
    az = -mu * DZ / r3;
    ct[0] = G * m * DX / r3;
    ct[1] = G * m * DY / r3;
    ct[2] = G * m * DZ / r3;
}

// A caller's start value of body k (vector_field's value before its mutual
// terms: the central, J2 and perturber terms, physics.c:121-172), in place of
// the central term.
__device__ inline void start_value(const double *init, int k, double &ax, double &ay, double &az)
{
    ax = init[3 * (size_t)k];
    ay = init[3 * (size_t)k + 1];
    az = init[3 * (size_t)k + 2];
}

// acc += start value for every existing body (the fast forms).
__global__ __launch_bounds__(MT) void k_add_start(int n, const int32_t *__restrict__ ex,
                                                  const double *__restrict__ init,
                                                  double *__restrict__ acc)
{
    const int k = blockIdx.x * MT + threadIdx.x;
    if (k >= n || (ex && !ex[k])) return;
    for (int c = 0; c < 3; ++c) acc[3 * (size_t)k + c] += init[3 * (size_t)k + c];
}

// Bodies the sorted kernel does not cover: k <= 2 (central term only; their
// mutual part is added by k_big_final) and k > 2 not binned, appended ones
// (k >= nkey) included (central term and bodies 0..2).
__global__ __launch_bounds__(MT) void k_mgrav_rest(const DBody *__restrict__ b, int n,
                                                   const int32_t *__restrict__ ex,
                                                   const unsigned *__restrict__ key, int nkey,
                                                   const double *__restrict__ cen,
                                                   double G, double s, double *__restrict__ acc,
                                                   double *__restrict__ cterm)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    const int k = blockIdx.x * MT + threadIdx.x;
    if (k >= n || (ex && !ex[k])) return;
    if (k > 2 && k < nkey && key[k] != NOKEY) return;
    const DBody q = b[k];
    double ax = 0., ay = 0., az = 0.;
    if (cp) central_term(q.x, q.y, q.z, q.mass, cx, cy, cz, mu, G, ax, ay, az, cterm + 3 * (size_t)k);
    if (k > 2)
        for (int big = 0; big < 3 && big < n; ++big) {
            if (ex && !ex[big]) continue;
            const DBody g = b[big];
            term_exact(q.x, q.y, q.z, q.radius, make_double4(g.x, g.y, g.z, g.mass), g.radius, G, s,
                       ax, ay, az);
        }
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

// Neighbour records of the fast forms: {x, y, z, G m} of the sorted bodies,
// G m = 0 for the bodies that take no part in the sorted walk (absent ones,
// and bodies 0..2, whose terms are added apart), so the walk needs no test.
__global__ __launch_bounds__(MT) void k_gweights(const double4 *__restrict__ pm,
                                                 const int *__restrict__ sx, int ins, double G,
                                                 double4 *__restrict__ pw)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s >= ins) return;
    const double4 q = pm[s];
    pw[s] = make_double4(q.x, q.y, q.z, sx[s] ? G * q.w : 0.0);
}

// Binned bodies k > 2, one lane per sorted position t (bodies of one cell
// are adjacent lanes and walk the same cells).  Per body: central term,
// bodies 0..2, then the neighbour terms cell l = 0..26, ascending id within
// a cell (cell order: accelerations agree with the reference's pair-list
// order to rounding).  The three cells (i-1, i, i+1) of a stencil row have
// consecutive keys, so their members are one contiguous run of the sorted
// order: the walk is 9 runs.  The lane's own record enters with weight 0.
// FP64 with hardware 1/sqrt and a Newton step.  SOFT: softening lengths
// s (R_k + R_p) (else 0).
template <bool SOFT>
__global__ __launch_bounds__(MT) void k_mgrav(const DBody *__restrict__ b, int n,
                                              const int32_t *__restrict__ ex, int ins,
                                              const unsigned *__restrict__ cs,
                                              const unsigned *__restrict__ ce,
                                              const double4 *__restrict__ pw,
                                              const double *__restrict__ rr,
                                              const int *__restrict__ sx,
                                              const int *__restrict__ sid,
                                              const unsigned *__restrict__ skey, int d, const double *__restrict__ cen,
                                              double G, double s, double *__restrict__ acc,
                                              double *__restrict__ cterm)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    const int t = blockIdx.x * MT + threadIdx.x;
    if (t >= ins || !sx[t]) return;
    const int k = sid[t];
    const DBody me = b[k];
    const double Rk = SOFT ? rr[t] : 0.0;
    double ax = 0., ay = 0., az = 0.;
    if (cp) central_term(me.x, me.y, me.z, me.mass, cx, cy, cz, mu, G, ax, ay, az, cterm + 3 * (size_t)k);
    for (int big = 0; big < 3 && big < n; ++big) {
        if (ex && !ex[big]) continue;
        const DBody g = b[big];
        term_exact(me.x, me.y, me.z, me.radius, make_double4(g.x, g.y, g.z, g.mass), g.radius, G,
                   s, ax, ay, az);
    }
    const long kk = skey[t];
    for (int r = 0; r < 9; ++r) {
        const long kr = kk + (long)(r % 3 - 1) * d + (long)(r / 3 - 1) * d * d;
        const unsigned s0 = cs[kr - 1], e0 = ce[kr - 1], s1 = cs[kr], e1 = ce[kr];
        const unsigned s2 = cs[kr + 1], e2 = ce[kr + 1];
        // Empty cells hold [0, 0).
        const unsigned a = e0 > s0 ? s0 : (e1 > s1 ? s1 : s2);
        const unsigned e = e2 > s2 ? e2 : (e1 > s1 ? e1 : e0);
        {
#pragma unroll 2
            for (unsigned u = a; u < e; ++u) {
                const double4 p = pw[u];
                const double w = u == (unsigned)t ? 0.0 : p.w;
                const double dx = p.x - me.x, dy = p.y - me.y, dz = p.z - me.z;
                const double ee = SOFT ? s * (Rk + rr[u]) : 0.0;
                // 1e-200 in the chain: the lane's own record adds 0 * finite.
                const double r2 = __builtin_fma(
                    dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, __builtin_fma(ee, ee, 1e-200))));
                double y = __builtin_amdgcn_rsq(r2);
                y = y * __builtin_fma(-0.5 * r2, y * y, 1.5);
                const double f = w * y * (y * y);
                ax = __builtin_fma(f, dx, ax);
                ay = __builtin_fma(f, dy, ay);
                az = __builtin_fma(f, dz, az);
            }
        }
    }
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

// Bodies 0..2 receive += G m_k (X_k - X_big)/D^3 from every k > 2
// (physics.c:221-224, 237-240, 253-256), and CM_acc the central terms of
// every body (physics.c:143-145): per-block partial sums, 12 per block.
__global__ __launch_bounds__(MT) void k_big_partial(const DBody *__restrict__ b, int n,
                                                    const int32_t *__restrict__ ex, double G,
                                                    double s, const double *__restrict__ cterm,
                                                    const double *__restrict__ cen,
                                                    double *__restrict__ part)
{
    const int cp = cen && cen[7] != 0.0;
    __shared__ double sh[MT];
    double a[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = blockIdx.x * MT + threadIdx.x; k < n; k += gridDim.x * MT) {
        if (ex && !ex[k]) continue;
        if (cp)
            for (int c = 0; c < 3; ++c) a[9 + c] += cterm[3 * (size_t)k + c];
        if (k < 3) continue;
        const DBody q = b[k];
        for (int big = 0; big < 3; ++big) {
            if (ex && !ex[big]) continue;
            const DBody g = b[big];
            const double dx = q.x - g.x, dy = q.y - g.y, dz = q.z - g.z;
            const double soft = s * (q.radius + g.radius);
            const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
            const double D3 = D * D * D;
            const double gm = G * q.mass;
            {
                double q0, q1, q2;
                div3(gm * dx, gm * dy, gm * dz, D3, q0, q1, q2);
                a[3 * big + 0] += q0;
                a[3 * big + 1] += q1;
                a[3 * big + 2] += q2;
            }
        }
    }
    for (int c = 0; c < 12; ++c) {
        const double r = bsum(a[c], sh);
        if (threadIdx.x == 0) part[12 * blockIdx.x + c] = r;
    }
}

// Bodies 0..2: add the partials, then the pairs (0,1), (0,2), (1,2)
// (physics.c:300-341); CM_acc from its partials.  One block.
__global__ __launch_bounds__(MT) void k_big_final(const DBody *__restrict__ b, int n,
                                                  const int32_t *__restrict__ ex,
                                                  const double *__restrict__ part, int np,
                                                  double G, double s, double *__restrict__ acc,
                                                  const double *__restrict__ cen, double *__restrict__ cm)
{
    const int cp = cen && cen[7] != 0.0;
    __shared__ double sh[MT];
    double a[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < np; q += MT)
        for (int c = 0; c < 12; ++c) a[c] += part[12 * q + c];
    double tot[12];
    for (int c = 0; c < 12; ++c) tot[c] = bsum(a[c], sh);
    if (threadIdx.x != 0) return;
    if (cp && cm)
        for (int c = 0; c < 3; ++c) cm[c] = tot[9 + c];
    for (int big = 0; big < 3 && big < n; ++big) {
        if (ex && !ex[big]) continue;
        for (int c = 0; c < 3; ++c) acc[3 * big + c] += tot[3 * big + c];
    }
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    for (int q = 0; q < 3; ++q) {
        const int i = pa[q], j = pb[q];
        if (j >= n || (ex && (!ex[i] || !ex[j]))) continue;
        const DBody A = b[i], B = b[j];
        const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
        const double soft = s * (A.radius + B.radius);
        const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
        const double D3 = D * D * D;
        {
            double q0, q1, q2;
            div3(G * B.mass * dx, G * B.mass * dy, G * B.mass * dz, D3, q0, q1, q2);
            acc[3 * i + 0] -= q0;
            acc[3 * i + 1] -= q1;
            acc[3 * i + 2] -= q2;
        }
        {
            double q0, q1, q2;
            div3(G * A.mass * dx, G * A.mass * dy, G * A.mass * dz, D3, q0, q1, q2);
            acc[3 * j + 0] += q0;
            acc[3 * j + 1] += q1;
            acc[3 * j + 2] += q2;
        }
    }
}

// Pairs registered by visits of bodies larger than the mesh size (both > 2),
// in registration order; one thread.
__global__ void k_extra_pairs(const DBody *__restrict__ b, const int32_t *__restrict__ ex,
                              const int2 *__restrict__ pr, int np, double G, double s,
                              double *__restrict__ acc)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int q = 0; q < np; ++q) {
        const int k = pr[q].x, p = pr[q].y;
        if (ex && (!ex[k] || !ex[p])) continue;
        const DBody A = b[k], B = b[p];
        const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
        const double soft = s * (A.radius + B.radius);
        const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
        const double D3 = D * D * D;
        {
            double q0, q1, q2;
            div3(G * B.mass * dx, G * B.mass * dy, G * B.mass * dz, D3, q0, q1, q2);
            acc[3 * k + 0] -= q0;
            acc[3 * k + 1] -= q1;
            acc[3 * k + 2] -= q2;
        }
        {
            double q0, q1, q2;
            div3(G * A.mass * dx, G * A.mass * dy, G * A.mass * dz, D3, q0, q1, q2);
            acc[3 * p + 0] += q0;
            acc[3 * p + 1] += q1;
            acc[3 * p + 2] += q2;
        }
    }
}

// Registered pairs of the small bodies, written at their offsets in the
// reference's order (k ascending; cells 26..0; p ascending).
__global__ __launch_bounds__(MT) void k_mpairs(const unsigned *__restrict__ skey,
                                               const int *__restrict__ sid,
                                               const unsigned *__restrict__ cs,
                                               const unsigned *__restrict__ ce, int ins, int d,
                                               const long long *__restrict__ off,
                                               int2 *__restrict__ out)
{
    const int s = blockIdx.x * MT + threadIdx.x;
    if (s >= ins) return;
    const int k = sid[s];
    const long key = skey[s];
    long long o = off[k];
    for (int l = 26; l >= 0; --l) {
        const unsigned ck = (unsigned)(key + stencil_offset(l, d));
        for (unsigned t = cs[ck], e = ce[ck]; t < e; ++t) {
            const int p = sid[t];
            if (p >= k) break;
            out[o++] = make_int2(k, p);
        }
    }
}

// ---- ordered FP64 mesh gravity (NCG_F64_EXACT) ---------------------------------
// vector_field's mesh branch (physics.c:204-261, 296-380) gives body k > 2, in
// this order: the central term, bodies 0..2, then its terms of the pair list
// (physics.c:346-379), which mesh() / get_neighbours_mesh() fill visit by
// visit in ascending visiting id (collision.c:339-415, physics.c:1104-1116):
//   (1) visits v < k that registered k: only visits of bodies larger than the
//       mesh size (a small visit pairs only with bodies already inserted);
//   (2) its own visit: cells l = 26..0, ascending id < k (small k), or its
//       recorded list (big k);
//   (3) visits v > k that registered k, ascending v: the bodies of its 27
//       cells with id > k (a merge of 27 id-sorted ranges) and the recorded
//       big visits that list k.
// A pair's term has the same bits whichever body is `fst` (negating the
// separation is exact), so each body needs only this order.  Bodies 0..2 and
// CM_acc are sums in body order (k_mseq).  Big visits: bvk[v] ascending
// visiting ids, the ids each one paired with in bvo[bvoff[v], bvoff[v + 1]),
// ascending.

// Is id `x` in the ascending list o[a, e)?
__device__ inline bool bv_has(const int *__restrict__ o, int a, int e, int x)
{
    while (a < e) {
        const int m = (a + e) >> 1;
        const int v = o[m];
        if (v == x) return true;
        if (v < x) a = m + 1;
        else e = m;
    }
    return false;
}

// Term of body v (by id) on the body at (X, Y, Z) of radius Rk; pairs with a
// body 0..2 or an absent body are skipped as in physics.c:351.
__device__ inline void term_id(const DBody *__restrict__ b, const int32_t *__restrict__ ex, int v,
                               double X, double Y, double Z, double Rk, double G, double s,
                               double &ax, double &ay, double &az)
{
    if (v <= 2 || (ex && !ex[v])) return;
    const DBody q = b[v];
    term_exact(X, Y, Z, Rk, make_double4(q.x, q.y, q.z, q.mass), q.radius, G, s, ax, ay, az);
}

// Next recorded big visit from index vi on that lists body k (NOKEY if none).
__device__ inline unsigned next_big(const int *__restrict__ bvk, const int *__restrict__ bvoff,
                                    const int *__restrict__ bvo, int nbv, int &vi, int k)
{
    for (; vi < nbv; ++vi)
        if (bvk[vi] > 2 && bv_has(bvo, bvoff[vi], bvoff[vi + 1], k)) return (unsigned)bvk[vi];
    return NOKEY;
}

// Id-ordered neighbourhood lists, one per occupied cell c (ukey[i], i < *nu):
// the sorted positions of every body in c's 27 cells, ascending id (a cell
// that appears under two stencil indexes, only with 2 cells per axis, is
// listed twice, as the reference registers it twice).  For a body of c these
// are, after its own entry, the later visits that register it.  usz[i]: list
// length (0 for i >= *nu, up to ins).
__global__ __launch_bounds__(MT) void k_usize(const unsigned *__restrict__ ukey,
                                              const int *__restrict__ nu, int ins,
                                              const unsigned *__restrict__ cs,
                                              const unsigned *__restrict__ ce, int d,
                                              unsigned *__restrict__ usz)
{
    const int i = blockIdx.x * MT + threadIdx.x;
    if (i > ins) return;
    unsigned z = 0;
    if (i < *nu) {
        const long kk = ukey[i];
        for (int l = 0; l < 27; ++l) {
            const unsigned ck = (unsigned)(kk + stencil_offset(l, d));
            z += ce[ck] - cs[ck];
        }
    }
    usz[i] = z;
}

// One lane per occupied cell: 27-way merge by id into L[uoff[i], uoff[i+1]);
// each body of the cell gets lw = index of its own (last) entry, le = end.
__global__ __launch_bounds__(MT) void k_umerge(const unsigned *__restrict__ ukey,
                                               const int *__restrict__ nu,
                                               const unsigned *__restrict__ uoff,
                                               const unsigned *__restrict__ cs,
                                               const unsigned *__restrict__ ce,
                                               const int *__restrict__ sid, int d,
                                               unsigned *__restrict__ L, unsigned *__restrict__ lw,
                                               unsigned *__restrict__ le)
{
    const int i = blockIdx.x * MT + threadIdx.x;
    if (i >= *nu) return;
    const long kk = ukey[i];
    unsigned o = uoff[i];
    const unsigned oe = uoff[i + 1];
    unsigned hd[27], ps[27], en[27];
#pragma unroll
    for (int l = 0; l < 27; ++l) {
        const unsigned ck = (unsigned)(kk + stencil_offset(l, d));
        ps[l] = cs[ck];
        en[l] = ce[ck];
        hd[l] = ps[l] < en[l] ? (unsigned)sid[ps[l]] : NOKEY;
    }
    while (true) {
        unsigned m = hd[0];
#pragma unroll
        for (int l = 1; l < 27; ++l) m = min(m, hd[l]);
        if (m == NOKEY) break;
        unsigned sel = 0, sen = 0;
        int rep = 0;
        bool own = false;
#pragma unroll
        for (int l = 0; l < 27; ++l)
            if (hd[l] == m) {
                sel = ps[l];
                sen = en[l];
                ++rep;
                own |= l == 13;
            }
        for (int r = 0; r < rep; ++r) L[o++] = sel;
        if (own) {
            lw[sel] = o - 1;
            le[sel] = oe;
        }
        const unsigned nx = sel + 1;
        const unsigned nh = nx < sen ? (unsigned)sid[nx] : NOKEY;
#pragma unroll
        for (int l = 0; l < 27; ++l)
            if (hd[l] == m) {
                ps[l] = nx;
                hd[l] = nh;
            }
    }
}

// Binned bodies k > 2, one lane per sorted position t.
__global__ __launch_bounds__(MT) void k_mord(const DBody *__restrict__ b, int n,
                                             const int32_t *__restrict__ ex, int ins,
                                             const unsigned *__restrict__ cs,
                                             const unsigned *__restrict__ ce,
                                             const double4 *__restrict__ pm,
                                             const double *__restrict__ rr,
                                             const int *__restrict__ sx,
                                             const int *__restrict__ sid,
                                             const unsigned *__restrict__ skey, int d, const double *__restrict__ cen, double G,
                                             double s, const int *__restrict__ bvk,
                                             const int *__restrict__ bvoff,
                                             const int *__restrict__ bvo, int nbv,
                                             const unsigned *__restrict__ L,
                                             const unsigned *__restrict__ lw,
                                             const unsigned *__restrict__ le,
                                             double *__restrict__ acc,
                                             const double *__restrict__ init)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    const int t = blockIdx.x * MT + threadIdx.x;
    if (t >= ins || !sx[t]) return;
    const int k = sid[t];
    const double4 me = pm[t];
    const double Rk = rr[t];
    double ax = 0., ay = 0., az = 0., ct[3];
    if (init) start_value(init, k, ax, ay, az);
    else if (cp) central_term(me.x, me.y, me.z, me.w, cx, cy, cz, mu, G, ax, ay, az, ct);
    for (int big = 0; big < 3 && big < n; ++big) {
        if (ex && !ex[big]) continue;
        const DBody g = b[big];
        term_exact(me.x, me.y, me.z, Rk, make_double4(g.x, g.y, g.z, g.mass), g.radius, G, s, ax,
                   ay, az);
    }
    // (1) big visits before k.
    int vi = 0;
    for (; vi < nbv && bvk[vi] < k; ++vi)
        if (bvk[vi] > 2 && bv_has(bvo, bvoff[vi], bvoff[vi + 1], k))
            term_id(b, ex, bvk[vi], me.x, me.y, me.z, Rk, G, s, ax, ay, az);
    // The lanes of one cell walk the same positions in step (so that their
    // loads coincide), each lane adding only the terms that are its own.
    const long kk = skey[t];
    // (2) own visit: cells l = 26..0, ascending id < k; the walk of each cell
    // stops at the largest id of k's cell among this wavefront's lanes.
    const unsigned t1 = min(ce[(unsigned)kk] - 1u, (unsigned)(t | 63));
    const int kmax = sid[t1];
    for (int l = 26; l >= 0; --l) {
        const unsigned ck = (unsigned)(kk + stencil_offset(l, d));
        const unsigned a = cs[ck];
        unsigned lo = a, hi = ce[ck];
        while (lo < hi) {
            const unsigned mid = (lo + hi) >> 1;
            if (sid[mid] < kmax) lo = mid + 1;
            else hi = mid;
        }
        if (a == lo) continue;
        // The next member's record is loaded while the current term is added.
        int nid = sid[a], nx = sx[a];
        double4 np = pm[a];
        double nr = rr[a];
        for (unsigned u = a; u < lo; ++u) {
            const int idu = nid, xu = nx;
            const double4 pu = np;
            const double ru = nr;
            if (u + 1 < lo) {
                nid = sid[u + 1];
                nx = sx[u + 1];
                np = pm[u + 1];
                nr = rr[u + 1];
            }
            if (idu < k && xu) term_exact(me.x, me.y, me.z, Rk, pu, ru, G, s, ax, ay, az);
        }
    }
    // (3) later visits: the cell's id-ordered neighbourhood list after k's own
    // entry, merged with the recorded big visits that list k; the walk starts
    // after the entry of the cell's first body in this wavefront.
    const unsigned t0 = max(cs[(unsigned)kk], (unsigned)(t & ~63));
    const unsigned mine = lw[t];
    unsigned nb = next_big(bvk, bvoff, bvo, nbv, vi, k);
    // The list entry two ahead and the record one ahead are loaded while the
    // current term is added.
    const unsigned j0 = lw[t0] + 1, e = le[t];
    unsigned cu = j0 < e ? L[j0] : 0u, nu2 = j0 + 1 < e ? L[j0 + 1] : 0u;
    int cxs = j0 < e ? sx[cu] : 0;
    double4 cpm = j0 < e ? pm[cu] : make_double4(0., 0., 0., 0.);
    double crr = j0 < e ? rr[cu] : 0.;
    for (unsigned j = j0; j < e; ++j) {
        const unsigned u = cu;
        const int xu = cxs;
        const double4 pu = cpm;
        const double ru = crr;
        cu = nu2;
        if (j + 1 < e) {
            cxs = sx[cu];
            cpm = pm[cu];
            crr = rr[cu];
        }
        if (j + 2 < e) nu2 = L[j + 2];
        if (j <= mine) continue;
        if (nb != NOKEY) {
            const unsigned id = (unsigned)sid[u];
            while (nb < id) {
                term_id(b, ex, (int)nb, me.x, me.y, me.z, Rk, G, s, ax, ay, az);
                ++vi;
                nb = next_big(bvk, bvoff, bvo, nbv, vi, k);
            }
        }
        if (xu) term_exact(me.x, me.y, me.z, Rk, pu, ru, G, s, ax, ay, az);
    }
    while (nb != NOKEY) {
        term_id(b, ex, (int)nb, me.x, me.y, me.z, Rk, G, s, ax, ay, az);
        ++vi;
        nb = next_big(bvk, bvoff, bvo, nbv, vi, k);
    }
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

// Existing bodies k > 2 that are not binned (outside the collision cube,
// larger than the mesh size in mesh(), or appended during the pass: slots
// from nkey on): central term, bodies 0..2, then the recorded big visits in
// visiting order (their own one included).
__global__ __launch_bounds__(MT) void k_mrest_ord(const DBody *__restrict__ b, int n,
                                                  const int32_t *__restrict__ ex,
                                                  const unsigned *__restrict__ key, int nkey,
                                                  const double *__restrict__ cen,
                                                  double G, double s, const int *__restrict__ bvk,
                                                  const int *__restrict__ bvoff,
                                                  const int *__restrict__ bvo, int nbv,
                                                  double *__restrict__ acc,
                                                  const double *__restrict__ init)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    const int k = 3 + blockIdx.x * MT + threadIdx.x;
    if (k >= n || (ex && !ex[k]) || (k < nkey && key[k] != NOKEY)) return;
    const DBody q = b[k];
    double ax = 0., ay = 0., az = 0., ct[3];
    if (init) start_value(init, k, ax, ay, az);
    else if (cp) central_term(q.x, q.y, q.z, q.mass, cx, cy, cz, mu, G, ax, ay, az, ct);
    for (int big = 0; big < 3; ++big) {
        if (ex && !ex[big]) continue;
        const DBody g = b[big];
        term_exact(q.x, q.y, q.z, q.radius, make_double4(g.x, g.y, g.z, g.mass), g.radius, G, s, ax,
                   ay, az);
    }
    for (int vi = 0; vi < nbv; ++vi) {
        const int v = bvk[vi];
        if (v == k) {
            for (int u = bvoff[vi]; u < bvoff[vi + 1]; ++u)
                term_id(b, ex, bvo[u], q.x, q.y, q.z, q.radius, G, s, ax, ay, az);
        } else if (v > 2 && bv_has(bvo, bvoff[vi], bvoff[vi + 1], k)) {
            term_id(b, ex, v, q.x, q.y, q.z, q.radius, G, s, ax, ay, az);
        }
    }
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

// Bodies whose term order the host lists (tg: id, first, count triples into
// pl): central term, bodies 0..2, then the listed partners in order, each
// term as in k_mord.  Runs after k_mord / k_mrest_ord and replaces theirs.
__global__ __launch_bounds__(MT) void k_mexplicit(const DBody *__restrict__ b,
                                                  const int32_t *__restrict__ ex,
                                                  const double *__restrict__ cen, double G, double s,
                                                  const int *__restrict__ tg, int ntg,
                                                  const int *__restrict__ pl,
                                                  double *__restrict__ acc,
                                                  const double *__restrict__ init)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    const int i = blockIdx.x * MT + threadIdx.x;
    if (i >= ntg) return;
    const int k = tg[3 * i], first = tg[3 * i + 1], cnt = tg[3 * i + 2];
    if (k <= 2 || (ex && !ex[k])) return;
    const DBody q = b[k];
    double ax = 0., ay = 0., az = 0., ct[3];
    if (init) start_value(init, k, ax, ay, az);
    else if (cp) central_term(q.x, q.y, q.z, q.mass, cx, cy, cz, mu, G, ax, ay, az, ct);
    for (int big = 0; big < 3; ++big) {
        if (ex && !ex[big]) continue;
        const DBody g = b[big];
        term_exact(q.x, q.y, q.z, q.radius, make_double4(g.x, g.y, g.z, g.mass), g.radius, G, s, ax,
                   ay, az);
    }
    for (int u = first; u < first + cnt; ++u)
        term_id(b, ex, pl[u], q.x, q.y, q.z, q.radius, G, s, ax, ay, az);
    acc[3 * (size_t)k + 0] = ax;
    acc[3 * (size_t)k + 1] = ay;
    acc[3 * (size_t)k + 2] = az;
}

constexpr int SQ_CH = 768;  // bodies per chunk of the body-order sums

// Terms of the body-order sums (physics.c:121-146, 215-256), one thread per
// body slot k < nch * SQ_CH: chain j < 9 is component j % 3 of body j / 3 and
// receives G m_k (X_k - X_j) / D^3 from every existing k > 2; chains 9..11
// are CM_acc's components, G m_k (X_k - X_c) / r^3 from every existing k.
// Absent terms are -0.0, which leaves every sum (-0.0 included) unchanged.
// Layout: chunk c is a [12][SQ_CH] block, so that each chain's terms are
// contiguous.
__global__ __launch_bounds__(MT) void k_mterms(const DBody *__restrict__ b, int n,
                                               const int32_t *__restrict__ ex, const double *__restrict__ cen, double G, double s,
                                               int nslots, double *__restrict__ T)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const int k = blockIdx.x * MT + threadIdx.x;
    if (k >= nslots) return;
    double v[12];
    for (int j = 0; j < 12; ++j) v[j] = -0.0;
    if (k < n && (!ex || ex[k])) {
        const DBody q = b[k];
        if (cp) {
            const double DX = q.x - cx, DY = q.y - cy, DZ = q.z - cz;
            const double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
            const double r3 = rk * rk * rk;
            v[9] = G * q.mass * DX / r3;
            v[10] = G * q.mass * DY / r3;
            v[11] = G * q.mass * DZ / r3;
        }
        if (k > 2)
            for (int i = 0; i < 3; ++i) {
                if (i >= n || (ex && !ex[i])) continue;
                const DBody g = b[i];
                const double dx = q.x - g.x, dy = q.y - g.y, dz = q.z - g.z;
                const double soft = s * (q.radius + g.radius);
                const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
                const double D3 = D * D * D;
                {
                    double q0, q1, q2;
                    div3(G * q.mass * dx, G * q.mass * dy, G * q.mass * dz, D3, q0, q1, q2);
                    v[3 * i + 0] = q0;
                    v[3 * i + 1] = q1;
                    v[3 * i + 2] = q2;
                }
            }
    }
    double *dst = T + (size_t)(k / SQ_CH) * (12 * SQ_CH) + (k % SQ_CH);
    for (int j = 0; j < 12; ++j) dst[(size_t)j * SQ_CH] = v[j];
}

// Staging of k_mseq: producer thread pt (waves 1..3) holds double2 number
// pt + 192 q (q < 24) of a chunk of terms in named registers (an indexed array
// would be placed in memory) and stores them into an LDS buffer of 12 rows of
// SQ_CH + 2 doubles.
static_assert(12 * SQ_CH == 2 * 192 * 24, "k_mseq stages 24 double2 per producer thread");
// Consumer groups: 12 paired reads (at most 15 LDS reads may be outstanding).
constexpr int SQ_G = 12;
static_assert((SQ_CH / 2) % (2 * SQ_G) == 0, "whole ping-pong rounds per chunk");
#define SEQ_EACH(X)                                                                        \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) \
    X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23)
#define SEQ_DECL(q) double2 ra##q, rb##q;
#define SEQ_LOAD_A(q) ra##q = src_[pt + 192 * (q)];
#define SEQ_LOAD_B(q) rb##q = src_[pt + 192 * (q)];
#define SEQ_STAGE_A(q)                                                       \
    {                                                                        \
        const int e_ = 2 * (pt + 192 * (q)), j_ = e_ / SQ_CH, c_ = e_ % SQ_CH; \
        *reinterpret_cast<double2 *>(&buf_[j_][c_]) = ra##q;                 \
    }
#define SEQ_STAGE_B(q)                                                       \
    {                                                                        \
        const int e_ = 2 * (pt + 192 * (q)), j_ = e_ / SQ_CH, c_ = e_ % SQ_CH; \
        *reinterpret_cast<double2 *>(&buf_[j_][c_]) = rb##q;                 \
    }

// The sums themselves, in ascending k.  Waves 1..3 stage the next chunk into
// LDS from registers, with loads issued two chunks ahead; lanes 0..11 of wave 0 add the
// current chunk from LDS, reading each group of 24 terms while adding the
// previous one.  Chains 0..8 start at the central term of their body,
// CM_acc's at 0.  Thread 0 then applies the pairs (0,1), (0,2), (1,2)
// (physics.c:296-341).  One block of 256 threads; producer and consumer run
// separate loops with the same barriers.
__global__ __launch_bounds__(256) void k_mseq(const DBody *__restrict__ b, int n,
                                              const int32_t *__restrict__ ex, const double *__restrict__ cen, double G, double s,
                                              const double2 *__restrict__ T, int nch,
                                              double *__restrict__ acc, double *__restrict__ cm,
                                              const double *__restrict__ init)
{
    const int cp = cen && cen[7] != 0.0;
    const double cx = cp ? cen[0] : 0.0, cy = cp ? cen[1] : 0.0, cz = cp ? cen[2] : 0.0;
    const double mu = cp ? G * cen[6] : 0.0;
    // Rows padded by two doubles: 16-byte aligned for paired reads, and the
    // 12 summing lanes start on different banks.
    __shared__ double S[2][12][SQ_CH + 2];
    __shared__ double fin[12];
    __shared__ DBody g[3];
    __shared__ int bex[3];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid < 3) {
        bex[tid] = tid < n && (!ex || ex[tid]);
        if (bex[tid]) g[tid] = b[tid];
    }
    __syncthreads();
    if (w == 0) {
        double sum = 0.0;
        if (lane < 9) {
            const int i = lane / 3, c = lane % 3;
            if (init && bex[i]) {
                sum = init[3 * i + c];
            } else if (cp && bex[i]) {
                double a[3], ct[3];
                central_term(g[i].x, g[i].y, g[i].z, g[i].mass, cx, cy, cz, mu, G, a[0], a[1], a[2],
                             ct);
                sum = a[c];
            }
        }
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            if (lane < 12) {
                const double2 *row = reinterpret_cast<const double2 *>(S[c & 1][lane]);
                double2 x[SQ_G], y[SQ_G];
#pragma unroll
                for (int m = 0; m < SQ_G; ++m) x[m] = row[m];
#pragma unroll
                for (int i = SQ_G; i < SQ_CH / 2; i += 2 * SQ_G) {
#pragma unroll
                    for (int m = 0; m < SQ_G; ++m) y[m] = row[i + m];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < SQ_G; ++m) {
                        sum += x[m].x;
                        sum += x[m].y;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    if (i + SQ_G < SQ_CH / 2) {
#pragma unroll
                        for (int m = 0; m < SQ_G; ++m) x[m] = row[i + SQ_G + m];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < SQ_G; ++m) {
                        sum += y[m].x;
                        sum += y[m].y;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __syncthreads();
        }
        if (lane < 12) fin[lane] = sum;
    } else {
        // Two register sets: chunk c + 1 is staged from one while the loads of
        // chunk c + 2 (issued an iteration earlier) land in the other; the
        // staged set is then refilled with chunk c + 3.
        const int pt = tid - 64;
        SEQ_EACH(SEQ_DECL)
        {
            const double2 *src_ = T;
            SEQ_EACH(SEQ_LOAD_A)
            double(*buf_)[SQ_CH + 2] = S[0];
            SEQ_EACH(SEQ_STAGE_A)
        }
        if (nch > 1) {
            const double2 *src_ = T + 6 * SQ_CH;
            SEQ_EACH(SEQ_LOAD_B)
        }
        if (nch > 2) {
            const double2 *src_ = T + 12 * SQ_CH;
            SEQ_EACH(SEQ_LOAD_A)
        }
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            if (c + 1 < nch) {
                double(*buf_)[SQ_CH + 2] = S[(c + 1) & 1];
                const double2 *src_ = T + (size_t)(c + 3) * (6 * SQ_CH);
                if ((c & 1) == 0) {
                    SEQ_EACH(SEQ_STAGE_B)
                    if (c + 3 < nch) {
                        SEQ_EACH(SEQ_LOAD_B)
                    }
                } else {
                    SEQ_EACH(SEQ_STAGE_A)
                    if (c + 3 < nch) {
                        SEQ_EACH(SEQ_LOAD_A)
                    }
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (tid != 0) return;
    double a[9];
    for (int j = 0; j < 9; ++j) a[j] = fin[j];
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    for (int q = 0; q < 3; ++q) {
        const int i = pa[q], j = pb[q];
        if (!bex[i] || !bex[j]) continue;
        const double dx = g[i].x - g[j].x, dy = g[i].y - g[j].y, dz = g[i].z - g[j].z;
        const double soft = s * (g[i].radius + g[j].radius);
        const double D = sqrt(dx * dx + dy * dy + dz * dz + soft * soft);
        const double D3 = D * D * D;
        {
            double q0, q1, q2;
            div3(G * g[j].mass * dx, G * g[j].mass * dy, G * g[j].mass * dz, D3, q0, q1, q2);
            a[3 * i + 0] -= q0;
            a[3 * i + 1] -= q1;
            a[3 * i + 2] -= q2;
        }
        {
            double q0, q1, q2;
            div3(G * g[i].mass * dx, G * g[i].mass * dy, G * g[i].mass * dz, D3, q0, q1, q2);
            a[3 * j + 0] += q0;
            a[3 * j + 1] += q1;
            a[3 * j + 2] += q2;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (bex[i])
            for (int c = 0; c < 3; ++c) acc[3 * i + c] = a[3 * i + c];
    if (cp && cm)
        for (int c = 0; c < 3; ++c) cm[c] = fin[9 + c];
}

#undef SEQ_EACH
#undef SEQ_DECL
#undef SEQ_LOAD_A
#undef SEQ_LOAD_B
#undef SEQ_STAGE_A
#undef SEQ_STAGE_B

}  // namespace

// ---- host side --------------------------------------------------------------

namespace {

// Host view of the binning, with the bodies a collision moved to another
// cell before they were inserted (see the replay below).
struct HostCells {
    std::vector<unsigned> key;   // per body, key of the GPU binning
    std::vector<unsigned> skey;  // per sorted position
    std::vector<int> sid;
    std::map<int, unsigned> reloc;               // body -> key at its insertion
    std::map<unsigned, std::vector<int>> added;  // cell -> relocated bodies (ascending)
    bool loaded = false;
    long d = 0;
    // Key at insertion time; NOKEY for bodies appended during the pass,
    // which are never inserted.
    unsigned eff(int b) const
    {
        if (b < 0 || b >= (int)key.size()) return NOKEY;
        auto it = reloc.find(b);
        return it == reloc.end() ? key[b] : it->second;
    }
    // Bodies of cell `ck` (ascending id) at insertion time.
    void members(unsigned ck, std::vector<int> &out) const
    {
        out.clear();
        auto lo = std::lower_bound(skey.begin(), skey.end(), ck);
        auto hi = std::upper_bound(lo, skey.end(), ck);
        for (auto it = lo; it != hi; ++it) {
            const int q = sid[it - skey.begin()];
            if (!reloc.count(q)) out.push_back(q);
        }
        auto ad = added.find(ck);
        if (ad != added.end()) {
            out.insert(out.end(), ad->second.begin(), ad->second.end());
            std::sort(out.begin(), out.end());
        }
    }
    void relocate(int b, unsigned nk)
    {
        const unsigned ok = eff(b);
        if (ok == nk) return;
        if (reloc.count(b) && ok != NOKEY) {
            auto &v = added[ok];
            v.erase(std::remove(v.begin(), v.end(), b), v.end());
        }
        reloc[b] = nk;
        if (nk != NOKEY) {
            auto &v = added[nk];
            v.insert(std::upper_bound(v.begin(), v.end(), b), b);
        }
    }
    int stencil_index(long delta) const
    {
        for (int l = 0; l < 27; ++l)
            if (stencil_offset(l, d) == delta) return l;
        return -1;
    }
};

// Key of a body at position (x, y, z) as hash_table_cell_index computes it
// (collision.c:197-207), NOKEY outside the collision cube.
unsigned host_key(double x, double y, double z, double gam, double cube, long d)
{
    if (fabs(x) > cube / 2.0 || fabs(y) > cube / 2.0 || fabs(z) > cube / 2.0) return NOKEY;
    const long i = d / 2 + 1 + (int)floor(x / gam);
    const long j = d / 2 + 1 + (int)floor(y / gam);
    const long k = d / 2 + 1 + (int)floor(z / gam);
    return (unsigned)(i + j * d + k * d * d);
}

constexpr int MESH_BIG_CAP = 1 << 16;

// Device buffers of one binning (kept in the context between the pass and
// the field).
struct MeshBin {
    unsigned *key, *skey, *cs, *ce, *pk;
    int *idx, *sid, *cbi, *cnt, *bigl;
    double4 *pr, *vm;
    unsigned long long *tot;
    void *tmp;
    size_t tmp_bytes;
};

// Bins the bodies db (device) in the reference's cells: keys, (key, id) radix
// sort, cell ranges, sorted copies; counters cnt (inserted, small, big,
// big-list length) and tot stay on the device.  Clears the cells of the
// previous pass first.
ncg_status mesh_bin(ncg_ctx *ctx, const DBody *db, int32_t n, const int32_t *dex, double gam,
                    double cube, long d, long long nkeys, int mode, MeshBin &B, hipStream_t st)
{
    MeshState &M = ctx->mesh;
    unsigned *key = dbuf<unsigned>(ctx, "m_key", n);
    int *idx = dbuf<int>(ctx, "m_idx", n);
    unsigned *skey = dbuf<unsigned>(ctx, "m_skey", n);
    int *sid = dbuf<int>(ctx, "m_sid", n);
    int *cbi = dbuf<int>(ctx, "m_count", n);
    double4 *pr = dbuf<double4>(ctx, "m_pr", n);
    double4 *vm = dbuf<double4>(ctx, "m_vm", n);
    int *cnt = dbuf<int>(ctx, "m_cnt", 4);
    unsigned long long *tot = dbuf<unsigned long long>(ctx, "m_tot", 2);
    const int big_cap = MESH_BIG_CAP;
    int *bigl = dbuf<int>(ctx, "m_big", big_cap);
    if (!key || !idx || !skey || !sid || !cbi || !pr || !vm || !cnt || !tot || !bigl)
        return NCG_ERR_NOMEM;
    // Dense cell table; cleared once when (re)allocated, afterwards only the
    // keys of the previous pass are reset.
    if (M.table_keys < nkeys) {
        M.table_keys = 0;
        M.prev_ins = 0;
        unsigned *cs0 = dbuf<unsigned>(ctx, "m_cs", nkeys);
        unsigned *ce0 = dbuf<unsigned>(ctx, "m_ce", nkeys);
        if (!cs0 || !ce0) return NCG_ERR_NOMEM;
        NCG_HIP(hipMemsetAsync(cs0, 0, sizeof(unsigned) * nkeys, st));
        NCG_HIP(hipMemsetAsync(ce0, 0, sizeof(unsigned) * nkeys, st));
        M.table_keys = nkeys;
    }
    unsigned *cs = dbuf<unsigned>(ctx, "m_cs", nkeys);
    unsigned *ce = dbuf<unsigned>(ctx, "m_ce", nkeys);
    if (M.prev_ins > 0) {
        hipLaunchKernelGGL(k_clear_cells, dim3(cdiv(M.prev_ins, MT)), dim3(MT), 0, st,
                           (const unsigned *)dbuf<unsigned>(ctx, "m_prevkeys", M.prev_ins),
                           M.prev_ins, cs, ce);
        NCG_CHECK_LAUNCH();
        M.prev_ins = 0;
    }
    // Sized for every body (growth happens here, after the clear above).
    unsigned *pk = dbuf<unsigned>(ctx, "m_prevkeys", n);
    size_t tmp_bytes = 0;
    NCG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, skey, idx, sid, n, 0, 32, st));
    void *tmp = ncg_dbuf(ctx, "m_sorttmp", tmp_bytes);
    if (!pk || !tmp) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * 4, st));
    NCG_HIP(hipMemsetAsync(tot, 0, sizeof(unsigned long long) * 2, st));
    const unsigned g = (unsigned)cdiv(n, MT);
    {
        // Binning: keys, (key, id) radix sort, cell ranges, sorted copies.
        // The inserted count stays on the device (cnt[0]): no host wait.
        ProfScope ps(ctx, st, "mesh_bin");
        hipLaunchKernelGGL(k_mkey, dim3(g), dim3(MT), 0, st, db, (int)n, dex, gam, cube,
                           (int)d, (int)mode, key, idx, cnt, bigl, big_cap);
        NCG_CHECK_LAUNCH();
        // NOKEY (all ones) sorts last; every real key is below 2^32 - 1.
        NCG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, skey, idx, sid, n, 0, 32, st));
        hipLaunchKernelGGL(k_bounds, dim3(g), dim3(MT), 0, st, skey, 0, cnt, cs, ce);
        hipLaunchKernelGGL(k_msorted, dim3(g), dim3(MT), 0, st, db, sid, 0, cnt, pr, vm);
        hipLaunchKernelGGL(k_copy_keys, dim3(g), dim3(MT), 0, st, skey, cnt, pk);
        NCG_CHECK_LAUNCH();
    }
    B = MeshBin{key, skey, cs, ce, pk, idx, sid, cbi, cnt, bigl, pr, vm, tot, tmp, tmp_bytes};
    return NCG_OK;
}

// Forget the previous pass (its binning stays until the next mesh_bin).
void reset_pass(MeshState &M, int32_t n)
{
    M.valid = false;
    M.pairs_known = false;
    M.ord_lists = false;
    M.extra_pairs.clear();
    M.big_visits.clear();
    M.big_pairs_all.clear();
    M.big_scan.clear();
    M.events.clear();
    M.frag_ids.clear();
    M.removed.clear();
    M.x_targets.clear();
    M.x_partners.clear();
    M.n_end = n;
    M.rm_drop = 0;
}

// Number of stencil indexes l' in (l, 26] naming the same cell as l (only
// with 2 cells per axis): a body listed twice in a visit is entry `occ` of it.
int stencil_occ(int l, long d)
{
    int c = 0;
    for (int q = l + 1; q < 27; ++q) c += stencil_offset(q, d) == stencil_offset(l, d);
    return c;
}

ncg_status load_cells(ncg_ctx *ctx, HostCells &hc, int32_t n, int ins, const unsigned *dkey,
                      const unsigned *dskey, const int *dsid)
{
    if (hc.loaded) return NCG_OK;
    hipStream_t st = ctx->stream;
    // Through pinned staging (one copy of each array, then host copies).
    unsigned *pk = pbuf<unsigned>(ctx, "m_hkey", (size_t)n + 2 * (size_t)ins);
    if (!pk) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(pk, dkey, sizeof(unsigned) * n, hipMemcpyDeviceToHost, st));
    if (ins) {
        NCG_HIP(hipMemcpyAsync(pk + n, dskey, sizeof(unsigned) * ins, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipMemcpyAsync(pk + n + ins, dsid, sizeof(int) * ins, hipMemcpyDeviceToHost, st));
    }
    NCG_HIP(hipStreamSynchronize(st));
    hc.key.assign(pk, pk + n);
    hc.skey.assign(pk + n, pk + n + ins);
    hc.sid.assign(reinterpret_cast<const int *>(pk + n + ins), reinterpret_cast<const int *>(pk + n + 2 * (size_t)ins));
    hc.loaded = true;
    return NCG_OK;
}

// Term lists of the bodies that full fragmentations touched, for the exact
// field (physics.c:347-378 sums each body's pairs in list order).  For each
// body X of a fragmentation block (survivor and appended bodies, ids > 2),
// X's entries of the pair list in order, i.e. by visit, then by position in
// the visit:
//   * visits of bodies larger than the mesh size: their recorded pairs
//     (block pairs included) that hold X;
//   * when X was inserted: its own visit (cells 26..0, ascending id below
//     it) and the later visits that gather it (stencil index 26..0 of their
//     own walk), each fragmentation's entry replaced by the block's other
//     bodies (physics.c:1002-1010: pairs (id[p], id[q]), q < p);
//   * for an appended body, the entry of the small visit whose block made it:
//     the other bodies of the block in block order.
// Absent partners stay listed: the field skips them as physics.c:351 does.
// FP64-fast and mixed forms add the block pairs of small visits after the
// rest (big visits' pairs are among the extra pairs already).
ncg_status explicit_lists(ncg_ctx *ctx, MeshState &M, HostCells &hc, int32_t n, int ins,
                          const unsigned *key, const unsigned *skey, const int *sid)
{
    std::map<std::tuple<int, int, int>, const MeshEvent *> blk;  // small visits: (visit, l, partner)
    std::set<int> targets;
    for (const MeshEvent &e : M.events)
        if (e.kind == MeshEvent::BLOCK) {
            if (e.l >= 0) blk[std::make_tuple(e.k, e.l, e.p)] = &e;
            for (int q = 0; q < e.count; ++q)
                if (M.frag_ids[e.first + q] > 2) targets.insert(M.frag_ids[e.first + q]);
        }
    if (targets.empty()) return NCG_OK;
    ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
    if (r != NCG_OK) return r;
    const long d = hc.d;
    std::vector<int> &P = M.x_partners;
    // (visit, position in the visit, order among equal positions, partner)
    std::vector<std::tuple<int, long, int, int>> ent;
    auto emit = [&](int v, int l, int p, int other, int self) {
        const long pos = (long)(26 - l) * (1L << 31) + p;
        auto it = blk.find(std::make_tuple(v, l, p));
        if (it == blk.end()) {
            ent.emplace_back(v, pos, 0, other);
            return;
        }
        const MeshEvent &e = *it->second;
        for (int q = 0; q < e.count; ++q)
            if (M.frag_ids[e.first + q] != self)
                ent.emplace_back(v, pos, q, M.frag_ids[e.first + q]);
    };
    std::vector<int> mem;
    for (int X : targets) {
        ent.clear();
        for (const BigVisit &bv : M.big_visits)
            for (int q = 0; q < bv.count; ++q) {
                const ncg_pair &pq = M.big_pairs_all[bv.first + q];
                if (pq.fst == X) ent.emplace_back(bv.k, (long)q, 0, pq.snd);
                else if (pq.snd == X) ent.emplace_back(bv.k, (long)q, 0, pq.fst);
            }
        const unsigned kx = hc.eff(X);
        if (kx != NOKEY) {
            for (int l = 26; l >= 0; --l) {
                hc.members((unsigned)((long)kx + stencil_offset(l, d)), mem);
                for (int p : mem) {
                    if (p >= X) break;
                    emit(X, l, p, p, X);
                }
            }
            for (int l = 0; l < 27; ++l) {
                hc.members((unsigned)((long)kx + stencil_offset(l, d)), mem);
                for (int q : mem)
                    if (q > X) emit(q, 26 - l, X, q, X);  // X at index 26 - l of q's visit
            }
        }
        for (const auto &kv : blk) {
            const MeshEvent &e = *kv.second;
            for (int a = 1; a < e.count; ++a)
                if (M.frag_ids[e.first + a] == X) {
                    const long pos = (long)(26 - e.l) * (1L << 31) + e.p;
                    for (int b = 0; b < e.count; ++b)
                        if (b != a) ent.emplace_back(e.k, pos, b, M.frag_ids[e.first + b]);
                }
        }
        std::stable_sort(ent.begin(), ent.end(), [](const std::tuple<int, long, int, int> &u,
                                                    const std::tuple<int, long, int, int> &w) {
            if (std::get<0>(u) != std::get<0>(w)) return std::get<0>(u) < std::get<0>(w);
            if (std::get<1>(u) != std::get<1>(w)) return std::get<1>(u) < std::get<1>(w);
            return std::get<2>(u) < std::get<2>(w);
        });
        const int first = (int)P.size();
        for (const auto &t : ent) P.push_back(std::get<3>(t));
        M.x_targets.push_back(X);
        M.x_targets.push_back(first);
        M.x_targets.push_back((int)P.size() - first);
    }
    for (const auto &kv : blk) {
        const MeshEvent &e = *kv.second;
        for (int a = 1; a < e.count; ++a)
            for (int b = 0; b < a; ++b) {
                const int u = M.frag_ids[e.first + a], w = M.frag_ids[e.first + b];
                if (u > 2 && w > 2) M.extra_pairs.push_back({u, w});
            }
    }
    return NCG_OK;
}

}  // namespace

extern "C" ncg_status ncg_mesh_pass(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                    const int32_t *exists, double timestep, double gam,
                                    double collision_cube, int32_t cells, int32_t mode,
                                    const ncg_replay_ops *ops, ncg_mesh_stats *stats)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || bodies) && (mode == 0 || mode == 1) && stats,
                "invalid arguments");
    NCG_REQUIRE(mode == 1 || (ops && ops->test && ops->resolve), "mesh() needs replay ops");
    NCG_REQUIRE(cells >= 2 && gam > 0.0 && collision_cube > 0.0, "invalid mesh geometry");
    const long d = cells;
    const long long nkeys = (long long)(d + 2) * (1 + d + d * d) + 1;
    NCG_REQUIRE(nkeys < 0xffffffffLL, "collision_cube_cells %d too large for 32-bit cell keys",
                (int)cells);
    memset(stats, 0, sizeof *stats);
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    MeshState &M = ctx->mesh;
    reset_pass(M, n);
    if (n == 0) {
        M.valid = true;
        M.pairs_known = true;
        M.n = 0;
        M.inserted = 0;
        return NCG_OK;
    }

    DBody *db = dbuf<DBody>(ctx, "m_b", n);
    int32_t *dex = exists ? dbuf<int32_t>(ctx, "m_ex", n) : nullptr;
    if (!db || (exists && !dex)) return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(db, bodies, sizeof(DBody) * (size_t)n, hipMemcpyHostToDevice, st));
    if (exists)
        NCG_HIP(hipMemcpyAsync(dex, exists, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    MeshBin B;
    ncg_status rb = mesh_bin(ctx, db, n, dex, gam, collision_cube, d, nkeys, mode, B, st);
    if (rb != NCG_OK) return rb;
    unsigned *key = B.key, *skey = B.skey, *cs = B.cs, *ce = B.ce, *pk = B.pk;
    int *idx = B.idx, *sid = B.sid, *cbi = B.cbi, *cnt = B.cnt, *bigl = B.bigl;
    double4 *pr = B.pr, *vm = B.vm;
    unsigned long long *tot = B.tot;
    void *tmp = B.tmp;
    size_t tmp_bytes = B.tmp_bytes;
    const int big_cap = MESH_BIG_CAP;
    unsigned long long cap = 1 << 16;
    int4 *dc = mode == 0 ? dbuf<int4>(ctx, "m_cand", cap) : nullptr;
    if (mode == 0 && !dc) return NCG_ERR_NOMEM;
    const unsigned g = (unsigned)cdiv(n, MT);
    int hc4[4];
    NCG_HIP(hipMemcpyAsync(hc4, cnt, sizeof hc4, hipMemcpyDeviceToHost, st));
    std::vector<int4> cand;
    unsigned long long total = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        {
            ProfScope ps(ctx, st, "mesh_count");
            if (mode == 0)
                hipLaunchKernelGGL(k_mcount<0>, dim3(g), dim3(MT), 0, st, skey, sid, pr, vm, cs, ce,
                                   0, cnt, (int)d, timestep, cbi, dc, cap, tot + 1, tot);
            else
                hipLaunchKernelGGL(k_mcount<1>, dim3(g), dim3(MT), 0, st, skey, sid, pr, vm, cs, ce,
                                   0, cnt, (int)d, timestep, cbi, (int4 *)nullptr, 0ull, tot + 1, tot);
            NCG_CHECK_LAUNCH();
        }
        unsigned long long ht[2];
        NCG_HIP(hipMemcpyAsync(ht, tot, sizeof ht, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        total = ht[0];
        if (ht[1] <= cap) {
            cand.resize(ht[1]);
            if (ht[1]) {
                NCG_HIP(hipMemcpyAsync(cand.data(), dc, sizeof(int4) * ht[1], hipMemcpyDeviceToHost, st));
                NCG_HIP(hipStreamSynchronize(st));
            }
            break;
        }
        cap = ht[1];
        dc = dbuf<int4>(ctx, "m_cand", cap);
        if (!dc) return NCG_ERR_NOMEM;
        NCG_HIP(hipMemsetAsync(tot, 0, sizeof(unsigned long long) * 2, st));
    }
    const int ins = hc4[0];
    M.prev_ins = ins;
    const int nbig = std::min(hc4[3], big_cap);
    NCG_REQUIRE(hc4[3] <= big_cap, "more than %d bodies larger than the mesh size", big_cap);
    std::vector<int> big(nbig);
    if (nbig) {
        NCG_HIP(hipMemcpyAsync(big.data(), bigl, sizeof(int) * nbig, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        std::sort(big.begin(), big.end());
    }
    // Visits of bodies larger than the mesh size still ahead (a body's class
    // is its radius at its visit, collision.c:342-344: an outcome may move an
    // unvisited body across gam).
    std::set<int> bigset(big.begin(), big.end());
    stats->total_neighbours = (int64_t)total;
    stats->how_many_small = mode == 0 ? hc4[1] : 0;
    stats->how_many_big = mode == 0 ? hc4[2] : 0;
    stats->candidates = (int64_t)cand.size();
    M.n = n;
    M.cells = cells;
    M.mode = mode;
    M.inserted = ins;
    M.gam = gam;
    M.cube = collision_cube;
    M.valid = true;
    M.pairs_known = true;
    if (mode == 1) return NCG_OK;

    // ---- host replay of collision.c:339-415 ----------------------------------
    // Candidates in visiting order: k ascending, cell 26..0, p ascending.
    std::sort(cand.begin(), cand.end(), [](const int4 &a, const int4 &b) {
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y > b.y;
        return a.z < b.z;
    });
    HostCells hc;
    hc.d = d;
    std::vector<char> dirty((size_t)n, 0);
    // Re-tests owed to later visits: (visit k, modified neighbour p < k), a
    // min-heap on k (entries are only ever added for visits still ahead).
    std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                        std::greater<std::pair<int, int>>>
        pending;
    std::set<int> full_visit;                 // small bodies modified before their visit
    std::vector<int> owed;                    // the current visit's entries of `pending`
    int64_t ncoll = 0, htests = 0;
    ncg_status err = NCG_OK;
    std::vector<int> mem;
    auto live = [&](int32_t k) { return !exists || exists[k] != 0; };
    // Visit during which each body was removed (sized at the first removal).
    // A visit gathers its neighbours when it starts (collision.c:283-300): a
    // body removed during an earlier visit is not among them, one removed
    // during this visit still is.
    std::vector<int> rm_at;
    auto gathered = [&](int p, int kv) { return rm_at.empty() || rm_at[p] >= kv; };
    // Bodies created by an outcome are appended in mesh mode (get_free_index(1),
    // physics.c:994-996, structure.c:857-861): slots n, n + 1, ...
    int32_t ncol = n;
    // Later visits q > kv that have b as a neighbour re-test it on the host.
    auto schedule = [&](int b, int kv) -> ncg_status {
        ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
        if (r != NCG_OK) return r;
        const unsigned kb = hc.eff(b);
        if (kb == NOKEY) return NCG_OK;  // not in the hash table: nobody's neighbour
        for (int l = 0; l < 27; ++l) {
            hc.members((unsigned)((long)kb + stencil_offset(l, d)), mem);
            for (int q : mem)
                if (q > kv && q > b) pending.push({q, b});
        }
        return NCG_OK;
    };
    auto mark_dirty = [&](int b, int kv) -> ncg_status {
        // Bodies appended during the pass are never inserted: nobody's neighbour.
        if (b >= n || dirty[b]) return NCG_OK;
        dirty[b] = 1;
        return schedule(b, kv);
    };
    int32_t born[64];
    int64_t small_adj = 0, big_adj = 0;
    // Class at the visit of the bodies an outcome changed before it: 0 small,
    // 1 big, 2 removed (not visited); others keep their start-of-pass class.
    std::map<int, int> cls;
    auto class_of = [&](int b) {
        auto it = cls.find(b);
        if (it != cls.end()) return it->second;
        return bigset.count(b) ? 1 : 0;
    };
    // Body b (kv < b < n, not yet visited) was changed by an outcome of visit
    // kv.  At its visit the reference reads its live state: skipped when it no
    // longer exists, inserted at its current cell when small, a scan of every
    // body when big (collision.c:339-344).  Its visit is then decided on the
    // host, and the later visits that gather it re-test it.
    auto unvisited_changed = [&](int b, int kv) -> ncg_status {
        ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
        if (r != NCG_OK) return r;
        const int was = class_of(b);
        if (!live(b)) {
            if (was == 2) return NCG_OK;
            if (hc.eff(b) != NOKEY) hc.relocate(b, NOKEY);
            (was == 0 ? small_adj : big_adj) -= 1;
            bigset.erase(b);
            full_visit.erase(b);
            cls[b] = 2;
            return NCG_OK;
        }
        if (bodies[b].radius <= gam) {
            const unsigned nk = host_key(bodies[b].x, bodies[b].y, bodies[b].z, gam,
                                         collision_cube, d);
            const bool moved_cell = nk != hc.eff(b);
            if (moved_cell) hc.relocate(b, nk);
            if (was == 1) {
                bigset.erase(b);
                ++small_adj;
                --big_adj;
            }
            cls[b] = 0;
            full_visit.insert(b);
            if (moved_cell || was == 1) return schedule(b, kv);
            return NCG_OK;
        }
        if (hc.eff(b) != NOKEY) hc.relocate(b, NOKEY);
        if (was == 0) {
            bigset.insert(b);
            --small_adj;
            ++big_adj;
        }
        cls[b] = 1;
        full_visit.erase(b);
        return NCG_OK;
    };
    // Resolves entry (k, p) of visit kv, found at stencil index l (-1: a big
    // visit), and records what the outcome leaves in the pair list.
    auto collide = [&](int p, int k, int kv, int l) -> ncg_status {
        MeshEvent e{kv, l, p, l >= 0 ? stencil_occ(l, d) : 0, MeshEvent::PAIR, {p, k}, 0, 0};
        if (!live(p) || !live(k)) {
            // collision_treatment returns at once (physics.c:1053-1055): the
            // visitor, or a partner listed twice, went earlier in this visit.
            e.kind = MeshEvent::NONE;
            M.events.push_back(e);
            return NCG_OK;
        }
        const double mp = bodies[p].mass, mk = bodies[k].mass;
        ops->resolve(ops->user, p, k);
        ++ncoll;
        const int32_t nborn = ops->created ? ops->created(ops->user, born, 64) : 0;
        if (nborn < 0)
            return ncg_fail(ctx, NCG_ERR_CAPACITY,
                            "mesh(): more than 64 bodies created by the outcome of collision (%d, %d)",
                            p, k);
        for (int q = 0; q < nborn; ++q)
            if (born[q] != ncol + q)
                return ncg_fail(ctx, NCG_ERR_INVALID,
                                "mesh(): body %d created by collision (%d, %d) is not appended at "
                                "slot %d (mesh() appends, structure.c:857-861)",
                                born[q], p, k, ncol + q);
        const bool lp = live(p), lk = live(k);
        if ((int32_t)rm_at.size() < ncol + nborn && !rm_at.empty())
            rm_at.resize((size_t)(ncol + nborn), INT32_MAX);
        if (lp && lk) {
            if (nborn != 0)
                return ncg_fail(ctx, NCG_ERR_INVALID,
                                "mesh(): the outcome of collision (%d, %d) kept both bodies and "
                                "created %d (no outcome of the reference does)",
                                p, k, nborn);
            // A bounce registers (p, k) (physics.c:1063-1080); a partial
            // fragmentation, which changes the masses, (heavier, lighter)
            // (physics.c:1049-1051, 886-890).
            if (bodies[p].mass != mp || bodies[k].mass != mk)
                e.reg = mp > mk ? ncg_pair{p, k} : ncg_pair{k, p};
            M.events.push_back(e);
            ncg_status r = mark_dirty(p, kv);
            if (r != NCG_OK) return r;
            return mark_dirty(k, kv);
        }
        if (rm_at.empty()) rm_at.assign((size_t)(ncol + nborn), INT32_MAX);
        if (!lp && !lk) {
            // A super-catastrophic disruption that leaves no body
            // (physics.c:819-823) registers nothing.
            if (nborn != 0)
                return ncg_fail(ctx, NCG_ERR_INVALID,
                                "mesh(): the outcome of collision (%d, %d) removed both bodies and "
                                "created %d (no outcome of the reference does)",
                                p, k, nborn);
            rm_at[p] = rm_at[k] = kv;
            M.removed.push_back({p, kv});
            M.removed.push_back({k, kv});
            e.kind = MeshEvent::NONE;
            M.events.push_back(e);
            return NCG_OK;
        }
        // A merger (physics.c:628-690) or a super-catastrophic disruption
        // (physics.c:797-840) registers nothing; a full fragmentation
        // (physics.c:899-1013) registers the pairs of its bodies, survivor
        // first, then the appended ones.
        const int A = lp ? p : k, gone = lp ? k : p;
        rm_at[gone] = kv;
        M.removed.push_back({gone, kv});
        if (nborn != 0) {
            e.kind = MeshEvent::BLOCK;
            e.first = (int)M.frag_ids.size();
            e.count = nborn + 1;
            M.frag_ids.push_back(A);
            for (int q = 0; q < nborn; ++q) M.frag_ids.push_back(born[q]);
            ncol += nborn;
        } else {
            e.kind = MeshEvent::NONE;
        }
        M.events.push_back(e);
        return mark_dirty(A, kv);
    };
    // Pairs of visit k from traversal position (cell from_l, id >= from_p)
    // onwards, all decided on the host.
    auto full_rest = [&](int k, int from_l, int from_p) -> ncg_status {
        ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
        if (r != NCG_OK) return r;
        const unsigned kk = hc.eff(k);
        if (kk == NOKEY) return NCG_OK;
        std::vector<int> cell;
        for (int l = from_l; l >= 0; --l) {
            hc.members((unsigned)((long)kk + stencil_offset(l, d)), cell);
            for (int p : cell) {
                if (p >= k) break;
                if ((l == from_l && p < from_p) || !gathered(p, k)) continue;
                ++htests;
                if (ops->test(ops->user, p, k)) {
                    r = collide(p, k, k, l);
                    if (r != NCG_OK) return r;
                }
            }
        }
        return NCG_OK;
    };
    size_t ci = 0;
    while (true) {
        const int kc = ci < cand.size() ? cand[ci].x : INT32_MAX;
        const int kb = bigset.empty() ? INT32_MAX : *bigset.begin();
        const int kp = pending.empty() ? INT32_MAX : pending.top().first;
        const int kf = full_visit.empty() ? INT32_MAX : *full_visit.begin();
        const int k = std::min(std::min(kc, kb), std::min(kp, kf));
        if (k == INT32_MAX) break;
        if (k == kb) {
            // Body larger than the mesh size (collision.c:381-414): every
            // existing body, small or a larger id, up to the live largest_id
            // (bodies appended during the pass included), at its live radius.
            bigset.erase(bigset.begin());
            while (ci < cand.size() && cand[ci].x == k) ++ci;
            while (!pending.empty() && pending.top().first == k) pending.pop();
            if (!live(k)) continue;
            std::vector<ncg_pair> visit;
            std::vector<char> scan;
            for (int32_t p = 0; p < ncol; ++p) {
                if (!(live(p) && p != k)) continue;
                const double Rp = bodies[p].radius;
                if (!(Rp < gam || (Rp >= gam && p > k))) continue;
                ++htests;
                visit.push_back({k, p});
                scan.push_back(1);
                if (ops->test(ops->user, p, k)) {
                    ncg_status r = collide(p, k, k, -1);
                    if (r != NCG_OK) return r;
                    const MeshEvent &ev = M.events.back();
                    if (ev.kind == MeshEvent::NONE) {
                        visit.pop_back();
                        scan.pop_back();
                    } else if (ev.kind == MeshEvent::PAIR) {
                        visit.back() = ev.reg;
                    } else {
                        // Full fragmentation: the pairs of its bodies
                        // (physics.c:1002-1010), survivor first.
                        visit.pop_back();
                        scan.pop_back();
                        for (int a = 1; a < ev.count; ++a)
                            for (int b2 = 0; b2 < a; ++b2) {
                                visit.push_back({M.frag_ids[ev.first + a], M.frag_ids[ev.first + b2]});
                                scan.push_back(0);
                            }
                    }
                    if (p > k && p < n) {
                        r = unvisited_changed(p, k);
                        if (r != NCG_OK) return r;
                    }
                }
            }
            M.big_visits.push_back({k, (int)M.big_pairs_all.size(), (int)visit.size()});
            for (size_t q = 0; q < visit.size(); ++q) {
                const ncg_pair &pq = visit[q];
                M.big_pairs_all.push_back(pq);
                M.big_scan.push_back(scan[q]);
                if (pq.fst > 2 && pq.snd > 2) M.extra_pairs.push_back(pq);
            }
            continue;
        }
        // Small body k.
        owed.clear();
        while (!pending.empty() && pending.top().first == k) {
            owed.push_back(pending.top().second);
            pending.pop();
        }
        if (!live(k) || class_of(k) != 0) {
            // Removed before its visit, or now larger than the mesh size: no
            // small visit (collision.c:340-344).
            full_visit.erase(k);
            while (ci < cand.size() && cand[ci].x == k) ++ci;
            continue;
        }
        if (full_visit.count(k)) {
            full_visit.erase(k);
            while (ci < cand.size() && cand[ci].x == k) ++ci;
            if (live(k)) {
                err = full_rest(k, 26, 0);
                if (err != NCG_OK) return err;
            }
            continue;
        }
        // Candidates and pending re-tests, merged in traversal order
        // ((26 - l) * 2^31 + p).  Pairs whose cells no longer match (a body
        // moved before insertion) are dropped.
        std::vector<std::pair<long, int>> todo;
        const bool moved = !hc.reloc.empty();
        while (ci < cand.size() && cand[ci].x == k) {
            const int p = cand[ci].z;
            int l = cand[ci].y;
            if (moved && (hc.reloc.count(p) || hc.reloc.count(k)))
                l = hc.eff(p) == NOKEY || hc.eff(k) == NOKEY
                        ? -1
                        : hc.stencil_index((long)hc.eff(p) - (long)hc.eff(k));
            if (l >= 0 && gathered(p, k)) todo.push_back({(long)(26 - l) * (1L << 31) + p, p});
            ++ci;
        }
        if (!owed.empty()) {
            ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
            if (r != NCG_OK) return r;
            for (int p : owed) {
                if (hc.eff(p) == NOKEY || hc.eff(k) == NOKEY || !gathered(p, k)) continue;
                const int l = hc.stencil_index((long)hc.eff(p) - (long)hc.eff(k));
                if (l >= 0) todo.push_back({(long)(26 - l) * (1L << 31) + p, p});
            }
        }
        std::sort(todo.begin(), todo.end());
        todo.erase(std::unique(todo.begin(), todo.end()), todo.end());
        for (size_t q = 0; q < todo.size(); ++q) {
            const int p = todo[q].second;
            ++htests;
            if (ops->test(ops->user, p, k)) {
                const int l = 26 - (int)(todo[q].first >> 31);
                err = collide(p, k, k, l);
                if (err != NCG_OK) return err;
                // The rest of this visit involves k, which changed: host.
                err = full_rest(k, l, p + 1);
                if (err != NCG_OK) return err;
                break;
            }
        }
    }
    if (!hc.reloc.empty()) {
        // Re-bin with the insertion-time cells so that total_neighbours, the
        // pair list and the mesh gravity see what the reference's hash saw.
        std::vector<unsigned> nk(hc.key);
        for (const auto &kv : hc.reloc) nk[kv.first] = kv.second;
        int nins = 0;
        for (unsigned v : nk) nins += v != NOKEY;
        hipLaunchKernelGGL(k_clear_cells, dim3(cdiv(std::max(ins, 1), MT)), dim3(MT), 0, st,
                           (const unsigned *)pk, ins, cs, ce);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipMemcpyAsync(key, nk.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_iota, dim3(cdiv(n, MT)), dim3(MT), 0, st, idx, (int)n);
        NCG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, skey, idx, sid, n, 0, 32, st));
        pk = dbuf<unsigned>(ctx, "m_prevkeys", std::max(nins, 1));
        if (!pk) return NCG_ERR_NOMEM;
        if (nins > 0) {
            hipLaunchKernelGGL(k_bounds, dim3(cdiv(nins, MT)), dim3(MT), 0, st, skey, nins,
                               (const int *)nullptr, cs, ce);
            NCG_HIP(hipMemcpyAsync(pk, skey, sizeof(unsigned) * nins, hipMemcpyDeviceToDevice, st));
            NCG_HIP(hipMemsetAsync(tot, 0, sizeof(unsigned long long) * 2, st));
            hipLaunchKernelGGL(k_mcount<1>, dim3(cdiv(nins, MT)), dim3(MT), 0, st, skey, sid, pr,
                               vm, cs, ce, nins, (const int *)nullptr, (int)d, timestep, cbi,
                               (int4 *)nullptr, 0ull, tot + 1, tot);
            NCG_CHECK_LAUNCH();
        }
        unsigned long long ht[2] = {0, 0};
        NCG_HIP(hipMemcpyAsync(ht, tot, sizeof ht, hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        M.prev_ins = nins;
        M.inserted = nins;
        stats->total_neighbours = (int64_t)ht[0];
    }
    M.n_end = ncol;
    stats->how_many_small += small_adj;
    stats->how_many_big += big_adj;
    if (!M.removed.empty()) {
        // Later visits do not gather a removed body (collision.c:296): one
        // neighbour fewer for every later visit whose cells hold it.
        ncg_status r = load_cells(ctx, hc, n, ins, key, skey, sid);
        if (r != NCG_OK) return r;
        int64_t drop = 0;
        for (const auto &rv : M.removed) {
            const unsigned kb = hc.eff(rv.first);
            if (kb == NOKEY) continue;
            for (int l = 0; l < 27; ++l) {
                hc.members((unsigned)((long)kb + stencil_offset(l, d)), mem);
                for (int q : mem) drop += q > rv.second;
            }
        }
        M.rm_drop = drop;
        stats->total_neighbours -= drop;
    }
    err = explicit_lists(ctx, M, hc, n, ins, key, skey, sid);
    if (err != NCG_OK) return err;
    stats->collisions = ncoll;
    stats->host_tests = htests;
    return NCG_OK;
}

static ncg_status build_ord_lists(ncg_ctx *ctx, int ins, long d, const unsigned *skey,
                                  const unsigned *cs, const unsigned *ce, const int *sid,
                                  hipStream_t st);
static ncg_status ensure_aux(ncg_ctx *ctx);

ncg_status ncg_mesh_pass_dev_flags(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                   const int32_t *d_exists, int32_t *d_did_collide,
                                   double timestep, double gam, double collision_cube,
                                   int32_t cells, int32_t mode, double collision_f,
                                   int32_t one_collision_only, ncg_mesh_stats *stats,
                                   unsigned *flags, hipStream_t st)
{
    *flags = 0u;
    NCG_REQUIRE(n >= 0 && (n == 0 || d_bodies) && (mode == 0 || mode == 1) && stats,
                "invalid arguments");
    NCG_REQUIRE(cells >= 3 && gam > 0.0 && collision_cube > 0.0, "invalid mesh geometry");
    NCG_REQUIRE(mode == 1 || collision_f > 0.0, "mesh() on the device needs collision_f > 0");
    NCG_REQUIRE(n < (1 << 29), "the device replay takes fewer than 2^29 body slots");
    NCG_REQUIRE(!(mode == 0 && one_collision_only && !d_did_collide),
                "one_collision_only needs a did_collide array");
    const long d = cells;
    const long long nkeys = (long long)(d + 2) * (1 + d + d * d) + 1;
    NCG_REQUIRE(nkeys < 0xffffffffLL, "collision_cube_cells %d too large for 32-bit cell keys",
                (int)cells);
    memset(stats, 0, sizeof *stats);
    MeshState &M = ctx->mesh;
    reset_pass(M, n);
    if (n == 0) {
        M.valid = true;
        M.pairs_known = true;
        M.n = 0;
        M.inserted = 0;
        return NCG_OK;
    }
    DBody *db = reinterpret_cast<DBody *>(d_bodies);
    MeshBin B;
    ncg_status r = mesh_bin(ctx, db, n, d_exists, gam, collision_cube, d, nkeys, mode, B, st);
    if (r != NCG_OK) return r;
    const unsigned cap = ncg_replay_cap(ctx, 1u << 16);
    int4 *dc = mode == 0 ? dbuf<int4>(ctx, "m_cand", cap) : nullptr;
    unsigned *stat = dbuf<unsigned>(ctx, "m_rstat", 2);
    unsigned *hs = pbuf<unsigned>(ctx, "m_hstat", 16);
    if ((mode == 0 && !dc) || !stat || !hs) return NCG_ERR_NOMEM;
    const unsigned g = (unsigned)cdiv(n, MT);
    {
        ProfScope ps(ctx, st, "mesh_count");
        // The device pass keeps no per-body counts (ncg_mesh_pairs is not
        // available after it): the total alone.
        if (mode == 0)
            hipLaunchKernelGGL(k_mcand_staged, dim3(g), dim3(MT), 0, st, B.skey, B.sid, B.pr,
                               B.vm, B.cs, B.ce, B.cnt, (int)d, timestep, dc,
                               (unsigned long long)cap, B.tot + 1, B.tot);
        else
            hipLaunchKernelGGL(k_mcount_total<1>, dim3(g), dim3(MT), 0, st, B.skey, B.sid, B.pr,
                               B.vm, B.cs, B.ce, B.cnt, (int)d, timestep, (int4 *)nullptr, 0ull,
                               B.tot + 1, B.tot);
        NCG_CHECK_LAUNCH();
    }
    if (M.want_lists) {
        // The exact field's lists depend on the binning only: built on the
        // second stream while the replay (one workgroup) runs here.
        int hins = 0;
        NCG_HIP(hipMemcpyAsync(&hins, B.cnt, sizeof(int), hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        if (hins > 0) {
            r = ensure_aux(ctx);
            if (r != NCG_OK) return r;
            NCG_HIP(hipEventRecord(ctx->ev_fork, st));
            NCG_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
            r = build_ord_lists(ctx, hins, d, B.skey, B.cs, B.ce, B.sid, ctx->aux);
            if (r != NCG_OK) return r;
            M.ord_lists = true;
        }
    }
    hipLaunchKernelGGL(k_lf_fill_u32, dim3(1), dim3(64), 0, st, stat, 2u);
    if (mode == 0) {
        // Candidates in visiting order, then the replay.
        unsigned long long *kin = dbuf<unsigned long long>(ctx, "m_ckey", cap);
        unsigned long long *kout = dbuf<unsigned long long>(ctx, "m_ckey2", cap);
        unsigned *vin = dbuf<unsigned>(ctx, "m_cval", cap);
        unsigned *vout = dbuf<unsigned>(ctx, "m_cval2", cap);
        size_t tb = 0;
        NCG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, (int)cap, 0, 64, st));
        void *tmp = ncg_dbuf(ctx, "m_csorttmp", tb);
        if (!kin || !kout || !vin || !vout || !tmp) return NCG_ERR_NOMEM;
        ProfScope ps(ctx, st, "mesh_replay");
        hipLaunchKernelGGL(k_mcand_keys, dim3(cdiv(cap, MT)), dim3(MT), 0, st, dc, B.tot + 1, cap, kin,
                           vin);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, (int)cap, 0, 64, st));
        int4 *csorted = dbuf<int4>(ctx, "m_csorted", cap);
        if (!csorted) return NCG_ERR_NOMEM;
        hipLaunchKernelGGL(k_mcand_gather, dim3(cdiv(cap, MT)), dim3(MT), 0, st, dc, vout, B.tot + 1,
                           cap, csorted);
        // Clusters replayed in parallel, from a snapshot of the bodies and
        // did_collide that a conflict between clusters restores.
        int4 *cscl = dbuf<int4>(ctx, "m_cscl", cap);
        int *coff = dbuf<int>(ctx, "m_coff", MCL_MAXM + 2);
        int *ncl = dbuf<int>(ctx, "m_ncl", 1);
        unsigned long long *ctab = dbuf<unsigned long long>(ctx, "m_ctab", MCL_TAB);
        unsigned *confl = dbuf<unsigned>(ctx, "m_confl", 1);
        DBody *snapb = dbuf<DBody>(ctx, "m_snapb", n);
        int32_t *snapdc = d_did_collide ? dbuf<int32_t>(ctx, "m_snapdc", n) : nullptr;
        if (!cscl || !coff || !ncl || !ctab || !confl || !snapb || (d_did_collide && !snapdc))
            return NCG_ERR_NOMEM;
        hipLaunchKernelGGL(k_mcluster, dim3(1), dim3(MCL_T), 0, st, csorted, B.tot + 1, cap, B.key,
                           (int)d, ctx->replay_cluster_dist, cscl, coff, ncl);
        hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(MCL_TAB, MT)), dim3(MT), 0, st, ctab, MCL_TAB);
        hipLaunchKernelGGL(k_lf_fill_u32, dim3(1), dim3(64), 0, st, confl, 1u);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipMemcpyAsync(snapb, db, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToDevice, st));
        if (d_did_collide)
            NCG_HIP(hipMemcpyAsync(snapdc, d_did_collide, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_mesh_replay, dim3(MCL_MAXM), dim3(MR_T), 0, st, db, d_did_collide,
                           (int)one_collision_only, collision_f, timestep, B.key, B.sid, B.cs, B.ce,
                           B.pr, B.vm, (int)d, cscl, B.tot + 1, cap, B.cnt, stat, coff, ncl, ctab,
                           confl);
        NCG_CHECK_LAUNCH();
        unsigned hconf = 0u;
        NCG_HIP(hipMemcpyAsync(hs + 10, confl, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        NCG_HIP(hipStreamSynchronize(st));
        hconf = hs[10];
        if (hconf) {
            // Two clusters met: the one-workgroup replay from the snapshot.
            NCG_HIP(hipMemcpyAsync(db, snapb, sizeof(DBody) * (size_t)n, hipMemcpyDeviceToDevice, st));
            if (d_did_collide)
                NCG_HIP(hipMemcpyAsync(d_did_collide, snapdc, sizeof(int32_t) * (size_t)n,
                                       hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_lf_fill_u32, dim3(1), dim3(64), 0, st, stat, 2u);
            hipLaunchKernelGGL(k_mesh_replay, dim3(1), dim3(MR_T), 0, st, db, d_did_collide,
                               (int)one_collision_only, collision_f, timestep, B.key, B.sid, B.cs,
                               B.ce, B.pr, B.vm, (int)d, csorted, B.tot + 1, cap, B.cnt, stat,
                               (const int *)nullptr, (const int *)nullptr,
                               (unsigned long long *)nullptr, (unsigned *)nullptr);
            NCG_CHECK_LAUNCH();
        }
        stats->replay_conflicts = hconf ? 1 : 0;
    }
    // Counters: inserted, small, big, big-list length; neighbours,
    // candidates; collisions, error bits.
    NCG_HIP(hipMemcpyAsync(hs, B.cnt, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(hs + 4, B.tot, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(hs + 8, stat, sizeof(unsigned) * 2, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    const int *hc4 = reinterpret_cast<const int *>(hs);
    unsigned long long ht[2];
    memcpy(ht, hs + 4, sizeof ht);
    const int ins = hc4[0];
    M.prev_ins = ins;
    M.n = n;
    M.cells = cells;
    M.mode = mode;
    M.inserted = ins;
    M.gam = gam;
    M.cube = collision_cube;
    stats->total_neighbours = (int64_t)ht[0];
    stats->how_many_small = mode == 0 ? hc4[1] : 0;
    stats->how_many_big = mode == 0 ? hc4[2] : 0;
    stats->candidates = (int64_t)ht[1];
    stats->collisions = hs[8];
    *flags = hs[9];
    if (hs[9]) return NCG_OK;
    M.valid = true;
    M.pairs_known = mode == 1;
    return NCG_OK;
}

extern "C" ncg_status ncg_mesh_pass_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                        const int32_t *d_exists, int32_t *d_did_collide,
                                        double timestep, double gam, double collision_cube,
                                        int32_t cells, int32_t mode, double collision_f,
                                        int32_t one_collision_only, ncg_mesh_stats *stats,
                                        void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_HIP(hipSetDevice(ctx->device));
    unsigned flags = 0u;
    ncg_status r = ncg_mesh_pass_dev_flags(ctx, d_bodies, n, d_exists, d_did_collide, timestep, gam,
                                           collision_cube, cells, mode, collision_f,
                                           one_collision_only, stats, &flags,
                                           pick_stream(ctx, stream));
    if (r != NCG_OK) return r;
    if (flags & 4u)
        return ncg_fail(ctx, NCG_ERR_INVALID,
                        "ncg_mesh_pass_dev: %lld bodies are larger than the mesh size (gam = %g); "
                        "use ncg_mesh_pass", (long long)stats->how_many_big, gam);
    if (flags)
        return ncg_fail(ctx, NCG_ERR_CAPACITY,
                        "ncg_mesh_pass_dev: replay capacity exceeded (flags %u: 1 = more than %u "
                        "candidates, 2 = more than %d pending re-tests); the state was not changed "
                        "only if flag 1 is set alone",
                        flags, ncg_replay_cap(ctx, 1u << 16), MR_PCAP);
    return NCG_OK;
}

// Id-ordered neighbourhood lists of the binning (k_usize, k_umerge), on
// stream st: L (27 entries per binned body), lw, le.
static ncg_status build_ord_lists(ncg_ctx *ctx, int ins, long d, const unsigned *skey,
                                  const unsigned *cs, const unsigned *ce, const int *sid,
                                  hipStream_t st)
{
    NCG_REQUIRE(27.0 * ins < 4294967295.0, "%d binned bodies: neighbourhood lists too long", ins);
    const size_t lcap = std::max<size_t>(27 * (size_t)ins, 1);
    unsigned *L = dbuf<unsigned>(ctx, "g_L", lcap);
    unsigned *lw = dbuf<unsigned>(ctx, "g_lw", std::max(ins, 1));
    unsigned *le = dbuf<unsigned>(ctx, "g_le", std::max(ins, 1));
    unsigned *ukey = dbuf<unsigned>(ctx, "g_ukey", ins);
    unsigned *ucnt = dbuf<unsigned>(ctx, "g_ucnt", ins);
    int *nu = dbuf<int>(ctx, "g_nu", 1);
    unsigned *usz = dbuf<unsigned>(ctx, "g_usz", (size_t)ins + 1);
    unsigned *uoff = dbuf<unsigned>(ctx, "g_uoff", (size_t)ins + 1);
    if (!L || !lw || !le || !ukey || !ucnt || !nu || !usz || !uoff) return NCG_ERR_NOMEM;
    size_t b1 = 0, b2 = 0;
    NCG_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, b1, skey, ukey, ucnt, nu, ins, st));
    NCG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, usz, uoff, ins + 1, st));
    void *tmp = ncg_dbuf(ctx, "g_utmp", std::max(b1, b2));
    if (!tmp) return NCG_ERR_NOMEM;
    ProfScope pl(ctx, st, "mesh_lists");
    NCG_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp, b1, skey, ukey, ucnt, nu, ins, st));
    hipLaunchKernelGGL(k_usize, dim3(cdiv(ins + 1, MT)), dim3(MT), 0, st, ukey, nu, ins, cs, ce,
                       (int)d, usz);
    NCG_CHECK_LAUNCH();
    NCG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b2, usz, uoff, ins + 1, st));
    hipLaunchKernelGGL(k_umerge, dim3(cdiv(ins, MT)), dim3(MT), 0, st, ukey, nu, uoff, cs, ce, sid,
                       (int)d, L, lw, le);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}

static ncg_status ensure_aux(ncg_ctx *ctx)
{
    if (ctx->aux) return NCG_OK;
    NCG_HIP(hipStreamCreate(&ctx->aux));
    NCG_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    NCG_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    return NCG_OK;
}

// Mesh field of the bodies db (device), central body dcen (8-double record
// {x, y, z, vx, vy, vz, M, present} or null), over the binning of the last
// pass, into dacc (3 per slot) and dcm (CM_acc).
// dinit (3n, or null): start values in place of the central term (dcen is
// then null).
static ncg_status mesh_gravity_core(ncg_ctx *ctx, const DBody *db, int32_t n, const int32_t *dex,
                                    const double *dcen, double G, double s, ncg_precision prec,
                                    double *dacc, double *dcm, hipStream_t st,
                                    const double *dinit = nullptr)
{
    MeshState &M = ctx->mesh;
    const long d = M.cells;
    const int ins = M.inserted;
    double *dct = dbuf<double>(ctx, "g_cterm", 3 * (size_t)n);
    double4 *pm = dbuf<double4>(ctx, "g_pm", std::max(ins, 1));
    double *rr = dbuf<double>(ctx, "g_rr", std::max(ins, 1));
    int *sx = dbuf<int>(ctx, "g_sx", std::max(ins, 1));
    const int nbp = std::max(1, std::min(512, (int)cdiv(n, MT)));
    double *bpart = dbuf<double>(ctx, "g_bpart", 12 * (size_t)nbp);
    const size_t nex = std::max<size_t>(M.extra_pairs.size(), 1);
    int2 *dextra = dbuf<int2>(ctx, "g_extra", nex);
    if (!dct || !pm || !rr || !sx || !bpart || !dextra) return NCG_ERR_NOMEM;
    // The pass's arrays cover its M.n slots; bodies appended during the pass
    // (slots M.n .. n - 1) are not binned.
    const int32_t nkey = M.n;
    const unsigned *key = dbuf<unsigned>(ctx, "m_key", nkey);
    const int *sid = dbuf<int>(ctx, "m_sid", nkey);
    const unsigned *cs = dbuf<unsigned>(ctx, "m_cs", M.table_keys);
    const unsigned *ce = dbuf<unsigned>(ctx, "m_ce", M.table_keys);
    if (!M.extra_pairs.empty())
        NCG_HIP(hipMemcpyAsync(dextra, M.extra_pairs.data(), sizeof(int2) * M.extra_pairs.size(),
                               hipMemcpyHostToDevice, st));
    if (ins > 0) {
        hipLaunchKernelGGL(k_gsorted, dim3(cdiv(ins, MT)), dim3(MT), 0, st, db, dex, sid, ins, pm,
                           rr, sx);
        NCG_CHECK_LAUNCH();
    }
    if (prec == NCG_F64_EXACT) {
        // Pair-list order, bit for bit: recorded big visits as (visiting id,
        // ascending partner ids); the per-body walks on a second stream.
        std::vector<int> hbk, hbo(1, 0), hbp;
        // Only the visitor's own pairs (ascending partner ids); the bodies of
        // fragmentation blocks recorded there have explicit term lists.
        for (const BigVisit &bv : M.big_visits) {
            hbk.push_back(bv.k);
            for (int q = 0; q < bv.count; ++q) {
                if (!M.big_scan[bv.first + q]) continue;
                const ncg_pair &pq = M.big_pairs_all[bv.first + q];
                hbp.push_back(pq.fst == bv.k ? pq.snd : pq.fst);
            }
            hbo.push_back((int)hbp.size());
        }
        const int nbv = (int)hbk.size();
        int *bvk = dbuf<int>(ctx, "g_bvk", std::max(nbv, 1));
        int *bvoff = dbuf<int>(ctx, "g_bvoff", nbv + 1);
        int *bvo = dbuf<int>(ctx, "g_bvo", std::max<size_t>(hbp.size(), 1));
        if (!bvk || !bvoff || !bvo) return NCG_ERR_NOMEM;
        if (nbv) {
            NCG_HIP(hipMemcpyAsync(bvk, hbk.data(), sizeof(int) * nbv, hipMemcpyHostToDevice, st));
            NCG_HIP(hipMemcpyAsync(bvoff, hbo.data(), sizeof(int) * (nbv + 1), hipMemcpyHostToDevice, st));
            if (!hbp.empty())
                NCG_HIP(hipMemcpyAsync(bvo, hbp.data(), sizeof(int) * hbp.size(),
                                       hipMemcpyHostToDevice, st));
        }
        ncg_status ra = ensure_aux(ctx);
        if (ra != NCG_OK) return ra;
        // Id-ordered neighbourhood lists, once per pass (every gravity call
        // of the pass shares them).
        const unsigned *skey = dbuf<unsigned>(ctx, "m_skey", nkey);
        NCG_REQUIRE(27.0 * ins < 4294967295.0, "%d binned bodies: neighbourhood lists too long", ins);
        const size_t lcap = std::max<size_t>(27 * (size_t)ins, 1);
        unsigned *L = dbuf<unsigned>(ctx, "g_L", lcap);
        unsigned *lw = dbuf<unsigned>(ctx, "g_lw", std::max(ins, 1));
        unsigned *le = dbuf<unsigned>(ctx, "g_le", std::max(ins, 1));
        if (!L || !lw || !le) return NCG_ERR_NOMEM;
        if (!M.ord_lists && ins > 0) {
            ncg_status rl = build_ord_lists(ctx, ins, d, skey, cs, ce, sid, st);
            if (rl != NCG_OK) return rl;
            M.ord_lists = true;
        }
        // Terms of the body-order sums; the sums themselves stay on the
        // caller's stream and the per-body walks go to the second stream.
        // Dispatched right behind k_mterms in the same queue, the single
        // latency-bound k_mseq workgroup is placed before the thousands of
        // k_mord workgroups; placed after them it runs about twice as long.
        const int nch = cdiv(n, SQ_CH);
        double *T = dbuf<double>(ctx, "g_seqT", (size_t)nch * 12 * SQ_CH);
        const int ntg = (int)M.x_targets.size() / 3;
        int *tg = dbuf<int>(ctx, "g_xtg", std::max<size_t>(M.x_targets.size(), 1));
        int *pl = dbuf<int>(ctx, "g_xpl", std::max<size_t>(M.x_partners.size(), 1));
        if (!T || !tg || !pl) return NCG_ERR_NOMEM;
        if (ntg) {
            NCG_HIP(hipMemcpyAsync(tg, M.x_targets.data(), sizeof(int) * M.x_targets.size(),
                                   hipMemcpyHostToDevice, st));
            NCG_HIP(hipMemcpyAsync(pl, M.x_partners.data(), sizeof(int) * M.x_partners.size(),
                                   hipMemcpyHostToDevice, st));
        }
        ProfScope ps(ctx, st, "mesh_gravity");
        hipLaunchKernelGGL(k_mterms, dim3(cdiv(nch * SQ_CH, MT)), dim3(MT), 0, st, db, (int)n, dex,
                           dcen, G, s, nch * SQ_CH, T);
        NCG_CHECK_LAUNCH();
        NCG_HIP(hipEventRecord(ctx->ev_fork, st));
        NCG_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
        hipLaunchKernelGGL(k_mseq, dim3(1), dim3(256), 0, st, db, (int)n, dex, dcen, G, s,
                           reinterpret_cast<const double2 *>(T), nch, dacc, dcm, dinit);
        NCG_CHECK_LAUNCH();
        hipStream_t ax = ctx->aux;
        if (ins > 0)
            hipLaunchKernelGGL(k_mord, dim3(cdiv(ins, MT)), dim3(MT), 0, ax, db, (int)n, dex, ins,
                               cs, ce, pm, rr, sx, sid, skey, (int)d, dcen, G, s,
                               bvk, bvoff, bvo, nbv, L, lw, le, dacc, dinit);
        if (n > 3)
            hipLaunchKernelGGL(k_mrest_ord, dim3(cdiv(n - 3, MT)), dim3(MT), 0, ax, db, (int)n, dex,
                               key, (int)nkey, dcen, G, s, bvk, bvoff, bvo, nbv, dacc, dinit);
        NCG_CHECK_LAUNCH();
        // Survivors and fragments of full fragmentations: host-listed order.
        if (ntg) {
            hipLaunchKernelGGL(k_mexplicit, dim3(cdiv(ntg, MT)), dim3(MT), 0, ax, db, dex, dcen, G, s,
                               (const int *)tg, ntg, (const int *)pl, dacc, dinit);
            NCG_CHECK_LAUNCH();
        }
        NCG_HIP(hipEventRecord(ctx->ev_join, ax));
        NCG_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
    } else {
        ProfScope ps(ctx, st, "mesh_gravity");
        const unsigned *skey = dbuf<unsigned>(ctx, "m_skey", nkey);
        double4 *pw = dbuf<double4>(ctx, "g_pw", std::max(ins, 1));
        if (!pw) return NCG_ERR_NOMEM;
        const dim3 gs(cdiv(std::max(ins, 1), MT));
        if (ins > 0)
            hipLaunchKernelGGL(k_gweights, gs, dim3(MT), 0, st, (const double4 *)pm, (const int *)sx,
                               ins, G, pw);
        const bool soft = s != 0.0;
        // NCG_MIXED runs this FP64 form too: every neighbour term is a near
        // one, and FP32 terms lost ~1e-5 to the cancellation of the sums (2.3e-5
        // at 2^20 bodies) for 7 % less time.
        if (soft)
            hipLaunchKernelGGL((k_mgrav<true>), gs, dim3(MT), 0, st, db, (int)n, dex, ins, cs, ce,
                               pw, rr, sx, sid, skey, (int)d, dcen, G, s, dacc, dct);
        else
            hipLaunchKernelGGL((k_mgrav<false>), gs, dim3(MT), 0, st, db, (int)n, dex, ins, cs, ce,
                               pw, rr, sx, sid, skey, (int)d, dcen, G, s, dacc, dct);
        hipLaunchKernelGGL(k_mgrav_rest, dim3(cdiv(n, MT)), dim3(MT), 0, st, db, (int)n, dex, key,
                           (int)nkey, dcen,
                           G, s, dacc, dct);
        NCG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_big_partial, dim3(nbp), dim3(MT), 0, st, db, (int)n, dex, G, s,
                           (const double *)dct, dcen, bpart);
        hipLaunchKernelGGL(k_big_final, dim3(1), dim3(MT), 0, st, db, (int)n, dex, bpart, nbp, G, s,
                           dacc, dcen, dcm);
        if (!M.extra_pairs.empty())
            hipLaunchKernelGGL(k_extra_pairs, dim3(1), dim3(1), 0, st, db, dex, dextra,
                               (int)M.extra_pairs.size(), G, s, dacc);
        NCG_CHECK_LAUNCH();
        // The fast forms add the start values after the mutual sums.
        if (dinit) {
            hipLaunchKernelGGL(k_add_start, dim3(cdiv(n, MT)), dim3(MT), 0, st, n, dex, dinit, dacc);
            NCG_CHECK_LAUNCH();
        }
    }
    return NCG_OK;
}

// ncg_mesh_gravity and ncg_mesh_gravity_from (start: 3n start values, or null).
static ncg_status mesh_gravity_host(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                    const int32_t *exists, const ncg_central *central,
                                    const double *start, double grav_const, double softening,
                                    ncg_precision prec, double *acc, double *cm_acc)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (bodies && acc)), "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    MeshState &M = ctx->mesh;
    if (!M.valid || M.n_end != n)
        return ncg_fail(ctx, NCG_ERR_INVALID,
                        "ncg_mesh_gravity needs a ncg_mesh_pass over the same %d body slots "
                        "(those of the pass and the bodies it appended)", n);
    if (n == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int cp = (central && central->present) ? 1 : 0;
    DBody *db = dbuf<DBody>(ctx, "g_b", n);
    int32_t *dex = exists ? dbuf<int32_t>(ctx, "g_ex", n) : nullptr;
    double *dacc = dbuf<double>(ctx, "g_acc", 3 * (size_t)n);
    double *dcm = dbuf<double>(ctx, "g_cm", 4);
    double *dcen = dbuf<double>(ctx, "g_cen", 8);
    double *hcen = pbuf<double>(ctx, "g_hcen", 8);
    double *dinit = start ? dbuf<double>(ctx, "g_start", 3 * (size_t)n) : nullptr;
    if (!db || (exists && !dex) || !dacc || !dcm || !dcen || !hcen || (start && !dinit))
        return NCG_ERR_NOMEM;
    NCG_HIP(hipMemcpyAsync(db, bodies, sizeof(DBody) * (size_t)n, hipMemcpyHostToDevice, st));
    if (start)
        NCG_HIP(hipMemcpyAsync(dinit, start, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice,
                               st));
    if (exists)
        NCG_HIP(hipMemcpyAsync(dex, exists, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    if (cp) {
        // Central body as the device record the kernels read (through pinned
        // memory; every call ends with a synchronisation, so it is free).
        const double rec[8] = {central->x, central->y, central->z, 0.0, 0.0, 0.0, central->mass, 1.0};
        for (int q = 0; q < 8; ++q) hcen[q] = rec[q];
        NCG_HIP(hipMemcpyAsync(dcen, hcen, sizeof(double) * 8, hipMemcpyHostToDevice, st));
    }
    ncg_status r = mesh_gravity_core(ctx, db, n, dex, cp ? dcen : nullptr, grav_const, softening,
                                     prec, dacc, dcm, st, dinit);
    if (r != NCG_OK) return r;
    double *hacc = hbuf<double>(ctx, "macc", 3 * (size_t)n);
    double hcm[4] = {0, 0, 0, 0};
    NCG_HIP(hipMemcpyAsync(hacc, dacc, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, st));
    if (cp) NCG_HIP(hipMemcpyAsync(hcm, dcm, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    for (int32_t k = 0; k < n; ++k) {
        if (exists && !exists[k]) continue;
        acc[3 * (size_t)k] = hacc[3 * (size_t)k];
        acc[3 * (size_t)k + 1] = hacc[3 * (size_t)k + 1];
        acc[3 * (size_t)k + 2] = hacc[3 * (size_t)k + 2];
    }
    if (cp && cm_acc) {
        cm_acc[0] = hcm[0];
        cm_acc[1] = hcm[1];
        cm_acc[2] = hcm[2];
    }
    return NCG_OK;
}

extern "C" ncg_status ncg_mesh_gravity(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                       const int32_t *exists, const ncg_central *central,
                                       double grav_const, double softening, ncg_precision prec,
                                       double *acc, double *cm_acc)
{
    return mesh_gravity_host(ctx, bodies, n, exists, central, nullptr, grav_const, softening, prec,
                             acc, cm_acc);
}

extern "C" ncg_status ncg_mesh_gravity_from(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                            const int32_t *exists, const double *start,
                                            double grav_const, double softening,
                                            ncg_precision prec, double *acc)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n == 0 || start, "invalid arguments");
    return mesh_gravity_host(ctx, bodies, n, exists, nullptr, start, grav_const, softening, prec,
                             acc, nullptr);
}

extern "C" ncg_status ncg_mesh_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                           const int32_t *d_exists, const double *d_central,
                                           double grav_const, double softening, ncg_precision prec,
                                           double *d_acc, double *d_cm_acc, void *stream)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(n >= 0 && (n == 0 || (d_bodies && d_acc)), "invalid arguments");
    NCG_REQUIRE(prec == NCG_F64_EXACT || prec == NCG_F64_FAST || prec == NCG_MIXED,
                "unknown precision %d", (int)prec);
    MeshState &M = ctx->mesh;
    if (!M.valid || M.n_end != n)
        return ncg_fail(ctx, NCG_ERR_INVALID,
                        "ncg_mesh_gravity_dev needs a mesh pass over the same %d body slots", n);
    if (n == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    double *dcm = d_cm_acc ? d_cm_acc : dbuf<double>(ctx, "g_cm", 4);
    if (!dcm) return NCG_ERR_NOMEM;
    return mesh_gravity_core(ctx, reinterpret_cast<const DBody *>(d_bodies), n, d_exists,
                             d_central, grav_const, softening, prec, d_acc, dcm, st);
}

extern "C" ncg_status ncg_mesh_pairs(ncg_ctx *ctx, ncg_pair *out, int64_t capacity, int64_t *count)
{
    if (!ctx) return NCG_ERR_INVALID;
    NCG_REQUIRE(count && capacity >= 0 && (capacity == 0 || out), "invalid arguments");
    MeshState &M = ctx->mesh;
    NCG_REQUIRE(M.valid, "no mesh pass");
    NCG_REQUIRE(M.pairs_known, "the pair list of a device-replayed mesh() pass is not kept");
    *count = 0;
    const int32_t n = M.n;
    if (n == 0) return NCG_OK;
    NCG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int ins = M.inserted;
    const long d = M.cells;
    // Per-body pair counts: small bodies from the GPU count, big bodies from
    // the host visits; offsets in visiting order.
    std::vector<int> cnt(n, 0);
    std::vector<unsigned> hkey(n);
    const int *cbi = dbuf<int>(ctx, "m_count", n);
    const unsigned *key = dbuf<unsigned>(ctx, "m_key", n);
    NCG_HIP(hipMemcpyAsync(cnt.data(), cbi, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipMemcpyAsync(hkey.data(), key, sizeof(unsigned) * n, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    for (int32_t k = 0; k < n; ++k)
        if (hkey[k] == NOKEY) cnt[k] = 0;
    for (const auto &bv : M.big_visits) cnt[bv.k] = bv.count;
    std::vector<long long> off(n);
    long long tot = 0;
    for (int32_t k = 0; k < n; ++k) {
        off[k] = tot;
        tot += cnt[k];
    }
    // Registered pairs: the gathered entries less those of bodies removed
    // before the visit, each collision's entry replaced by what its outcome
    // registered (physics.c:1043-1091).
    long long fin = tot - M.rm_drop;
    bool simple = M.rm_drop == 0;
    for (const MeshEvent &e : M.events) {
        if (e.l < 0) continue;  // big visits: recorded as registered
        if (e.kind == MeshEvent::NONE) --fin;
        if (e.kind == MeshEvent::BLOCK) fin += (long long)e.count * (e.count - 1) / 2 - 1;
        simple &= e.kind == MeshEvent::PAIR;
    }
    *count = fin;
    if (fin > capacity)
        return ncg_fail(ctx, NCG_ERR_CAPACITY, "%lld pairs exceed capacity %lld", fin,
                        (long long)capacity);
    if (tot == 0) return NCG_OK;
    long long *doff = dbuf<long long>(ctx, "p_off", n);
    int2 *dout = dbuf<int2>(ctx, "p_out", (size_t)tot);
    if (!doff || !dout) return NCG_ERR_NOMEM;
    std::vector<ncg_pair> all;
    if (!simple) all.resize((size_t)tot);
    ncg_pair *gl = simple ? out : all.data();
    NCG_HIP(hipMemcpyAsync(doff, off.data(), sizeof(long long) * n, hipMemcpyHostToDevice, st));
    if (ins > 0) {
        hipLaunchKernelGGL(k_mpairs, dim3(cdiv(ins, MT)), dim3(MT), 0, st,
                           (const unsigned *)dbuf<unsigned>(ctx, "m_skey", n),
                           (const int *)dbuf<int>(ctx, "m_sid", n),
                           (const unsigned *)dbuf<unsigned>(ctx, "m_cs", M.table_keys),
                           (const unsigned *)dbuf<unsigned>(ctx, "m_ce", M.table_keys), ins,
                           (int)d, doff, dout);
        NCG_CHECK_LAUNCH();
    }
    NCG_HIP(hipMemcpyAsync(gl, dout, sizeof(int2) * (size_t)tot, hipMemcpyDeviceToHost, st));
    NCG_HIP(hipStreamSynchronize(st));
    for (const auto &bv : M.big_visits)
        for (int q = 0; q < bv.count; ++q) gl[off[bv.k] + q] = M.big_pairs_all[bv.first + q];
    // The entry of each collision of a small visit: the occ-th entry of its
    // partner in the visit.
    std::map<long long, const MeshEvent *> at;
    for (const MeshEvent &e : M.events) {
        if (e.l < 0) continue;
        int seen = 0;
        long long pos = -1;
        for (long long q = off[e.k]; q < off[e.k] + cnt[e.k]; ++q)
            if (gl[q].snd == e.p && seen++ == e.occ) {
                pos = q;
                break;
            }
        if (pos < 0)
            return ncg_fail(ctx, NCG_ERR_INVALID, "collision (%d, %d) not among the pairs of visit %d",
                            e.p, e.k, e.k);
        at[pos] = &e;
    }
    if (simple) {
        for (const auto &pe : at) out[pe.first] = pe.second->reg;
        return NCG_OK;
    }
    // Removal visit of every id a registered pair can hold: the pass's slots
    // and the ones appended during it (big visits pair with those).
    const int32_t nall = std::max(n, M.n_end);
    std::vector<int> rmv((size_t)nall, INT32_MAX);
    for (const auto &rv : M.removed)
        if (rv.first >= 0 && rv.first < nall) rmv[rv.first] = rv.second;
    long long o = 0;
    for (int32_t k = 0; k < n; ++k)
        for (long long q = off[k]; q < off[k] + cnt[k]; ++q) {
            auto it = at.find(q);
            if (it != at.end()) {
                const MeshEvent &e = *it->second;
                if (e.kind == MeshEvent::PAIR) out[o++] = e.reg;
                if (e.kind == MeshEvent::BLOCK)
                    for (int a = 1; a < e.count; ++a)
                        for (int b = 0; b < a; ++b)
                            out[o++] = {M.frag_ids[e.first + a], M.frag_ids[e.first + b]};
                continue;
            }
            const int ps = gl[q].snd;
            if (ps >= 0 && ps < nall && rmv[ps] < k) continue;  // removed before visit k gathered
            out[o++] = gl[q];
        }
    if (o != fin)
        return ncg_fail(ctx, NCG_ERR_INVALID, "pair list of %lld entries, %lld expected", o, fin);
    return NCG_OK;
}
