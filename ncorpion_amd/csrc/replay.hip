// Device replay of brute_force's collision pass (collision.c:425-456) for
// callers that keep the bodies on the GPU: the ensemble driver (one system
// per wavefront) and the GPU-resident leapfrog (one system).
//
// A GPU pass has listed candidate pairs (i < j) whose closest approach may
// fall within dt.  The replay decides them in the reference's order -- rows i
// ascending, j > i ascending -- with closest_approach (collision.c:63-172)
// and the elastic / inelastic bounce of collision (physics.c:545-605):
//   - a pair is tested when the pass listed it or when one of its bodies
//     changed earlier in this pass;
//   - a row whose body changed is scanned in full, 64 pairs per step, and the
//     first hit in j order wins (the tests after it are redone with the new
//     state);
//   - rows with nothing to do are found 64 at a time: a row whose tests all
//     miss changes nothing, so every row before the first row with a hit (or
//     a changed body) is skipped.
// did_collide (collision.c:438, physics.c:1063-1064) and holes are honoured
// when given.  This translation unit is compiled without FMA contraction:
// the arithmetic is the reference's, operation by operation.
#include "ncg_internal.hpp"
#include "collide_dev.hpp"

#pragma clang fp contract(off)

namespace {

// One wavefront per system (blockIdx.x).  err bit 0: more candidates than
// `cap`, bit 1: more than RCAP for one system, bit 2: more than DCAP bodies
// changed in one system.
__global__ __launch_bounds__(64) void k_bf_replay(DBody *__restrict__ y0, const int32_t *__restrict__ ex0,
                                                  int32_t *__restrict__ dc0, int one_coll,
                                                  const int *__restrict__ off,
                                                  const int2 *__restrict__ cand,
                                                  const unsigned *__restrict__ ccnt32,
                                                  const unsigned long long *__restrict__ ccnt64,
                                                  unsigned cap, double dt, double f,
                                                  unsigned char *__restrict__ dirty,
                                                  unsigned *__restrict__ ncoll,
                                                  unsigned *__restrict__ err)
{
    __shared__ int2 lc[NCG_REPLAY_RCAP];
    __shared__ int dl[NCG_REPLAY_DCAP];
    __shared__ int s_nc, s_nd, s_row, s_from, s_hit;
    __shared__ double s_ap[6], s_tc;
    const int lane = threadIdx.x;
    const int sys = blockIdx.x;
    const int base = off[sys], ns = off[sys + 1] - base;
    DBody *b = y0 + base;
    const int32_t *ex = ex0 ? ex0 + base : nullptr;
    int32_t *dc = dc0 ? dc0 + base : nullptr;
    unsigned char *dt8 = dirty + base;
    const unsigned long long total = ccnt64 ? *ccnt64 : (unsigned long long)*ccnt32;
    if (lane == 0) {
        s_nc = 0;
        s_nd = 0;
        if (total > cap) atomicOr(err, 1u);
    }
    __syncthreads();
    const unsigned m = (unsigned)(total < cap ? total : cap);
    for (unsigned k = lane; k < m; k += 64) {
        const int2 c = cand[k];
        if (c.x >= base && c.x < base + ns) {
            const int slot = atomicAdd(&s_nc, 1);
            if (slot < NCG_REPLAY_RCAP) lc[slot] = make_int2(c.x - base, c.y - base);
        }
    }
    __syncthreads();
    if (s_nc == 0) return;
    if (lane == 0) {
        if (s_nc > NCG_REPLAY_RCAP) {
            atomicOr(err, 2u);
            s_nc = NCG_REPLAY_RCAP;
        }
        // (i, j) order; lists are short.
        for (int a = 1; a < s_nc; ++a) {
            const int2 v = lc[a];
            int q = a - 1;
            while (q >= 0 && (lc[q].x > v.x || (lc[q].x == v.x && lc[q].y > v.y))) {
                lc[q + 1] = lc[q];
                --q;
            }
            lc[q + 1] = v;
        }
    }
    __syncthreads();
    unsigned hits = 0;
    auto mark = [&](int q) {  // lane 0 only
        if (dt8[q]) return;
        dt8[q] = 1;
        if (s_nd >= NCG_REPLAY_DCAP) {
            atomicOr(err, 4u);
            return;
        }
        int p = s_nd++;
        while (p > 0 && dl[p - 1] > q) {
            dl[p] = dl[p - 1];
            --p;
        }
        dl[p] = q;
    };
    auto resolve = [&](int i, int j, const double *ap, double tc) {  // lane 0 only
        bf_bounce(b, i, j, f, ap, tc);
        if (dc) {
            dc[i] = one_coll;
            dc[j] = one_coll;
        }
        ++hits;
    };
    // First candidate of row r (candidates are in (i, j) order).
    auto first_cand = [&](int r) {
        int lo = 0, hi = s_nc;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (lc[mid].x < r) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    int cur = 0;  // every row below cur is done
    while (true) {
        // Rows past the largest changed body only have candidate pairs.
        if (lane == 0) {
            const int maxd = s_nd ? dl[s_nd - 1] : -1;
            int start = cur;
            if (start > maxd) {
                const int c = first_cand(start);
                start = c < s_nc ? lc[c].x : ns;
            }
            s_row = start < ns ? start : -1;
        }
        __syncthreads();
        const int start = s_row;
        if (start < 0) break;
        // 64 rows at once: does row r have anything to do?
        const int r = start + lane;
        bool ev = false;
        if (r < ns && (!ex || ex[r])) {
            if (dt8[r]) {
                ev = true;
            } else {
                double ap[6], tc;
                for (int q = 0; q < s_nd && !ev; ++q)
                    if (dl[q] > r) ev = bf_hit(b, ex, dc, r, dl[q], dt, ap, &tc);
                for (int c = first_cand(r); !ev && c < s_nc && lc[c].x == r; ++c)
                    ev = bf_hit(b, ex, dc, r, lc[c].y, dt, ap, &tc);
            }
        }
        const unsigned long long emask = __ballot(ev);
        if (emask == 0ull) {
            cur = start + 64;
            __syncthreads();
            continue;
        }
        const int row = start + __ffsll((long long)emask) - 1;
        // That row in the reference's order, on lane 0 until its first hit.
        if (lane == 0) {
            s_from = -1;
            if (dt8[row]) {
                s_from = row + 1;  // changed row: scan it in full
            } else {
                int ci = first_cand(row);
                int jc = row + 1;
                while (jc < ns) {
                    while (ci < s_nc && lc[ci].x == row && lc[ci].y < jc) ++ci;
                    const int jcand = (ci < s_nc && lc[ci].x == row) ? lc[ci].y : ns;
                    int jdirty = ns;
                    for (int q = 0; q < s_nd; ++q)
                        if (dl[q] >= jc) {
                            jdirty = dl[q];
                            break;
                        }
                    const int j = min(jcand, jdirty);
                    if (j >= ns) break;
                    double ap[6], tc;
                    if (bf_hit(b, ex, dc, row, j, dt, ap, &tc)) {
                        resolve(row, j, ap, tc);
                        mark(row);
                        mark(j);
                        s_from = j + 1;  // row changed: the rest of it in full
                        break;
                    }
                    jc = j + 1;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        int jc = s_from;
        while (jc >= 0 && jc < ns) {
            const int j = jc + lane;
            double ap[6], tc;
            const bool hit = j < ns && bf_hit(b, ex, dc, row, j, dt, ap, &tc);
            const unsigned long long mask = __ballot(hit);
            if (mask == 0ull) {
                jc += 64;
                continue;
            }
            const int first = __ffsll((long long)mask) - 1;
            if (lane == first) {
                for (int q = 0; q < 6; ++q) s_ap[q] = ap[q];
                s_tc = tc;
                s_hit = j;
            }
            __syncthreads();
            if (lane == 0) {
                resolve(row, s_hit, s_ap, s_tc);
                mark(s_hit);
            }
            jc = s_hit + 1;
            __threadfence_block();
            __syncthreads();
        }
        cur = row + 1;
        __syncthreads();
    }
    if (lane == 0) {
        for (int q = 0; q < s_nd; ++q) dt8[dl[q]] = 0;
        if (hits) atomicAdd(ncoll, hits);
    }
}

}  // namespace

ncg_status ncg_bf_replay_launch(ncg_ctx *ctx, DBody *d_b, const int32_t *d_ex, int32_t *d_dc,
                                int one_collision_only, const int *d_off, int nsys,
                                const int2 *d_cand, const unsigned *d_count32,
                                const unsigned long long *d_count64, unsigned cap, double dt,
                                double f, unsigned char *d_dirty, unsigned *d_stat,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_bf_replay, dim3(nsys), dim3(64), 0, st, d_b, d_ex, d_dc,
                       one_collision_only, d_off, d_cand, d_count32, d_count64, cap, dt, f,
                       d_dirty, d_stat, d_stat + 1);
    NCG_CHECK_LAUNCH();
    return NCG_OK;
}
