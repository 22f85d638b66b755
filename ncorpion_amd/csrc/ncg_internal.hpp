// Internal declarations shared by the libncg translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ncg.h"

// Device-side view of the reference's struct moonlet (structure.h:74-84).
struct alignas(64) DBody {
    double x, y, z, vx, vy, vz, mass, radius;
};
static_assert(sizeof(DBody) == sizeof(ncg_body), "body layout");

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Visit of a body larger than the mesh size: its registered pairs are
// big_pairs_all[first, first + count).
struct BigVisit {
    int k, first, count;
};

// What a resolved collision of mesh() left at its entry (k, p) of visit k
// (stencil index l, occ-th entry of p in that visit) of the pair list.
struct MeshEvent {
    enum Kind { PAIR, NONE, BLOCK };
    int k, l, p, occ;
    Kind kind;
    ncg_pair reg;       // PAIR: the registered pair (physics.c:1065-1080, 886-890)
    int first, count;   // BLOCK: frag_ids[first, first + count), survivor first
};

// Mesh binning kept between ncg_mesh_pass and ncg_mesh_gravity.
struct MeshState {
    bool valid = false;
    bool ord_lists = false;  // id-ordered neighbourhood lists built for this pass
    bool pairs_known = false;  // the host knows the registered pair list (host replay)
    bool want_lists = false;   // device passes start the exact field's lists (mesh loop)
    int32_t n = 0;           // body slots of the pass
    int32_t cells = 0;       // collision_cube_cells
    int64_t table_keys = 0;  // allocated (and zeroed) entries of the cell tables
    int32_t prev_ins = 0;    // keys of the previous pass still set in the tables
    int32_t mode = 0;        // 0 = mesh(), 1 = get_neighbours_mesh()
    int32_t inserted = 0;    // bodies inserted in the hash
    double gam = 0.0, cube = 0.0;
    std::vector<ncg_pair> extra_pairs;    // big-body pairs with both ids > 2
    std::vector<ncg_pair> big_pairs_all;  // every pair registered by big visits
    std::vector<char> big_scan;           // per entry: 1 the visitor's own pair, 0 a block pair
    std::vector<BigVisit> big_visits;
    std::vector<MeshEvent> events;        // collisions in resolution order
    std::vector<int32_t> frag_ids;        // bodies of the fragmentation blocks
    std::vector<std::pair<int, int>> removed;  // (body, visit during which it was removed)
    int32_t n_end = 0;                    // n + bodies appended during the pass
    int64_t rm_drop = 0;                  // gathered entries that removed bodies left out
    // Bodies whose exact-field sums the host lists term by term (survivors and
    // fragments of full fragmentations): (id, first, count) into x_partners.
    std::vector<int> x_targets, x_partners;
};

struct ProfRec {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double ms = 0.0;
    int64_t launches = 0;
};

struct Rk4Graph;  // defined in rk4.hip

// Launch shape of the tiled gravity forms.
struct TiledCfg {
    int ipl = 1;          // bodies per lane
    int nsplit = 0;       // j-splits per i-block; 0 = from the CU count
    double theta = 0.35;  // near/far threshold
    int src = 0;          // 0 scalar loads, 1 LDS staging
};

// One row of the tuning table (ncg_tuning.txt next to libncg.so): the shape
// to use for `prec` from n_min bodies up.
struct TuneRow {
    int prec;
    long n_min;
    TiledCfg cfg;
};

struct ncg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second stream for work that runs beside the main stream within one call
    // (forked and joined with the two events); created on first use.
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    std::unordered_map<std::string, DevBuf> bufs;
    std::unordered_map<std::string, std::vector<unsigned char>> hbufs;
    std::unordered_map<std::string, DevBuf> pbufs;  // pinned host buffers
    MeshState mesh;
    TiledCfg tiled;
    bool tiled_explicit = false;     // ncg_set_tiling overrides the table
    std::vector<TuneRow> tune;       // from the tuning table, sorted by n_min
    int num_cus = 0;
    bool profiling = false;
    std::unordered_map<std::string, ProfRec> prof;
    Rk4Graph *rk4 = nullptr;
    uint64_t buf_gen = 0;  // bumped whenever a device buffer is (re)allocated
    bool capturing = false;  // a stream capture is in progress
    // Candidate capacity of the device collision replays (GPU-resident
    // loops); NCG_REPLAY_CAP in the environment lowers it (tests of the
    // overflow fallback).  0 = each loop's default.
    unsigned replay_cap = 0;
    // Cells between visits of one cluster of the parallel mesh replay
    // (NCG_REPLAY_CLUSTER_DIST; a test knob: 0 makes clusters meet).
    int replay_cluster_dist = 4;
};

void ncg_rk4_release(ncg_ctx *ctx);

// Looks up the address of the static scratch array of
// ncg_gravity_allpairs_static_dev on the context's device (current), once
// per device.
ncg_status ncg_static_scratch_init(ncg_ctx *ctx);

// Candidate capacity of a device replay whose default is `dflt`.
static inline unsigned ncg_replay_cap(const ncg_ctx *ctx, unsigned dflt)
{
    return ctx->replay_cap > 0 && ctx->replay_cap < dflt ? ctx->replay_cap : dflt;
}

// ncg_mesh_pass_dev without the error mapping: *flags receives the device
// replay's status bits (1 more candidates than its capacity, 2 more pending
// re-tests than it holds, 4 a body larger than the mesh size); with a
// nonzero *flags the pass's results are not valid and the bodies may have
// been changed.
ncg_status ncg_mesh_pass_dev_flags(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n,
                                   const int32_t *d_exists, int32_t *d_did_collide,
                                   double timestep, double gam, double collision_cube,
                                   int32_t cells, int32_t mode, double collision_f,
                                   int32_t one_collision_only, ncg_mesh_stats *stats,
                                   unsigned *flags, hipStream_t st);

// Exact brute-force field with the central body in a device record (8
// doubles {x, y, z, vx, vy, vz, M, present}); CM_acc summed in body order.
ncg_status ncg_grav_exact_dev(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                              const double *d_cen, double G, double soft, double *d_acc,
                              double *d_cterm, double *d_cm_acc, hipStream_t st);

// Candidates of brute_force's pass (collide_allpairs.hip), appended to
// d_out[0 .. cap) with the count in *d_count (zeroed by the caller).
ncg_status ncg_candidates_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                 double dt, int2 *d_out, unsigned long long cap,
                                 unsigned long long *d_count, hipStream_t st);

// Device replay of a brute-force collision pass (replay.hip): one wavefront
// per system (systems d_off[0..nsys]); candidates d_cand[0 .. *count) with
// the count in d_count32 or d_count64; d_ex, d_dc may be null.  d_stat[0] +=
// collisions, d_stat[1] |= capacity error bits.  d_dirty: one zeroed byte
// per body, left zeroed.
// end_of_timestep's host half (rk4.c:155-221, 395-397): removals by the
// disruption / high-dumping distances (r_hd = +inf: none), the reduction of
// the centre of mass, tidy_up; returns the bodies left (how_many_moonlets).
// central: 8 doubles {x, y, z, vx, vy, vz, M, present}; mesh: the slot
// rules of mesh mode (no list of free slots).
int32_t ncg_host_step_end(ncg_body *X, int32_t *exists, int32_t *free_indexes, int32_t &largest_id,
                          int32_t &how_many_free, double *central, double r_dis, double r_hd,
                          int reduce_to_com, int32_t &need_to_reduce, int32_t &removed,
                          int32_t &tidied, int mesh);

constexpr int NCG_REPLAY_RCAP = 1024;  // candidates of one system per pass
constexpr int NCG_REPLAY_DCAP = 256;   // bodies changed in one system per pass
ncg_status ncg_bf_replay_launch(ncg_ctx *ctx, DBody *d_b, const int32_t *d_ex, int32_t *d_dc,
                                int one_collision_only, const int *d_off, int nsys,
                                const int2 *d_cand, const unsigned *d_count32,
                                const unsigned long long *d_count64, unsigned cap, double dt,
                                double f, unsigned char *d_dirty, unsigned *d_stat,
                                hipStream_t st);
// Launch shape for `prec` at n bodies: ncg_set_tiling's, else the tuning table's.
TiledCfg ncg_tiled_cfg(const ncg_ctx *ctx, int prec, long n);

// Brackets one kernel launch with HIP events on its stream when profiling.
struct ProfScope {
    ncg_ctx *c;
    hipStream_t s;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(ncg_ctx *c_, hipStream_t s_, const char *n_) : c(c_), s(s_), name(n_)
    {
        // No timing inside a stream capture: event records do not survive
        // as queryable events in a replayed graph.
        if (c->profiling && !c->capturing && hipEventCreate(&a) == hipSuccess &&
            hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, s);
        else
            a = b = nullptr;
    }
    ~ProfScope()
    {
        if (a && b) {
            (void)hipEventRecord(b, s);
            c->prof[name].pending.emplace_back(a, b);
        }
    }
};

ncg_status ncg_fail(ncg_ctx *ctx, ncg_status code, const char *fmt, ...);

// Device scratch buffer `name`, grown to at least `bytes` (contents are not
// preserved across growth).  Returns nullptr and records an error on failure.
void *ncg_dbuf(ncg_ctx *ctx, const char *name, size_t bytes);

template <class T>
static inline T *dbuf(ncg_ctx *ctx, const char *name, size_t count)
{
    return static_cast<T *>(ncg_dbuf(ctx, name, count * sizeof(T) + 16));
}

// Pinned host buffer `name` (hipHostMalloc), grown like ncg_dbuf.
void *ncg_pbuf(ncg_ctx *ctx, const char *name, size_t bytes);
template <class T>
static inline T *pbuf(ncg_ctx *ctx, const char *name, size_t count)
{
    return static_cast<T *>(ncg_pbuf(ctx, name, count * sizeof(T) + 16));
}

// Host staging vector `name` of at least `count` elements.
template <class T>
static inline T *hbuf(ncg_ctx *ctx, const char *name, size_t count)
{
    auto &v = ctx->hbufs[name];
    if (v.size() < count * sizeof(T)) v.resize(count * sizeof(T) + 64);
    return reinterpret_cast<T *>(v.data());
}

#define NCG_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ncg_fail(ctx, NCG_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,        \
                            hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define NCG_CHECK_LAUNCH()                                                              \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return ncg_fail(ctx, NCG_ERR_DEVICE, "kernel launch failed: %s (%s:%d)",     \
                            hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define NCG_REQUIRE(cond, ...)                                                          \
    do {                                                                                \
        if (!(cond)) return ncg_fail(ctx, NCG_ERR_INVALID, __VA_ARGS__);                \
    } while (0)

static inline hipStream_t pick_stream(ncg_ctx *ctx, void *stream)
{
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Internal launchers used across translation units.
// Order of the tiled forms: bounding box, keys, radix sort -- Morton keys
// for NCG_MIXED (spatially compact tiles), body order for NCG_F64_FAST.
// Leaves the permutation and the count of existing bodies on the device.
ncg_status ncg_tiled_order(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                           ncg_precision prec, hipStream_t s);
// One evaluation with the order of the last ncg_tiled_order (no allocation,
// no host synchronisation: safe under stream capture once buffers exist).
ncg_status ncg_tiled_eval(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                          const ncg_central *central, const double *d_central, double G,
                          double soft, ncg_precision prec, double *d_acc, double *d_cterm,
                          double *d_cm_acc, hipStream_t s);
// Stage s (0..3) update of the RK4 driver (rk4.hip), see k_rk4_stage there.
ncg_status ncg_rk4_stage_launch(ncg_ctx *ctx, int s, double dt, int32_t n, const int32_t *ex,
                                DBody *y0, DBody *Y, const double *acc, double *sx, double *sv,
                                double *c0, double *Yc, const double *cm, double *sc,
                                hipStream_t st);
ncg_status ncg_gravity_dev_impl(ncg_ctx *ctx, const DBody *d_bodies, int32_t n,
                                const int32_t *d_exists, const ncg_central *central, double G,
                                double soft, ncg_precision prec, double *d_acc,
                                double *d_cterm, double *d_cm_acc, hipStream_t s,
                                const double *d_init = nullptr);
// Direct form of the fast gravity forms for 1..4096 bodies (k_small): one
// wavefront per body; with d_cm_acc, CM_acc as a fixed-shape tree sum of the
// terms d_cterm (3n) through d_p3 (3 * ncg_small_partials(n) doubles).
int ncg_small_partials(int32_t n);
ncg_status ncg_small_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                            const ncg_central *central, const double *d_central, double G,
                            double soft, double *d_acc, double *d_cterm, double *d_p3,
                            double *d_cm_acc, hipStream_t st);
// Exact form (gravity_exact.hip): the field with per-body CM_acc terms d_cterm
// (3n, written when the central body is present), and CM_acc summed from
// them in body order.
// d_init (3n, or null): each body's sum starts from it instead of the
// central term.
ncg_status ncg_grav_exact_launch(ncg_ctx *ctx, const DBody *d_b, int32_t n, const int32_t *d_ex,
                                 const ncg_central *c, double G, double soft, double *d_acc,
                                 double *d_cterm, hipStream_t st, const double *d_init);
ncg_status ncg_cm_seq_launch(ncg_ctx *ctx, const double *d_cterm, int32_t n, const int32_t *d_ex,
                             int cp, double *d_cm_acc, hipStream_t st);
