/*
 * ncg.h -- C ABI of the MI355X mutual-interaction library (libncg.so).
 *
 * The library replaces NcorpiON's mutual-interaction step on one GPU:
 * brute-force O(N^2) gravity with pairwise collision detection, and the mesh
 * module's cell binning with neighbour-cell gravity and collision search.
 * Reference entry points it stands behind (NcorpiON src/):
 *
 *   void vector_field(struct moonlet *)          physics.h:55, physics.c:63-441
 *   void brute_force(struct moonlet *)           collision.h:70, collision.c:425-456
 *   void mesh(struct moonlet *)                  collision.h:67, collision.c:324-417
 *   void get_neighbours_mesh(struct moonlet *)   physics.h:76, physics.c:1094-1118
 *
 * The reference functions work on globals; this ABI takes everything as
 * arguments (plain pointers and sizes).  The host glue in
 * ncorpion_amd/glue/ncg_dropin.c marshals the reference globals into these
 * calls.  Collision outcomes stay on the host: the library finds candidates
 * on the GPU and drives the caller's own closest_approach/collision_treatment
 * through ncg_replay_ops, in the reference's visiting order.
 *
 * Conventions
 *   - Every function returns ncg_status; NCG_OK == 0.  On failure a message
 *     is available from ncg_last_error(ctx) and outputs are left untouched.
 *   - Host-pointer entry points copy inputs to the device and results back;
 *     the *_dev entry points take device pointers and a hipStream_t passed as
 *     void* (NULL = the context's own stream) and do not synchronise.
 *   - A context is not re-entrant: one context per host thread.
 *   - Indices are int32 like the reference (int); n is largest_id + 1.
 */
#ifndef NCG_H
#define NCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NCG_ABI_VERSION 5

typedef struct ncg_ctx ncg_ctx;

typedef enum {
    NCG_OK = 0,
    NCG_ERR_INVALID = 1,     /* bad argument (null pointer, negative size, ...) */
    NCG_ERR_DEVICE = 2,      /* a HIP runtime call failed */
    NCG_ERR_NOMEM = 3,       /* device or host allocation failed */
    NCG_ERR_UNSUPPORTED = 4, /* a configuration the GPU path does not cover */
    NCG_ERR_CAPACITY = 5     /* an output buffer is too small; *count says how many */
} ncg_status;

/* Arithmetic forms of the gravity kernels.
 *   NCG_F64_EXACT  FP64, the reference's expression tree and summation order:
 *                  bit-identical to vector_field (physics.c:120-269).
 *   NCG_F64_FAST   FP64, hardware reciprocal square root refined by Newton
 *                  steps, FMA, tile order (not the reference's order).
 *   NCG_MIXED      FP64 positions and tile anchors; far tiles evaluate
 *                  1/r^3 and pair forces in FP32 with FP32 tile partials
 *                  added into FP64 accumulators; near tiles take the FP64
 *                  form.  */
typedef enum { NCG_F64_EXACT = 0, NCG_F64_FAST = 1, NCG_MIXED = 2 } ncg_precision;

/* Same 64-byte layout as the reference's struct moonlet (structure.h:74-84). */
typedef struct {
    double x, y, z;
    double vx, vy, vz;
    double mass, radius;
} ncg_body;

/* The central body of central_mass_bool (CM, structure.c:82; physics.c:77-80). */
typedef struct {
    double x, y, z;
    double mass;
    int32_t present;
    int32_t pad_;
} ncg_central;

/* Same layout as the reference's struct pair (structure.h:138-143). */
typedef struct {
    int32_t fst, snd;
} ncg_pair;

/* Host callbacks that make the collision decisions and outcomes, in the
 * reference's order.  `test` must return nonzero exactly when the reference
 * would call collision_treatment for the pair at the current host state
 * (closest_approach(a, b) != NULL and neither did_collide flag set,
 * collision.c:436-437 / 351-352).  `resolve` applies the outcome
 * (collision_treatment, physics.c:1043).  Both are called with the argument
 * order of the reference: (i, j) for brute_force, (p, k) for mesh.
 *   Outcomes that remove bodies (merger, fragmentation: lose_moonlet,
 * structure.c:1077) clear their `exists` entries, which the library reads
 * live.  Outcomes that create bodies (fragmentation: get_free_index,
 * structure.c:849) must be reported by `created`, which the library calls
 * after each resolve: it writes the slots brought into existence by that
 * resolve (at most cap) and returns their number; the caller's `exists`
 * array must cover them.  `created` may be NULL when no outcome creates
 * bodies (elastic and inelastic bounces, mergers). */
typedef struct {
    void *user;
    int (*test)(void *user, int32_t a, int32_t b);
    void (*resolve)(void *user, int32_t a, int32_t b);
    int32_t (*created)(void *user, int32_t *ids, int32_t cap);
} ncg_replay_ops;

/* Built-in decision and outcome for callers other than NcorpiON itself
 * (NcorpiON's glue passes its own closest_approach and collision_treatment).
 * They restate, on a host body array, closest_approach (collision.c:63-172)
 * and the elastic (collision_f = 2) or inelastic bounce of
 * collision_treatment (physics.c:545-605, 1043-1070), including the
 * did_collide rule of one_collision_only_bool.  Fill the public fields, then
 * ncg_builtin_replay_ops gives the ops to pass to ncg_brute_force_pass or
 * ncg_mesh_pass; `collisions` counts resolved pairs. */
typedef struct {
    ncg_body *bodies;         /* host array, updated in place */
    const int32_t *exists;    /* may be NULL */
    int32_t *did_collide;     /* may be NULL */
    double timestep;          /* dt of the pass (sign and size as the caller's) */
    double collision_f;
    int32_t one_collision_only;
    int32_t pad_;
    int64_t collisions;
    double approach_[6];      /* set by test, read by resolve */
    double t_c_;
} ncg_builtin_collider;

/* Per-pass counters of the mesh module (structure.h:130-137). */
typedef struct {
    int64_t total_neighbours;
    int64_t how_many_small;
    int64_t how_many_big;
    int64_t collisions;
    int64_t candidates;     /* GPU candidate pairs handed to the host */
    int64_t host_tests;     /* pairs the host re-tested after a collision */
    int64_t replay_conflicts; /* device mesh() passes: 1 when clusters of the parallel
                                 replay met and the pass was replayed by one workgroup */
} ncg_mesh_stats;

int ncg_abi_version(void);

/* Context: one GPU (device ordinal), its stream and grow-on-demand buffers. */
ncg_status ncg_create(ncg_ctx **ctx, int device);
void ncg_destroy(ncg_ctx *ctx);
const char *ncg_last_error(const ncg_ctx *ctx);
ncg_status ncg_synchronize(ncg_ctx *ctx);
/* Device stream of the context, as hipStream_t. */
void *ncg_stream(ncg_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Brute-force gravity: replaces the brute-force and central branches of
 * vector_field (physics.c:63-82, 120-202, 263-266).
 *   acc[3k..3k+2] receives the acceleration of every body with exists[k]
 *   (what the reference writes into .vx/.vy/.vz); entries of holes are left
 *   untouched.  cm_acc (may be NULL) receives CM_acc (physics.c:80, 140-142),
 *   the central body's acceleration, summed in the reference's order for
 *   NCG_F64_EXACT.  exists may be NULL (all present).
 */
ncg_status ncg_gravity_allpairs(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                const int32_t *exists, const ncg_central *central,
                                double grav_const, double softening, ncg_precision prec,
                                double *acc, double *cm_acc);

/* ncg_gravity_allpairs for a vector_field with terms beyond the central
 * point mass (the J2 bulge and the point-mass perturber, physics.c:148-172,
 * and the inner fluid disk's mass in mu, physics.c:79): start (3n) holds,
 * for every existing body, the value vector_field has accumulated before the
 * mutual terms (the central, J2 and perturber terms, in the reference's
 * order), and each body's mutual terms are added to it in the reference's
 * order (NCG_F64_EXACT: bit for bit; the other forms add it after their
 * sums).  The caller also forms CM_acc, which those terms feed. */
ncg_status ncg_gravity_allpairs_from(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                     const int32_t *exists, const double *start,
                                     double grav_const, double softening, ncg_precision prec,
                                     double *acc);

/* Same on device buffers: bodies (n records), exists (n or NULL), acc (3n),
 * cm_acc (3 doubles or NULL; in the reference's order for NCG_F64_EXACT, as
 * above). */
ncg_status ncg_gravity_allpairs_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                    const int32_t *d_exists, const ncg_central *central,
                                    double grav_const, double softening, ncg_precision prec,
                                    double *d_acc, double *d_cm_acc, void *stream);

/* No counterpart in the reference.  ncg_gravity_allpairs_dev for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or
 * allocating library routine.  Its scratch (sort keys, tiles, j-split
 * partials, CM_acc terms) is one static device array of
 * NCG_STATIC_SCRATCH_BYTES per GPU, held from the first context on that GPU
 * until the process ends; n above NCG_STATIC_MAX_N is refused
 * (NCG_ERR_INVALID).  The caller passes inputs and outputs as for
 * ncg_gravity_allpairs_dev, with the same results: NCG_F64_EXACT bit for bit;
 * the fast forms alike up to the grouping of the j-split sums (the split
 * count is lowered when the partials would not fit).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (from several
 * streams or threads) race.  Nothing is carried from one call to the next. */
#define NCG_STATIC_SCRATCH_BYTES (32 << 20)
#define NCG_STATIC_MAX_N 32768
ncg_status ncg_gravity_allpairs_static_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                           const int32_t *d_exists, const ncg_central *central,
                                           double grav_const, double softening,
                                           ncg_precision prec, double *d_acc, double *d_cm_acc,
                                           void *stream);

/* ---------------------------------------------------------------------------
 * Brute-force collision search: replaces brute_force (collision.c:425-456).
 * The GPU tests every pair i<j of existing bodies at the given state with the
 * closest_approach predicate (collision.c:63-172) and hands the candidates to
 * the host, which replays the reference's lexicographic (i, j) order through
 * ops: candidate pairs are confirmed with ops->test and resolved with
 * ops->resolve; after a collision every later pair that involves a modified
 * body is re-tested on the host through ops->test.  The GPU reads `bodies`
 * once, at the start of the pass; the callbacks see and mutate the caller's
 * live array.  Rows run over the n slots given (collision.c:432), columns
 * also over slots created during the pass (collision.c:435 reads the live
 * largest_id); a body removed during the pass takes part in no further pair,
 * and a created one is tested on the host against every later row.
 * *n_collisions (may be NULL) receives the number of resolved collisions.
 */
ncg_status ncg_brute_force_pass(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                const int32_t *exists, double timestep,
                                const ncg_replay_ops *ops, int64_t *n_collisions);

/* Candidate pairs only (the GPU half of ncg_brute_force_pass), sorted
 * lexicographically.  Returns NCG_ERR_CAPACITY when more than `capacity`
 * exist, with *count set to the number found. */
ncg_status ncg_collide_candidates_allpairs(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                           const int32_t *exists, double timestep,
                                           ncg_pair *out, int64_t capacity, int64_t *count);

/* ---------------------------------------------------------------------------
 * Mesh module.
 *
 * ncg_mesh_pass replaces mesh() (mode 0, collision.c:324-417) and
 * get_neighbours_mesh() (mode 1, physics.c:1094-1118).  It bins the bodies in
 * the reference's hash-table cells (collision.c:175-261: cell side gam,
 * collision_cube_cells cells per axis plus one guard layer, key i + j*d + k*d^2)
 * by a (key, id) radix sort on the GPU, counts total_neighbours exactly, finds the
 * collision candidates among neighbour pairs (mode 0) and replays them on the
 * host through ops in the reference's visiting order (bodies by ascending id;
 * for each, neighbour cells 26..0 and ascending id within a cell).  Bodies
 * with radius > gam (mode 0) are visited on the host against every body, as
 * in collision.c:381-414.  The binning is kept in the context for the next
 * ncg_mesh_gravity call, which evaluates the pairs registered by the pass at
 * the positions it is given (physics.c:296-378).
 *   Outcomes may remove bodies (mergers; seen through the live `exists`, a
 * removed body leaves the later visits' neighbours, collision.c:296) and
 * create bodies, which mesh mode appends at slots n, n + 1, ... (fragmentation,
 * physics.c:994-996 and structure.c:857-861; reported through ops->created,
 * the arrays must cover them) and whose pairs it registers in the pair list
 * (physics.c:1002-1010); a big body's later visit scans every body up to
 * the live largest id, appended ones included, and a body an outcome moves
 * across gam is visited in its new class.  The next ncg_mesh_gravity then
 * takes the n slots of the pass plus the appended ones.
 *   ops may be NULL in mode 1.
 */
ncg_status ncg_mesh_pass(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                         const int32_t *exists, double timestep, double gam,
                         double collision_cube, int32_t cells, int32_t mode,
                         const ncg_replay_ops *ops, ncg_mesh_stats *stats);

/* ncg_mesh_pass on device arrays for callers that keep the bodies on the GPU:
 * d_bodies (n records, modified in place by mode 0), d_exists (n, or NULL),
 * d_did_collide (n, or NULL unless one_collision_only).  Mode 0 replays
 * mesh()'s visiting order on the device with the library's own outcome, the
 * bounce of collision (physics.c:545-605) with restitution collision_f (2:
 * elastic), and needs every body to be at most the mesh size gam (bodies
 * larger than gam: NCG_ERR_INVALID, use ncg_mesh_pass).  The replay runs
 * clusters of candidate visits (cells within 4 of each other) in parallel
 * workgroups; when two clusters touch one cell the pass is replayed by one
 * workgroup from a snapshot (stats->replay_conflicts = 1), so the result is
 * always that of the visiting order.  Enqueued on `stream` (NULL: the
 * context's stream); returns after the counters are read back (one
 * synchronisation, two after a conflict).  stats->host_tests is 0.  The binning is
 * kept for ncg_mesh_gravity / ncg_mesh_gravity_dev; ncg_mesh_pairs is not
 * available after it. */
ncg_status ncg_mesh_pass_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n, const int32_t *d_exists,
                             int32_t *d_did_collide, double timestep, double gam,
                             double collision_cube, int32_t cells, int32_t mode,
                             double collision_f, int32_t one_collision_only,
                             ncg_mesh_stats *stats, void *stream);

/* Mesh geometry of NcorpiON's mesh integrator (rk4.c:662-837). */
typedef struct {
    double mesh_size;              /* gam: updated after every step as end_of_timestep does
                                      (rk4.c:370-377) and written back */
    double mesh_size_min;          /* gam_min = collision_cube_min / collision_cube_cells */
    double target_neighbours;      /* how_many_neighbours: target average of neighbours */
    double r_disruption;           /* disruption_threshold: checked, no body may come this
                                      close to the central body */
    double r_high_dumping;         /* high_dumping_threshold: checked, no body may go this far */
    int32_t cells;                 /* collision_cube_cells */
    int32_t host_passes;           /* out: mesh() passes of the call replayed on the host */
    int32_t conflict_passes;       /* out: device mesh() passes whose replay clusters met */
} ncg_mesh_loop_params;

/* GPU-resident mesh loop: NcorpiON's integration_mesh (SBAB1, rk4.c:662-837)
 * with bodies and central body on the device.  Every step: three_largest_moonlets
 * and three_largest_three_first (structure.c:914-1064) on the device; with
 * collision_f > 0, mesh() (ncg_mesh_pass_dev, the library's bounce) then
 * drift; without, drift then get_neighbours_mesh(); kick with the mesh field
 * of `prec` (NCG_F64_EXACT: the reference's arithmetic); end_of_timestep's
 * did_collide reset (one_collision_only) and mesh-size update.  flags bit 1:
 * the half kick before the loop (rk4.c:697-717); bit 2: the backward half
 * kick after it (rk4.c:807-812, the central record taking CM_buffer's kick).
 * A mesh() pass the device replay cannot take (a body larger than the mesh
 * size, or more candidates or pending re-tests than the replay holds) is
 * undone from a snapshot and replayed on the host with the same decision and
 * bounce (ncg_mesh_pass with the built-in ops): the results are the same,
 * the pass slower; params->host_passes counts them.  No body may cross the
 * disruption or high-dumping thresholds (end_of_timestep's removals are not
 * applied: NCG_ERR_INVALID, reported with the next pass's counters or at the
 * end of the call; ncg_mesh_leapfrog_model_dev applies them).  One
 * synchronisation per mesh pass (its counters set the next mesh size; the
 * census of the bodies, whose count cannot change here, is read with them),
 * two with NCG_F64_EXACT (the binned count sizes the field's neighbourhood
 * lists, built on a second stream during the pass).  d_exists may be written (a body moved into an empty slot 0..2).
 * *n_collisions (may be NULL): collisions resolved. */
ncg_status ncg_mesh_leapfrog_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n, int32_t *d_exists,
                                 int32_t *d_did_collide, double *d_central, double grav_const,
                                 double softening, ncg_precision prec, double dt, int32_t nsteps,
                                 ncg_mesh_loop_params *params, double collision_f,
                                 int32_t one_collision_only, int32_t flags, int64_t *n_collisions,
                                 void *stream);

/* Mesh branch of vector_field (physics.c:63-109, 120-146, 204-261, 296-380):
 * central term, bodies 0..2 against every body, and the neighbour pairs of
 * the last ncg_mesh_pass (bodies > 2 only).  Outputs as ncg_gravity_allpairs.
 * NCG_F64_EXACT is bit-identical to the reference: each body's terms are added
 * in the pass's pair-list order and bodies 0..2 and CM_acc in body order (the
 * first exact call after a pass builds id-ordered neighbourhood lists, 108
 * bytes per binned body, reused by later calls of the pass).  NCG_F64_FAST
 * sums per neighbour cell; NCG_MIXED is the same FP64 form here (every
 * neighbour term is a near one, where FP32 terms lose ~1e-5 to cancellation).  n: the slots of the pass plus the
 * bodies it appended; pairs of absent bodies are skipped (physics.c:351). */
ncg_status ncg_mesh_gravity(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                            const int32_t *exists, const ncg_central *central, double grav_const,
                            double softening, ncg_precision prec, double *acc,
                            double *cm_acc);

/* ncg_mesh_gravity with the caller's start values, as
 * ncg_gravity_allpairs_from: body k's sum starts from start[3k .. 3k + 2]
 * instead of the central term, then takes bodies 0..2 and the pass's pairs
 * in the reference's order. */
ncg_status ncg_mesh_gravity_from(ncg_ctx *ctx, const ncg_body *bodies, int32_t n,
                                 const int32_t *exists, const double *start, double grav_const,
                                 double softening, ncg_precision prec, double *acc);

/* ncg_mesh_gravity on device arrays, enqueued on `stream` (NULL: the
 * context's stream) without synchronisation: d_bodies (n records), d_exists
 * (n, or NULL), d_central (8 doubles {x, y, z, vx, vy, vz, M, present}, or
 * NULL), outputs d_acc (3 per slot; absent slots untouched) and d_cm_acc (3,
 * may be NULL).  Uses the binning of the last mesh pass over the same n. */
ncg_status ncg_mesh_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                                const int32_t *d_exists, const double *d_central,
                                double grav_const, double softening, ncg_precision prec,
                                double *d_acc, double *d_cm_acc, void *stream);

/* Neighbour pairs of the last ncg_mesh_pass in the reference's registration
 * order, (k, p) with k the visited body (what mesh()/get_neighbours_mesh()
 * append to `pairs`, collision.c:353-358, physics.c:1108-1111), each
 * collision's entry replaced by what its outcome registered
 * (physics.c:1043-1091).  Returns
 * NCG_ERR_CAPACITY with *count when the buffer is too small. */
ncg_status ncg_mesh_pairs(ncg_ctx *ctx, ncg_pair *out, int64_t capacity, int64_t *count);

/* ---------------------------------------------------------------------------
 * Classic RK4 on device buffers (NcorpiON integrates with a leapfrog,
 * rk4.c:68-127; this driver is an addition).  Advances d_bodies (n records,
 * x/v updated in place, mass and radius kept) and the central body
 * d_central = {x, y, z, vx, vy, vz, M, present} by nsteps steps of dt; the
 * central body moves under CM_acc as kick() applies it (rk4.c:103-105).
 * prec: NCG_F64_FAST or NCG_MIXED.  flags bit 0: capture the four stages of
 * a step in a hipGraph once and replay it every step (the spatial sort is
 * refreshed eagerly before each step).  Asynchronous on `stream`.
 */
ncg_status ncg_rk4_steps_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n, const int32_t *d_exists,
                             double *d_central, double grav_const, double softening,
                             ncg_precision prec, double dt, int32_t nsteps, int32_t flags,
                             void *stream);

/* ---------------------------------------------------------------------------
 * GPU-resident leapfrog of NcorpiON's brute-force integrator
 * (integration_brute_force_SABA1, rk4.c:840-954): bodies and central body
 * (d_central: 8 doubles {x, y, z, vx, vy, vz, M, present}) stay on the
 * device; each of nsteps steps is kick(dt) with the field of vector_field
 * (rk4.c:68-107), a brute_force collision pass with dt (collision.c:425-456:
 * candidates on the GPU, then the reference's visiting order replayed on the
 * GPU with closest_approach and the bounce of collision, physics.c:545-605),
 * drift(dt) (rk4.c:110-127), and, with one_collision_only, did_collide
 * cleared (rk4.c:348-353).  collision_f <= 0: no collisions; 2: elastic;
 * otherwise the inelastic bounce with that parameter.  d_exists and
 * d_did_collide may be NULL (no holes; no did_collide unless
 * one_collision_only).  flags bit 0: replay steps from a hipGraph; bit 1:
 * start with the half drift of rk4.c:871-879 (collision pass and drift with
 * dt/2); bit 2: end with the backward half drift of rk4.c:938-942 (the
 * central body included).  With NCG_F64_EXACT every step is the reference's
 * arithmetic.  *n_collisions (may be NULL) receives the collisions resolved.
 * No counterpart as one call in NcorpiON: it moves the reference's loop to
 * the device.
 */
ncg_status ncg_leapfrog_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n, const int32_t *d_exists,
                            int32_t *d_did_collide, double *d_central, double grav_const,
                            double softening, ncg_precision prec, double dt, int32_t nsteps,
                            double collision_f, int32_t one_collision_only, int32_t flags,
                            int64_t *n_collisions, void *stream);

/* ---------------------------------------------------------------------------
 * Batched ensembles: nsys independent systems of uneven size, concatenated in
 * d_bodies (system s = records offsets[s] .. offsets[s+1]-1; offsets is a HOST
 * array of nsys+1 ints, offsets[0] = 0), each with its own central body
 * (d_centrals: 8 doubles per system, layout of ncg_rk4_steps_dev).  Every
 * force evaluation of all systems is one launch.  NCG_F64_FAST sums every
 * body's terms in body order; NCG_MIXED puts each system's bodies in Hilbert
 * order, cuts them into 128-body tiles, and one workgroup takes one
 * (system, body group) unit against every tile of its system: FP64 terms
 * for near tiles, the scaled FP32 far term with FP64 tile anchors for far
 * ones.
 *
 * ncg_ensemble_rk4_dev advances every system by nsteps RK4 steps of dt.
 * flags bit 0: replay each step from a hipGraph; bit 1: collisions -- each
 * step starts with a brute-force collision pass per system (tile pairs that
 * cannot meet within dt skipped), replayed on the device in the reference's
 * lexicographic order (collision.c:425-456) with the library's own
 * restatement of closest_approach and of the elastic / inelastic outcome
 * with parameter collision_f (physics.c:545-605); a 16-step window whose
 * replay overflowed its buffers runs again step by step, such passes
 * replayed on the host.
 * *n_collisions (may be NULL) receives the collisions resolved.
 * ncg_ensemble_gravity_dev evaluates every system once (d_acc: 3 per body,
 * d_cm_acc: 3 per system).  prec: NCG_F64_FAST or NCG_MIXED.
 */
ncg_status ncg_ensemble_rk4_dev(ncg_ctx *ctx, ncg_body *d_bodies, const int32_t *offsets,
                                int32_t nsys, double *d_centrals, double grav_const,
                                double softening, ncg_precision prec, double dt, int32_t nsteps,
                                int32_t flags, double collision_f, int64_t *n_collisions,
                                void *stream);
ncg_status ncg_ensemble_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies, const int32_t *offsets,
                                    int32_t nsys, const double *d_centrals, double grav_const,
                                    double softening, ncg_precision prec, double *d_acc,
                                    double *d_cm_acc, void *stream);

/* ---------------------------------------------------------------------------
 * Adjoints (FP64).  No counterpart in NcorpiON; collision-free dynamics only
 * (contact events are not differentiable).
 *
 * ncg_vjp_gravity_dev: vector-Jacobian product of the brute-force field
 * (the mutual accelerations of ncg_gravity_allpairs plus the central term
 * and the central body's acceleration CM_acc) at bodies d_bodies (n) and
 * central body d_central (8 doubles, layout of ncg_rk4_steps_dev).  Inputs:
 * d_lam (3n) cotangent of the body accelerations, d_lam_central (3, may be
 * NULL) cotangent of CM_acc.  Outputs: d_gx (3n) gradient w.r.t. positions,
 * d_gm (n) w.r.t. masses, d_gcentral (4) w.r.t. the central position and
 * mass.
 *
 * ncg_rk4_adjoint_dev: discrete adjoint of nsteps FP64 RK4 steps of
 * ncg_rk4_steps_dev (NCG_F64_FAST) from d_bodies0 / d_central0.  The forward
 * pass keeps the state every checkpoint_every steps; the backward pass
 * recomputes each segment from its checkpoint and reverses it stage by stage.
 * d_gfinal (6n: per body x, y, z, vx, vy, vz) and d_gcentral_final (6, may
 * be NULL) are the cotangents of the final state; outputs: d_g0 (6n)
 * gradient w.r.t. the initial state, d_gm (n) w.r.t. masses, d_gcentral0
 * (7: central x, y, z, vx, vy, vz, M).  d_final (n records, may be NULL)
 * receives the final state.
 */
ncg_status ncg_vjp_gravity_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                               const double *d_central, double grav_const, double softening,
                               const double *d_lam, const double *d_lam_central, double *d_gx,
                               double *d_gm, double *d_gcentral, void *stream);
ncg_status ncg_rk4_adjoint_dev(ncg_ctx *ctx, const ncg_body *d_bodies0, int32_t n,
                               const double *d_central0, double grav_const, double softening,
                               double dt, int32_t nsteps, int32_t checkpoint_every,
                               const double *d_gfinal, const double *d_gcentral_final,
                               double *d_g0, double *d_gm, double *d_gcentral0, ncg_body *d_final,
                               void *stream);

/* The same adjoint split in two, so that the forward run the caller needs
 * anyway keeps the checkpoints:
 * ncg_rk4_forward_ckpt_dev advances d_bodies / d_central by nsteps FP64 RK4
 * steps (the ncg_rk4_steps_dev driver, NCG_F64_FAST, graph replay) and
 * stores the state before step q * checkpoint_every in d_ckpt (n records per
 * checkpoint) and d_ckpt_central (8 doubles per checkpoint), q = 0 ..
 * (nsteps - 1) / checkpoint_every.
 * ncg_rk4_adjoint_ckpt_dev runs the backward pass of ncg_rk4_adjoint_dev
 * from those checkpoints (same cotangent inputs and gradient outputs). */
ncg_status ncg_rk4_forward_ckpt_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t n, double *d_central,
                                    double grav_const, double softening, double dt, int32_t nsteps,
                                    int32_t checkpoint_every, ncg_body *d_ckpt,
                                    double *d_ckpt_central, void *stream);
ncg_status ncg_rk4_adjoint_ckpt_dev(ncg_ctx *ctx, const ncg_body *d_ckpt, int32_t n,
                                    const double *d_ckpt_central, double grav_const,
                                    double softening, double dt, int32_t nsteps,
                                    int32_t checkpoint_every, const double *d_gfinal,
                                    const double *d_gcentral_final, double *d_g0, double *d_gm,
                                    double *d_gcentral0, void *stream);

ncg_status ncg_builtin_replay_ops(ncg_builtin_collider *state, ncg_replay_ops *ops);

/* ---------------------------------------------------------------------------
 * Collision outcome models of collision_treatment (physics.c:1043-1091):
 * the bounce (types 0 and 1: collision, physics.c:545-605), the instant
 * merger (type 2: merger, physics.c:628-707) and the fragmentation model
 * (type 3: fragmentation, physics.c:710-1040 -- a merger below the ejected
 * mass threshold, a super-catastrophic disruption, a partial fragmentation
 * into one tail body, or a full one into n_tilde tail fragments).  The
 * fields are the reference's compile-time parameters (parameters.h). */
typedef enum {
    NCG_OUTCOME_BOUNCE = 0,
    NCG_OUTCOME_MERGER = 1,
    NCG_OUTCOME_FRAGMENTATION = 2
} ncg_outcome_kind;

#define NCG_MAX_TILDE 63

typedef struct {
    int32_t kind;               /* ncg_outcome_kind */
    int32_t n_tilde;            /* N_tilde (parameters.h:297), 1 .. NCG_MAX_TILDE */
    int32_t pq[4];              /* pq_min_max (parameters.h:307):
                                   (pq[1] - pq[0] + 1) (pq[3] - pq[2] + 1) = n_tilde */
    double collision_f;         /* bounce: 2 elastic, else collision_parameter */
    double mu, nu, c1, k;       /* mu_parameter, nu_parameter, C1_parameter, k_parameter
                                   (parameters.h:298-301) */
    double merge_threshold;     /* merging_threshold (parameters.h:302) */
    double frag_threshold;      /* fragment_threshold (parameters.h:303) */
    double grav_const;          /* G (the escape speed, physics.c:754) */
} ncg_outcome_model;

/* Built-in ops with an outcome model, on host arrays.  Bodies are removed
 * and created with the reference's slot rules (lose_moonlet,
 * structure.c:1077-1087; get_free_index, structure.c:849-875): a removed
 * slot is pushed on free_indexes, a created body takes the last freed slot
 * (brute force) or the slot after largest_id (mesh mode, physics.c:994-996).
 * The decision is that of ncg_builtin_collider (c.exists is pointed at
 * `exists`).  Fill the public fields, then pass the ops to
 * ncg_brute_force_pass or ncg_mesh_pass. */
typedef struct {
    ncg_builtin_collider c;      /* bodies (n_max records), did_collide, timestep,
                                    one_collision_only; collision_f is model.collision_f */
    ncg_outcome_model model;
    int32_t *exists;             /* n_max entries, updated by the outcomes */
    int32_t *free_indexes;       /* n_max entries: the stack of free slots */
    double *central;             /* 8 doubles {x, y, z, vx, vy, vz, M, present} or NULL: a
                                    disruption's discarded mass is added to M when present
                                    (physics.c:826-829) */
    int32_t n_max;               /* slots of bodies / exists / did_collide (N_max) */
    int32_t largest_id;          /* in/out */
    int32_t how_many_free;       /* in/out */
    int32_t mesh;                /* mesh mode: fragments appended (get_free_index(1)); the
                                    list of free slots is then never read and not kept */
    int32_t need_to_reduce_com;  /* out: set by a super-catastrophic disruption (physics.c:815) */
    int32_t error;               /* out: 1 N_max exceeded, 2 an ejection speed below the escape
                                    speed (where the reference aborts; the pass's state is then
                                    not valid) */
    int64_t mergers, super_catastrophic, partial, full;  /* out: outcome counts */
    int32_t created_[NCG_MAX_TILDE + 1];
    int32_t n_created_;
    int32_t pad_;
} ncg_outcome_state;

ncg_status ncg_builtin_outcome_ops(ncg_outcome_state *state, ncg_replay_ops *ops);

/* GPU-resident brute-force leapfrog (ncg_leapfrog_dev) with an outcome
 * model and end_of_timestep's removals (rk4.c:155-221, 348-353, 395-397):
 * after each step, bodies closer to the central body than r_disruption are
 * merged into it (central_mass_bool: d_central[7] != 0) and bodies farther
 * than r_high_dumping are removed; the centre of mass is reduced when a
 * super-catastrophic disruption or a removal asked for it (reduce_to_com);
 * did_collide is cleared (one_collision_only or the fragmentation model);
 * and the array is compacted (tidy_up, structure.c:881-911) when fewer than
 * 9/10 of largest_id slots hold bodies.  Steps without such events run on the
 * device as in ncg_leapfrog_dev (64-step windows, graph replay with flags
 * bit 0); a window in which a collision pass under the merger or
 * fragmentation model resolved a pair, or a body crossed a threshold, runs
 * again from its snapshot step by step, and each such pass or step end is
 * replayed on the host with the reference's arithmetic (the outcome's `pow`
 * is the host's).  d_bodies, d_exists and d_did_collide (may be NULL unless
 * one_collision_only) hold n_max slots; the call works on slots
 * 0 .. largest_id and updates largest_id, how_many_free and free_indexes.
 * flags: as ncg_leapfrog_dev.  No counterpart as one call in NcorpiON. */
typedef struct {
    ncg_outcome_model model;
    double r_disruption;         /* disruption_threshold; <= 0: none */
    double r_high_dumping;       /* high_dumping_threshold; <= 0: none */
    int32_t reduce_to_com;       /* reduce_to_COM_bool */
    int32_t n_max;               /* slots of the device arrays and of free_indexes (N_max) */
    int32_t largest_id;          /* in/out */
    int32_t how_many_free;       /* in/out */
    int32_t *free_indexes;       /* host array of n_max, in/out */
    int32_t need_to_reduce_com;  /* in/out (need_to_reduce_COM_bool) */
    int32_t host_passes;         /* out: collision passes replayed on the host */
    int32_t host_step_ends;      /* out: step ends (removals, reduction, tidy_up) on the host */
    int32_t removed;             /* out: bodies removed by the thresholds */
    int32_t tidy_ups;            /* out */
    int32_t pad_;
    int64_t mergers, super_catastrophic, partial, full;  /* out: outcome counts of the call */
} ncg_leapfrog_model_params;

ncg_status ncg_leapfrog_model_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t *d_exists,
                                  int32_t *d_did_collide, double *d_central, double grav_const,
                                  double softening, ncg_precision prec, double dt,
                                  int32_t nsteps, int32_t one_collision_only, int32_t flags,
                                  ncg_leapfrog_model_params *params, int64_t *n_collisions,
                                  void *stream);

/* GPU-resident mesh loop (ncg_mesh_leapfrog_dev) with an outcome model and
 * end_of_timestep's removals, as ncg_leapfrog_model_dev: a mesh() pass in
 * which the device replay resolved a pair under the merger or fragmentation
 * model is undone and replayed on the host (ncg_mesh_pass with the model's
 * ops in mesh mode: fragments appended after largest_id with their pairs
 * registered, physics.c:994-1010), and a step end with a body across a
 * removal distance, a pending centre-of-mass reduction or due compaction
 * runs on the host, before the mesh-size update, which counts the bodies
 * left.  The loop synchronises once per step anyway (the census sets the
 * next mesh size), so no windows are needed.  The removal distances are
 * params' (mesh's r_disruption / r_high_dumping are not read); in mesh mode
 * free_indexes is neither read nor written (fragments are appended). */
ncg_status ncg_mesh_leapfrog_model_dev(ncg_ctx *ctx, ncg_body *d_bodies, int32_t *d_exists,
                                       int32_t *d_did_collide, double *d_central,
                                       double grav_const, double softening, ncg_precision prec,
                                       double dt, int32_t nsteps, ncg_mesh_loop_params *mesh,
                                       ncg_leapfrog_model_params *params,
                                       int32_t one_collision_only, int32_t flags,
                                       int64_t *n_collisions, void *stream);

/* ---------------------------------------------------------------------------
 * Total energy (FP64) of bodies d_bodies (n, d_exists may be NULL) and the
 * central body d_central (8 doubles as in ncg_rk4_steps_dev, may be NULL):
 * d_out[4] = {kinetic (bodies + central body), mutual potential
 * -sum_{i<j} G m_i m_j / sqrt(r^2 + (s (R_i + R_j))^2), central potential
 * -sum_i G M m_i / |x_i - C|, total}.  No counterpart in NcorpiON on this
 * path; it is the energy conserved by the field of vector_field
 * (physics.c:127-146, 175-201), for drift measurements.
 */
ncg_status ncg_energy_dev(ncg_ctx *ctx, const ncg_body *d_bodies, int32_t n,
                          const int32_t *d_exists, const double *d_central, double grav_const,
                          double softening, double *d_out, void *stream);

/* ---------------------------------------------------------------------------
 * Output path: what display() (display.c:239-451) writes for every body.
 *
 * ncg_format_fixed: the text fprintf(file, "%.<decimals>lf ", v) writes for
 * each of `count` values (display.c:323-351 use 14 decimals), concatenated
 * without terminator into `text`: the exact binary value rounded to nearest
 * with ties to even, a '-' for every negative sign bit (-0.0 included),
 * "nan" / "inf" for non-finite values, one space after each value, as the
 * C library prints them.  decimals 0..17, count < 2^30.  *length receives the
 * byte count; NCG_ERR_CAPACITY when it exceeds `capacity` (nothing written).
 */
ncg_status ncg_format_fixed(ncg_ctx *ctx, const double *values, int64_t count, int32_t decimals,
                            char *text, int64_t capacity, int64_t *length);

/* cart2ell (structure.c:172-284) of every existing body k of the n slots
 * with mu + G m_k (display.c:340: mu = G (central mass [+ inner fluid disk
 * mass])), relative to `central` = {x, y, z, vx, vy, vz} of CM_buffer (NULL:
 * the origin at rest, central_mass_bool 0): alkhqp[6k .. 6k+5] = a, lambda,
 * k, h, q, p, bit-identical to the reference (J2 correction not applied).
 * warn (n entries, or NULL): 1 where the reference prints its (k, h)
 * consistency warning (structure.c:245-247).  Absent slots are untouched. */
ncg_status ncg_cart2ell(ncg_ctx *ctx, const ncg_body *bodies, int32_t n, const int32_t *exists,
                        const double *central, double mu, double grav_const, double *alkhqp,
                        int32_t *warn);

/* ---------------------------------------------------------------------------
 * Launch shape of the tiled gravity forms (NCG_F64_FAST, NCG_MIXED): bodies
 * per lane (1, 2, 4; 0 keeps), j-splits per i-block (0 = from the CU count,
 * -1 keeps), near/far threshold theta (<= 0 keeps), tile source (0 scalar
 * loads, 1 LDS, -1 keeps).  Results do not depend on the shape beyond FP
 * rounding of the split sums.  The autotuner (ncorpion_amd/autotune.py)
 * picks these per N.
 */
ncg_status ncg_set_tiling(ncg_ctx *ctx, int32_t ipl, int32_t nsplit, double theta, int32_t source);

/* ---------------------------------------------------------------------------
 * Kernel timing with HIP events recorded on the stream each kernel is
 * launched on.  Names: "exact", "pairs_fast64", "pairs_mixed",
 * "candidates", "mesh_bin", "mesh_count", "mesh_gravity", "ens_mixed",
 * "ens_fast64", "vjp", "energy".
 * ncg_profile_read waits for the recorded events and returns the summed
 * duration and launch count (reset != 0 clears them).
 */
ncg_status ncg_profile_enable(ncg_ctx *ctx, int enable);
ncg_status ncg_profile_read(ncg_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches,
                            int reset);

#ifdef __cplusplus
}
#endif

#endif /* NCG_H */
