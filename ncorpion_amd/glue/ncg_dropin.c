/*
 * ncg_dropin.c -- host glue that puts libncg behind NcorpiON's own entry points.
 *
 * Compiled together with an NcorpiON source tree (its parameters.h,
 * structure.h, physics.h and collision.h on the include path) and linked with
 * -Wl,--wrap=vector_field,--wrap=brute_force,--wrap=mesh,--wrap=get_neighbours_mesh,
 * --wrap=get_free_index plus ncg_wrap.c, so that the integrator's calls from
 * rk4.o reach the GPU without editing any NcorpiON file.  The parameters.h switches are honoured
 * as compile-time constants, exactly as the reference reads them.
 *
 * What each replaced function does here:
 *   vector_field         physics.c:63-441 -> ncg_gravity_allpairs (brute force,
 *                        or N < switch_to_brute_force) / ncg_mesh_gravity (mesh),
 *                        writes X[k].vx/vy/vz and CM_acc
 *   brute_force          collision.c:425-456 -> ncg_brute_force_pass, with the
 *                        reference's closest_approach and collision_treatment as
 *                        the host decision and outcome (every outcome: bounces,
 *                        mergers, fragmentation; created slots are reported
 *                        from the get_free_index wrap)
 *   mesh                 collision.c:324-417 -> ncg_mesh_pass (mode 0), with the
 *                        same host decision and outcomes (every outcome;
 *                        fragments are appended, as mesh mode requires)
 *   get_neighbours_mesh  physics.c:1094-1118 -> ncg_mesh_pass (mode 1)
 * The hash table of the reference is left untouched (all NULL) and
 * how_many_modified stays 0, so end_of_timestep's clearing loop
 * (rk4.c:361-364) has nothing to do.  The `pairs` array holds only what
 * collision_treatment registers: the replaced vector_field evaluates the
 * pass's pairs from the GPU binning and the library's record of each
 * collision's outcome.
 *
 * Environment:
 *   NCG_DEVICE     GPU ordinal (default 0)
 *   NCG_PRECISION  exact (default; bit-identical to the reference), fast64, mixed
 *
 * Errors follow the reference convention: message on stderr, then abort().
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parameters.h"
#include "structure.h"
#include "physics.h"
#include "collision.h"
#include "ncg.h"

#if !(brute_force_bool || mesh_bool)
#error "ncg drop-in: set brute_force_bool or mesh_bool in parameters.h (tree modules are not accelerated)"
#endif
#if viscoelastic_bool || J2_bool || inner_fluid_disk_bool
#error "ncg drop-in: viscoelastic, J2 and inner-fluid-disk physics are not supported on the GPU path"
#endif

static ncg_ctx *g_ctx = NULL;
static ncg_precision g_prec = NCG_F64_EXACT;

static void ncg_die(const char *what, ncg_status st)
{
    fprintf(stderr, "Error : ncg drop-in: %s failed (status %d): %s\n", what, (int)st,
            g_ctx ? ncg_last_error(g_ctx) : "no context");
    abort();
}

static void ncg_release(void)
{
    if (g_ctx) ncg_destroy(g_ctx);
    g_ctx = NULL;
}

static ncg_ctx *ctx(void)
{
    if (g_ctx) return g_ctx;
    if (pert_mass > 0.) {
        fprintf(stderr, "Error : ncg drop-in: the point-mass perturber is not supported on the GPU path.\n");
        abort();
    }
    if (!mutual_bool) {
        fprintf(stderr, "Error : ncg drop-in: mutual_bool 0 is not supported on the GPU path.\n");
        abort();
    }
    const char *dev = getenv("NCG_DEVICE");
    ncg_status st = ncg_create(&g_ctx, dev ? atoi(dev) : 0);
    if (st != NCG_OK) ncg_die("ncg_create", st);
    const char *p = getenv("NCG_PRECISION");
    if (p && !strcmp(p, "fast64")) g_prec = NCG_F64_FAST;
    else if (p && !strcmp(p, "mixed")) g_prec = NCG_MIXED;
    else if (p && strcmp(p, "exact")) {
        fprintf(stderr, "Error : ncg drop-in: NCG_PRECISION must be exact, fast64 or mixed.\n");
        abort();
    }
    atexit(ncg_release);
    return g_ctx;
}

static ncg_central central(void)
{
    ncg_central c;
    memset(&c, 0, sizeof c);
    c.present = central_mass_bool ? 1 : 0;
    if (central_mass_bool) {
        c.x = CM.x;
        c.y = CM.y;
        c.z = CM.z;
        c.mass = CM.mass;  /* physics.c:79 without the inner fluid disk */
    }
    return c;
}

/* Staging buffer for accelerations (3 doubles per slot). */
static double *g_acc = NULL;
static size_t g_acc_cap = 0;

static double *acc_buffer(int n)
{
    if ((size_t)n * 3 > g_acc_cap) {
        free(g_acc);
        g_acc_cap = (size_t)n * 3 + 1024;
        g_acc = (double *)malloc(sizeof(double) * g_acc_cap);
        if (!g_acc) {
            fprintf(stderr, "Error : ncg drop-in: cannot allocate the acceleration buffer.\n");
            abort();
        }
    }
    return g_acc;
}

void ncg_dropin_vector_field(struct moonlet *X)
{
    ncg_ctx *c = ctx();
    const int n = largest_id + 1;
    if (n <= 0) return;
    double *acc = acc_buffer(n);
    double cm[3] = {0., 0., 0.};
    ncg_central cen = central();
    ncg_status st;
#if mesh_bool
    if (!force_naive_bool)
        st = ncg_mesh_gravity(c, (const ncg_body *)X, n, exists, &cen, G, softening_parameter,
                              g_prec, acc, cm);
    else
#endif
        st = ncg_gravity_allpairs(c, (const ncg_body *)X, n, exists, &cen, G, softening_parameter,
                                  g_prec, acc, cm);
    if (st != NCG_OK) ncg_die("vector_field", st);
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        X[k].vx = acc[3 * k];
        X[k].vy = acc[3 * k + 1];
        X[k].vz = acc[3 * k + 2];
    }
    if (central_mass_bool) {
        CM_acc[0] = cm[0];
        CM_acc[1] = cm[1];
        CM_acc[2] = cm[2];
    }
}

/* Collision outcome selected by the switches (collision.c:438-449). */
static int collision_type(void)
{
    if (elastic_collision_bool) return 0;
    if (inelastic_collision_bool) return 1;
    if (instant_merger_bool) return 2;
    return 3;
}

struct replay_user {
    struct moonlet *X;
    int type;
};

static int replay_test(void *u, int32_t a, int32_t b)
{
    struct replay_user *r = (struct replay_user *)u;
    typ *ap = closest_approach(r->X, a, b);
    return ap != NULL && !did_collide[a] && !did_collide[b];
}

static void replay_resolve(void *u, int32_t a, int32_t b)
{
    struct replay_user *r = (struct replay_user *)u;
    collision_treatment(r->X, a, b, r->type);
}

/* Slots brought into existence by fragmentation (get_free_index,
 * structure.c:849, called from physics.c:994-1001), recorded here through
 * -Wl,--wrap=get_free_index and reported to the library after each outcome. */
int __real_get_free_index(int should_be_put_at_the_end);
#define NCG_BORN_CAP 64
static int32_t g_born[NCG_BORN_CAP];
static int g_nborn = 0;

int __wrap_get_free_index(int should_be_put_at_the_end)
{
    int index = __real_get_free_index(should_be_put_at_the_end);
    if (g_nborn < NCG_BORN_CAP) g_born[g_nborn] = index;
    g_nborn++;
    return index;
}

static int32_t replay_created(void *u, int32_t *ids, int32_t cap)
{
    (void)u;
    int nb = g_nborn;
    g_nborn = 0;
    if (nb > NCG_BORN_CAP || nb > cap) return -1; /* the library reports the error */
    memcpy(ids, g_born, sizeof(int32_t) * (size_t)nb);
    return nb;
}

void ncg_dropin_brute_force(struct moonlet *X)
{
    ncg_ctx *c = ctx();
    const int n = largest_id + 1;
    struct replay_user u = {X, collision_type()};
    ncg_replay_ops ops = {&u, replay_test, replay_resolve, replay_created};
    g_nborn = 0;
    int64_t ncoll = 0;
    ncg_status st = ncg_brute_force_pass(c, (const ncg_body *)X, n, exists, timestep, &ops, &ncoll);
    if (st != NCG_OK) ncg_die("brute_force", st);
}

#if mesh_bool
static void mesh_common(struct moonlet *X, int mode)
{
    ncg_ctx *c = ctx();
    const int n = largest_id + 1;
    struct replay_user u = {X, collision_type()};
    ncg_replay_ops ops = {&u, replay_test, replay_resolve, replay_created};
    ncg_mesh_stats stats;
    g_nborn = 0;
    memset(&stats, 0, sizeof stats);
    ncg_status st = ncg_mesh_pass(c, (const ncg_body *)X, n, exists, timestep, gam, collision_cube,
                                  collision_cube_cells, mode, mode == 0 ? &ops : NULL, &stats);
    if (st != NCG_OK) ncg_die(mode == 0 ? "mesh" : "get_neighbours_mesh", st);
    total_neighbours += (int)stats.total_neighbours;
    if (mode == 0) {
        how_many_small += (int)stats.how_many_small;
        how_many_big += (int)stats.how_many_big;
    }
}
#endif

void ncg_dropin_mesh(struct moonlet *X)
{
#if mesh_bool
    mesh_common(X, 0);
#else
    (void)X;
    fprintf(stderr, "Error : ncg drop-in: mesh called without mesh_bool.\n");
    abort();
#endif
}

void ncg_dropin_get_neighbours_mesh(struct moonlet *X)
{
#if mesh_bool
    mesh_common(X, 1);
#else
    (void)X;
    fprintf(stderr, "Error : ncg drop-in: get_neighbours_mesh called without mesh_bool.\n");
    abort();
#endif
}
