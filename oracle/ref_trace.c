/*
 * ref_trace.c -- call tracer for whole NcorpiON runs.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/build_ref.py links it with the reference's
 * own objects (main.c included) and -Wl,--wrap=vector_field,--wrap=brute_force,
 * --wrap=mesh,--wrap=get_neighbours_mesh,--wrap=collision_treatment.  Every
 * hot-path call is forwarded to its target and logged with FNV-1a hashes of
 * the body array before and after, so two runs can be compared call by call:
 *
 *   built without TRACE_DROPIN: target = the reference's own function
 *   built with    TRACE_DROPIN: target = the GPU drop-in (ncg_dropin_*)
 *
 * The trace goes to $NCG_TRACE (default ./trace.txt).  Per-function wall time
 * (CLOCK_MONOTONIC around the target call) is appended at exit.
 *
 * With $NCG_TRACE_DUMP set, the state at the start of the integration is
 * written there: for brute force the input of the first brute_force call (the
 * half drift before the loop, rk4.c:871-879), for the mesh the input of the
 * first three_largest_moonlets call (the half kick, rk4.c:697-700; linked
 * with --wrap=three_largest_moonlets): int32 n = largest_id + 1; n 64-byte
 * bodies; int32 exists[n]; int32 did_collide[n]; double CM {x, y, z, vx, vy,
 * vz, mass, radius}; double timestep; double gam; double collision_cube;
 * int32 how_many_moonlets.  A device run can start from it.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "parameters.h"
#include "structure.h"
#include "physics.h"
#include "collision.h"

#ifdef TRACE_DROPIN
void ncg_dropin_vector_field(struct moonlet *X);
void ncg_dropin_brute_force(struct moonlet *X);
void ncg_dropin_mesh(struct moonlet *X);
void ncg_dropin_get_neighbours_mesh(struct moonlet *X);
#define T_VF ncg_dropin_vector_field
#define T_BF ncg_dropin_brute_force
#define T_MESH ncg_dropin_mesh
#define T_GNM ncg_dropin_get_neighbours_mesh
void ncg_dropin_display(struct moonlet *X);
#define T_DISP ncg_dropin_display
#else
void __real_vector_field(struct moonlet *X);
void __real_brute_force(struct moonlet *X);
void __real_mesh(struct moonlet *X);
void __real_get_neighbours_mesh(struct moonlet *X);
#define T_VF __real_vector_field
#define T_BF __real_brute_force
#define T_MESH __real_mesh
#define T_GNM __real_get_neighbours_mesh
void __real_display(struct moonlet *X);
#define T_DISP __real_display
#endif
void __real_collision_treatment(struct moonlet *moonlets, int a, int b, int type_of_collision);

static FILE *trace_fp = NULL;
static long call_no = 0;
static double t_acc[5] = {0, 0, 0, 0, 0};
static long n_acc[5] = {0, 0, 0, 0, 0};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void trace_close(void)
{
    if (!trace_fp) return;
    static const char *nm[5] = {"vector_field", "brute_force", "mesh", "get_neighbours_mesh",
                                "display"};
    for (int q = 0; q < 5; q++)
        if (n_acc[q]) fprintf(trace_fp, "TIME %s calls=%ld total_s=%.6f\n", nm[q], n_acc[q], t_acc[q]);
    fclose(trace_fp);
    trace_fp = NULL;
}

static FILE *tf(void)
{
    if (!trace_fp) {
        const char *p = getenv("NCG_TRACE");
        trace_fp = fopen(p ? p : "trace.txt", "w");
        if (!trace_fp) abort();
        atexit(trace_close);
    }
    return trace_fp;
}

static uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= c[i];
        h *= 1099511628211ULL;
    }
    return h;
}

/* Hash of every existing body (index, 64-byte record) up to largest_id. */
static uint64_t state_hash(const struct moonlet *X)
{
    uint64_t h = 1469598103934665603ULL;
    for (int k = 0; k <= largest_id; k++) {
        if (!exists[k]) continue;
        h = fnv(h, &k, sizeof k);
        h = fnv(h, X + k, sizeof(struct moonlet));
    }
    return h;
}

static uint64_t accel_hash(const struct moonlet *X)
{
    uint64_t h = 1469598103934665603ULL;
    for (int k = 0; k <= largest_id; k++) {
        if (!exists[k]) continue;
        h = fnv(h, &k, sizeof k);
        h = fnv(h, &X[k].vx, 3 * sizeof(double));
    }
    return h;
}

void __wrap_collision_treatment(struct moonlet *moonlets, int a, int b, int type_of_collision)
{
    fprintf(tf(), "C %ld %d %d\n", call_no, a, b);
    __real_collision_treatment(moonlets, a, b, type_of_collision);
}

void __wrap_vector_field(struct moonlet *X)
{
    call_no++;
    uint64_t hin = state_hash(X);
    double t0 = now_s();
    T_VF(X);
    t_acc[0] += now_s() - t0;
    n_acc[0]++;
    fprintf(tf(), "VF %ld %016llx %016llx %a %a %a\n", call_no, (unsigned long long)hin,
            (unsigned long long)accel_hash(X), CM_acc[0], CM_acc[1], CM_acc[2]);
}

static void dump_state(const struct moonlet *X)
{
    const char *p = getenv("NCG_TRACE_DUMP");
    if (!p) return;
    FILE *f = fopen(p, "wb");
    if (!f) abort();
    int n = largest_id + 1;
    fwrite(&n, sizeof n, 1, f);
    fwrite(X, sizeof(struct moonlet), (size_t)n, f);
    fwrite(exists, sizeof(int), (size_t)n, f);
    fwrite(did_collide, sizeof(int), (size_t)n, f);
    double cm[8] = {CM.x, CM.y, CM.z, CM.vx, CM.vy, CM.vz, CM.mass, CM.radius};
    fwrite(cm, sizeof(double), 8, f);
    double ts = timestep;
    fwrite(&ts, sizeof ts, 1, f);
    double g2[2] = {gam, collision_cube};
    fwrite(g2, sizeof(double), 2, f);
    int hm = how_many_moonlets;
    fwrite(&hm, sizeof hm, 1, f);
    fclose(f);
}

static int dumped = 0;

void __real_three_largest_moonlets(struct moonlet *moonlets);
void __wrap_three_largest_moonlets(struct moonlet *X)
{
    if (!dumped) {
        dump_state(X);
        dumped = 1;
    }
    __real_three_largest_moonlets(X);
}

static void collide_call(int which, void (*fn)(struct moonlet *), const char *tag, struct moonlet *X)
{
    if (which == 1 && !dumped) {
        dump_state(X);
        dumped = 1;
    }
    call_no++;
    uint64_t hin = state_hash(X);
    fprintf(tf(), "%s %ld in %016llx dt %a\n", tag, call_no, (unsigned long long)hin, timestep);
    double t0 = now_s();
    fn(X);
    t_acc[which] += now_s() - t0;
    n_acc[which]++;
    fprintf(tf(), "%s %ld out %016llx", tag, call_no, (unsigned long long)state_hash(X));
    if (!brute_force_bool)
        fprintf(tf(), " tn %d small %d big %d", total_neighbours, how_many_small, how_many_big);
    fprintf(tf(), "\n");
}

void __wrap_brute_force(struct moonlet *X) { collide_call(1, T_BF, "BF", X); }
void __wrap_mesh(struct moonlet *X) { collide_call(2, T_MESH, "MESH", X); }
void __wrap_get_neighbours_mesh(struct moonlet *X) { collide_call(3, T_GNM, "GNM", X); }

/* display() (display.c:239): the output files are compared as files. */
void __wrap_display(struct moonlet *X)
{
    call_no++;
    fprintf(tf(), "DISP %ld in %016llx\n", call_no, (unsigned long long)state_hash(X));
    double t0 = now_s();
    T_DISP(X);
    t_acc[4] += now_s() - t0;
    n_acc[4]++;
}
