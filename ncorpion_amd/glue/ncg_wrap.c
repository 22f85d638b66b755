/*
 * ncg_wrap.c -- link-time redirection of NcorpiON's hot-path calls.
 *
 * Linked with -Wl,--wrap=vector_field,--wrap=brute_force,--wrap=mesh,
 * --wrap=get_neighbours_mesh,--wrap=display: every call from rk4.o
 * (rk4.c:85, 593, 700, 741, 762, 767, 776, 806, 876, 909; display:
 * rk4.c:506, 570, 643, 692, 745, 812, 869, 902, 938) lands here and goes to
 * the GPU glue.
 *
 * With NCG_TIMING set in the environment, the wall time (CLOCK_MONOTONIC)
 * spent in each replaced function is summed and printed to stderr at exit,
 * one line per function:
 *   ncg_timing call <name> calls=<count> seconds=<sum> first=<first call>
 * (the first call of a process also creates the GPU context and loads the
 * kernels it uses).
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

struct moonlet;

void ncg_dropin_vector_field(struct moonlet *X);
void ncg_dropin_brute_force(struct moonlet *X);
void ncg_dropin_mesh(struct moonlet *X);
void ncg_dropin_get_neighbours_mesh(struct moonlet *X);
void ncg_dropin_display(struct moonlet *X);

enum { W_VF, W_BF, W_MESH, W_GNM, W_DISP, W_COUNT };
static const char *const w_names[W_COUNT] = {"vector_field", "brute_force", "mesh",
                                             "get_neighbours_mesh", "display"};
static double w_seconds[W_COUNT], w_first[W_COUNT];
static long w_calls[W_COUNT];
static int w_timing = -1;  /* -1: environment not read yet */

static double w_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void w_report(void)
{
    int k;
    for (k = 0; k < W_COUNT; ++k)
        if (w_calls[k])
            fprintf(stderr, "ncg_timing call %s calls=%ld seconds=%.9f first=%.9f\n", w_names[k],
                    w_calls[k], w_seconds[k], w_first[k]);
}

static void w_call(int k, void (*f)(struct moonlet *), struct moonlet *X)
{
    double t0;
    if (w_timing < 0) {
        w_timing = getenv("NCG_TIMING") != NULL;
        if (w_timing) atexit(w_report);
    }
    if (!w_timing) {
        f(X);
        return;
    }
    t0 = w_now();
    f(X);
    t0 = w_now() - t0;
    if (w_calls[k] == 0) w_first[k] = t0;
    w_seconds[k] += t0;
    w_calls[k]++;
}

void __wrap_vector_field(struct moonlet *X) { w_call(W_VF, ncg_dropin_vector_field, X); }
void __wrap_brute_force(struct moonlet *X) { w_call(W_BF, ncg_dropin_brute_force, X); }
void __wrap_mesh(struct moonlet *X) { w_call(W_MESH, ncg_dropin_mesh, X); }
void __wrap_get_neighbours_mesh(struct moonlet *X)
{
    w_call(W_GNM, ncg_dropin_get_neighbours_mesh, X);
}
void __wrap_display(struct moonlet *X) { w_call(W_DISP, ncg_dropin_display, X); }
