/*
 * ref_probe.c -- function-level probe of the compiled reference.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked by oracle/build_ref.py against NcorpiON's
 * own objects (compiled from the sources where they lie, with a generated
 * parameters.h) to produce golden vectors for one call of a hot-path function
 * on a given state.  It never ships with the product.
 *
 *   ref_probe <in.bin> <out.bin>
 *
 * in.bin  : int32 magic 0x4e434750, op, n, npairs_in;
 *           double timestep, gam, collision_cube, pad[5];
 *           double CM[8]; double bodies[8n]; int32 exists[n];
 *           int32 did_collide[n]; int32 pairs_in[2*npairs_in]
 * out.bin : int32 op, n, n_coll, n_pairs, total_neighbours, how_many_small,
 *           how_many_big, n_ca; double CM_acc[3], pad;
 *           double bodies[8n]; int32 coll[2*n_coll]; int32 pairs[2*n_pairs];
 *           then for op CA, n_ca records {int32 a, b, hit, pad; double t_c,
 *           approach[6]}.
 *
 * ops: 0 vector_field (physics.c:63), 1 brute_force (collision.c:425),
 *      2 mesh (collision.c:324), 3 get_neighbours_mesh (physics.c:1094),
 *      4 closest_approach on every i<j (collision.c:63), 5 get_neighbours_mesh
 *      then vector_field with the pairs it registered (each timed: gnm_seconds
 *      and op_seconds on stderr; no pairs written); op | 0x100 writes only the
 *      header and CM_acc (timing runs); op | 0x200 leaves the pair list out
 *      (n_pairs 0: large mesh passes, whose list runs to gigabytes).
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "parameters.h"
#include "structure.h"
#include "physics.h"
#include "collision.h"
#include "display.h"

static int *coll_log = NULL;
static int n_coll_log = 0, cap_coll_log = 0;

void __real_collision_treatment(struct moonlet *moonlets, int a, int b, int type_of_collision);

void __wrap_collision_treatment(struct moonlet *moonlets, int a, int b, int type_of_collision)
{
    if (n_coll_log == cap_coll_log) {
        cap_coll_log = cap_coll_log ? 2 * cap_coll_log : 1024;
        coll_log = (int *)realloc(coll_log, sizeof(int) * 2 * (size_t)cap_coll_log);
        if (!coll_log) abort();
    }
    coll_log[2 * n_coll_log] = a;
    coll_log[2 * n_coll_log + 1] = b;
    n_coll_log++;
    __real_collision_treatment(moonlets, a, b, type_of_collision);
}

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (n && fread(p, sz, n, f) != n) {
        fprintf(stderr, "ref_probe: short read\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: ref_probe in.bin out.bin\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[4];
    rd(hdr, sizeof(int), 4, f);
    if (hdr[0] != 0x4e434750) return 2;
    /* op | 0x100: write the header and CM_acc only (timing runs). */
    const int header_only = (hdr[1] & 0x100) != 0;
    const int no_pairs = (hdr[1] & 0x200) != 0;
    int op = hdr[1] & 0xff, n = hdr[2], npin = hdr[3];
    double prm[8], cm[8];
    rd(prm, sizeof(double), 8, f);
    rd(cm, sizeof(double), 8, f);
    if (n > N_max) {
        fprintf(stderr, "ref_probe: n=%d exceeds N_max=%d\n", n, (int)N_max);
        return 2;
    }
    variable_initialization();
    array_initialization();
    struct moonlet *bodies = (struct moonlet *)malloc(sizeof(struct moonlet) * N_max);
    memset(bodies, 0, sizeof(struct moonlet) * N_max);
    rd(bodies, sizeof(double), 8 * (size_t)n, f);
    int *ex = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *dc = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    rd(ex, sizeof(int), n, f);
    rd(dc, sizeof(int), n, f);
    int count = 0;
    for (int k = 0; k < N_max; k++) exists[k] = k < n ? ex[k] : 0;
    for (int k = 0; k < n; k++) count += ex[k] != 0;
    if (collision_bool)
        for (int k = 0; k < N_max; k++) did_collide[k] = k < n ? dc[k] : 0;
    largest_id = n - 1;
    how_many_moonlets = count;
    force_naive_bool = 0;
    timestep = prm[0];
    if (!brute_force_bool) {
        gam = prm[1];
        collision_cube = prm[2];
    }
    CM.x = cm[0]; CM.y = cm[1]; CM.z = cm[2];
    CM.vx = cm[3]; CM.vy = cm[4]; CM.vz = cm[5];
    CM.mass = cm[6]; CM.radius = cm[7];
    if (mesh_bool && npin > 0) {
        int *pin = (int *)malloc(sizeof(int) * 2 * (size_t)npin);
        rd(pin, sizeof(int), 2 * (size_t)npin, f);
        for (int j = 0; j < npin; j++) {
            pairs[j].fst = pin[2 * j];
            pairs[j].snd = pin[2 * j + 1];
        }
        how_many_pairs = npin;
        free(pin);
    }
    fclose(f);

    int n_ca = 0;
    int *ca_ab = NULL, *ca_hit = NULL;
    double *ca_rec = NULL;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    switch (op) {
    case 0: vector_field(bodies); break;
    case 1: brute_force(bodies); break;
    case 2: mesh(bodies); break;
    case 3: get_neighbours_mesh(bodies); break;
    case 5: {
        get_neighbours_mesh(bodies);
        struct timespec tg;
        clock_gettime(CLOCK_MONOTONIC, &tg);
        fprintf(stderr, "gnm_seconds=%.6f\n",
                (double)(tg.tv_sec - t0.tv_sec) + 1e-9 * (double)(tg.tv_nsec - t0.tv_nsec));
        t0 = tg;
        vector_field(bodies);
        break;
    }
    case 4: {
        long np = (long)n * (n - 1) / 2;
        ca_ab = (int *)malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
        ca_hit = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
        ca_rec = (double *)malloc(sizeof(double) * 7 * (size_t)(np > 0 ? np : 1));
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                time_until_collision = -1234.5;
                double *ap = closest_approach(bodies, i, j);
                ca_ab[2 * n_ca] = i;
                ca_ab[2 * n_ca + 1] = j;
                ca_hit[n_ca] = ap != NULL;
                ca_rec[7 * n_ca] = time_until_collision;
                for (int q = 0; q < 6; q++) ca_rec[7 * n_ca + 1 + q] = ap ? ap[q] : 0.0;
                n_ca++;
            }
        break;
    }
    case 6: {
        /* cart2ell (structure.c:172) of every existing body with mu = prm[3] + G m_k
           (display.c:340), relative to CM; the six elements replace x .. vz. */
        CM_buffer = CM;
        double *el = (double *)malloc(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1));
        for (int k = 0; k < n; k++)
            if (exists[k]) cart2ell(bodies, k, el + 6 * (size_t)k, prm[3] + G * bodies[k].mass);
        for (int k = 0; k < n; k++)
            if (exists[k]) memcpy(&bodies[k].x, el + 6 * (size_t)k, 6 * sizeof(double));
        free(el);
        break;
    }
    case 7: {
        /* display() (display.c:239) of the bodies into the output files of
           the working directory, then file_closing() (its final flush). */
#if write_to_files_bool
        file_opening();
        CM_buffer = CM;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        display(bodies);
        file_closing();
#else
        return 2;
#endif
        break;
    }
    default: return 2;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    fprintf(stderr, "op_seconds=%.6f\n",
            (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));

    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    int np_out = (mesh_bool && (op == 2 || op == 3) && !no_pairs) ? how_many_pairs : 0;
    int ohdr[8] = {op, n, n_coll_log, np_out, 0, 0, 0, n_ca};
    if (!brute_force_bool) {
        ohdr[4] = total_neighbours;
        ohdr[5] = how_many_small;
        ohdr[6] = how_many_big;
    }
    fwrite(ohdr, sizeof(int), 8, o);
    double ca4[4] = {CM_acc[0], CM_acc[1], CM_acc[2], 0.0};
    fwrite(ca4, sizeof(double), 4, o);
    if (header_only) {
        fclose(o);
        return 0;
    }
    fwrite(bodies, sizeof(double), 8 * (size_t)n, o);
    fwrite(coll_log, sizeof(int), 2 * (size_t)n_coll_log, o);
    for (int j = 0; j < np_out; j++) {
        int pr[2] = {pairs[j].fst, pairs[j].snd};
        fwrite(pr, sizeof(int), 2, o);
    }
    for (int q = 0; q < n_ca; q++) {
        int rec[4] = {ca_ab[2 * q], ca_ab[2 * q + 1], ca_hit[q], 0};
        fwrite(rec, sizeof(int), 4, o);
        fwrite(ca_rec + 7 * q, sizeof(double), 7, o);
    }
    fclose(o);
    return 0;
}
