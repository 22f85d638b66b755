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
 *                        writes X[k].vx/vy/vz and CM_acc; with the J2 bulge or
 *                        the point-mass perturber, the terms before the mutual
 *                        ones are formed here (start_terms) and the library adds
 *                        the mutual terms to them (ncg_*_from); the inner fluid
 *                        disk's mass enters the central mass
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
 *   NCG_TIMING     when set: wall time per replaced function (ncg_wrap.c) and
 *                  kernel time per kernel group, printed to stderr at exit
 *
 * Errors follow the reference convention: message on stderr, then abort().
 */
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "parameters.h"
#include "structure.h"
#include "physics.h"
#include "collision.h"
#include "spring.h"
#include "ncg.h"

#if !(brute_force_bool || mesh_bool)
#error "ncg drop-in: set brute_force_bool or mesh_bool in parameters.h (tree modules are not accelerated)"
#endif
#if viscoelastic_bool
#error "ncg drop-in: the viscoelastic body is not supported on the GPU path"
#endif

static ncg_ctx *g_ctx = NULL;
static ncg_precision g_prec = NCG_F64_EXACT;

static void ncg_die(const char *what, ncg_status st)
{
    fprintf(stderr, "Error : ncg drop-in: %s failed (status %d): %s\n", what, (int)st,
            g_ctx ? ncg_last_error(g_ctx) : "no context");
    abort();
}

/* With NCG_TIMING set, the library times its kernels with HIP events and the
 * totals are printed to stderr at exit, one line per kernel group:
 *   ncg_timing kernel <name> launches=<count> ms=<sum> */
static int g_timing = 0;

static void ncg_release(void)
{
    static const char *const kernels[] = {"exact", "pairs_fast64", "pairs_mixed",
                                          "candidates", "mesh_bin", "mesh_count",
                                          "mesh_gravity", "mesh_lists", "mesh_replay"};
    size_t k;
    if (g_ctx && g_timing) {
        for (k = 0; k < sizeof kernels / sizeof kernels[0]; ++k) {
            double ms = 0.0;
            int64_t launches = 0;
            if (ncg_profile_read(g_ctx, kernels[k], &ms, &launches, 1) == NCG_OK && launches > 0)
                fprintf(stderr, "ncg_timing kernel %s launches=%lld ms=%.6f\n", kernels[k],
                        (long long)launches, ms);
        }
    }
    if (g_ctx) ncg_destroy(g_ctx);
    g_ctx = NULL;
}

static ncg_ctx *ctx(void)
{
    if (g_ctx) return g_ctx;
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
    if (getenv("NCG_TIMING")) {
        g_timing = 1;
        st = ncg_profile_enable(g_ctx, 1);
        if (st != NCG_OK) ncg_die("ncg_profile_enable", st);
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
        /* physics.c:79, the inner fluid disk's mass included */
        c.mass = inner_fluid_disk_bool ? CM.mass + fluid_disk_Sigma*M_PI*(Rout*Rout - R_unit*R_unit)
                                       : CM.mass;
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

/* Start values with the J2 bulge or the perturber (3 doubles per slot). */
static double *g_start = NULL;
static size_t g_start_cap = 0;

/* vector_field's terms before the mutual ones (physics.c:77-80, 110-172):
 * the central point mass, the J2 bulge and the point-mass perturber, with
 * the reference's expressions in its order, and CM_acc from the same loop.
 * Compiled like the reference (no FMA contraction), so the values are its. */
static void start_terms(const struct moonlet *X, int n, double *st, double *cm)
{
    typ XX = central_mass_bool ? CM.x : 0., YY = central_mass_bool ? CM.y : 0.,
        ZZ = central_mass_bool ? CM.z : 0.;
    typ M = 0., mu, x, y, z, mk, DX, DY, DZ, rk, r2, r3, r5, K, aX, aY, aZ;
    typ X_sun = 0., Y_sun = 0., Z_sun = 0., r_sun, KKK, aX_sun = 0., aY_sun = 0., aZ_sun = 0.;
    typ Xp, Yp, Zp, Rp, KK;
    int k;
    if (central_mass_bool)
        M = inner_fluid_disk_bool ? CM.mass + fluid_disk_Sigma*M_PI*(Rout*Rout - R_unit*R_unit)
                                  : CM.mass;
    cm[0] = 0.;  cm[1] = 0.;  cm[2] = 0.;
    if (pert_mass > 0.) {
        need_to_reduce_COM_bool = 1;
        get_pert_coordinates(t_init + time_elapsed + 0.5*timestep, &X_sun, &Y_sun, &Z_sun);
        r_sun  = sqrt(X_sun*X_sun + Y_sun*Y_sun + Z_sun*Z_sun);
        KKK    = G*pert_mass/(r_sun*r_sun*r_sun);
        aX_sun = KKK*X_sun;  aY_sun = KKK*Y_sun;  aZ_sun = KKK*Z_sun;
    }
    for (k = 0; k < n; k++) {
        if (!exists[k]) continue;
        x  = X[k].x;  y = X[k].y;  z = X[k].z;  mk = X[k].mass;
        DX = x - XX;  DY = y - YY;  DZ = z - ZZ;
        rk = sqrt(DX*DX + DY*DY + DZ*DZ);
        r3 = rk*rk*rk;
        mu = G*M;
        if (central_mass_bool) {
            aX =  -mu*DX/r3;  aY = -mu*DY/r3;  aZ = -mu*DZ/r3;
            cm[0] += G*mk*DX/r3;  cm[1] += G*mk*DY/r3;  cm[2] += G*mk*DZ/r3;
        } else {
            aX = 0.;  aY = 0.;  aZ = 0.;
        }
        if (J2_bool) {
            r2 = rk*rk;
            r5 = r3*r2;
            K  = mu*R_unit*R_unit*J2/r5;
            aX += K*(-1.5*DX + 7.5/r2*DZ*DZ*DX);
            aY += K*(-1.5*DY + 7.5/r2*DZ*DZ*DY);
            aZ += K*(-4.5*DZ + 7.5/r2*DZ*DZ*DZ);
            cm[0] += K*( 1.5*DX - 7.5/r2*DZ*DZ*DX)*mk/M;
            cm[1] += K*( 1.5*DY - 7.5/r2*DZ*DZ*DY)*mk/M;
            cm[2] += K*( 4.5*DZ - 7.5/r2*DZ*DZ*DZ)*mk/M;
        }
        if (pert_mass > 0.) {
            Xp = X_sun - x;  Yp = Y_sun - y;  Zp = Z_sun - z;
            Rp = sqrt(Xp*Xp + Yp*Yp + Zp*Zp);
            KK = G*pert_mass/(Rp*Rp*Rp);
            aX += KK*Xp - aX_sun;  aY += KK*Yp - aY_sun;  aZ += KK*Zp - aZ_sun;
        }
        st[3 * k] = aX;  st[3 * k + 1] = aY;  st[3 * k + 2] = aZ;
    }
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
    if (J2_bool || pert_mass > 0.) {
        /* The library adds the mutual terms to the reference's own start
         * values; CM_acc comes from the same host loop. */
        if ((size_t)n * 3 > g_start_cap) {
            free(g_start);
            g_start_cap = (size_t)n * 3 + 1024;
            g_start = (double *)malloc(sizeof(double) * g_start_cap);
            if (!g_start) {
                fprintf(stderr, "Error : ncg drop-in: cannot allocate the start values.\n");
                abort();
            }
        }
        start_terms(X, n, g_start, cm);
#if mesh_bool
        if (!force_naive_bool)
            st = ncg_mesh_gravity_from(c, (const ncg_body *)X, n, exists, g_start, G,
                                       softening_parameter, g_prec, acc);
        else
#endif
            st = ncg_gravity_allpairs_from(c, (const ncg_body *)X, n, exists, g_start, G,
                                           softening_parameter, g_prec, acc);
    } else {
#if mesh_bool
        if (!force_naive_bool)
            st = ncg_mesh_gravity(c, (const ncg_body *)X, n, exists, &cen, G, softening_parameter,
                                  g_prec, acc, cm);
        else
#endif
            st = ncg_gravity_allpairs(c, (const ncg_body *)X, n, exists, &cen, G,
                                      softening_parameter, g_prec, acc, cm);
    }
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

/* ---- display() (display.c:239-451) -------------------------------------------
 * The per-body columns (x ... vz, a ... p, radius, mass) are formatted on the
 * GPU (ncg_format_fixed: the bytes fprintf's "%.14lf " writes) and their
 * elliptic elements computed there (ncg_cart2ell, cart2ell's arithmetic; the
 * bodies it warns about go through the reference's cart2ell again so that
 * the warning is printed as it would be).  The central body's values, the
 * statistics line and the collision lines are few and written here with the
 * reference's own format strings. */
#if write_to_files_bool
#include "display.h"

static char *g_text = NULL;
static int64_t g_text_cap = 0;
static double *g_cols = NULL;
static size_t g_cols_cap = 0;

static void write_column(FILE *f, const double *v, int64_t count)
{
    if (count == 0) return;
    ncg_ctx *c = ctx();
    int64_t len = 0;
    ncg_status st = ncg_format_fixed(c, v, count, 14, g_text, g_text_cap, &len);
    if (st == NCG_ERR_CAPACITY) {
        free(g_text);
        g_text_cap = len + 4096;
        g_text = (char *)malloc((size_t)g_text_cap);
        if (!g_text) {
            fprintf(stderr, "Error : ncg drop-in: cannot allocate the output text buffer.\n");
            abort();
        }
        st = ncg_format_fixed(c, v, count, 14, g_text, g_text_cap, &len);
    }
    if (st != NCG_OK) ncg_die("display", st);
    fwrite(g_text, 1, (size_t)len, f);
}
#endif

void ncg_dropin_display(struct moonlet *moonlets)
{
#if write_to_files_bool
    FILE *filex = NULL, *filey = NULL, *filez = NULL, *filevx = NULL, *filevy = NULL, *filevz = NULL;
    FILE *filer = NULL, *filem = NULL, *filea = NULL, *filel = NULL, *filek = NULL, *fileh = NULL;
    FILE *fileq = NULL, *filep = NULL, *filestat, *filecol = NULL;
    const int collision = collision_bool && write_collisions_bool ? 1 : 0;
    filestat = *files;
    if (collision) filecol = *(files + 1);
    if (write_cartesian_bool) {
        filem = files[1 + collision];
        filer = files[2 + collision];
        filex = files[3 + collision];
        filey = files[4 + collision];
        filez = files[5 + collision];
        filevx = files[6 + collision];
        filevy = files[7 + collision];
        filevz = files[8 + collision];
        if (write_elliptic_bool) {
            filea = files[9 + collision];
            filel = files[10 + collision];
            filek = files[11 + collision];
            fileh = files[12 + collision];
            fileq = files[13 + collision];
            filep = files[14 + collision];
        }
    } else if (write_elliptic_bool) {
        filem = files[1 + collision];
        filer = files[2 + collision];
        filea = files[3 + collision];
        filel = files[4 + collision];
        filek = files[5 + collision];
        fileh = files[6 + collision];
        fileq = files[7 + collision];
        filep = files[8 + collision];
    }
    typ alkhqp[6];
    typ total_mass = 0.0, mu;
    typ M1 = 0., M2 = 0., R1, R2, a1, a2, l1, l2, K1, K2, H1, H2, Q1, Q2, P1, P2;
    int i1 = 0, i2 = 0, p;
    mu = central_mass_bool ? CM_buffer.mass : M_unit;
    mu *= G;
    if (central_mass_bool && write_cartesian_bool) {
        fprintf(filex, "%.14lf ", CM_buffer.x);
        fprintf(filey, "%.14lf ", CM_buffer.y);
        fprintf(filez, "%.14lf ", CM_buffer.z);
        fprintf(filevx, "%.14lf ", CM_buffer.vx);
        fprintf(filevy, "%.14lf ", CM_buffer.vy);
        fprintf(filevz, "%.14lf ", CM_buffer.vz);
    }
    const int n = largest_id + 1;
    int cnt = 0;
    for (p = 0; p < n; p++) cnt += exists[p] != 0;
    /* Columns of the existing bodies: 6 cartesian, 6 elliptic, radius, mass. */
    if ((size_t)cnt * 14 > g_cols_cap) {
        free(g_cols);
        g_cols_cap = (size_t)cnt * 14 + 1024;
        g_cols = (double *)malloc(sizeof(double) * g_cols_cap);
        if (!g_cols) {
            fprintf(stderr, "Error : ncg drop-in: cannot allocate the output columns.\n");
            abort();
        }
    }
    double *col[14];
    for (int c = 0; c < 14; c++) col[c] = g_cols + (size_t)c * (size_t)cnt;
    int q = 0;
    for (p = 0; p < n; p++) {
        if (!exists[p]) continue;
        const struct moonlet *b = moonlets + p;
        col[0][q] = b->x;
        col[1][q] = b->y;
        col[2][q] = b->z;
        col[3][q] = b->vx;
        col[4][q] = b->vy;
        col[5][q] = b->vz;
        col[12][q] = b->radius;
        col[13][q] = b->mass;
        const typ m = b->mass;
        total_mass += m;
        if (m > M1) {
            M2 = M1;
            i2 = i1;
            M1 = m;
            i1 = p;
        } else if (m > M2) {
            M2 = m;
            i2 = p;
        }
        q++;
    }
    if (write_cartesian_bool) {
        write_column(filex, col[0], cnt);
        write_column(filey, col[1], cnt);
        write_column(filez, col[2], cnt);
        write_column(filevx, col[3], cnt);
        write_column(filevy, col[4], cnt);
        write_column(filevz, col[5], cnt);
    }
    if (write_elliptic_bool && n > 0) {
        double *el = (double *)malloc(sizeof(double) * 6 * (size_t)n);
        int32_t *warn = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        if (!el || !warn) {
            fprintf(stderr, "Error : ncg drop-in: cannot allocate the elliptic elements.\n");
            abort();
        }
        const double cm6[6] = {CM_buffer.x, CM_buffer.y, CM_buffer.z,
                               CM_buffer.vx, CM_buffer.vy, CM_buffer.vz};
        ncg_status st = ncg_cart2ell(ctx(), (const ncg_body *)moonlets, n, exists,
                                     central_mass_bool ? cm6 : NULL, mu, G, el, warn);
        if (st != NCG_OK) ncg_die("display", st);
        q = 0;
        for (p = 0; p < n; p++) {
            if (!exists[p]) continue;
            if (warn[p]) cart2ell(moonlets, p, alkhqp, mu + G * moonlets[p].mass);
            for (int c = 0; c < 6; c++) col[6 + c][q] = el[6 * (size_t)p + c];
            q++;
        }
        free(el);
        free(warn);
        write_column(filea, col[6], cnt);
        write_column(filel, col[7], cnt);
        write_column(filek, col[8], cnt);
        write_column(fileh, col[9], cnt);
        write_column(fileq, col[10], cnt);
        write_column(filep, col[11], cnt);
    }
    if (write_cartesian_bool || write_elliptic_bool) {
        write_column(filer, col[12], cnt);
        write_column(filem, col[13], cnt);
    }
    /* Statistics regarding the two most massive bodies (display.c:373-396). */
    if (how_many_moonlets == 0) {
        M1 = 0.; M2 = 0.; R1 = 0.; R2 = 0.; a1 = 0.; a2 = 0.; l1 = 0.; l2 = 0.;
        K1 = 0.; K2 = 0.; H1 = 0.; H2 = 0.; Q1 = 0.; Q2 = 0.; P1 = 0.; P2 = 0.;
    } else if (how_many_moonlets == 1) {
        M2 = 0.; R2 = 0.; a2 = 0.; l2 = 0.; K2 = 0.; H2 = 0.; Q2 = 0.; P2 = 0.;
        R1 = moonlets[i1].radius;
        cart2ell(moonlets, i1, alkhqp, mu + G * M1);
        a1 = alkhqp[0]; l1 = alkhqp[1]; K1 = alkhqp[2]; H1 = alkhqp[3]; Q1 = alkhqp[4]; P1 = alkhqp[5];
    } else {
        R1 = moonlets[i1].radius;
        R2 = moonlets[i2].radius;
        cart2ell(moonlets, i1, alkhqp, mu + G * M1);
        a1 = alkhqp[0]; l1 = alkhqp[1]; K1 = alkhqp[2]; H1 = alkhqp[3]; Q1 = alkhqp[4]; P1 = alkhqp[5];
        cart2ell(moonlets, i2, alkhqp, mu + G * M2);
        a2 = alkhqp[0]; l2 = alkhqp[1]; K2 = alkhqp[2]; H2 = alkhqp[3]; Q2 = alkhqp[4]; P2 = alkhqp[5];
    }
    /* The statistics line (display.c:398-414). */
    fprintf(filestat, "%.14lf %d %d", time_elapsed + t_init, how_many_moonlets, collision_count);
    if (central_mass_bool) fprintf(filestat, " %.14lf", CM_buffer.mass);
    else fprintf(filestat, " %.1lf", 0.);
    fprintf(filestat, " %.14lf", total_mass);
    fprintf(filestat, " %.1lf %.1lf", 0., 0.);
    if (central_tides_bool) fprintf(filestat, " %.14lf", SideralOmega);
    else fprintf(filestat, " %.1lf", 0.);
    fprintf(filestat, " %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf "
                      "%.14lf %.14lf %.14lf %.14lf %.14lf",
            M1, M2, R1, R2, a1, a2, l1, l2, K1, K2, H1, H2, Q1, Q2, P1, P2);
    fprintf(filestat, "\n");
    if (write_cartesian_bool) {
        fprintf(filex, "\n");
        fprintf(filey, "\n");
        fprintf(filez, "\n");
        fprintf(filevx, "\n");
        fprintf(filevy, "\n");
        fprintf(filevz, "\n");
    }
    if (write_elliptic_bool) {
        fprintf(filea, "\n");
        fprintf(filel, "\n");
        fprintf(filek, "\n");
        fprintf(fileh, "\n");
        fprintf(fileq, "\n");
        fprintf(filep, "\n");
    }
    if (write_cartesian_bool || write_elliptic_bool) {
        fprintf(filer, "\n");
        fprintf(filem, "\n");
    }
    /* Collision data (display.c:433-447). */
    if (collision) {
        for (p = 0; p < indexCollision; p++) {
            const struct collisionData *cd = collisionDatas + p;
            fprintf(filecol, "%.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf %.14lf\n", cd->time,
                    cd->m1, cd->m2, cd->R1, cd->R2, cd->DeltaV, cd->impact_angle, cd->m_tilde);
        }
        indexCollision = 0;
    }
#else
    (void)moonlets;
    fprintf(stderr, "Error : ncg drop-in: display called without write_to_files_bool.\n");
    abort();
#endif
}
