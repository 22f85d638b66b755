/*
 * ncg_oracle.c -- CPU restatement of NcorpiON's mutual-interaction step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * ncorpion_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function restates one piece of the reference in plain C, with the
 * same expression trees and the same visiting order, so that on identical
 * inputs it produces bit-identical outputs (x86-64 SSE2, no FMA contraction:
 * built with -O2 -ffp-contract=off).  Reference citations are to
 * NcorpiON src/ files.
 *
 * Body records are the reference's struct moonlet (structure.h:74-84): eight
 * doubles x,y,z,vx,vy,vz,mass,radius, 64 bytes, indexed 0..n-1 with holes
 * marked by exists[k] == 0.
 *
 * Pinning (what checks each function beyond retyping it):
 *   oracle_vector_field_brute  golden vectors from the compiled reference
 *                              (tests/golden), Newton's third law, two-body
 *                              closed form.
 *   oracle_vector_field_rows   equal, bit for bit, to the rows of
 *                              oracle_vector_field_brute (tests/test_oracle_golden.py).
 *   oracle_vector_field_mesh   golden vectors from the compiled reference.
 *   oracle_closest_approach    analytic head-on / tangent / receding cases and
 *                              golden vectors.
 *   oracle_brute_force         collision sequences from the compiled reference.
 *   oracle_mesh / oracle_get_neighbours_mesh
 *                              pair lists and total_neighbours from the
 *                              compiled reference; brute-force stencil count.
 *   oracle_rk4_step            4th-order convergence slope, two-body orbit
 *                              (no reference counterpart: NcorpiON has no RK4).
 *   oracle_energy              two-body closed form.
 *   oracle_state_hash          the FNV-1a of oracle/ref_trace.c (a helper for
 *                              comparing with traced runs, not a restatement).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BX(b, k) ((b)[8 * (size_t)(k) + 0])
#define BY(b, k) ((b)[8 * (size_t)(k) + 1])
#define BZ(b, k) ((b)[8 * (size_t)(k) + 2])
#define BVX(b, k) ((b)[8 * (size_t)(k) + 3])
#define BVY(b, k) ((b)[8 * (size_t)(k) + 4])
#define BVZ(b, k) ((b)[8 * (size_t)(k) + 5])
#define BM(b, k) ((b)[8 * (size_t)(k) + 6])
#define BR(b, k) ((b)[8 * (size_t)(k) + 7])

/* ------------------------------------------------------------------------ */
/* Gravity: brute-force branch of vector_field, physics.c:63-82, 120-202,
 * 263-266.  `acc` receives what the reference writes into .vx/.vy/.vz; only
 * entries with exists[k] are written.  central = {x, y, z, M}; when
 * central_present is 0 the central term and cm_acc are skipped
 * (physics.c:77-81, 137-146). */
void oracle_vector_field_brute(const double *bodies, int n, const int *exists,
                               int central_present, const double *central,
                               double G, double softening_parameter,
                               double *acc, double *cm_acc)
{
    double XX = central_present ? central[0] : 0.;
    double YY = central_present ? central[1] : 0.;
    double ZZ = central_present ? central[2] : 0.;
    double M = central_present ? central[3] : 0.;
    if (central_present) {
        cm_acc[0] = 0.;
        cm_acc[1] = 0.;
        cm_acc[2] = 0.;
    }
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        double X = BX(bodies, k), Y = BY(bodies, k), Z = BZ(bodies, k);
        double mk = BM(bodies, k);
        double DX = X - XX, DY = Y - YY, DZ = Z - ZZ;
        double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
        double r3 = rk * rk * rk;
        double mu = G * M;
        double aX, aY, aZ;
        if (central_present) {
            aX = -mu * DX / r3;
            aY = -mu * DY / r3;
            aZ = -mu * DZ / r3;
            cm_acc[0] += G * mk * DX / r3;
            cm_acc[1] += G * mk * DY / r3;
            cm_acc[2] += G * mk * DZ / r3;
        } else {
            aX = 0.;
            aY = 0.;
            aZ = 0.;
        }
        for (int p = 0; p < k; p++) {
            if (!exists[p]) continue;
            double mp = BM(bodies, p);
            double Xp = BX(bodies, p), Yp = BY(bodies, p), Zp = BZ(bodies, p);
            double Rp = BR(bodies, p), Rk = BR(bodies, k);
            double dx = X - Xp, dy = Y - Yp, dz = Z - Zp;
            double softening = softening_parameter * (Rk + Rp);
            double D = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
            double D3 = D * D * D;
            aX -= G * mp * dx / D3;
            aY -= G * mp * dy / D3;
            aZ -= G * mp * dz / D3;
            acc[3 * p + 0] += G * mk * dx / D3;
            acc[3 * p + 1] += G * mk * dy / D3;
            acc[3 * p + 2] += G * mk * dz / D3;
        }
        acc[3 * k + 0] = aX;
        acc[3 * k + 1] = aY;
        acc[3 * k + 2] = aZ;
    }
}

/* One symmetric pair update of the mesh branch (physics.c:208-221 pattern):
 * body `a` (outer) -= G*m_b*D/D3, body `b` += G*m_a*D/D3 with D = X_a - X_b. */
/* The same field for the bodies rows[0..nrows) only, each summed over all n
 * bodies: acc[3 r ..] receives row r's acceleration, bit-identical to what
 * oracle_vector_field_brute writes for that body (the sum vector_field forms
 * for body k: its central term, minus the terms of p < k in ascending p while
 * k is the outer body, then plus the terms it receives from every q > k in
 * ascending q, physics.c:175-201).  O(nrows n): the checker of large
 * fields on a sample of bodies.  Rows of absent bodies are left untouched. */
void oracle_vector_field_rows(const double *bodies, int n, const int *exists,
                              int central_present, const double *central, double G,
                              double softening_parameter, const int *rows, int nrows, double *acc)
{
    const double mu = G * (central_present ? central[3] : 0.);
    for (int r = 0; r < nrows; r++) {
        const int k = rows[r];
        if (k < 0 || k >= n || !exists[k]) continue;
        const double X = BX(bodies, k), Y = BY(bodies, k), Z = BZ(bodies, k);
        const double Rk = BR(bodies, k);
        double aX = 0., aY = 0., aZ = 0.;
        if (central_present) {
            double DX = X - central[0], DY = Y - central[1], DZ = Z - central[2];
            double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
            double r3 = rk * rk * rk;
            aX = -mu * DX / r3;
            aY = -mu * DY / r3;
            aZ = -mu * DZ / r3;
        }
        for (int p = 0; p < k; p++) {
            if (!exists[p]) continue;
            double mp = BM(bodies, p);
            double dx = X - BX(bodies, p), dy = Y - BY(bodies, p), dz = Z - BZ(bodies, p);
            double softening = softening_parameter * (Rk + BR(bodies, p));
            double D = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
            double D3 = D * D * D;
            aX -= G * mp * dx / D3;
            aY -= G * mp * dy / D3;
            aZ -= G * mp * dz / D3;
        }
        for (int q = k + 1; q < n; q++) {
            if (!exists[q]) continue;
            double mq = BM(bodies, q);
            double dx = BX(bodies, q) - X, dy = BY(bodies, q) - Y, dz = BZ(bodies, q) - Z;
            double softening = softening_parameter * (BR(bodies, q) + Rk);
            double D = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
            double D3 = D * D * D;
            aX += G * mq * dx / D3;
            aY += G * mq * dy / D3;
            aZ += G * mq * dz / D3;
        }
        acc[3 * (size_t)r + 0] = aX;
        acc[3 * (size_t)r + 1] = aY;
        acc[3 * (size_t)r + 2] = aZ;
    }
}

static void pair_update(const double *bodies, int a, int b, double G, double s,
                        double *acc)
{
    double dx = BX(bodies, a) - BX(bodies, b);
    double dy = BY(bodies, a) - BY(bodies, b);
    double dz = BZ(bodies, a) - BZ(bodies, b);
    double softening = s * (BR(bodies, a) + BR(bodies, b));
    double D = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
    double D3 = D * D * D;
    double ma = BM(bodies, a), mb = BM(bodies, b);
    acc[3 * a + 0] -= G * mb * dx / D3;
    acc[3 * a + 1] -= G * mb * dy / D3;
    acc[3 * a + 2] -= G * mb * dz / D3;
    acc[3 * b + 0] += G * ma * dx / D3;
    acc[3 * b + 1] += G * ma * dy / D3;
    acc[3 * b + 2] += G * ma * dz / D3;
}

/* Gravity: mesh branch of vector_field, physics.c:84-109, 204-261, 296-380.
 * Bodies 0..2 are the three largest; every k > 2 feels them, they feel each
 * other, and all other mutual terms come from the pair list
 * pairs[2*j], pairs[2*j+1] (struct pair fst, snd; structure.h:138-143). */
void oracle_vector_field_mesh(const double *bodies, int n, const int *exists,
                              int central_present, const double *central,
                              double G, double softening_parameter,
                              const int *pairs, long npairs,
                              double *acc, double *cm_acc)
{
    double XX = central_present ? central[0] : 0.;
    double YY = central_present ? central[1] : 0.;
    double ZZ = central_present ? central[2] : 0.;
    double M = central_present ? central[3] : 0.;
    if (central_present) {
        cm_acc[0] = 0.;
        cm_acc[1] = 0.;
        cm_acc[2] = 0.;
    }
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        double X = BX(bodies, k), Y = BY(bodies, k), Z = BZ(bodies, k);
        double mk = BM(bodies, k);
        double DX = X - XX, DY = Y - YY, DZ = Z - ZZ;
        double rk = sqrt(DX * DX + DY * DY + DZ * DZ);
        double r3 = rk * rk * rk;
        double mu = G * M;
        double aX, aY, aZ;
        if (central_present) {
            aX = -mu * DX / r3;
            aY = -mu * DY / r3;
            aZ = -mu * DZ / r3;
            cm_acc[0] += G * mk * DX / r3;
            cm_acc[1] += G * mk * DY / r3;
            cm_acc[2] += G * mk * DZ / r3;
        } else {
            aX = 0.;
            aY = 0.;
            aZ = 0.;
        }
        if (k > 2) {
            double Rk = BR(bodies, k);
            for (int big = 0; big < 3; big++) {
                if (big >= n || !exists[big]) continue;
                double dx = X - BX(bodies, big), dy = Y - BY(bodies, big), dz = Z - BZ(bodies, big);
                double softening = softening_parameter * (Rk + BR(bodies, big));
                double D = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
                double D3 = D * D * D;
                double mb = BM(bodies, big);
                aX -= G * mb * dx / D3;
                aY -= G * mb * dy / D3;
                aZ -= G * mb * dz / D3;
                acc[3 * big + 0] += G * mk * dx / D3;
                acc[3 * big + 1] += G * mk * dy / D3;
                acc[3 * big + 2] += G * mk * dz / D3;
            }
        }
        acc[3 * k + 0] = aX;
        acc[3 * k + 1] = aY;
        acc[3 * k + 2] = aZ;
    }
    if (n > 1 && exists[0] && exists[1]) pair_update(bodies, 0, 1, G, softening_parameter, acc);
    if (n > 2 && exists[0] && exists[2]) pair_update(bodies, 0, 2, G, softening_parameter, acc);
    if (n > 2 && exists[1] && exists[2]) pair_update(bodies, 1, 2, G, softening_parameter, acc);
    for (long j = 0; j < npairs; j++) {
        int k = pairs[2 * j], p = pairs[2 * j + 1];
        if (exists[k] && exists[p] && k > 2 && p > 2)
            pair_update(bodies, k, p, G, softening_parameter, acc);
    }
}

/* ------------------------------------------------------------------------ */
/* closest_approach, collision.c:63-172.  Returns 1 on a collision within
 * `timestep` and fills approach[6] (heavier body first) and *t_c; 0 otherwise. */
int oracle_closest_approach(const double *bodies, const int *exists, int a, int b,
                            double timestep, double *approach, double *t_c)
{
    if (!(exists[a] && exists[b])) return 0;
    double R_a = BR(bodies, a), R_b = BR(bodies, b);
    double vx_a = BVX(bodies, a), vy_a = BVY(bodies, a), vz_a = BVZ(bodies, a);
    double vx_b = BVX(bodies, b), vy_b = BVY(bodies, b), vz_b = BVZ(bodies, b);
    double xa = BX(bodies, a), ya = BY(bodies, a), za = BZ(bodies, a);
    double xb = BX(bodies, b), yb = BY(bodies, b), zb = BZ(bodies, b);
    double dx = xa - xb, dy = ya - yb, dz = za - zb;
    double dvx = vx_a - vx_b, dvy = vy_a - vy_b, dvz = vz_a - vz_b;
    double dr_dot_dv = dx * dvx + dy * dvy + dz * dvz;
    if (dr_dot_dv >= 0.0) return 0;
    double dr2 = dx * dx + dy * dy + dz * dz;
    double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    double R = R_a + R_b;
    double discriminant = dr_dot_dv * dr_dot_dv + dv2 * (R * R - dr2);
    if (discriminant <= 0.0) return 0;
    double t = -(dr_dot_dv + sqrt(discriminant)) / dv2;
    *t_c = t;
    if (t > timestep) return 0;
    xa += t * vx_a;
    ya += t * vy_a;
    za += t * vz_a;
    xb += t * vx_b;
    yb += t * vy_b;
    zb += t * vz_b;
    if (BM(bodies, a) > BM(bodies, b)) {
        approach[0] = xa; approach[1] = ya; approach[2] = za;
        approach[3] = xb; approach[4] = yb; approach[5] = zb;
    } else {
        approach[0] = xb; approach[1] = yb; approach[2] = zb;
        approach[3] = xa; approach[4] = ya; approach[5] = za;
    }
    return 1;
}

/* collision, physics.c:545-605: impulse with parameter f, then both bodies
 * rewound by t_c along their post-impact velocity.  a is the heavier body. */
static void oracle_collision(double *bodies, int a, int b, double f,
                             const double *approach, double t_c)
{
    double R = BR(bodies, a) + BR(bodies, b);
    double m_a = BM(bodies, a), m_b = BM(bodies, b);
    double vx_a = BVX(bodies, a), vy_a = BVY(bodies, a), vz_a = BVZ(bodies, a);
    double vx_b = BVX(bodies, b), vy_b = BVY(bodies, b), vz_b = BVZ(bodies, b);
    double xa = approach[0], ya = approach[1], za = approach[2];
    double xb = approach[3], yb = approach[4], zb = approach[5];
    double dr_dot_dv = (xa - xb) * (vx_a - vx_b) + (ya - yb) * (vy_a - vy_b) + (za - zb) * (vz_a - vz_b);
    double alpha = f * m_a * m_b / ((m_a + m_b) * R * R);
    double Jx = alpha * dr_dot_dv * (xb - xa);
    double Jy = alpha * dr_dot_dv * (yb - ya);
    double Jz = alpha * dr_dot_dv * (zb - za);
    vx_a += Jx / m_a;
    vy_a += Jy / m_a;
    vz_a += Jz / m_a;
    vx_b -= Jx / m_b;
    vy_b -= Jy / m_b;
    vz_b -= Jz / m_b;
    BX(bodies, a) = xa - t_c * vx_a;
    BY(bodies, a) = ya - t_c * vy_a;
    BZ(bodies, a) = za - t_c * vz_a;
    BVX(bodies, a) = vx_a;
    BVY(bodies, a) = vy_a;
    BVZ(bodies, a) = vz_a;
    BX(bodies, b) = xb - t_c * vx_b;
    BY(bodies, b) = yb - t_c * vy_b;
    BZ(bodies, b) = zb - t_c * vz_b;
    BVX(bodies, b) = vx_b;
    BVY(bodies, b) = vy_b;
    BVZ(bodies, b) = vz_b;
}

/* collision_treatment for types 0 (elastic, f = 2) and 1 (inelastic, f =
 * collision_parameter), physics.c:1043-1070.  The caller passes the approach
 * and t_c that the preceding closest_approach call produced. */
static void oracle_treat(double *bodies, const int *exists, int *did_collide,
                         int a, int b, int type, double collision_parameter,
                         int one_collision_only, const double *approach, double t_c)
{
    double m_a = BM(bodies, a), m_b = BM(bodies, b);
    int A = (m_a > m_b ? a : b);
    int B = (m_a > m_b ? b : a);
    if (!(exists[a] && exists[b])) return;
    oracle_collision(bodies, A, B, type == 0 ? 2.0 : collision_parameter, approach, t_c);
    did_collide[a] = one_collision_only;
    did_collide[b] = one_collision_only;
}

/* brute_force, collision.c:425-456, for elastic (type 0) and inelastic
 * (type 1) outcomes.  Mutates bodies/did_collide in place; writes the
 * resolved pairs (i, j) in the order the reference resolves them.
 * Returns the number of collisions (may exceed max_pairs; only the first
 * max_pairs are stored). */
long oracle_brute_force(double *bodies, int n, const int *exists, int *did_collide,
                        double timestep, int type, double collision_parameter,
                        int one_collision_only, int *out_pairs, long max_pairs)
{
    long count = 0;
    double approach[6], t_c = 0.;
    for (int i = 0; i < n; i++) {
        if (!exists[i]) continue;
        for (int j = i + 1; j < n; j++) {
            if (!exists[j]) continue;
            int hit = oracle_closest_approach(bodies, exists, i, j, timestep, approach, &t_c);
            if (hit && !did_collide[i] && !did_collide[j]) {
                oracle_treat(bodies, exists, did_collide, i, j, type, collision_parameter,
                             one_collision_only, approach, t_c);
                if (count < max_pairs) {
                    out_pairs[2 * count] = i;
                    out_pairs[2 * count + 1] = j;
                }
                count++;
            }
        }
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* Mesh.  The hash of chains (structure.h:116-120, structure.c:712-737) is
 * restated as head[key] = newest id in the cell and next[id] = the id that was
 * newest before it, which gives the same newest-to-oldest traversal as the
 * chain walk in neighbours() (collision.c:285-302). */

/* hash_table_cell_index, collision.c:175-261.  Returns 0 when the body is
 * outside the collision cube; otherwise fills idx[27] in the reference order. */
int oracle_cell_indexes(double X, double Y, double Z, double gam, double collision_cube,
                        int cells, long *idx)
{
    if (fabs(X) > collision_cube / 2.0 || fabs(Y) > collision_cube / 2.0 ||
        fabs(Z) > collision_cube / 2.0)
        return 0;
    long d = cells, d2 = (long)cells * cells;
    long i = d / 2 + 1 + ((int)(floor(X / gam)));
    long j = d / 2 + 1 + ((int)(floor(Y / gam)));
    long k = d / 2 + 1 + ((int)(floor(Z / gam)));
    int l = 0;
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++)
                idx[l++] = (i + di) + (j + dj) * d + (k + dk) * d2;
    return 1;
}

typedef struct {
    int *head;      /* per key, -1 = empty */
    int *next;      /* per body */
    long nkeys;
    long *touched;  /* keys set since last clear */
    long ntouched;
} oracle_grid;

static int grid_init(oracle_grid *g, int cells, int n)
{
    long d = cells;
    g->nkeys = (d + 2) * (d + 2) * (d + 2);
    g->head = (int *)malloc(sizeof(int) * g->nkeys);
    g->next = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    g->touched = (long *)malloc(sizeof(long) * (n > 0 ? n : 1));
    g->ntouched = 0;
    if (!g->head || !g->next || !g->touched) return -1;
    for (long q = 0; q < g->nkeys; q++) g->head[q] = -1;
    return 0;
}

static void grid_free(oracle_grid *g)
{
    free(g->head);
    free(g->next);
    free(g->touched);
}

/* neighbours(), collision.c:264-316: gathers the existing ids of the 27 cells
 * into `stack` (gather order: cell 0..26, newest to oldest) and inserts `a`.
 * Returns the number gathered, or -1 when `a` is outside the cube. */
static long gather_and_insert(oracle_grid *g, const double *bodies, const int *exists, int a,
                              double gam, double cube, int cells, int *stack)
{
    long idx[27];
    if (!oracle_cell_indexes(BX(bodies, a), BY(bodies, a), BZ(bodies, a), gam, cube, cells, idx))
        return -1;
    long cnt = 0;
    for (int l = 0; l < 27; l++) {
        for (int id = g->head[idx[l]]; id >= 0; id = g->next[id])
            if (exists[id]) stack[cnt++] = id;
    }
    long own = idx[13];
    if (g->head[own] < 0) g->touched[g->ntouched++] = own;
    g->next[a] = g->head[own];
    g->head[own] = a;
    return cnt;
}

/* mesh, collision.c:324-417, elastic/inelastic outcomes.  Writes the gravity
 * pairs (k, p) in the order the reference appends them -- including the pair
 * appended by collision_treatment (physics.c:1065-1080) for a collision --
 * the resolved collisions (p, k) in order, and the counters.  Returns 0, or -1
 * on allocation failure / pair-buffer overflow. */
int oracle_mesh(double *bodies, int n, const int *exists, int *did_collide, double timestep,
                double gam, double collision_cube, int cells, int type, double collision_parameter,
                int one_collision_only, int mutual,
                int *out_pairs, long max_pairs, long *how_many_pairs,
                int *out_coll, long max_coll, long *n_coll,
                long *total_neighbours, long *how_many_small, long *how_many_big)
{
    oracle_grid g;
    if (grid_init(&g, cells, n)) { grid_free(&g); return -1; }
    int *stack = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!stack) { grid_free(&g); return -1; }
    long np = 0, nc = 0, tn = 0, hs = 0, hb = 0;
    int overflow = 0;
    double approach[6], t_c = 0.;
#define APPEND_PAIR(A, B)                                      \
    do {                                                       \
        if (np < max_pairs) {                                  \
            out_pairs[2 * np] = (A);                           \
            out_pairs[2 * np + 1] = (B);                       \
        } else overflow = 1;                                   \
        np++;                                                  \
    } while (0)
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        double R = BR(bodies, k);
        if (R <= gam) {
            hs++;
            long cnt = gather_and_insert(&g, bodies, exists, k, gam, collision_cube, cells, stack);
            while (cnt > 0) {
                int p = stack[--cnt];
                int hit = oracle_closest_approach(bodies, exists, p, k, timestep, approach, &t_c);
                if (!hit || did_collide[k] || did_collide[p]) {
                    if (mutual) APPEND_PAIR(k, p);
                } else {
                    oracle_treat(bodies, exists, did_collide, p, k, type, collision_parameter,
                                 one_collision_only, approach, t_c);
                    if (mutual) APPEND_PAIR(p, k);
                    if (nc < max_coll) {
                        out_coll[2 * nc] = p;
                        out_coll[2 * nc + 1] = k;
                    }
                    nc++;
                }
                tn++;
            }
        } else {
            hb++;
            for (int p = 0; p < n; p++) {
                if (!(exists[p] && p != k)) continue;
                double Rp = BR(bodies, p);
                if (Rp < gam || (Rp >= gam && p > k)) {
                    int hit = oracle_closest_approach(bodies, exists, p, k, timestep, approach, &t_c);
                    if (!hit || did_collide[k] || did_collide[p]) {
                        if (mutual) APPEND_PAIR(k, p);
                    } else {
                        oracle_treat(bodies, exists, did_collide, p, k, type, collision_parameter,
                                     one_collision_only, approach, t_c);
                        if (mutual) APPEND_PAIR(p, k);
                        if (nc < max_coll) {
                            out_coll[2 * nc] = p;
                            out_coll[2 * nc + 1] = k;
                        }
                        nc++;
                    }
                }
            }
        }
    }
#undef APPEND_PAIR
    *how_many_pairs = np;
    *n_coll = nc;
    *total_neighbours = tn;
    *how_many_small = hs;
    *how_many_big = hb;
    free(stack);
    grid_free(&g);
    return overflow ? -1 : 0;
}

/* get_neighbours_mesh, physics.c:1094-1118: every existing body is gathered
 * and inserted (no radius test); pairs (k, p) in LIFO order. */
int oracle_get_neighbours_mesh(const double *bodies, int n, const int *exists, double gam,
                               double collision_cube, int cells, int *out_pairs, long max_pairs,
                               long *how_many_pairs, long *total_neighbours)
{
    oracle_grid g;
    if (grid_init(&g, cells, n)) { grid_free(&g); return -1; }
    int *stack = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!stack) { grid_free(&g); return -1; }
    long np = 0;
    int overflow = 0;
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        long cnt = gather_and_insert(&g, bodies, exists, k, gam, collision_cube, cells, stack);
        while (cnt > 0) {
            int p = stack[--cnt];
            if (np < max_pairs) {
                out_pairs[2 * np] = k;
                out_pairs[2 * np + 1] = p;
            } else overflow = 1;
            np++;
        }
    }
    *how_many_pairs = np;
    *total_neighbours = np;
    free(stack);
    grid_free(&g);
    return overflow ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Classic RK4 on (x, v) with the brute-force field above (central body held in
 * the state as well, moving under cm_acc).  NcorpiON has no RK4 (its rk4.c
 * integrates with a leapfrog, rk4.c:68-127); this is the FP64 definition the
 * GPU RK4 driver is checked against.  state: bodies (x, v read/written), central
 * {x,y,z,vx,vy,vz,M} read/written when central_present. */
static void rk4_field(const double *b, int n, const int *exists, int cp, const double *c,
                      double G, double s, double *acc, double *cacc)
{
    double cen[4] = {c[0], c[1], c[2], c[6]};
    oracle_vector_field_brute(b, n, exists, cp, cen, G, s, acc, cacc);
}

void oracle_rk4_step(double *bodies, int n, const int *exists, int central_present,
                     double *central, double G, double softening_parameter, double dt)
{
    size_t N = (size_t)n;
    double *Y = (double *)malloc(sizeof(double) * 8 * N);
    double *kx = (double *)calloc(4 * 3 * N, sizeof(double));
    double *kv = (double *)calloc(4 * 3 * N, sizeof(double));
    double *acc = (double *)calloc(3 * N, sizeof(double));
    double C[7], kc[4][6], cacc[3] = {0, 0, 0};
    memcpy(Y, bodies, sizeof(double) * 8 * N);
    memcpy(C, central, sizeof(double) * 7);
    const double cf[4] = {0.0, 0.5, 0.5, 1.0};
    for (int s = 0; s < 4; s++) {
        if (s > 0) {
            double h = cf[s] * dt;
            for (size_t k = 0; k < N; k++) {
                if (!exists[k]) continue;
                for (int d = 0; d < 3; d++) {
                    Y[8 * k + d] = bodies[8 * k + d] + h * kx[3 * N * (s - 1) + 3 * k + d];
                    Y[8 * k + 3 + d] = bodies[8 * k + 3 + d] + h * kv[3 * N * (s - 1) + 3 * k + d];
                }
            }
            if (central_present)
                for (int d = 0; d < 3; d++) {
                    C[d] = central[d] + h * kc[s - 1][d];
                    C[3 + d] = central[3 + d] + h * kc[s - 1][3 + d];
                }
        }
        rk4_field(Y, n, exists, central_present, C, G, softening_parameter, acc, cacc);
        for (size_t k = 0; k < N; k++) {
            if (!exists[k]) continue;
            for (int d = 0; d < 3; d++) {
                kx[3 * N * s + 3 * k + d] = Y[8 * k + 3 + d];
                kv[3 * N * s + 3 * k + d] = acc[3 * k + d];
            }
        }
        for (int d = 0; d < 3; d++) {
            kc[s][d] = C[3 + d];
            kc[s][3 + d] = central_present ? cacc[d] : 0.0;
        }
    }
    for (size_t k = 0; k < N; k++) {
        if (!exists[k]) continue;
        for (int d = 0; d < 3; d++) {
            double sx = kx[3 * k + d] + 2.0 * kx[3 * N + 3 * k + d] + 2.0 * kx[6 * N + 3 * k + d] + kx[9 * N + 3 * k + d];
            double sv = kv[3 * k + d] + 2.0 * kv[3 * N + 3 * k + d] + 2.0 * kv[6 * N + 3 * k + d] + kv[9 * N + 3 * k + d];
            bodies[8 * k + d] += dt / 6.0 * sx;
            bodies[8 * k + 3 + d] += dt / 6.0 * sv;
        }
    }
    if (central_present)
        for (int d = 0; d < 3; d++) {
            central[d] += dt / 6.0 * (kc[0][d] + 2.0 * kc[1][d] + 2.0 * kc[2][d] + kc[3][d]);
            central[3 + d] += dt / 6.0 * (kc[0][3 + d] + 2.0 * kc[1][3 + d] + 2.0 * kc[2][3 + d] + kc[3][3 + d]);
        }
    free(Y);
    free(kx);
    free(kv);
    free(acc);
}

/* Total energy (kinetic + central + mutual potential with the same relative
 * softening), summed with Neumaier compensation. */
double oracle_energy(const double *bodies, int n, const int *exists, int central_present,
                     const double *central, double G, double softening_parameter)
{
    double sum = 0., c = 0.;
#define NADD(v)                                         \
    do {                                                \
        double _v = (v), _t = sum + _v;                 \
        if (fabs(sum) >= fabs(_v)) c += (sum - _t) + _v; \
        else c += (_v - _t) + sum;                      \
        sum = _t;                                       \
    } while (0)
    for (int k = 0; k < n; k++) {
        if (!exists[k]) continue;
        double m = BM(bodies, k);
        NADD(0.5 * m * (BVX(bodies, k) * BVX(bodies, k) + BVY(bodies, k) * BVY(bodies, k) + BVZ(bodies, k) * BVZ(bodies, k)));
        if (central_present) {
            double dx = BX(bodies, k) - central[0], dy = BY(bodies, k) - central[1], dz = BZ(bodies, k) - central[2];
            NADD(-G * central[6] * m / sqrt(dx * dx + dy * dy + dz * dz));
        }
        for (int p = 0; p < k; p++) {
            if (!exists[p]) continue;
            double dx = BX(bodies, k) - BX(bodies, p), dy = BY(bodies, k) - BY(bodies, p), dz = BZ(bodies, k) - BZ(bodies, p);
            double eps = softening_parameter * (BR(bodies, k) + BR(bodies, p));
            NADD(-G * m * BM(bodies, p) / sqrt(dx * dx + dy * dy + dz * dz + eps * eps));
        }
    }
    if (central_present)
        NADD(0.5 * central[6] * (central[3] * central[3] + central[4] * central[4] + central[5] * central[5]));
#undef NADD
    return sum + c;
}

/* collision_treatment for types 0/1 (physics.c:1043-1070) given the approach
 * and t_c of the closest_approach call that detected the pair; exported so a
 * test can drive the library's replay with the oracle's own outcome. */
void oracle_resolve(double *bodies, const int *exists, int *did_collide, int a, int b, int type,
                    double collision_parameter, int one_collision_only, const double *approach,
                    double t_c)
{
    oracle_treat(bodies, exists, did_collide, a, b, type, collision_parameter, one_collision_only,
                 approach, t_c);
}

/* FNV-1a 64 hash of every existing body (int32 index, then its 64-byte
 * record), as oracle/ref_trace.c logs the state at each traced call; lets a
 * device run be compared with a traced reference run step by step. */
uint64_t oracle_state_hash(const double *bodies, int n, const int *exists)
{
    uint64_t h = 1469598103934665603ULL;
    for (int k = 0; k < n; k++) {
        if (exists && !exists[k]) continue;
        const unsigned char *c = (const unsigned char *)&k;
        for (size_t i = 0; i < sizeof k; i++) {
            h ^= c[i];
            h *= 1099511628211ULL;
        }
        c = (const unsigned char *)(bodies + 8 * (size_t)k);
        for (size_t i = 0; i < 64; i++) {
            h ^= c[i];
            h *= 1099511628211ULL;
        }
    }
    return h;
}
