// Built-in host collision decision and outcomes for callers that are not
// NcorpiON itself (Python users, tests, the GPU-resident loops).  It
// restates, for one body array, closest_approach (collision.c:63-172) and
// collision_treatment (physics.c:1043-1091) with the elastic / inelastic
// bounce (physics.c:545-605), the instant merger (physics.c:628-707) and the
// fragmentation model (physics.c:710-1040), the slot rules of
// get_free_index / lose_moonlet (structure.c:849-875, 1077-1087), and the
// host half of end_of_timestep (removals, centre-of-mass reduction and
// tidy_up: rk4.c:155-221, structure.c:881-911, 1193-1227), in the same FP64
// operation order.  The drop-in glue does not use it: there the reference's
// own functions decide.
#include <cmath>

#include "ncg_internal.hpp"

#pragma clang fp contract(off)

namespace {

using Builtin = ncg_builtin_collider;

int builtin_test(void *u, int32_t a, int32_t bb)
{
    Builtin *s = static_cast<Builtin *>(u);
    if (s->exists && !(s->exists[a] && s->exists[bb])) return 0;
    const ncg_body &A = s->bodies[a], &B = s->bodies[bb];
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double dvx = A.vx - B.vx, dvy = A.vy - B.vy, dvz = A.vz - B.vz;
    const double dr_dot_dv = dx * dvx + dy * dvy + dz * dvz;
    if (dr_dot_dv >= 0.0) return 0;
    const double dr2 = dx * dx + dy * dy + dz * dz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double R = A.radius + B.radius;
    const double disc = dr_dot_dv * dr_dot_dv + dv2 * (R * R - dr2);
    if (disc <= 0.0) return 0;
    const double t = -(dr_dot_dv + sqrt(disc)) / dv2;
    s->t_c_ = t;
    if (t > s->timestep) return 0;
    const double xa = A.x + t * A.vx, ya = A.y + t * A.vy, za = A.z + t * A.vz;
    const double xb = B.x + t * B.vx, yb = B.y + t * B.vy, zb = B.z + t * B.vz;
    if (A.mass > B.mass) {
        s->approach_[0] = xa; s->approach_[1] = ya; s->approach_[2] = za;
        s->approach_[3] = xb; s->approach_[4] = yb; s->approach_[5] = zb;
    } else {
        s->approach_[0] = xb; s->approach_[1] = yb; s->approach_[2] = zb;
        s->approach_[3] = xa; s->approach_[4] = ya; s->approach_[5] = za;
    }
    if (s->did_collide && (s->did_collide[a] || s->did_collide[bb])) return 0;
    return 1;
}

void builtin_resolve(void *u, int32_t a, int32_t bb)
{
    Builtin *s = static_cast<Builtin *>(u);
    const int A = s->bodies[a].mass > s->bodies[bb].mass ? a : bb;
    const int B = A == a ? bb : a;
    ncg_body &P = s->bodies[A], &Q = s->bodies[B];
    const double R = P.radius + Q.radius;
    const double m_a = P.mass, m_b = Q.mass;
    double vx_a = P.vx, vy_a = P.vy, vz_a = P.vz, vx_b = Q.vx, vy_b = Q.vy, vz_b = Q.vz;
    const double *ap = s->approach_;
    const double xa = ap[0], ya = ap[1], za = ap[2], xb = ap[3], yb = ap[4], zb = ap[5];
    const double dr_dot_dv = (xa - xb) * (vx_a - vx_b) + (ya - yb) * (vy_a - vy_b) + (za - zb) * (vz_a - vz_b);
    const double alpha = s->collision_f * m_a * m_b / ((m_a + m_b) * R * R);
    const double Jx = alpha * dr_dot_dv * (xb - xa);
    const double Jy = alpha * dr_dot_dv * (yb - ya);
    const double Jz = alpha * dr_dot_dv * (zb - za);
    vx_a += Jx / m_a;
    vy_a += Jy / m_a;
    vz_a += Jz / m_a;
    vx_b -= Jx / m_b;
    vy_b -= Jy / m_b;
    vz_b -= Jz / m_b;
    const double t = s->t_c_;
    P.x = xa - t * vx_a; P.y = ya - t * vy_a; P.z = za - t * vz_a;
    P.vx = vx_a; P.vy = vy_a; P.vz = vz_a;
    Q.x = xb - t * vx_b; Q.y = yb - t * vy_b; Q.z = zb - t * vz_b;
    Q.vx = vx_b; Q.vy = vy_b; Q.vz = vz_b;
    if (s->did_collide) {
        s->did_collide[a] = s->one_collision_only;
        s->did_collide[bb] = s->one_collision_only;
    }
    ++s->collisions;
}

// ---- outcome models -----------------------------------------------------------
using Outcome = ncg_outcome_state;

// lose_moonlet (structure.c:1077-1087).  In mesh mode the list of free
// slots is never read (fragments are appended, get_free_index(1)), so it is
// not kept.
void lose(Outcome *s, int a)
{
    s->exists[a] = 0;
    if (s->mesh) return;
    s->free_indexes[s->how_many_free] = a;
    ++s->how_many_free;
}

// get_free_index (structure.c:849-875); -1 (error 1) where the reference aborts.
int take_slot(Outcome *s, bool at_end)
{
    int index;
    if (s->how_many_free == 0 || at_end) {
        index = s->largest_id + 1;
        if (index >= s->n_max) {
            s->error = 1;
            return -1;
        }
        s->largest_id = index;
    } else {
        --s->how_many_free;
        index = s->free_indexes[s->how_many_free];
        if (index > s->largest_id) s->largest_id = index;
    }
    if (s->exists[index]) {
        s->error = 1;
        return -1;
    }
    s->exists[index] = 1;
    s->created_[s->n_created_++] = index;
    return index;
}

void set_dc(Outcome *s, int a)
{
    if (s->c.did_collide) s->c.did_collide[a] = s->c.one_collision_only;
}

// merger (physics.c:628-707); a is the heavier body.
void merger(Outcome *s, int a, int b)
{
    ncg_body *X = s->c.bodies;
    const double m_a = X[a].mass, m_b = X[b].mass;
    const double R_a = X[a].radius, R_b = X[b].radius;
    const double rho_a = 3. * m_a / (4. * M_PI * R_a * R_a * R_a);
    const double rho_b = 3. * m_b / (4. * M_PI * R_b * R_b * R_b);
    const double m = m_a + m_b;
    const double average_density = (m_a * rho_a + m_b * rho_b) / m;
    const double vx_a = X[a].vx, vy_a = X[a].vy, vz_a = X[a].vz;
    const double vx_b = X[b].vx, vy_b = X[b].vy, vz_b = X[b].vz;
    const double *ap = s->c.approach_;
    const double xa = ap[0], ya = ap[1], za = ap[2], xb = ap[3], yb = ap[4], zb = ap[5];
    const double vt0 = (m_a * vx_a + m_b * vx_b) / m;
    const double vt1 = (m_a * vy_a + m_b * vy_b) / m;
    const double vt2 = (m_a * vz_a + m_b * vz_b) / m;
    const double rt0 = (m_a * xa + m_b * xb) / m;
    const double rt1 = (m_a * ya + m_b * yb) / m;
    const double rt2 = (m_a * za + m_b * zb) / m;
    const double t = s->c.t_c_;
    X[a].x = rt0 - t * vt0;
    X[a].y = rt1 - t * vt1;
    X[a].z = rt2 - t * vt2;
    X[a].vx = vt0;
    X[a].vy = vt1;
    X[a].vz = vt2;
    X[a].mass = m;
    X[a].radius = pow(3.0 * m / (4.0 * M_PI * average_density), 1.0 / 3.0);
    lose(s, b);
    set_dc(s, a);
}

void cross(double u1, double u2, double u3, double v1, double v2, double v3, double *w)
{
    w[0] = u2 * v3 - u3 * v2;
    w[1] = u3 * v1 - u1 * v3;
    w[2] = u1 * v2 - u2 * v1;
}

// Position of the largest of (i, j, k), ties to the first (structure.c:1090-1111).
int argmax3(double i, double j, double k)
{
    int m;
    double mx;
    if (i >= j) {
        m = 0;
        mx = i;
    } else {
        m = 1;
        mx = j;
    }
    return mx >= k ? m : 2;
}

// fragmentation (physics.c:710-1040); a is the heavier body (the target).
void fragmentation(Outcome *s, int a, int b)
{
    const ncg_outcome_model &P = s->model;
    ncg_body *X = s->c.bodies;
    const double t = s->c.t_c_;
    const double stigma = (3. * P.mu - 1.) / (3. * P.mu);
    const double R_2 = X[a].radius, R_1 = X[b].radius, R = R_1 + R_2;
    const double m_2 = X[a].mass, m_1 = X[b].mass, M = m_1 + m_2;
    const double rho_1 = 3. * m_1 / (4. * M_PI * R_1 * R_1 * R_1);
    const double rho_2 = 3. * m_2 / (4. * M_PI * R_2 * R_2 * R_2);
    const double average_density = (m_1 * rho_1 + m_2 * rho_2) / M;
    const double rho_12 = rho_1 / rho_2;
    const double rho_12_3nu = pow(rho_12, 3. * P.nu);
    const double vx_a = X[a].vx, vy_a = X[a].vy, vz_a = X[a].vz;
    const double vx_b = X[b].vx, vy_b = X[b].vy, vz_b = X[b].vz;
    const double *ap = s->c.approach_;
    const double xa = ap[0], ya = ap[1], za = ap[2], xb = ap[3], yb = ap[4], zb = ap[5];
    const double dx = xa - xb, dy = ya - yb, dz = za - zb;
    const double dvx = vx_a - vx_b, dvy = vy_a - vy_b, dvz = vz_a - vz_b;
    const double dr_dot_dv = dx * dvx + dy * dvy + dz * dvz;
    const double dv_norm = sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
    const double costheta = -dr_dot_dv / (R * dv_norm);
    const double costheta3mu = pow(costheta, 3. * P.mu);
    const double R_eq = pow(3. * M / (4. * M_PI * average_density), 1. / 3.);
    const double vesc = sqrt(2. * P.grav_const * M / R_eq);
    const double K = 3. * P.k * m_1 / (4. * M_PI) / rho_12 * costheta3mu;
    const double m_check = K * (pow(P.c1 * dv_norm / vesc, 3. * P.mu) * rho_12_3nu - 1.);
    double m_tilde = M - m_check;
    const double m_tilde_2 = m_check / (double)P.n_tilde;
    double v_cm[3], r_cm[3], r_tilde[3], v_tilde[3];
    v_cm[0] = (m_2 * vx_a + m_1 * vx_b) / M;
    v_cm[1] = (m_2 * vy_a + m_1 * vy_b) / M;
    v_cm[2] = (m_2 * vz_a + m_1 * vz_b) / M;
    r_cm[0] = (m_2 * xa + m_1 * xb) / M;
    r_cm[1] = (m_2 * ya + m_1 * yb) / M;
    r_cm[2] = (m_2 * za + m_1 * zb) / M;

    if (m_check <= P.merge_threshold * M) {  // no ejecta worth keeping: merger
        merger(s, a, b);
        ++s->mergers;
        return;
    }
    if (m_tilde < 0.1 * M) {  // super-catastrophic: only the largest fragment stays
        m_tilde = 0.1 * M * pow(1.0 / 0.9 * m_check / M, -1.5);
        X[a].vx = v_cm[0];
        X[a].vy = v_cm[1];
        X[a].vz = v_cm[2];
        X[a].x = r_cm[0] - t * v_cm[0];
        X[a].y = r_cm[1] - t * v_cm[1];
        X[a].z = r_cm[2] - t * v_cm[2];
        X[a].mass = m_tilde;
        X[a].radius = pow(3. * m_tilde / (4. * M_PI * average_density), 1. / 3.);
        s->need_to_reduce_com = 1;
        lose(s, b);
        set_dc(s, a);
        ++s->super_catastrophic;
        double discarded_mass = M - m_tilde;
        if (m_tilde < P.frag_threshold / 4.) {
            lose(s, a);
            discarded_mass = M;
        }
        if (s->central && s->central[7] != 0.0) s->central[6] += discarded_mass;
        return;
    }
    if (m_tilde_2 < P.frag_threshold && m_check <= m_tilde) {  // partial: one tail body
        double r_k[3], v_k[3];
        const double v_k_scalar =
            vesc / stigma * (m_check + K) / m_check * (1.0 - pow(K / (K + m_check), stigma));
        if (v_k_scalar < vesc) {
            s->error = 2;
            return;
        }
        const double in_front_of = v_k_scalar / R;
        r_k[0] = -dx;
        r_k[1] = -dy;
        r_k[2] = -dz;
        v_k[0] = in_front_of * r_k[0];
        v_k[1] = in_front_of * r_k[1];
        v_k[2] = in_front_of * r_k[2];
        const double R_tilde = pow(3.0 * m_tilde / (4.0 * M_PI * average_density), 1.0 / 3.0);
        const double R_check = pow(3.0 * m_check / (4.0 * M_PI * average_density), 1.0 / 3.0);
        const double R_tilde12 = R_tilde + R_check;
        const double dr_norm = sqrt(r_k[0] * r_k[0] + r_k[1] * r_k[1] + r_k[2] * r_k[2]);
        for (int q = 0; q < 3; ++q) r_k[q] *= R_tilde12 / dr_norm;
        for (int q = 0; q < 3; ++q) {
            r_tilde[q] = r_cm[q] - m_check / M * r_k[q];
            v_tilde[q] = v_cm[q] - m_check / M * v_k[q];
        }
        X[a].x = r_tilde[0] - t * v_tilde[0];
        X[a].y = r_tilde[1] - t * v_tilde[1];
        X[a].z = r_tilde[2] - t * v_tilde[2];
        X[a].vx = v_tilde[0];
        X[a].vy = v_tilde[1];
        X[a].vz = v_tilde[2];
        X[a].mass = m_tilde;
        X[a].radius = R_tilde;
        X[b].x = r_tilde[0] + r_k[0] - t * (v_tilde[0] + v_k[0]);
        X[b].y = r_tilde[1] + r_k[1] - t * (v_tilde[1] + v_k[1]);
        X[b].z = r_tilde[2] + r_k[2] - t * (v_tilde[2] + v_k[2]);
        X[b].vx = v_tilde[0] + v_k[0];
        X[b].vy = v_tilde[1] + v_k[1];
        X[b].vz = v_tilde[2] + v_k[2];
        X[b].mass = m_check;
        X[b].radius = R_check;
        set_dc(s, a);
        set_dc(s, b);
        ++s->partial;
        return;
    }
    // Full fragmentation: a tail of n_tilde fragments on the (p, q) lattice.
    const int NT = P.n_tilde;
    double r_k[3 * (NCG_MAX_TILDE + 1)], v_k[3 * (NCG_MAX_TILDE + 1)];
    double dr[3], dv[3], u[3], v[3];
    const double R_tilde = pow(3.0 * m_tilde / (4.0 * M_PI * average_density), 1.0 / 3.0);
    const double R_tilde_2 = pow(3.0 * m_tilde_2 / (4.0 * M_PI * average_density), 1.0 / 3.0);
    const double R_tilde12 = R_tilde + R_tilde_2;
    dr[0] = -dx;
    dr[1] = -dy;
    dr[2] = -dz;
    dv[0] = -dvx;
    dv[1] = -dvy;
    dv[2] = -dvz;
    const double dr_norm = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    for (int q = 0; q < 3; ++q) dr[q] *= R_tilde12 / dr_norm;
    double dr_x_dv[3];
    cross(dr[0], dr[1], dr[2], dv[0], dv[1], dv[2], dr_x_dv);
    const double dr_x_dv_norm =
        sqrt(dr_x_dv[0] * dr_x_dv[0] + dr_x_dv[1] * dr_x_dv[1] + dr_x_dv[2] * dr_x_dv[2]);
    if (dr_x_dv_norm > 1.0e-5) {  // oblique: v along dr x dv
        for (int q = 0; q < 3; ++q) v[q] = dr_x_dv[q] / dr_x_dv_norm;
    } else {  // nearly frontal: any unit vector orthogonal to dr
        const int m = argmax3(fabs(dr[0]), fabs(dr[1]), fabs(dr[2]));
        double numerator;
        if (m == 0) {
            numerator = dr[1] + dr[2];
            v[1] = 1.0 / sqrt(2.0 + numerator * numerator / (dr[0] * dr[0]));
            v[2] = v[1];
            v[0] = -v[1] * numerator / dr[0];
        } else if (m == 1) {
            numerator = dr[0] + dr[2];
            v[0] = 1.0 / sqrt(2.0 + numerator * numerator / (dr[1] * dr[1]));
            v[2] = v[0];
            v[1] = -v[0] * numerator / dr[1];
        } else {
            numerator = dr[0] + dr[1];
            v[0] = 1.0 / sqrt(2.0 + numerator * numerator / (dr[2] * dr[2]));
            v[1] = v[0];
            v[2] = -v[0] * numerator / dr[2];
        }
    }
    double v_x_dr[3];
    cross(v[0], v[1], v[2], dr[0], dr[1], dr[2], v_x_dr);
    for (int q = 0; q < 3; ++q) u[q] = v_x_dr[q] / R_tilde12;
    int n = 0;
    for (int p = P.pq[0]; p <= P.pq[1]; ++p) {
        for (int q = P.pq[2]; q <= P.pq[3]; ++q) {
            const double two_p_R_tilde_2 = ((double)(2 * p)) * R_tilde_2;
            const double two_q_R_tilde_2 = ((double)(2 * q)) * R_tilde_2;
            for (int c = 0; c < 3; ++c)
                r_k[3 * n + c] = dr[c] + two_p_R_tilde_2 * u[c] + two_q_R_tilde_2 * v[c];
            const double zk_stg = pow(1.0 - ((double)n) * m_tilde_2 / (K + m_check), stigma);
            const double zk1_stg = pow(1.0 - ((double)n + 1.) * m_tilde_2 / (K + m_check), stigma);
            const double v_k_scalar = vesc / stigma * (m_check + K) / m_tilde_2 * (zk_stg - zk1_stg);
            if (v_k_scalar < vesc) {
                s->error = 2;
                return;
            }
            const double in_front_of =
                v_k_scalar / sqrt(1.0 + ((double)p) * ((double)p) + ((double)q) * ((double)q));
            for (int c = 0; c < 3; ++c)
                v_k[3 * n + c] =
                    in_front_of * (dr[c] / R_tilde12 + ((double)p) * u[c] + ((double)q) * v[c]);
            ++n;
        }
    }
    double corr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (n = 0; n < NT; ++n) {
        corr[0] += r_k[3 * n];
        corr[1] += r_k[3 * n + 1];
        corr[2] += r_k[3 * n + 2];
        corr[3] += v_k[3 * n];
        corr[4] += v_k[3 * n + 1];
        corr[5] += v_k[3 * n + 2];
    }
    for (int q = 0; q < 3; ++q) {
        r_tilde[q] = r_cm[q] - m_tilde_2 / M * corr[q];
        v_tilde[q] = v_cm[q] - m_tilde_2 / M * corr[3 + q];
    }
    int id[NCG_MAX_TILDE + 1];
    id[0] = a;
    for (n = 1; n <= NT; ++n) {
        id[n] = take_slot(s, s->mesh != 0);
        if (id[n] < 0) return;
    }
    lose(s, b);
    for (n = 0; n <= NT; ++n) set_dc(s, id[n]);
    X[id[0]].x = r_tilde[0] - t * v_tilde[0];
    X[id[0]].y = r_tilde[1] - t * v_tilde[1];
    X[id[0]].z = r_tilde[2] - t * v_tilde[2];
    X[id[0]].vx = v_tilde[0];
    X[id[0]].vy = v_tilde[1];
    X[id[0]].vz = v_tilde[2];
    X[id[0]].mass = m_tilde;
    X[id[0]].radius = R_tilde;
    for (n = 0; n < NT; ++n) {
        ncg_body &F = X[id[n + 1]];
        F.x = r_tilde[0] + r_k[3 * n] - t * (v_tilde[0] + v_k[3 * n]);
        F.y = r_tilde[1] + r_k[3 * n + 1] - t * (v_tilde[1] + v_k[3 * n + 1]);
        F.z = r_tilde[2] + r_k[3 * n + 2] - t * (v_tilde[2] + v_k[3 * n + 2]);
        F.vx = v_tilde[0] + v_k[3 * n];
        F.vy = v_tilde[1] + v_k[3 * n + 1];
        F.vz = v_tilde[2] + v_k[3 * n + 2];
        F.mass = m_tilde_2;
        F.radius = R_tilde_2;
    }
    ++s->full;
}

int outcome_test(void *u, int32_t a, int32_t b)
{
    return builtin_test(&static_cast<Outcome *>(u)->c, a, b);
}

// collision_treatment (physics.c:1043-1091) with the state's model.
void outcome_resolve(void *u, int32_t a, int32_t b)
{
    Outcome *s = static_cast<Outcome *>(u);
    s->n_created_ = 0;
    if (s->error) return;
    if (!(s->exists[a] && s->exists[b])) return;
    const ncg_body *X = s->c.bodies;
    const int A = X[a].mass > X[b].mass ? a : b;
    const int B = X[a].mass > X[b].mass ? b : a;
    switch (s->model.kind) {
    case NCG_OUTCOME_BOUNCE:
        builtin_resolve(&s->c, a, b);
        return;  // builtin_resolve counts it
    case NCG_OUTCOME_MERGER:
        merger(s, A, B);
        ++s->mergers;
        break;
    default:
        fragmentation(s, A, B);
        break;
    }
    ++s->c.collisions;
}

int32_t outcome_created(void *u, int32_t *ids, int32_t cap)
{
    Outcome *s = static_cast<Outcome *>(u);
    const int32_t nb = s->n_created_ < cap ? s->n_created_ : cap;
    for (int32_t q = 0; q < nb; ++q) ids[q] = s->created_[q];
    return nb;
}

}  // namespace

extern "C" ncg_status ncg_builtin_outcome_ops(ncg_outcome_state *state, ncg_replay_ops *ops)
{
    if (!state || !ops || !state->c.bodies || !state->exists || !state->free_indexes)
        return NCG_ERR_INVALID;
    const ncg_outcome_model &P = state->model;
    if (P.kind < NCG_OUTCOME_BOUNCE || P.kind > NCG_OUTCOME_FRAGMENTATION) return NCG_ERR_INVALID;
    if (P.kind == NCG_OUTCOME_FRAGMENTATION &&
        (P.n_tilde < 1 || P.n_tilde > NCG_MAX_TILDE || P.pq[1] < P.pq[0] || P.pq[3] < P.pq[2] ||
         (int64_t)(P.pq[1] - P.pq[0] + 1) * (P.pq[3] - P.pq[2] + 1) != P.n_tilde))
        return NCG_ERR_INVALID;
    if (state->largest_id >= state->n_max || state->how_many_free < 0) return NCG_ERR_INVALID;
    state->c.exists = state->exists;
    state->c.collision_f = P.collision_f;
    state->n_created_ = 0;
    ops->user = state;
    ops->test = outcome_test;
    ops->resolve = outcome_resolve;
    ops->created = outcome_created;
    return NCG_OK;
}

// ---- end_of_timestep on the host ----------------------------------------------
int32_t ncg_host_step_end(ncg_body *X, int32_t *exists, int32_t *free_indexes, int32_t &largest_id,
                          int32_t &how_many_free, double *central, double r_dis, double r_hd,
                          int reduce_to_com, int32_t &need_to_reduce, int32_t &removed,
                          int32_t &tidied, int mesh)
{
    Outcome s = {};
    s.mesh = mesh;
    s.exists = exists;
    s.free_indexes = free_indexes;
    s.largest_id = largest_id;
    s.how_many_free = how_many_free;
    const bool cm = central[7] != 0.0;
    // Removals (rk4.c:155-200): XX .. ZZ are the central position before any
    // merger into it.
    int32_t how_many_moonlets = 0;
    const double XX = cm ? central[0] : 0., YY = cm ? central[1] : 0., ZZ = cm ? central[2] : 0.;
    for (int32_t j = 0; j <= largest_id; ++j) {
        if (!exists[j]) continue;
        const double Xj = X[j].x, Yj = X[j].y, Zj = X[j].z, m = X[j].mass;
        const double DX = Xj - XX, DY = Yj - YY, DZ = Zj - ZZ;
        bool losing = false;
        if (DX * DX + DY * DY + DZ * DZ < r_dis * r_dis && cm) {
            const double vX = X[j].vx, vY = X[j].vy, vZ = X[j].vz;
            double *C = central;
            C[6] += m;
            const double M = C[6];
            C[0] *= (M - m) / M;
            C[1] *= (M - m) / M;
            C[2] *= (M - m) / M;
            C[3] *= (M - m) / M;
            C[4] *= (M - m) / M;
            C[5] *= (M - m) / M;
            C[0] += m * Xj / M;
            C[1] += m * Yj / M;
            C[2] += m * Zj / M;
            C[3] += m * vX / M;
            C[4] += m * vY / M;
            C[5] += m * vZ / M;
            lose(&s, j);
            losing = true;
        } else if (DX * DX + DY * DY + DZ * DZ > r_hd * r_hd) {
            lose(&s, j);
            losing = true;
            need_to_reduce = 1;
        }
        if (losing) ++removed;
        else ++how_many_moonlets;
    }
    how_many_free = s.how_many_free;
    // Centre-of-mass reduction (rk4.c:203-221 with total_momentum,
    // structure.c:1193-1227).
    if (need_to_reduce && reduce_to_com) {
        double com[6];
        const double M0 = cm ? central[6] : 0.;
        double M = M0;
        com[0] = M * (cm ? central[0] : 0.);
        com[1] = M * (cm ? central[1] : 0.);
        com[2] = M * (cm ? central[2] : 0.);
        com[3] = M * (cm ? central[3] : 0.);
        com[4] = M * (cm ? central[4] : 0.);
        com[5] = M * (cm ? central[5] : 0.);
        for (int32_t j = 0; j <= largest_id; ++j) {
            if (!exists[j]) continue;
            const double m = X[j].mass;
            M += m;
            com[0] += m * X[j].x;
            com[1] += m * X[j].y;
            com[2] += m * X[j].z;
            com[3] += m * X[j].vx;
            com[4] += m * X[j].vy;
            com[5] += m * X[j].vz;
        }
        for (int q = 0; q < 6; ++q) com[q] /= M;
        if (cm)
            for (int q = 0; q < 6; ++q) central[q] -= com[q];
        for (int32_t j = 0; j <= largest_id; ++j) {
            if (!exists[j]) continue;
            X[j].x -= com[0];
            X[j].y -= com[1];
            X[j].z -= com[2];
            X[j].vx -= com[3];
            X[j].vy -= com[4];
            X[j].vz -= com[5];
        }
        need_to_reduce = 0;
    }
    // tidy_up (structure.c:881-911) when holes exceed a tenth (rk4.c:396).
    if (how_many_moonlets < 9 * largest_id / 10) {
        int32_t i = 0;
        for (int32_t j = 0; j <= largest_id; ++j) {
            if (exists[j]) {
                if (j > i) X[i] = X[j];
                ++i;
            }
        }
        for (int32_t j = 0; j <= largest_id; ++j) exists[j] = j < i ? 1 : 0;
        largest_id = i - 1;
        how_many_free = 0;
        ++tidied;
    }
    return how_many_moonlets;
}

extern "C" ncg_status ncg_builtin_replay_ops(ncg_builtin_collider *state, ncg_replay_ops *ops)
{
    if (!state || !ops || !state->bodies) return NCG_ERR_INVALID;
    ops->user = state;
    ops->test = builtin_test;
    ops->resolve = builtin_resolve;
    ops->created = nullptr;  // bounces create no bodies
    return NCG_OK;
}
