// Built-in host collision decision and outcome for callers that are not
// NcorpiON itself (Python users, tests).  It restates, for one
// body array, closest_approach (collision.c:63-172) and collision_treatment
// with elastic (f = 2) or inelastic (f = collision_parameter) outcomes
// (physics.c:545-605, 1043-1070), in the same FP64 operation order.  The
// drop-in glue does not use it: there the reference's own functions decide.
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

}  // namespace

extern "C" ncg_status ncg_builtin_replay_ops(ncg_builtin_collider *state, ncg_replay_ops *ops)
{
    if (!state || !ops || !state->bodies) return NCG_ERR_INVALID;
    ops->user = state;
    ops->test = builtin_test;
    ops->resolve = builtin_resolve;
    ops->created = nullptr;  // bounces create no bodies
    return NCG_OK;
}
