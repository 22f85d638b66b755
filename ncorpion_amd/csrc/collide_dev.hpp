// Device restatement of the reference's collision decision and two-body
// outcome, shared by the device replays (replay.hip: brute_force's order,
// mesh.hip: mesh()'s order): closest_approach (collision.c:63-172) and the
// elastic / inelastic bounce of collision (physics.c:545-605).  Every
// translation unit that includes this header is compiled without FMA
// contraction: the arithmetic is the reference's, operation by operation.
#pragma once
#include "ncg_internal.hpp"

#pragma clang fp contract(off)

// closest_approach on bodies a < c: hit when the pair approaches
// (dr.dv < 0), meets (disc > 0) and meets within dt (t_c <= dt); ap receives
// the positions at closest approach, heavier body first.
__device__ inline bool bf_approach(const DBody *__restrict__ b, int a, int c, double dt, double *ap,
                            double *tc)
{
    const DBody A = b[a], B = b[c];
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double dvx = A.vx - B.vx, dvy = A.vy - B.vy, dvz = A.vz - B.vz;
    const double dr_dot_dv = dx * dvx + dy * dvy + dz * dvz;
    if (dr_dot_dv >= 0.0) return false;
    const double dr2 = dx * dx + dy * dy + dz * dz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double R = A.radius + B.radius;
    const double disc = dr_dot_dv * dr_dot_dv + dv2 * (R * R - dr2);
    if (disc <= 0.0) return false;
    const double t = -(dr_dot_dv + sqrt(disc)) / dv2;
    *tc = t;
    if (t > dt) return false;
    const double xa = A.x + t * A.vx, ya = A.y + t * A.vy, za = A.z + t * A.vz;
    const double xb = B.x + t * B.vx, yb = B.y + t * B.vy, zb = B.z + t * B.vz;
    if (A.mass > B.mass) {
        ap[0] = xa; ap[1] = ya; ap[2] = za; ap[3] = xb; ap[4] = yb; ap[5] = zb;
    } else {
        ap[0] = xb; ap[1] = yb; ap[2] = zb; ap[3] = xa; ap[4] = ya; ap[5] = za;
    }
    return true;
}

// The bounce from the approach of the same pair: impulse along the line of
// centres at contact, then positions rewound to x_c - t_c v'.
__device__ inline void bf_bounce(DBody *__restrict__ b, int a, int c, double f, const double *ap, double t)
{
    const int ia = b[a].mass > b[c].mass ? a : c;
    const int ib = ia == a ? c : a;
    DBody P = b[ia], Q = b[ib];
    const double R = P.radius + Q.radius;
    const double m_a = P.mass, m_b = Q.mass;
    double vx_a = P.vx, vy_a = P.vy, vz_a = P.vz, vx_b = Q.vx, vy_b = Q.vy, vz_b = Q.vz;
    const double xa = ap[0], ya = ap[1], za = ap[2], xb = ap[3], yb = ap[4], zb = ap[5];
    const double dr_dot_dv = (xa - xb) * (vx_a - vx_b) + (ya - yb) * (vy_a - vy_b) + (za - zb) * (vz_a - vz_b);
    const double alpha = f * m_a * m_b / ((m_a + m_b) * R * R);
    const double Jx = alpha * dr_dot_dv * (xb - xa);
    const double Jy = alpha * dr_dot_dv * (yb - ya);
    const double Jz = alpha * dr_dot_dv * (zb - za);
    vx_a += Jx / m_a;
    vy_a += Jy / m_a;
    vz_a += Jz / m_a;
    vx_b -= Jx / m_b;
    vy_b -= Jy / m_b;
    vz_b -= Jz / m_b;
    P.x = xa - t * vx_a; P.y = ya - t * vy_a; P.z = za - t * vz_a;
    P.vx = vx_a; P.vy = vy_a; P.vz = vz_a;
    Q.x = xb - t * vx_b; Q.y = yb - t * vy_b; Q.z = zb - t * vz_b;
    Q.vx = vx_b; Q.vy = vy_b; Q.vz = vz_b;
    b[ia] = P;
    b[ib] = Q;
}

// The reference's decision for pair (a, c): both bodies exist, the approach
// hits within dt and neither collided earlier in this step.
__device__ inline bool bf_hit(const DBody *__restrict__ b, const int32_t *ex, const int32_t *dc, int a,
                       int c, double dt, double *ap, double *tc)
{
    if (ex && !(ex[a] && ex[c])) return false;
    if (!bf_approach(b, a, c, dt, ap, tc)) return false;
    return !dc || (!dc[a] && !dc[c]);
}

