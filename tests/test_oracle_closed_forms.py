"""The oracle against what the mathematics fixes (closed forms and invariants).

These checks are independent of the reference build: a dropped term, a wrong
sign or a swapped index in the restatement fails one of them.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic


def test_two_body_central_field():
    # One body at distance r from the central mass: a = -G M x / r^3 exactly.
    b = np.array([[3.0, 4.0, 0.0, 0, 0, 0, 1e-6, 1e-3]])
    acc, cm = orc.vector_field_brute(b, central=(0, 0, 0, 2.0), G=1.5)
    r = 5.0
    assert np.allclose(acc[0], [-1.5 * 2.0 * 3.0 / r ** 3, -1.5 * 2.0 * 4.0 / r ** 3, 0.0],
                       rtol=1e-15, atol=0)
    assert np.allclose(cm, [1.5 * 1e-6 * 3.0 / r ** 3, 1.5 * 1e-6 * 4.0 / r ** 3, 0.0],
                       rtol=1e-15, atol=0)


def test_pair_force_with_softening():
    # Two bodies, no central mass: a_0 = G m_1 d / (d^2 + eps^2)^(3/2), eps = s (R0 + R1).
    b = np.array([[0.0, 0, 0, 0, 0, 0, 2.0, 0.1], [1.0, 0, 0, 0, 0, 0, 3.0, 0.2]])
    acc, _ = orc.vector_field_brute(b, G=0.5, soft=2.0)
    eps = 2.0 * 0.3
    D3 = (1.0 + eps * eps) ** 1.5
    assert acc[0, 0] == pytest.approx(0.5 * 3.0 / D3, rel=1e-15)
    assert acc[1, 0] == pytest.approx(-0.5 * 2.0 / D3, rel=1e-15)


def test_newton_third_law():
    # sum_k m_k a_k + M * CM_acc = 0 (the central body's reaction balances).
    b = ic.disk(300, seed=3, mass=1e-5, radius=1e-3)
    acc, cm = orc.vector_field_brute(b, central=(0.01, -0.02, 0.0, 1.0), G=1.0, soft=0.01)
    tot = (b[:, 6:7] * acc).sum(0) + 1.0 * cm
    assert np.max(np.abs(tot)) < 1e-13 * np.max(np.abs(b[:, 6:7] * acc))


def test_holes_are_ignored():
    b = ic.disk(50, seed=4, mass=1e-4, radius=1e-3)
    ex = np.ones(50, np.int32)
    ex[[3, 17, 40]] = 0
    acc, _ = orc.vector_field_brute(b, ex, central=(0, 0, 0, 1.0))
    keep = ex.astype(bool)
    acc2, _ = orc.vector_field_brute(b[keep], None, central=(0, 0, 0, 1.0))
    assert np.allclose(acc[keep], acc2, rtol=1e-13, atol=0)


def head_on(gap, v, R=0.1):
    return np.array([[0.0, 0, 0, v / 2, 0, 0, 2.0, R / 2], [gap + R, 0, 0, -v / 2, 0, 0, 1.0, R / 2]])


def test_closest_approach_head_on():
    # Head-on: contact after t = gap / |dv|.
    b = head_on(0.3, 2.0)
    hit, tc, ap = orc.closest_approach(b, 0, 1, 1.0)
    assert hit and tc == pytest.approx(0.15, rel=1e-14)
    assert ap[0] == pytest.approx(0.15, rel=1e-14)          # heavier body first
    assert ap[3] == pytest.approx(0.4 - 0.15, rel=1e-14)
    hit, _, _ = orc.closest_approach(b, 0, 1, 0.1)           # contact after the step
    assert not hit


def test_closest_approach_receding_tangent_overlap():
    b = head_on(0.3, -2.0)                                   # receding
    assert not orc.closest_approach(b, 0, 1, 10.0)[0]
    # Tangent fly-by: minimum distance exactly R -> discriminant 0 -> miss.
    b = np.array([[0.0, 0, 0, 0, 0, 0, 1.0, 0.25], [-1.0, 0.5, 0, 1.0, 0, 0, 1.0, 0.25]])
    assert not orc.closest_approach(b, 0, 1, 10.0)[0]
    # Already overlapping and approaching: t_c < 0, reported as a collision.
    b = np.array([[0.0, 0, 0, 0, 0, 0, 1.0, 0.2], [0.3, 0, 0, -1.0, 0, 0, 1.0, 0.2]])
    hit, tc, _ = orc.closest_approach(b, 0, 1, 0.01)
    assert hit and tc == pytest.approx(-0.1, rel=1e-14)


def test_inelastic_bounce_conserves_momentum():
    b = head_on(0.01, 2.0)
    out, _, pairs = orc.brute_force(b, 1.0, ctype=1, f=1.5)
    assert pairs.tolist() == [[0, 1]]
    p0 = (b[:, 6:7] * b[:, 3:6]).sum(0)
    p1 = (out[:, 6:7] * out[:, 3:6]).sum(0)
    assert np.allclose(p0, p1, rtol=0, atol=1e-15)
    # Restitution e = f - 1: relative normal velocity flips and scales by 0.5.
    assert (out[0, 3] - out[1, 3]) == pytest.approx(-0.5 * (b[0, 3] - b[1, 3]), rel=1e-14)


def test_mesh_neighbour_count_matches_stencil_bruteforce():
    # total_neighbours from get_neighbours_mesh = number of ordered (k, p<k)
    # pairs whose cell keys differ by a 27-stencil offset (brute force over pairs).
    rng = np.random.default_rng(5)
    n, cells, gam = 400, 8, 0.02
    cube = cells * gam
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-0.085, 0.085, (n, 3))
    b[:, 7] = 1e-3
    _, tn = orc.get_neighbours_mesh(b, gam, cube, cells)
    d = cells
    keys = {}
    for k in range(n):
        x = b[k, :3]
        if np.any(np.abs(x) > cube / 2):
            continue
        i, j, l = (d // 2 + 1 + int(math.floor(c / gam)) for c in x)
        keys[k] = i + j * d + l * d * d
    offs = {di + dj * d + dl * d * d for di in (-1, 0, 1) for dj in (-1, 0, 1) for dl in (-1, 0, 1)}
    cnt = sum(1 for k in keys for p in keys if p < k and keys[p] - keys[k] in offs)
    assert tn == cnt


def _kepler_state(a, e, mu, t):
    n = math.sqrt(mu / a ** 3)
    M = n * t
    E = M
    for _ in range(60):
        E -= (E - e * math.sin(E) - M) / (1 - e * math.cos(E))
    r = a * (1 - e * math.cos(E))
    x, y = a * (math.cos(E) - e), a * math.sqrt(1 - e * e) * math.sin(E)
    vx, vy = -a * n * math.sin(E) * a / r, a * math.sqrt(1 - e * e) * n * math.cos(E) * a / r
    return np.array([x, y, 0, vx, vy, 0])


def test_rk4_two_body_and_order():
    # A massless-ish body on a Kepler orbit around a fixed (huge) central mass:
    # RK4 error scales as dt^4.
    mu, a, e = 1.0, 1.0, 0.1
    T = 2 * math.pi
    errs = []
    for nsteps in (64, 128):
        dt = T / 4 / nsteps
        b = np.zeros((1, 8))
        b[0, :6] = _kepler_state(a, e, mu, 0.0)
        b[0, 6] = 1e-300
        b[0, 7] = 1e-3
        c7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
        for _ in range(nsteps):
            b, c7 = orc.rk4_step(b, c7, 1.0, 0.0, dt)
        ref = _kepler_state(a, e, mu, T / 4)
        errs.append(np.max(np.abs(b[0, :6] - ref)))
    order = math.log2(errs[0] / errs[1])
    assert 3.7 < order < 4.3, (errs, order)


def test_rk4_energy_conservation():
    # Light disk (no close encounters): energy error is small and shrinks by
    # about 2^4-2^5 when dt halves.
    drift = []
    for dt in (2 * math.pi / 256, 2 * math.pi / 512):
        b = ic.disk(40, seed=7, mass=1e-9, radius=1e-3)
        c7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
        E0 = orc.energy(b, c7, 1.0, 0.01)
        for _ in range(int(round(50 * (2 * math.pi / 256) / dt))):
            b, c7 = orc.rk4_step(b, c7, 1.0, 0.01, dt)
        drift.append(abs(orc.energy(b, c7, 1.0, 0.01) - E0) / abs(E0))
    assert drift[0] < 1e-8
    assert drift[0] / drift[1] > 12.0
