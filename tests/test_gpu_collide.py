"""GPU brute-force collision search + host replay against the reference.

The GPU candidate set must contain every pair closest_approach reports at the
start-of-pass state; the replay through the library must reproduce the
reference's collision sequence and final state bit for bit.
"""
import os

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def cluster(n, seed, box=0.05, radius=2e-3, vscale=1.0):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-box / 2, box / 2, (n, 3)) + np.array([1.0, 0.0, 0.0])
    b[:, 3:6] = rng.normal(0.0, vscale, (n, 3))
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = radius * rng.uniform(0.5, 1.5, n)
    return b


def test_candidates_equal_reference_hits(ctx):
    g = load("ca_n48")
    cand = ctx.collide_candidates(g["bodies"], None, float(g["dt"]))
    exact = g["ab"][g["hit"].astype(bool)]
    assert np.array_equal(cand, exact)


@pytest.mark.parametrize("seed,dt", [(1, 0.01), (2, -0.01), (3, 0.005)])
def test_candidates_superset_random(ctx, seed, dt):
    b = cluster(700, seed, box=0.08)
    ex = (np.random.default_rng(seed).uniform(size=700) > 0.1).astype(np.int32)
    cand = {tuple(p) for p in ctx.collide_candidates(b, ex, dt)}
    hits = set()
    for i in range(700):
        for j in range(i + 1, 700):
            if orc.closest_approach(b, i, j, dt, ex)[0]:
                hits.add((i, j))
    assert hits <= cand
    assert len(cand - hits) <= 2  # only last-bit boundary cases may be extra
    if dt > 0:
        assert len(hits) > 10


@pytest.mark.parametrize("name", ["bf_inelastic_n300", "bf_elastic_once_n200",
                                  "bf_inelastic125_n250"])
def test_replay_matches_reference(ctx, name):
    g = load(name)
    r = orc.Replay(g["bodies"], float(g["dt"]), int(g["ctype"]), float(g["f"]), int(g["one_coll"]),
                   g.get("exists"))
    n = ctx.brute_force_pass(r.b, r.e, float(g["dt"]), r.test, r.resolve)
    assert n == len(g["coll"])
    assert np.array_equal(np.array(r.log, dtype=np.int32).reshape(-1, 2), g["coll"])
    assert np.array_equal(bits(r.b), bits(g["out_bodies"]))


@pytest.mark.parametrize("seed", [21, 22])
def test_replay_matches_oracle_dense(ctx, seed):
    b = cluster(1200, seed, box=0.06, radius=2.5e-3)
    out, _, pairs = orc.brute_force(b, 0.01, 1, 1.5, 0)
    r = orc.Replay(b, 0.01, 1, 1.5, 0)
    ctx.brute_force_pass(r.b, r.e, 0.01, r.test, r.resolve)
    assert np.array_equal(np.array(r.log, dtype=np.int32).reshape(-1, 2), pairs)
    assert np.array_equal(bits(r.b), bits(out))
    assert len(pairs) > 100


@pytest.mark.parametrize("seed,f,one", [(23, 1.5, 0), (24, 2.0, 1)])
def test_replay_builtin_ops_match_oracle(ctx, seed, f, one):
    """ncg_builtin_replay_ops: the library's restatement of closest_approach
    and the bounce gives the reference's collisions and final state."""
    b = cluster(1200, seed, box=0.06, radius=2.5e-3)
    out, _, pairs = orc.brute_force(b, 0.01, 1 if f != 2.0 else 0, f, one)
    bb = b.copy()
    dc = np.zeros(len(b), np.int32)
    ops = ctx.builtin_ops(bb, 0.01, did_collide=dc, f=f, one_collision_only=one)
    nc = ctx.brute_force_pass(bb, None, 0.01, ops=ops)
    assert nc == ops.collisions == len(pairs) > 50
    assert np.array_equal(bits(bb), bits(out))


def test_replay_sparse_collisions_few_host_tests(ctx):
    # A handful of collisions: the host only re-tests rows/columns of the
    # bodies involved, not all n(n-1)/2 pairs.
    b = cluster(3000, 31, box=0.5, radius=2e-3)
    out, _, pairs = orc.brute_force(b, 0.01, 1, 1.5, 0)
    r = orc.Replay(b, 0.01, 1, 1.5, 0)
    ctx.brute_force_pass(r.b, r.e, 0.01, r.test, r.resolve)
    assert np.array_equal(np.array(r.log, dtype=np.int32).reshape(-1, 2), pairs)
    assert np.array_equal(bits(r.b), bits(out))
    assert 0 < len(pairs) < 50
    assert r.tests < 4 * 3000 * (len(pairs) + 1)


class ToyWorld:
    """A host world whose outcomes remove and create bodies (test-only rules:
    merge the second body into the first, swap velocities, or split off a new
    body at a free slot, reused last-freed-first, or appended), with arrays
    of `cap` slots like NcorpiON's N_max."""

    def __init__(self, b, cap, dt):
        n = b.shape[0]
        self.b = np.zeros((cap, 8))
        self.b[:n] = b
        self.e = np.zeros(cap, np.int32)
        self.e[:n] = 1
        self.largest = n - 1
        self.free = []
        self.born = []
        self.log = []
        self.dt = dt

    def test(self, a, c):
        return orc.closest_approach(self.b, a, c, self.dt, self.e)[0]

    def resolve(self, a, c):
        if not (self.e[a] and self.e[c]):
            return  # as collision_treatment (physics.c:1059-1061)
        self.log.append((a, c))
        b = self.b
        kind = (7 * a + c) % 3
        if kind == 0:
            m = b[a, 6] + b[c, 6]
            b[a, 3:6] = (b[a, 6] * b[a, 3:6] + b[c, 6] * b[c, 3:6]) / m
            b[a, 6] = m
            self.e[c] = 0
            self.free.append(c)
        elif kind == 1:
            va = b[a, 3:6].copy()
            b[a, 3:6] = b[c, 3:6]
            b[c, 3:6] = va
        else:
            s = self.free.pop() if self.free else self.largest + 1
            self.largest = max(self.largest, s)
            b[s] = b[a]
            b[s, 0] += 3.0 * b[a, 7]
            b[s, 3:6] = -b[a, 3:6]
            b[s, 6] *= 0.5
            b[a, 6] *= 0.5
            self.e[s] = 1
            self.born.append(s)

    def created(self):
        out, self.born = self.born, []
        return out


def toy_brute_force(w, n0):
    """collision.c:432-455 with the live largest_id for the columns."""
    for i in range(n0):
        if not w.e[i]:
            continue
        j = i + 1
        while j <= w.largest:
            if w.e[j] and w.test(i, j):
                w.resolve(i, j)
            j += 1


@pytest.mark.parametrize("seed", [11, 12])
def test_replay_with_removed_and_created_bodies(ctx, seed):
    """Outcomes that remove bodies and create new ones at free or appended
    slots, as mergers and fragmentation do: the replay gives the reference
    loop's collision sequence and final state."""
    n0, cap, dt = 400, 1200, 0.01
    b = cluster(n0, seed, box=0.06, radius=2.5e-3)
    ref = ToyWorld(b, cap, dt)
    toy_brute_force(ref, n0)
    got = ToyWorld(b, cap, dt)
    nc = ctx.brute_force_pass(got.b, got.e, dt, got.test, got.resolve, created=got.created, n=n0)
    assert got.log == ref.log
    assert nc == len(ref.log)
    assert np.array_equal(got.e, ref.e)
    assert np.array_equal(bits(got.b[got.e != 0]), bits(ref.b[ref.e != 0]))
    kinds = {(7 * a + c) % 3 for a, c in ref.log}
    assert kinds == {0, 1, 2} and ref.largest >= n0  # removals, reuse and appends happened
