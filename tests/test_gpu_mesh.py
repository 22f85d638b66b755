"""GPU mesh module (libncg.so) against the compiled reference and the oracle.

mesh() / get_neighbours_mesh(): collision sequences, total_neighbours,
how_many_small/big, final states and the registered pair list must be
identical; the neighbour-cell gravity must match vector_field's mesh branch.
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


def run_mesh(ctx, g):
    r = orc.Replay(g["bodies"], float(g["dt"]), int(g["ctype"]), float(g["f"]), int(g["one_coll"]),
                   g.get("exists"))
    st = ctx.mesh_pass(r.b, r.e if "exists" in g else None, float(g["dt"]), float(g["gam"]),
                       float(g["cube"]), int(g["cells"]), 0, r.test, r.resolve)
    return r, st


@pytest.mark.parametrize("name", ["mesh_inelastic_n1500", "mesh_elastic_once_n1200"])
def test_mesh_pass_matches_reference(ctx, name):
    g = load(name)
    r, st = run_mesh(ctx, g)
    assert np.array_equal(np.array(r.log, dtype=np.int32).reshape(-1, 2), g["coll"])
    assert st["total_neighbours"] == int(g["total_neighbours"])
    assert st["how_many_small"] == int(g["how_many_small"])
    assert st["how_many_big"] == int(g["how_many_big"])
    assert np.array_equal(bits(r.b), bits(g["out_bodies"]))
    pairs = ctx.mesh_pairs()
    assert np.array_equal(pairs, g["pairs"])
    assert len(g["coll"]) > 20


def test_get_neighbours_mesh_matches_reference(ctx):
    g = load("gnm_n1500")
    b = np.array(g["bodies"])
    st = ctx.mesh_pass(b, g["exists"], 0.0, float(g["gam"]), float(g["cube"]), int(g["cells"]), 1)
    assert st["total_neighbours"] == int(g["total_neighbours"])
    assert np.array_equal(ctx.mesh_pairs(), g["pairs"])


def test_mesh_gravity_matches_reference(ctx):
    # Binning of get_neighbours_mesh at the old positions, field at the new ones.
    gn = load("gnm_n1500")
    g = load("vf_mesh_n1500")
    b = np.array(gn["bodies"])
    ctx.mesh_pass(b, gn["exists"], 0.0, float(gn["gam"]), float(gn["cube"]), int(gn["cells"]), 1)
    c8 = g["central"]
    cen = (c8[0], c8[1], c8[2], c8[6])
    ex = g["exists"].astype(bool)
    # The exact form sums in the reference's pair-list order: bit for bit.
    acc, cm = ctx.mesh_gravity(g["bodies"], g["exists"], cen, float(g["G"]), float(g["soft"]),
                               "exact")
    assert np.array_equal(bits(acc[ex]), bits(g["acc"][ex]))
    assert np.array_equal(bits(cm), bits(g["cm_acc"]))
    acc, cm = ctx.mesh_gravity(g["bodies"], g["exists"], cen, float(g["G"]), float(g["soft"]),
                               "fast64")
    ref = g["acc"]
    assert np.abs(acc[ex] - ref[ex]).max() / np.abs(ref[ex]).max() < 1e-12
    assert np.allclose(cm, g["cm_acc"], rtol=1e-12, atol=0)
    # Mutual part alone against the oracle with the reference pair list.
    accn, _ = ctx.mesh_gravity(g["bodies"], g["exists"], None, float(g["G"]), float(g["soft"]),
                               "exact")
    refn, _ = orc.vector_field_mesh(g["bodies"], g["pairs"], g["exists"], None, float(g["G"]),
                                    float(g["soft"]))
    assert np.array_equal(bits(accn[ex]), bits(refn[ex]))
    # Mixed form on the mesh path: every neighbour term is a near one, where
    # FP32 terms lose ~1e-5 to the cancellation of the neighbour sums, so it
    # runs the FP64-fast form (same bits) and meets the 1e-6 bar by far.
    accm, _ = ctx.mesh_gravity(g["bodies"], g["exists"], None, float(g["G"]), float(g["soft"]),
                               "mixed")
    accf, _ = ctx.mesh_gravity(g["bodies"], g["exists"], None, float(g["G"]), float(g["soft"]),
                               "fast64")
    assert np.array_equal(bits(accm[ex]), bits(accf[ex]))
    dm = np.linalg.norm(accm[ex] - refn[ex], axis=1)
    rn = np.linalg.norm(refn[ex], axis=1)
    print("mesh mixed: max rel %.3e" % (dm / rn).max())
    assert (dm / rn).max() < 1e-11


def blob(n, seed, half=0.07, rmin=5e-4, rmax=2e-3, vs=0.3):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-half, half, (n, 3))
    b[:, 3:6] = rng.normal(0, vs, (n, 3))
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = rng.uniform(rmin, rmax, n)
    return b


@pytest.mark.parametrize("seed,cells,gam", [(1, 32, 0.004), (2, 64, 0.0025), (3, 16, 0.01)])
def test_mesh_pass_matches_oracle_random(ctx, seed, cells, gam):
    b = blob(6000, seed)
    cube = cells * gam
    ex = (np.random.default_rng(seed).uniform(size=6000) > 0.05).astype(np.int32)
    ref = orc.mesh(b, 0.01, gam, cube, cells, 1, 1.5, 0, 1, ex)
    r = orc.Replay(b, 0.01, 1, 1.5, 0, ex)
    st = ctx.mesh_pass(r.b, r.e, 0.01, gam, cube, cells, 0, r.test, r.resolve)
    assert np.array_equal(np.array(r.log, dtype=np.int32).reshape(-1, 2), ref["collisions"])
    assert st["total_neighbours"] == ref["total_neighbours"]
    assert st["how_many_small"] == ref["how_many_small"]
    assert np.array_equal(bits(r.b), bits(ref["bodies"]))
    assert np.array_equal(ctx.mesh_pairs(), ref["pairs"])


@pytest.mark.parametrize("seed,cells,gam,one", [(5, 32, 0.004, 0), (6, 16, 0.01, 1)])
def test_mesh_pass_builtin_ops_match_oracle(ctx, seed, cells, gam, one):
    """The library's own decision/outcome (ncg_builtin_replay_ops) gives the
    reference's collisions, counters, states and pairs."""
    b = blob(6000, seed)
    cube = cells * gam
    ex = (np.random.default_rng(seed).uniform(size=6000) > 0.05).astype(np.int32)
    ref = orc.mesh(b, 0.01, gam, cube, cells, 1, 1.5, one, 1, ex)
    bb = b.copy()
    dc = np.zeros(6000, np.int32)
    ops = ctx.builtin_ops(bb, 0.01, exists=ex, did_collide=dc, f=1.5, one_collision_only=one)
    st = ctx.mesh_pass(bb, ex, 0.01, gam, cube, cells, 0, ops=ops)
    assert ops.collisions == len(ref["collisions"]) > 10
    assert st["total_neighbours"] == ref["total_neighbours"]
    assert np.array_equal(bits(bb), bits(ref["bodies"]))
    assert np.array_equal(ctx.mesh_pairs(), ref["pairs"])


def test_mesh_pass_empty_and_outside(ctx):
    b = blob(50, 4, half=1.0)  # most bodies outside a small cube
    st = ctx.mesh_pass(b, None, 0.01, 0.01, 0.16, 16, 1)
    pairs, tn = orc.get_neighbours_mesh(b, 0.01, 0.16, 16)
    assert st["total_neighbours"] == tn
    st = ctx.mesh_pass(np.zeros((0, 8)), None, 0.01, 0.01, 0.16, 16, 1)
    assert st["total_neighbours"] == 0


@pytest.mark.parametrize("seed,mode,nbig,holes,cen,soft", [
    (11, 0, 0, False, True, 0.0), (12, 0, 5, True, True, 0.3), (13, 1, 0, True, False, 0.1),
    (14, 0, 3, False, False, 0.0), (15, 1, 2, False, True, 0.0)])
def test_mesh_gravity_exact_in_pair_list_order(ctx, seed, mode, nbig, holes, cen, soft):
    """Exact mesh gravity = vector_field's mesh branch over the pass's own pair
    list, bit for bit: bodies larger than the mesh size (their visits pair
    them with every body), bodies outside the collision cube, holes (body 0
    absent too), collisions during mesh(), with and without central body."""
    n, cells, gam = 5000, 32, 0.004
    cube = cells * gam  # the blob reaches past the cube: some bodies are not binned
    b = blob(n, seed)
    rng = np.random.default_rng(seed)
    if nbig:
        big = rng.choice(np.arange(1, n), nbig, replace=False)
        big[0] = 1  # a big body among 0..2 too
        b[big, 7] = rng.uniform(1.5 * gam, 3 * gam, nbig)
    ex = np.ones(n, np.int32)
    if holes:
        ex[rng.choice(n, 200, replace=False)] = 0
        ex[0] = 0
    if mode == 0:
        r = orc.Replay(b, 0.01, 1, 1.5, 0, ex)
        ctx.mesh_pass(r.b, ex, 0.01, gam, cube, cells, 0, r.test, r.resolve)
        bb = r.b
        assert len(r.log) > 5
    else:
        bb = b
        ctx.mesh_pass(bb, ex, 0.0, gam, cube, cells, 1)
    pairs = ctx.mesh_pairs()
    # The field is evaluated after a drift, as in the integrator.
    b2 = bb.copy()
    b2[:, 0:3] += 0.002 * b2[:, 3:6]
    central = (0.01, -0.02, 0.005, 1.0) if cen else None
    acc, cm = ctx.mesh_gravity(b2, ex, central, 1.0, soft, "exact")
    ref, rcm = orc.vector_field_mesh(b2, pairs, ex, central, 1.0, soft)
    live = ex != 0
    assert np.array_equal(bits(acc[live]), bits(ref[live]))
    if cen:
        assert np.array_equal(bits(cm), bits(rcm))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7])
def test_mesh_gravity_exact_few_bodies(ctx, n):
    b = blob(n, 20 + n, half=0.01)
    ex = np.ones(n, np.int32)
    ctx.mesh_pass(b, ex, 0.0, 0.004, 0.128, 32, 1)
    pairs = ctx.mesh_pairs()
    central = (0.0, 0.0, 0.0, 1.0)
    acc, cm = ctx.mesh_gravity(b, ex, central, 1.0, 0.05, "exact")
    ref, rcm = orc.vector_field_mesh(b, pairs, ex, central, 1.0, 0.05)
    assert np.array_equal(bits(acc), bits(ref))
    assert np.array_equal(bits(cm), bits(rcm))


class MeshToyWorld:
    """A host world for mesh(): arrays of `cap` slots like NcorpiON's N_max,
    closest_approach as the decision, and test-only outcomes in the four
    shapes of the reference's (physics.c:1043-1091): a bounce registering
    (a, b); a merger into the heavier body registering nothing; a partial
    fragmentation changing both masses and registering (heavier, lighter); a
    full fragmentation keeping the heavier body, removing the other and
    appending two bodies, registering the pairs of its three bodies
    (id[p], id[q]), q < p."""

    def __init__(self, b, cap, dt):
        n = b.shape[0]
        self.b = np.zeros((cap, 8))
        self.b[:n] = b
        self.e = np.zeros(cap, np.int32)
        self.e[:n] = 1
        self.largest = n - 1
        self.born = []
        self.log = []
        self.pairs = []
        self.kinds = set()
        self.dt = dt
        self.visits = []  # (id, 1 if visited as larger than the mesh size)

    def test(self, a, c):
        return orc.closest_approach(self.b, a, c, self.dt, self.e)[0]

    def resolve(self, a, c):
        b = self.b
        A, B = (a, c) if b[a, 6] > b[c, 6] else (c, a)
        if not (self.e[a] and self.e[c]):
            return
        self.log.append((a, c))
        kind = (7 * a + c) % 4
        self.kinds.add(kind)
        m = b[A, 6] + b[B, 6]
        if kind == 0:
            va = b[a, 3:6].copy()
            b[a, 3:6] = b[c, 3:6]
            b[c, 3:6] = va
            self.pairs.append((a, c))
        elif kind == 1:
            b[A, 3:6] = (b[A, 6] * b[A, 3:6] + b[B, 6] * b[B, 3:6]) / m
            b[A, 6] = m
            b[A, 7] = np.cbrt(b[A, 7] ** 3 + b[B, 7] ** 3)
            self.e[B] = 0
        elif kind == 2:
            # Radii follow the masses (constant density), so a body may move
            # across the mesh size before its own visit.
            b[A, 7] *= np.cbrt(0.75 * m / b[A, 6])
            b[B, 7] *= np.cbrt(0.25 * m / b[B, 6])
            b[A, 6], b[B, 6] = 0.75 * m, 0.25 * m
            b[A, 3:6], b[B, 3:6] = b[B, 3:6].copy(), b[A, 3:6].copy()
            self.pairs.append((A, B))
        else:
            ids = [A]
            for q in range(2):
                s = self.largest + 1
                self.largest = s
                b[s] = b[A]
                b[s, 0] += (3.0 + 3.0 * q) * b[A, 7]
                b[s, 3:6] = -b[A, 3:6]
                b[s, 6] = 0.1 * m
                self.e[s] = 1
                self.born.append(s)
                ids.append(s)
            b[A, 6] = 0.8 * m
            self.e[B] = 0
            for p in range(len(ids)):
                for q in range(p):
                    self.pairs.append((ids[p], ids[q]))

    def created(self):
        out, self.born = self.born, []
        return out


def toy_mesh(w, n0, gam, cube, cells, with_big=False):
    """mesh(), collision.c:324-417: bodies visited by ascending id with their
    live state.  A body at most the mesh size (radius read at its visit)
    joins the hash, which holds each cell's visited bodies in insertion
    order; its visit gathers the existing bodies of its 27 cells (newest
    first per cell) and pops them last-in first-out (collision.c:264-316,
    structure.c:712-841).  A larger one (with_big) tests every existing body
    up to the live largest id that is smaller than the mesh size or larger
    with a larger id (collision.c:381-414).  Returns (total_neighbours,
    how_many_small[, how_many_big])."""
    cells_of = {}
    total = small = big = 0
    for k in range(n0):
        if not w.e[k]:
            continue
        if with_big and w.b[k, 7] > gam:
            big += 1
            if hasattr(w, "visits"):
                w.visits.append((k, 1))
            p = 0
            while p <= w.largest:
                if w.e[p] and p != k:
                    Rp = w.b[p, 7]
                    if Rp < gam or (Rp >= gam and p > k):
                        if w.test(p, k):
                            w.resolve(p, k)
                        else:
                            w.pairs.append((k, p))
                p += 1
            continue
        small += 1
        if with_big and hasattr(w, "visits"):
            w.visits.append((k, 0))
        idx = orc.cell_indexes(w.b[k, 0], w.b[k, 1], w.b[k, 2], gam, cube, cells)
        if idx is None:
            continue
        nghb = []
        for l in range(27):
            for q in reversed(cells_of.get(int(idx[l]), [])):
                if w.e[q]:
                    nghb.append(q)
        cells_of.setdefault(int(idx[13]), []).append(k)
        while nghb:
            p = nghb.pop()
            if w.test(p, k):
                w.resolve(p, k)
            else:
                w.pairs.append((k, p))
            total += 1
    return (total, small, big) if with_big else (total, small)


@pytest.mark.parametrize("seed", [31, 32])
def test_mesh_pass_with_removed_and_appended_bodies(ctx, seed):
    """Outcomes that remove bodies and append new ones, as mergers and
    fragmentation do in mesh mode: the replay gives mesh()'s collision
    sequence, counters, final state and pair list, and the exact field over
    the appended slots is vector_field's over that list, bit for bit."""
    n0, cap, dt, cells, gam = 3000, 6000, 0.01, 16, 0.01
    cube = cells * gam
    rng = np.random.default_rng(seed)
    b = blob(n0, seed, half=0.06, rmin=1.5e-3, rmax=3e-3)
    b[:, 6] = rng.uniform(1e-9, 1e-8, n0)
    ref = MeshToyWorld(b, cap, dt)
    total, small = toy_mesh(ref, n0, gam, cube, cells)
    got = MeshToyWorld(b, cap, dt)
    st = ctx.mesh_pass(got.b, got.e, dt, gam, cube, cells, 0, got.test, got.resolve,
                       created=got.created, n=n0)
    assert got.log == ref.log
    assert ref.kinds == {0, 1, 2, 3} and ref.largest >= n0
    assert st["collisions"] == len(ref.log)
    assert st["total_neighbours"] == total
    assert st["how_many_small"] == small
    assert np.array_equal(got.e, ref.e)
    live = ref.e != 0
    assert np.array_equal(bits(got.b[live]), bits(ref.b[live]))
    pairs = ctx.mesh_pairs()
    assert np.array_equal(pairs, np.array(ref.pairs, np.int32).reshape(-1, 2))
    # The field after a drift over the n0 slots and the appended ones.
    ne = ref.largest + 1
    b2 = ref.b[:ne].copy()
    b2[:, 0:3] += 0.002 * b2[:, 3:6]
    ex = ref.e[:ne].copy()
    central = (0.01, -0.02, 0.005, 1.0)
    acc, cm = ctx.mesh_gravity(b2, ex, central, 1.0, 0.1, "exact")
    racc, rcm = orc.vector_field_mesh(b2, pairs, ex, central, 1.0, 0.1)
    lv = ex != 0
    assert np.array_equal(bits(acc[lv]), bits(racc[lv]))
    assert np.array_equal(bits(cm), bits(rcm))
    acc, cm = ctx.mesh_gravity(b2, ex, central, 1.0, 0.1, "fast64")
    assert np.abs(acc[lv] - racc[lv]).max() / np.abs(racc[lv]).max() < 1e-12


def test_mesh_pass_with_big_bodies_and_all_outcomes(ctx):
    """Bodies larger than the mesh size together with outcomes that remove,
    create and resize bodies (collision.c:381-414 with physics.c:628-1040's
    shapes): a big visit scans every body up to the live largest id, so bodies
    removed earlier in the pass leave it and appended ones join it; a body an
    outcome moves across the mesh size before its own visit is visited in its
    new class.  Collision sequence, counters, final state, pair list and the
    exact field over the pass's own list (appended slots included) equal the
    restatement's (pinned to the oracle's mesh() with big bodies on the CPU)."""
    n0, cap, dt, cells, gam = 3000, 9000, 0.01, 16, 0.01
    cube = cells * gam
    seen = {"big_coll": 0, "to_big": 0, "to_small": 0, "appended_in_big": 0}
    for seed in (41, 42, 43):
        rng = np.random.default_rng(seed)
        b = blob(n0, seed, half=0.06, rmin=1.5e-3, rmax=3e-3)
        b[:, 6] = rng.uniform(1e-9, 1e-8, n0)
        near = rng.choice(np.arange(n0), 80, replace=False)
        b[near, 7] = rng.uniform(0.85 * gam, 1.3 * gam, 80)
        b[near, 6] = rng.uniform(1e-8, 4e-8, 80)
        start_big = set(np.flatnonzero(b[:, 7] > gam).tolist())
        ref = MeshToyWorld(b, cap, dt)
        total, small, big = toy_mesh(ref, n0, gam, cube, cells, with_big=True)
        got = MeshToyWorld(b, cap, dt)
        print("big bodies, seed %d: %d collisions, %d big visits" % (seed, len(ref.log), big),
              flush=True)
        st = ctx.mesh_pass(got.b, got.e, dt, gam, cube, cells, 0, got.test, got.resolve,
                           created=got.created, n=n0)
        assert got.log == ref.log
        assert st["collisions"] == len(ref.log)
        assert (st["total_neighbours"], st["how_many_small"], st["how_many_big"]) == \
            (total, small, big)
        assert np.array_equal(got.e, ref.e)
        live = ref.e != 0
        assert np.array_equal(bits(got.b[live]), bits(ref.b[live]))
        pairs = ctx.mesh_pairs()
        assert np.array_equal(pairs, np.array(ref.pairs, np.int32).reshape(-1, 2))
        ne = ref.largest + 1
        b2 = ref.b[:ne].copy()
        b2[:, 0:3] += 0.002 * b2[:, 3:6]
        ex = ref.e[:ne].copy()
        central = (0.01, -0.02, 0.005, 1.0)
        for soft in (0.0, 0.1):
            acc, cm = ctx.mesh_gravity(b2, ex, central, 1.0, soft, "exact")
            racc, rcm = orc.vector_field_mesh(b2, pairs, ex, central, 1.0, soft)
            lv = ex != 0
            assert np.array_equal(bits(acc[lv]), bits(racc[lv]))
            assert np.array_equal(bits(cm), bits(rcm))
        acc, _ = ctx.mesh_gravity(b2, ex, central, 1.0, 0.1, "fast64")
        assert np.abs(acc[lv] - racc[lv]).max() / np.abs(racc[lv]).max() < 1e-12
        bigv = {k for k, c in ref.visits if c == 1}
        seen["big_coll"] += sum(1 for a, c in ref.log if c in bigv or a in bigv)
        seen["to_big"] += len(bigv - start_big)
        seen["to_small"] += len({k for k, c in ref.visits if c == 0} & start_big)
        seen["appended_in_big"] += sum(1 for a, c in ref.pairs if c >= n0 and a in bigv)
    print(seen)
    assert all(v > 0 for v in seen.values()), seen



def test_mesh_start_values_equal_central_term(ctx):
    """ncg_mesh_gravity_from with the central term as start values equals
    ncg_mesh_gravity with the central body: bit for bit in the exact form,
    to 1e-12 in the FP64-fast one."""
    gn = load("gnm_n1500")
    g = load("vf_mesh_n1500")
    b = np.array(gn["bodies"])
    ctx.mesh_pass(b, gn["exists"], 0.0, float(gn["gam"]), float(gn["cube"]), int(gn["cells"]), 1)
    c8 = g["central"]
    cen = (c8[0], c8[1], c8[2], c8[6])
    ex = g["exists"].astype(bool)
    G = float(g["G"])
    bb = np.array(g["bodies"])
    mu = G * cen[3]
    DX, DY, DZ = bb[:, 0] - cen[0], bb[:, 1] - cen[1], bb[:, 2] - cen[2]
    rk = np.sqrt(DX * DX + DY * DY + DZ * DZ)
    r3 = rk * rk * rk
    st = np.stack([-mu * DX / r3, -mu * DY / r3, -mu * DZ / r3], axis=1)
    st[~ex] = 0.0
    for prec in ("exact", "fast64"):
        ref, _ = ctx.mesh_gravity(g["bodies"], g["exists"], cen, G, float(g["soft"]), prec)
        got = ctx.mesh_gravity_from(g["bodies"], st, g["exists"], G, float(g["soft"]), prec)
        if prec == "exact":
            assert np.array_equal(bits(got[ex]), bits(ref[ex]))
        else:
            assert np.abs(got[ex] - ref[ex]).max() / np.abs(ref[ex]).max() < 1e-12
