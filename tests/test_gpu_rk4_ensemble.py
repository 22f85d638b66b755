"""Device RK4 driver and batched ensembles against the FP64 oracle.

NcorpiON has no RK4 and no ensembles; these are pinned by the oracle's RK4
(4th-order slope and two-body closed form, tests/test_oracle_closed_forms.py),
by per-system equality with single-system runs, and by graph replays being
bit-identical to eager launches.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic

pytestmark = pytest.mark.gpu


def torch_dev():
    import torch
    return torch


def small_disk(n, seed, mass=1e-7, radius=1e-3):
    return ic.disk(n, seed=seed, mass=mass, radius=radius)


@pytest.mark.parametrize("prec,tol", [("fast64", 1e-12), ("mixed", 2e-7)])
def test_rk4_driver_matches_oracle(ctx, prec, tol):
    torch = torch_dev()
    n, dt, steps = 600, 2 * math.pi / 256, 5
    b = small_disk(n, 3)
    c7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
    ref, rc = b.copy(), c7.copy()
    for _ in range(steps):
        ref, rc = orc.rk4_step(ref, rc, 1.0, 0.01, dt)
    bt = torch.from_numpy(b).cuda()
    ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.rk4_steps_dev(bt, ct, steps, dt, 1.0, 0.01, prec, graph=False)
    ctx.synchronize()
    out = bt.cpu().numpy()
    # Displacement over the run vs its error.
    moved = np.abs(ref[:, :6] - b[:, :6]).max()
    err = np.abs(out[:, :6] - ref[:, :6]).max()
    assert err < tol * moved, (err, moved)
    cen = ct.cpu().numpy()
    assert np.allclose(cen[:6], rc[:6], rtol=0, atol=tol * max(1e-12, np.abs(rc[:6]).max()) + 1e-18)


def test_rk4_graph_replays_bitwise_equal_to_eager(ctx):
    torch = torch_dev()
    n, dt = 3000, 2 * math.pi / 256
    b = small_disk(n, 4)
    res = []
    for graph in (False, True):
        bt = torch.from_numpy(b).cuda()
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(3):  # several replays with host synchronisation between them
            ctx.rk4_steps_dev(bt, ct, 2, dt, 1.0, 1e-3, "mixed", graph=graph)
            ctx.synchronize()
        res.append((bt.cpu().numpy(), ct.cpu().numpy()))
    assert np.array_equal(res[0][0].view(np.uint64), res[1][0].view(np.uint64))
    assert np.array_equal(res[0][1].view(np.uint64), res[1][1].view(np.uint64))


def make_ensemble(sizes, seed0=0, mass=1e-7, radius=1e-3):
    systems = [small_disk(int(nn), seed0 + s, mass, radius) for s, nn in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum([len(x) for x in systems])]).astype(np.int32)
    return systems, np.concatenate(systems), off


@pytest.mark.parametrize("prec", ["fast64", "mixed"])
def test_ensemble_gravity_matches_per_system_oracle(ctx, prec):
    torch = torch_dev()
    sizes = [1, 7, 64, 255, 256, 257, 513, 1000, 2049, 3001]
    systems, allb, off = make_ensemble(sizes)
    S = len(sizes)
    cen = np.zeros((S, 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    cen[::2, 7] = 0.0  # half the systems without a central body
    bt = torch.from_numpy(allb).cuda()
    ct = torch.from_numpy(cen).cuda()
    at = torch.zeros((len(allb), 3), dtype=torch.float64, device="cuda")
    cm = torch.zeros((S, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.ensemble_gravity_dev(bt, off, ct, at, cm, 1.0, 0.01, prec)
    ctx.synchronize()
    acc, cma = at.cpu().numpy(), cm.cpu().numpy()
    for s, sysb in enumerate(systems):
        present = cen[s, 7] != 0
        ref, rcm = orc.vector_field_brute(sysb, None, (0, 0, 0, 1.0) if present else None, 1.0, 0.01)
        got = acc[off[s]:off[s + 1]]
        if present:
            assert np.allclose(cma[s], rcm, rtol=1e-12, atol=0)
        if len(sysb) > 1:
            # Errors are measured against the mutual part of the field (the
            # central term is FP64 in every form), plus the rounding of the
            # oracle's own running sum, which adds each small mutual term to
            # an accumulator of the central term's size.
            mut, _ = orc.vector_field_brute(sysb, None, None, 1.0, 0.01)
            tol = 1e-12 if prec == "fast64" else 1e-6
            bound = (tol * np.linalg.norm(mut, axis=1)
                     + (4e-16 * math.sqrt(len(sysb)) + 1e-15) * np.linalg.norm(ref, axis=1))
            err = np.linalg.norm(got - ref, axis=1)
            assert np.all(err <= bound), (err / bound).max()


@pytest.mark.parametrize("prec,tol", [("fast64", 1e-12), ("mixed", 1e-6)])
def test_ensemble_no_softening_with_large_gm(ctx, prec, tol):
    # Without softening a body's pair with itself has r = 0; with G m up to
    # 1e3 times the FP32 form's overflow point it must still add nothing.
    torch = torch_dev()
    sizes = [7, 300, 1000, 2049]
    systems, allb, off = make_ensemble(sizes, seed0=50, mass=1e-6)
    S = len(sizes)
    cen = np.zeros((S, 8))
    G = 1e15
    bt = torch.from_numpy(allb).cuda()
    ct = torch.from_numpy(cen).cuda()
    at = torch.zeros((len(allb), 3), dtype=torch.float64, device="cuda")
    cm = torch.zeros((S, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.ensemble_gravity_dev(bt, off, ct, at, cm, G, 0.0, prec)
    ctx.synchronize()
    acc = at.cpu().numpy()
    assert np.all(np.isfinite(acc))
    for s, sysb in enumerate(systems):
        ref, _ = orc.vector_field_brute(sysb, None, None, G, 0.0)
        got = acc[off[s]:off[s + 1]]
        err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert err.max() <= tol, (s, err.max())


def test_ensemble_rk4_matches_single_systems_and_graph(ctx):
    torch = torch_dev()
    rng = np.random.default_rng(7)
    sizes = np.rint(np.exp(rng.uniform(np.log(16), np.log(1500), 24))).astype(int)
    systems, allb, off = make_ensemble(sizes, seed0=100)
    S = len(sizes)
    cen = np.zeros((S, 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    dt, steps = 2 * math.pi / 256, 3
    outs = []
    for graph in (False, True):
        bt = torch.from_numpy(allb).cuda()
        ct = torch.from_numpy(cen).cuda()
        torch.cuda.synchronize()
        ctx.ensemble_rk4_dev(bt, off, ct, steps, dt, 1.0, 0.01, "fast64", graph=graph)
        ctx.synchronize()
        outs.append((bt.cpu().numpy(), ct.cpu().numpy()))
    assert np.array_equal(outs[0][0].view(np.uint64), outs[1][0].view(np.uint64))
    out, cout = outs[1]
    for s, sysb in enumerate(systems):
        ref, rc = sysb.copy(), np.array([0, 0, 0, 0, 0, 0, 1.0])
        for _ in range(steps):
            ref, rc = orc.rk4_step(ref, rc, 1.0, 0.01, dt)
        got = out[off[s]:off[s + 1]]
        moved = np.abs(ref[:, :6] - sysb[:, :6]).max()
        assert np.abs(got[:, :6] - ref[:, :6]).max() < 1e-11 * moved
        assert np.allclose(cout[s, :6], rc[:6], rtol=1e-9, atol=1e-16)


def test_ensemble_collisions_match_oracle(ctx):
    torch = torch_dev()
    # Dense small systems so that collisions happen every step.
    rng = np.random.default_rng(11)
    sizes = rng.integers(50, 400, 12)
    systems = []
    for s, nn in enumerate(sizes):
        b = np.zeros((nn, 8))
        r2 = np.random.default_rng(200 + s)
        b[:, 0:3] = r2.uniform(-0.03, 0.03, (nn, 3)) + np.array([1.0, 0, 0])
        b[:, 3:6] = r2.normal(0, 0.3, (nn, 3)) + np.array([0, 1.0, 0])
        b[:, 6] = r2.uniform(1e-12, 1e-11, nn)
        b[:, 7] = r2.uniform(1e-3, 2e-3, nn)
        systems.append(b)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    allb = np.concatenate(systems)
    S = len(sizes)
    cen = np.zeros((S, 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    dt, steps = 0.002, 4
    bt = torch.from_numpy(allb).cuda()
    ct = torch.from_numpy(cen).cuda()
    torch.cuda.synchronize()
    nc = ctx.ensemble_rk4_dev(bt, off, ct, steps, dt, 1.0, 0.0, "fast64", graph=True,
                              collisions=True, collision_f=1.5)
    ctx.synchronize()
    out = bt.cpu().numpy()
    total = 0
    for s, sysb in enumerate(systems):
        ref, rc = sysb.copy(), np.array([0, 0, 0, 0, 0, 0, 1.0])
        for _ in range(steps):
            ref, _, pairs = orc.brute_force(ref, dt, 1, 1.5, 0)
            total += len(pairs)
            ref, rc = orc.rk4_step(ref, rc, 1.0, 0.0, dt)
        got = out[off[s]:off[s + 1]]
        assert np.abs(got[:, :6] - ref[:, :6]).max() < 1e-9
    assert nc == total and total > 20


def test_ensemble_collision_steps_graph_equals_eager(ctx):
    """The captured step (collision pass, device replay, four stages) replays
    bit for bit like eager launches."""
    torch = torch_dev()
    rng = np.random.default_rng(12)
    sizes = rng.integers(100, 700, 9)
    systems = []
    for s, nn in enumerate(sizes):
        b = np.zeros((nn, 8))
        r2 = np.random.default_rng(300 + s)
        b[:, 0:3] = r2.uniform(-0.04, 0.04, (nn, 3)) + np.array([0, 1.0, 0])
        b[:, 3:6] = r2.normal(0, 0.3, (nn, 3)) + np.array([-1.0, 0, 0])
        b[:, 6] = r2.uniform(1e-12, 1e-11, nn)
        b[:, 7] = r2.uniform(1e-3, 2e-3, nn)
        systems.append(b)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    allb = np.concatenate(systems)
    cen = np.zeros((len(sizes), 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    res = []
    for graph in (False, True):
        bt = torch.from_numpy(allb).cuda()
        ct = torch.from_numpy(cen).cuda()
        torch.cuda.synchronize()
        nc = ctx.ensemble_rk4_dev(bt, off, ct, 6, 0.002, 1.0, 0.0, "mixed", graph=graph,
                                  collisions=True, collision_f=1.5)
        ctx.synchronize()
        res.append((nc, bt.cpu().numpy(), ct.cpu().numpy()))
    assert res[0][0] == res[1][0] > 50
    assert np.array_equal(res[0][1].view(np.uint64), res[1][1].view(np.uint64))
    assert np.array_equal(res[0][2].view(np.uint64), res[1][2].view(np.uint64))


def test_ensemble_replay_overflow_replayed_on_host(ctx):
    """With the device replay's candidate capacity forced down to 3
    (NCG_REPLAY_CAP, read at context creation), the passes that overflow it
    are undone and replayed on the host: 20 steps (more than one 16-step
    check window) equal the unbounded device run bit for bit."""
    import os
    from ncorpion_amd import Context
    torch = torch_dev()
    rng = np.random.default_rng(13)
    sizes = rng.integers(100, 500, 7)
    systems = []
    for s, nn in enumerate(sizes):
        b = np.zeros((nn, 8))
        r2 = np.random.default_rng(400 + s)
        b[:, 0:3] = r2.uniform(-0.04, 0.04, (nn, 3)) + np.array([1.0, 0, 0])
        b[:, 3:6] = r2.normal(0, 0.3, (nn, 3)) + np.array([0, 1.0, 0])
        b[:, 6] = r2.uniform(1e-12, 1e-11, nn)
        b[:, 7] = r2.uniform(1e-3, 2e-3, nn)
        systems.append(b)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    allb = np.concatenate(systems)
    cen = np.zeros((len(sizes), 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    os.environ["NCG_REPLAY_CAP"] = "3"
    try:
        small = Context(0)
    finally:
        del os.environ["NCG_REPLAY_CAP"]
    res = []
    try:
        for c in (ctx, small):
            bt = torch.from_numpy(allb).cuda()
            ct = torch.from_numpy(cen).cuda()
            torch.cuda.synchronize()
            nc = c.ensemble_rk4_dev(bt, off, ct, 20, 0.002, 1.0, 0.0, "fast64", graph=True,
                                    collisions=True, collision_f=1.5)
            c.synchronize()
            res.append((nc, bt.cpu().numpy(), ct.cpu().numpy()))
    finally:
        small.close()
    assert res[0][0] == res[1][0] > 50
    assert np.array_equal(res[0][1].view(np.uint64), res[1][1].view(np.uint64))
    assert np.array_equal(res[0][2].view(np.uint64), res[1][2].view(np.uint64))


def _collision_run(ctx, systems, dt, steps, f=1.5):
    torch = torch_dev()
    off = np.concatenate([[0], np.cumsum([len(x) for x in systems])]).astype(np.int32)
    cen = np.zeros((len(systems), 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    bt = torch.from_numpy(np.concatenate(systems)).cuda()
    ct = torch.from_numpy(cen).cuda()
    torch.cuda.synchronize()
    nc = ctx.ensemble_rk4_dev(bt, off, ct, steps, dt, 1.0, 0.0, "fast64", graph=True,
                              collisions=True, collision_f=f)
    ctx.synchronize()
    out = bt.cpu().numpy()
    total = 0
    for s, sysb in enumerate(systems):
        ref, rc = sysb.copy(), np.array([0, 0, 0, 0, 0, 0, 1.0])
        for _ in range(steps):
            ref, _, pairs = orc.brute_force(ref, dt, 1, f, 0)
            total += len(pairs)
            ref, rc = orc.rk4_step(ref, rc, 1.0, 0.0, dt)
        got = out[off[s]:off[s + 1]]
        assert np.abs(got[:, :6] - ref[:, :6]).max() < 1e-9, s
    return nc, total


def test_ensemble_collisions_on_spread_disks_match_oracle(ctx):
    """Disks of 300-4000 bodies (the benchmark's kind, with R = 1e-3 for 42
    collisions): the collision pass tests only the tiles a group's position
    and velocity boxes can reach within dt, and the collisions equal the
    oracle's all-pairs brute_force over 3 steps."""
    rng = np.random.default_rng(21)
    sizes = np.rint(np.exp(rng.uniform(np.log(300), np.log(4000), 16))).astype(int)
    systems = [ic.disk(int(nn), seed=500 + s, mass=1e-8, radius=1e-3) for s, nn in enumerate(sizes)]
    nc, total = _collision_run(ctx, systems, 2 * math.pi / 256, 3)
    assert nc == total and total > 20, (nc, total)


def test_ensemble_collisions_of_clumps_crossing_within_dt(ctx):
    """Two clumps 0.5 apart moving through each other at relative speed 20:
    their tiles are far apart in position but meet within dt = 0.04, which
    only the velocity boxes of the collision pass's reach test can tell."""
    systems = []
    for s in range(3):
        r = np.random.default_rng(600 + s)
        n = 300
        b = np.zeros((2 * n, 8))
        b[:n, 0:3] = r.uniform(-0.05, 0.05, (n, 3)) + np.array([0.75, 0.0, 0.0])
        b[n:, 0:3] = r.uniform(-0.05, 0.05, (n, 3)) + np.array([1.25, 0.0, 0.0])
        b[:n, 3:6] = r.normal(0, 0.01, (n, 3)) + np.array([10.0, 1.0, 0.0])
        b[n:, 3:6] = r.normal(0, 0.01, (n, 3)) + np.array([-10.0, 1.0, 0.0])
        b[:, 6] = 1e-12
        b[:, 7] = 6e-4
        systems.append(b)
    nc, total = _collision_run(ctx, systems, 0.04, 1)
    assert nc == total and total > 100, (nc, total)


@pytest.mark.parametrize("prec", ["fast64", "mixed"])
def test_ensemble_with_empty_and_tiny_systems(ctx, prec):
    """Systems of 0, 1 and 2 bodies between ordinary ones (the padded slot
    layout of the mixed form and the collision pass give an empty system no
    slot): gravity equals the per-system oracle, and RK4 steps with
    collisions leave every body finite and match the single-system runs of
    the nonempty systems."""
    torch = torch_dev()
    sizes = [0, 300, 0, 1, 2, 0, 129, 1000, 0]
    systems = [small_disk(int(nn), 700 + s) if nn else np.zeros((0, 8))
               for s, nn in enumerate(sizes)]
    allb = np.concatenate(systems)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    S = len(sizes)
    cen = np.zeros((S, 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    bt = torch.from_numpy(allb).cuda()
    ct = torch.from_numpy(cen).cuda()
    at = torch.zeros((len(allb), 3), dtype=torch.float64, device="cuda")
    cm = torch.zeros((S, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.ensemble_gravity_dev(bt, off, ct, at, cm, 1.0, 0.01, prec)
    ctx.synchronize()
    acc, cma = at.cpu().numpy(), cm.cpu().numpy()
    for s, sysb in enumerate(systems):
        if len(sysb) == 0:
            assert np.all(cma[s] == 0.0)
            continue
        ref, rcm = orc.vector_field_brute(sysb, None, (0, 0, 0, 1.0), 1.0, 0.01)
        got = acc[off[s]:off[s + 1]]
        mut, _ = orc.vector_field_brute(sysb, None, None, 1.0, 0.01)
        tol = 1e-12 if prec == "fast64" else 1e-6
        bound = (tol * np.linalg.norm(mut, axis=1)
                 + (4e-16 * math.sqrt(len(sysb)) + 1e-15) * np.linalg.norm(ref, axis=1))
        assert np.all(np.linalg.norm(got - ref, axis=1) <= bound), s
        assert np.allclose(cma[s], rcm, rtol=1e-12, atol=0)
    dt = 2 * math.pi / 256
    ctx.ensemble_rk4_dev(bt, off, ct, 3, dt, 1.0, 0.0, prec, graph=True, collisions=True,
                         collision_f=1.5)
    ctx.synchronize()
    out = bt.cpu().numpy()
    assert np.all(np.isfinite(out))
    if prec == "fast64":
        for s, sysb in enumerate(systems):
            if len(sysb) == 0:
                continue
            ref, rc = sysb.copy(), np.array([0, 0, 0, 0, 0, 0, 1.0])
            for _ in range(3):
                ref, _, _ = orc.brute_force(ref, dt, 1, 1.5, 0)
                ref, rc = orc.rk4_step(ref, rc, 1.0, 0.0, dt)
            moved = np.abs(ref[:, :6] - sysb[:, :6]).max()
            assert np.abs(out[off[s]:off[s + 1], :6] - ref[:, :6]).max() < 1e-11 * max(moved, 1e-300), s
