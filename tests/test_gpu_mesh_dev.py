"""Device-resident mesh pass (ncg_mesh_pass_dev): mesh()'s visiting order
replayed on the GPU with the library's bounce must give what the host replay
gives with the same outcome (ncg_builtin_replay_ops, itself equal to the
oracle's mesh()): collisions, counters, did_collide and final state bit for
bit; the field on the device equals the host entry point's."""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def blob(n, seed, half=0.07, rmin=5e-4, rmax=2e-3, vs=0.3):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-half, half, (n, 3))
    b[:, 3:6] = rng.normal(0, vs, (n, 3))
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = rng.uniform(rmin, rmax, n)
    return b


def host_pass(ctx, b, ex, dt, gam, cube, cells, f, one):
    bb = b.copy()
    dc = np.zeros(b.shape[0], np.int32)
    ops = ctx.builtin_ops(bb, dt, exists=ex, did_collide=dc, f=f, one_collision_only=one)
    st = ctx.mesh_pass(bb, ex, dt, gam, cube, cells, 0, ops=ops)
    return bb, dc, st, ops.collisions


def dev_pass(ctx, b, ex, dt, gam, cube, cells, f, one):
    import torch
    bt = torch.from_numpy(b.copy()).cuda()
    et = None if ex is None else torch.from_numpy(ex).cuda()
    dct = torch.zeros(b.shape[0], dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = ctx.mesh_pass_dev(bt, dt, gam, cube, cells, 0, f, one, exists_t=et, did_collide_t=dct)
    ctx.synchronize()
    return bt.cpu().numpy(), dct.cpu().numpy(), st


@pytest.mark.parametrize("seed,cells,gam,rmax,f,one,holes", [
    (31, 32, 0.004, 2e-3, 1.5, 0, False),
    (32, 32, 0.004, 3.5e-3, 1.5, 0, True),    # dense: several collisions per visit
    (33, 16, 0.01, 6e-3, 2.0, 1, False),      # elastic, one collision per body
    (34, 64, 0.0025, 2.2e-3, 1.25, 0, True),
    (35, 32, 0.004, 3.9e-3, 1.5, 1, True)])
def test_mesh_pass_dev_matches_host_replay(ctx, seed, cells, gam, rmax, f, one, holes):
    n, dt = 6000, 0.01
    b = blob(n, seed, rmax=rmax)
    cube = cells * gam
    ex = None
    if holes:
        ex = (np.random.default_rng(seed).uniform(size=n) > 0.05).astype(np.int32)
    hb, hdc, hst, hcoll = host_pass(ctx, b, ex, dt, gam, cube, cells, f, one)
    db, ddc, dst = dev_pass(ctx, b, ex, dt, gam, cube, cells, f, one)
    assert dst["collisions"] == hcoll > 10
    assert dst["total_neighbours"] == hst["total_neighbours"]
    assert dst["how_many_small"] == hst["how_many_small"]
    live = np.ones(n, bool) if ex is None else ex != 0
    assert np.array_equal(bits(db[live]), bits(hb[live]))
    assert np.array_equal(ddc, hdc)


def test_mesh_pass_dev_cluster_conflicts_fall_back_to_one_workgroup():
    """The device replay runs clusters of visits in parallel workgroups and,
    when two clusters claim one cell, restores the bodies and replays the
    pass in one workgroup.  With the cluster distance forced to 0
    (NCG_REPLAY_CLUSTER_DIST, read at context creation) every cell is a
    cluster, so dense passes meet: the results still equal the host replay
    bit for bit, and the fallback ran."""
    import os
    from ncorpion_amd import Context
    os.environ["NCG_REPLAY_CLUSTER_DIST"] = "0"
    try:
        c0 = Context(0)
    finally:
        del os.environ["NCG_REPLAY_CLUSTER_DIST"]
    conflicts = 0
    try:
        for seed, cells, gam, rmax, f, one, holes in [
                (41, 32, 0.004, 3.5e-3, 1.5, 0, True), (42, 64, 0.0025, 2.2e-3, 1.25, 1, False),
                (43, 16, 0.01, 6e-3, 2.0, 1, False)]:
            n, dt = 6000, 0.01
            b = blob(n, seed, rmax=rmax)
            cube = cells * gam
            ex = None
            if holes:
                ex = (np.random.default_rng(seed).uniform(size=n) > 0.05).astype(np.int32)
            hb, hdc, hst, hcoll = host_pass(c0, b, ex, dt, gam, cube, cells, f, one)
            db, ddc, dst = dev_pass(c0, b, ex, dt, gam, cube, cells, f, one)
            conflicts += dst["replay_conflicts"]
            assert dst["collisions"] == hcoll > 10
            live = np.ones(n, bool) if ex is None else ex != 0
            assert np.array_equal(bits(db[live]), bits(hb[live]))
            assert np.array_equal(ddc, hdc)
    finally:
        c0.close()
    assert conflicts > 0


def test_mesh_pass_dev_against_oracle_and_field(ctx):
    """The device pass against the oracle's mesh() directly, then the field
    on the device after it equals the host entry point's (bit for bit)."""
    import torch
    n, dt, cells, gam = 5000, 0.01, 32, 0.004
    b = blob(n, 41, rmax=3e-3)
    ex = (np.random.default_rng(41).uniform(size=n) > 0.05).astype(np.int32)
    ref = orc.mesh(b, dt, gam, cells * gam, cells, 1, 1.5, 0, 1, ex)
    bt = torch.from_numpy(b.copy()).cuda()
    et = torch.from_numpy(ex).cuda()
    st = ctx.mesh_pass_dev(bt, dt, gam, cells * gam, cells, 0, 1.5, 0, exists_t=et)
    got = bt.cpu().numpy()
    assert st["collisions"] == len(ref["collisions"]) > 20
    assert st["total_neighbours"] == ref["total_neighbours"]
    assert np.array_equal(bits(got[ex != 0]), bits(ref["bodies"][ex != 0]))
    cen = (0.01, -0.02, 0.0, 1.0)
    for prec in ("exact", "fast64"):
        acc_h, cm_h = ctx.mesh_gravity(got, ex, cen, 1.0, 0.1, prec)
        acc_t = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        cm_t = torch.zeros(3, dtype=torch.float64, device="cuda")
        ct = torch.tensor([cen[0], cen[1], cen[2], 0, 0, 0, cen[3], 1.0], dtype=torch.float64,
                          device="cuda")
        ctx.mesh_gravity_dev(bt, acc_t, ct, cm_t, exists_t=et, G=1.0, softening=0.1,
                             precision=prec)
        ctx.synchronize()
        acc_d = acc_t.cpu().numpy()
        assert np.array_equal(bits(acc_d[ex != 0]), bits(acc_h[ex != 0])), prec
        assert np.array_equal(bits(cm_t.cpu().numpy()), bits(cm_h)), prec


def test_mesh_pass_dev_refuses_bodies_larger_than_the_mesh(ctx):
    import torch
    b = blob(500, 51)
    b[7, 7] = 0.05  # larger than gam
    bt = torch.from_numpy(b.copy()).cuda()
    with pytest.raises(Exception, match="larger than the mesh size"):
        ctx.mesh_pass_dev(bt, 0.01, 0.004, 0.128, 32, 0, 1.5, 0)
    assert np.array_equal(bt.cpu().numpy(), b)  # nothing changed
