"""GPU-resident leapfrog (ncg_leapfrog_dev) against NcorpiON's own integration
loop (integration_brute_force_SABA1, rk4.c:840-954).

`oracle_leapfrog` restates that loop on the CPU with the oracle's
vector_field and brute_force: half drift (brute_force and drift with dt/2),
then per step kick(dt), brute_force(dt), drift(dt).  The CPU test pins it to a
traced run of the compiled reference (state hash at every vector_field call);
the GPU tests compare the device loop with it bit for bit, and with the traced
reference runs directly.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def oracle_leapfrog(bodies, exists, central8, G, soft, dt, nsteps, f, one_coll, half_start,
                    did_collide=None, on_step=None):
    """rk4.c:871-879 (half drift) and 896-905 (kick, brute_force, drift) with
    kick/drift of rk4.c:68-127; did_collide cleared after each step when
    one_coll (rk4.c:348-353).  Returns (bodies, central8, collisions)."""
    b = np.array(bodies, dtype=np.float64)
    ex = np.ones(b.shape[0], np.int32) if exists is None else np.array(exists, np.int32)
    live = ex != 0
    c = np.array(central8, dtype=np.float64)
    dc = np.zeros(b.shape[0], np.int32) if did_collide is None else np.array(did_collide, np.int32)
    ctype = 0 if f == 2.0 else 1
    present = c[7] != 0
    ncoll = 0

    def drift(h):
        b[live, 0:3] += h * b[live, 3:6]
        if present:
            c[0:3] += h * c[3:6]

    if half_start:
        if f > 0:
            nb, dc, pairs = orc.brute_force(b, 0.5 * dt, ctype, f, one_coll, exists=ex, did_collide=dc)
            b[:] = nb
            ncoll += len(pairs)
        drift(0.5 * dt)
    for s in range(nsteps):
        if on_step:
            on_step(s, b)
        acc, cm = orc.vector_field_brute(b, ex, (c[0], c[1], c[2], c[6]) if present else None, G,
                                         soft)
        b[live, 3:6] += dt * acc[live]
        if present:
            c[3:6] += dt * cm
        if f > 0:
            nb, dc, pairs = orc.brute_force(b, dt, ctype, f, one_coll, exists=ex, did_collide=dc)
            b[:] = nb
            ncoll += len(pairs)
        drift(dt)
        if one_coll:
            dc[:] = 0
    return b, c, ncoll


def read_dump(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    o = 4
    bodies = np.frombuffer(raw[o:o + 64 * n], np.float64).reshape(n, 8).copy()
    o += 64 * n
    ex = np.frombuffer(raw[o:o + 4 * n], np.int32).copy()
    o += 4 * n
    dc = np.frombuffer(raw[o:o + 4 * n], np.int32).copy()
    o += 4 * n
    cm = np.frombuffer(raw[o:o + 64], np.float64).copy()
    o += 64
    ts = float(np.frombuffer(raw[o:o + 8], np.float64)[0])
    return bodies, ex, dc, cm, ts


def read_dump_mesh(path):
    """read_dump plus the mesh fields the tracer appends: gam, collision_cube,
    how_many_moonlets."""
    bodies, ex, dc, cm, ts = read_dump(path)
    raw = open(path, "rb").read()
    o = 4 + bodies.shape[0] * (64 + 8) + 64 + 8
    gam, cube = np.frombuffer(raw[o:o + 16], np.float64)
    hm = int(np.frombuffer(raw[o + 16:o + 20], np.int32)[0])
    return bodies, ex, dc, cm, ts, float(gam), float(cube), hm


def three_largest_first(b, ex):
    """three_largest_moonlets + three_largest_three_first (structure.c:914-1064)
    restated on a numpy state (moves bodies in place)."""
    L = [-1, -1, -1]
    R0 = R1 = R2 = 0.0
    i = 0
    for j in range(b.shape[0]):
        if not ex[j]:
            continue
        r = b[j, 7]
        if i == 0:
            R0, L[0] = r, j
        elif i == 1:
            R1 = r
            if R0 > R1:
                L[1] = j
            else:
                L[1], L[0] = L[0], j
                R0, R1 = R1, R0
        elif i == 2:
            R2 = r
            if R1 >= R2:
                L[2] = j
            elif R2 > R1 and R2 <= R0:
                L[2], L[1] = L[1], j
                R1, R2 = R2, R1
            else:
                L[2], L[1], L[0] = L[1], L[0], j
                R0, R1, R2 = R2, R0, R1
        else:
            if r > R0:
                L[2], L[1], L[0] = L[1], L[0], j
                R2, R1, R0 = R1, R0, r
            elif r > R1:
                L[2], L[1] = L[1], j
                R2, R1 = R1, r
            elif r > R2:
                L[2], R2 = j, r
        i += 1
    first, second, third = L

    def place(slot, frm):
        if ex[slot]:
            tmp = b[slot].copy()
            b[slot] = b[frm]
            b[frm] = tmp
        else:
            b[slot] = b[frm]
            ex[frm] = 0
            ex[slot] = 1

    if first != 0:
        place(0, first)
        if second == 0:
            second = first
        elif third == 0:
            third = first
    if second != 1:
        place(1, second)
        if third == 1:
            third = second
    if third != 2:
        place(2, third)


def traced_reference(cfg, bodies, tmp_path):
    """Runs the traced reference; returns (dump, VF input hashes, collisions)."""
    exe = os.path.join(REF, cfg, "ncorpion_trace")
    if not os.path.exists(exe):
        pytest.skip("reference builds missing (oracle/build_ref.py)")
    ic.write_init_txt(str(tmp_path / "init.txt"), bodies)
    env = dict(os.environ, NCG_TRACE=str(tmp_path / "trace.txt"),
               NCG_TRACE_DUMP=str(tmp_path / "dump.bin"))
    r = subprocess.run([exe], cwd=str(tmp_path), env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = open(tmp_path / "trace.txt").read().splitlines()
    vf = [int(l.split()[2], 16) for l in lines if l.startswith("VF ")]
    ncoll = sum(1 for l in lines if l.startswith("C "))
    return read_dump(str(tmp_path / "dump.bin")), vf, ncoll


def test_oracle_leapfrog_reproduces_reference_trace(tmp_path):
    """CPU: the restated loop hits the reference's state hash at every one of
    the 40 vector_field calls of a dense run with collisions."""
    b0 = ic.disk(2000, seed=1, mass=1e-8, radius=4e-3)
    (bodies, ex, dc, cm, ts), vf, ncoll = traced_reference("run_c1_dense", b0, tmp_path)
    dt = 2.0 * ts
    seen = []
    _, _, nc = oracle_leapfrog(bodies, ex, np.concatenate([cm[0:7], [1.0]]), 1.0, 0.0, dt,
                               len(vf), 1.5, 0, True, did_collide=dc,
                               on_step=lambda s, b: seen.append(orc.state_hash(b, ex)))
    assert seen == vf
    assert nc == ncoll and ncoll > 20


def tt():
    import torch
    return torch


def cluster(n, seed, radius=2e-3, box=0.05):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-box, box, (n, 3)) + np.array([1.0, 0, 0])
    b[:, 3:6] = rng.normal(0, 0.2, (n, 3)) + np.array([0, 1.0, 0])
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = rng.uniform(0.5 * radius, radius, n)
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("f,one,graph,holes", [(1.5, 0, False, False), (1.5, 0, True, True),
                                               (2.0, 1, True, False), (1.25, 1, False, True)])
def test_leapfrog_matches_oracle_leapfrog(ctx, f, one, graph, holes):
    torch = tt()
    n, dt, steps = 1500, 0.004, 6
    b = cluster(n, 40 + int(10 * f) + one)
    ex = np.ones(n, np.int32)
    if holes:
        ex[np.random.default_rng(5).choice(n, 100, replace=False)] = 0
    c8 = np.array([0.001, 0.0, 0.0, 0.0, 1e-3, 0.0, 1.0, 1.0])
    ref_b, ref_c, ref_n = oracle_leapfrog(b, ex, c8, 1.0, 0.01, dt, steps, f, one, True)
    bt = torch.from_numpy(b).cuda()
    ct = torch.from_numpy(c8).cuda()
    et = torch.from_numpy(ex).cuda()
    dct = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    nc = ctx.leapfrog_dev(bt, ct, steps, dt, 1.0, 0.01, "exact", f, one, exists_t=et,
                          did_collide_t=dct, graph=graph, half_drift_start=True)
    ctx.synchronize()
    got = bt.cpu().numpy()
    live = ex != 0
    assert nc == ref_n and ref_n > 10
    assert np.array_equal(got[live].view(np.uint64), ref_b[live].view(np.uint64))
    assert np.array_equal(ct.cpu().numpy().view(np.uint64), ref_c.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,seed,radius", [("run_c1", 0, 2e-4), ("run_c1_dense", 1, 4e-3)])
def test_leapfrog_reproduces_reference_run(ctx, tmp_path, cfg, seed, radius):
    """Start from the reference's own state before its half drift, then step
    on the device: the state hash after every step equals the reference's at
    the next vector_field call, and the collision counts agree."""
    torch = tt()
    b0 = ic.disk(2000, seed=seed, mass=1e-8, radius=radius)
    (bodies, ex, dc, cm, ts), vf, ncoll = traced_reference(cfg, b0, tmp_path)
    dt = 2.0 * ts
    bt = torch.from_numpy(bodies).cuda()
    ct = torch.from_numpy(np.concatenate([cm[0:7], [1.0]])).cuda()
    et = torch.from_numpy(ex).cuda()
    dct = torch.from_numpy(dc).cuda()
    total = ctx.leapfrog_dev(bt, ct, 0, dt, 1.0, 0.0, "exact", 1.5, 0, exists_t=et,
                             did_collide_t=dct, half_drift_start=True)
    for s, h in enumerate(vf):
        ctx.synchronize()
        assert orc.state_hash(bt.cpu().numpy(), ex) == h, "step %d" % s
        if s + 1 < len(vf):
            total += ctx.leapfrog_dev(bt, ct, 1, dt, 1.0, 0.0, "exact", 1.5, 0, exists_t=et,
                                      did_collide_t=dct)
    # The reference's last step (after its last vector_field call) is not hashed.
    total += ctx.leapfrog_dev(bt, ct, 1, dt, 1.0, 0.0, "exact", 1.5, 0, exists_t=et,
                              did_collide_t=dct)
    assert total == ncoll


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fast64", 1e-12), ("mixed", 1e-8)])
def test_leapfrog_tiled_forms(ctx, prec, tol):
    """The tiled fields in the same loop stay close to the exact one."""
    torch = tt()
    n, dt, steps = 4096, 2 * np.pi / 256, 8
    b = ic.disk(n, seed=3, mass=1e-9, radius=1e-4)
    c8 = np.array([0, 0, 0, 0, 0, 0, 1.0, 1.0])
    out = {}
    for p in ("exact", prec):
        bt = torch.from_numpy(b).cuda()
        ct = torch.from_numpy(c8).cuda()
        torch.cuda.synchronize()
        ctx.leapfrog_dev(bt, ct, steps, dt, 1.0, 1e-6, p, 0.0, 0, graph=True,
                         half_drift_start=True, half_drift_end=True)
        ctx.synchronize()
        out[p] = bt.cpu().numpy()
    d = np.abs(out[prec][:, 0:6] - out["exact"][:, 0:6]).max()
    assert d < tol, d


def big_body_disk():
    """16,384-body disk (m = 1e-8, R = 2e-3) with 24 bodies of radius
    0.03-0.06 (same density), half of them on fast crossing orbits."""
    n = 16384
    b = ic.disk(n, seed=6, mass=1e-8, radius=2e-3)
    rng = np.random.default_rng(6)
    big = rng.choice(np.arange(n), 24, replace=False)
    b[big, 7] = rng.uniform(0.03, 0.06, 24)
    b[big, 6] = 1e-8 * (b[big, 7] / 2e-3) ** 3
    b[big[:12], 3:6] += rng.normal(0.0, 0.05, (12, 3))
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("prec,cfg", [("exact", "run_mesh"), ("exact", "run_mesh_big")])
def test_mesh_loop_reproduces_reference_run(ctx, tmp_path, prec, cfg):
    """NcorpiON's mesh integrator (16,384 bodies, inelastic bounces, 10
    steps) from the reference's own starting state: after every device step,
    the state the reference's next mesh() sees (after its three_largest swap)
    has the same hash, and the collision counts agree.  run_mesh: radius
    1e-3; run_mesh_big: bodies larger than the mesh size, whose passes the
    loop replays on the host."""
    torch = tt()
    if cfg == "run_mesh_big":
        b0 = big_body_disk()
        nbig_min = 1
    else:
        b0 = ic.disk(16384, seed=2, mass=1e-10, radius=1e-3)
        nbig_min = 0
    exe = os.path.join(REF, cfg, "ncorpion_trace")
    if not os.path.exists(exe):
        pytest.skip("reference builds missing (oracle/build_ref.py)")
    ic.write_init_txt(str(tmp_path / "init.txt"), b0)
    env = dict(os.environ, NCG_TRACE=str(tmp_path / "trace.txt"),
               NCG_TRACE_DUMP=str(tmp_path / "dump.bin"))
    r = subprocess.run([exe], cwd=str(tmp_path), env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = open(tmp_path / "trace.txt").read().splitlines()
    mesh_in, coll = [], []
    for l in lines:
        f = l.split()
        if f[0] == "MESH" and f[2] == "in":
            mesh_in.append(int(f[3], 16))
            coll.append(0)
        elif f[0] == "C" and coll:
            coll[-1] += 1
    bodies, ex, dc, cm, ts, gam, cube, hm = read_dump_mesh(str(tmp_path / "dump.bin"))
    assert len(mesh_in) == 10 and sum(coll) > 10
    dt = 2.0 * ts  # dumped inside the half kick (timestep halved)
    params = dict(gam=gam, gam_min=2.56 / 256,
                  how_many_neighbours=30.0 if cfg == "run_mesh_big" else 100.0,
                  disruption_threshold=0.05, high_dumping_threshold=100.0, cells=256)
    bt = torch.from_numpy(bodies).cuda()
    ct = torch.from_numpy(np.concatenate([cm[0:7], [1.0]])).cuda()
    et = torch.from_numpy(ex).cuda()
    dct = torch.from_numpy(dc).cuda()
    ctx.mesh_leapfrog_dev(bt, ct, 0, dt, params, 1.0, 0.0, prec, 1.5, 0, exists_t=et,
                          did_collide_t=dct, half_kick_start=True)
    host_passes = 0
    for s, h in enumerate(mesh_in):
        ctx.synchronize()
        st = bt.cpu().numpy().copy()
        e = et.cpu().numpy().copy()
        three_largest_first(st, e)
        assert orc.state_hash(st, e) == h, "step %d" % s
        nc = ctx.mesh_leapfrog_dev(bt, ct, 1, dt, params, 1.0, 0.0, prec, 1.5, 0, exists_t=et,
                                   did_collide_t=dct)
        host_passes += params["host_passes"]
        assert nc == coll[s], "step %d" % s
    if nbig_min:
        assert host_passes > 0
    else:
        assert host_passes == 0


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_leapfrog_replay_overflow_replayed_on_host(tmp_path, graph):
    """With the device replay's candidate capacity forced down to 2
    (NCG_REPLAY_CAP, read at context creation), every collision pass that
    overflows it is undone and replayed on the host: 70 steps (more than one
    64-step check window) of a dense cluster equal the oracle's loop bit for
    bit."""
    torch = tt()
    from ncorpion_amd import Context
    os.environ["NCG_REPLAY_CAP"] = "2"
    try:
        small = Context(0)
    finally:
        del os.environ["NCG_REPLAY_CAP"]
    try:
        n, dt, steps = 600, 0.004, 70
        b = cluster(n, 77, radius=3e-3)
        c8 = np.array([0.001, 0.0, 0.0, 0.0, 1e-3, 0.0, 1.0, 1.0])
        ref_b, ref_c, ref_n = oracle_leapfrog(b, None, c8, 1.0, 0.01, dt, steps, 1.5, 0, True)
        bt = torch.from_numpy(b).cuda()
        ct = torch.from_numpy(c8).cuda()
        torch.cuda.synchronize()
        nc = small.leapfrog_dev(bt, ct, steps, dt, 1.0, 0.01, "exact", 1.5, 0, graph=graph,
                                half_drift_start=True)
        small.synchronize()
        assert nc == ref_n and ref_n > 20
        assert np.array_equal(bt.cpu().numpy().view(np.uint64), ref_b.view(np.uint64))
        assert np.array_equal(ct.cpu().numpy().view(np.uint64), ref_c.view(np.uint64))
    finally:
        small.close()


# Reference runs with the collision models that remove and create bodies,
# and with end_of_timestep's removals: (config, outcome model, N_max,
# disruption_threshold, high_dumping_threshold), as built by
# oracle/build_ref.py.  Initial conditions as in test_gpu_dropin.py.
MODEL_RUNS = [("run_c1_frag", "fragmentation", 4000, 0.05, 100.0),
              ("run_c1_merge", "merger", 2016, 0.05, 100.0),
              ("run_c1_rm", "fragmentation", 4000, 0.93, 1.07),
              ("run_c1_frag1", "fragmentation", 4000, 0.05, 100.0)]
ONE_COLLISION = {"run_c1_frag1": 1}


def _model_setup(ctx, tmp_path, cfg, kind, n_max):
    torch = tt()
    from ncorpion_amd import _abi
    b0 = ic.disk(2000, seed=1, mass=1e-8, radius=4e-3)
    (bodies, ex, dc, cm, ts), vf, ncoll = traced_reference(cfg, b0, tmp_path)
    n0 = bodies.shape[0]
    B = np.zeros((n_max, 8))
    B[:n0] = bodies
    E = np.zeros(n_max, np.int32)
    E[:n0] = ex
    DC = np.zeros(n_max, np.int32)
    DC[:n0] = dc
    t = dict(b=torch.from_numpy(B).cuda(), e=torch.from_numpy(E).cuda(),
             dc=torch.from_numpy(DC).cuda(),
             c=torch.from_numpy(np.concatenate([cm[0:7], [1.0]])).cuda())
    state = dict(largest_id=n0 - 1, how_many_free=0, free_indexes=np.zeros(n_max, np.int32),
                 need_to_reduce_com=0)
    return t, state, _abi.outcome_model(kind, G=1.0), 2.0 * ts, vf, ncoll


def _model_hash(t, state):
    n = state["largest_id"] + 1
    return orc.state_hash(t["b"].cpu().numpy()[:n], t["e"].cpu().numpy()[:n])


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,kind,n_max,rdis,rhd", MODEL_RUNS)
def test_leapfrog_model_reproduces_reference_run(ctx, tmp_path, cfg, kind, n_max, rdis, rhd):
    """The GPU-resident loop with the reference's outcome model (mergers,
    super-catastrophic, partial and full fragmentation; bodies removed and
    created at free or appended slots) and end_of_timestep's removals,
    centre-of-mass reduction and compaction: one step per call from the
    reference's own state before its half drift, the state hash (every slot
    up to largest_id) equal to the reference's at each vector_field call."""
    t, state, model, dt, vf, ncoll = _model_setup(ctx, tmp_path, cfg, kind, n_max)
    kw = dict(G=1.0, softening=0.0, precision="exact", did_collide_t=t["dc"],
              r_disruption=rdis, r_high_dumping=rhd,
              one_collision_only=ONE_COLLISION.get(cfg, 0))
    total = ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], 0, dt, state, model,
                                   half_drift_start=True, **kw)
    for s, h in enumerate(vf):
        ctx.synchronize()
        assert _model_hash(t, state) == h, "step %d" % s
        total += ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], 1, dt, state, model, **kw)
    print("%s: %d collisions (%d treatment calls), mergers %d, super-catastrophic %d, "
          "partial %d, full %d, removed %d, tidy_ups %d, largest_id %d" % (
              cfg, total, ncoll, state["mergers"], state["super_catastrophic"],
              state["partial"], state["full"], state["removed"], state["tidy_ups"],
              state["largest_id"]))
    # Treatment calls of pairs with a body removed earlier in the pass do
    # nothing (physics.c:1057-1059) and are not counted here.
    assert 100 < total <= ncoll
    if kind == "fragmentation":
        assert state["full"] > 0 and state["partial"] > 0
    if cfg == "run_c1_rm":
        assert state["removed"] > 100 and state["tidy_ups"] > 0 and state["super_catastrophic"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,kind,n_max,rdis,rhd", MODEL_RUNS)
def test_leapfrog_model_one_call_matches_reference(ctx, tmp_path, cfg, kind, n_max, rdis, rhd):
    """The same runs in one call of 39 steps replayed from a captured graph:
    windows whose passes resolved pairs or whose bodies crossed a threshold
    run again step by step with those passes and step ends on the host; the
    state then equals the reference's before its 40th step."""
    t, state, model, dt, vf, ncoll = _model_setup(ctx, tmp_path, cfg, kind, n_max)
    ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], len(vf) - 1, dt, state, model, G=1.0,
                           softening=0.0, precision="exact", did_collide_t=t["dc"],
                           r_disruption=rdis, r_high_dumping=rhd, graph=True,
                           half_drift_start=True, one_collision_only=ONE_COLLISION.get(cfg, 0))
    ctx.synchronize()
    assert _model_hash(t, state) == vf[-1]
    assert state["host_passes"] > 10


# Mesh runs with the collision models and removals: (config, mass, radius,
# N_max, disruption_threshold, high_dumping_threshold); 16,384-body disks
# (seed 2) as in test_gpu_dropin.py.
MESH_MODEL_RUNS = [("run_mesh_frag", "fragmentation", 1e-8, 2e-3, 40000, 0.05, 100.0),
                   ("run_mesh_merge", "merger", 1e-10, 1e-3, 16400, 0.05, 100.0),
                   ("run_mesh_rm", "fragmentation", 1e-8, 2e-3, 40000, 0.93, 1.07)]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,kind,mass,radius,n_max,rdis,rhd", MESH_MODEL_RUNS)
def test_mesh_loop_model_reproduces_reference_run(ctx, tmp_path, cfg, kind, mass, radius, n_max,
                                                  rdis, rhd):
    """NcorpiON's mesh integrator with its default collision model (mergers,
    partial, full and super-catastrophic fragmentation: fragments appended
    with their pairs) or instant mergers, and with end_of_timestep's
    removals: from the reference's own starting state, the state its next
    mesh() sees (every slot up to largest_id, after the three_largest swap)
    has the same hash after every device step."""
    torch = tt()
    from ncorpion_amd import _abi
    b0 = ic.disk(16384, seed=2, mass=mass, radius=radius)
    exe = os.path.join(REF, cfg, "ncorpion_trace")
    if not os.path.exists(exe):
        pytest.skip("reference builds missing (oracle/build_ref.py)")
    ic.write_init_txt(str(tmp_path / "init.txt"), b0)
    env = dict(os.environ, NCG_TRACE=str(tmp_path / "trace.txt"),
               NCG_TRACE_DUMP=str(tmp_path / "dump.bin"))
    r = subprocess.run([exe], cwd=str(tmp_path), env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = open(tmp_path / "trace.txt").read().splitlines()
    mesh_in, coll = [], []
    for l in lines:
        f = l.split()
        if f[0] == "MESH" and f[2] == "in":
            mesh_in.append(int(f[3], 16))
            coll.append(0)
        elif f[0] == "C" and coll:
            coll[-1] += 1
    bodies, ex, dc, cm, ts, gam, cube, hm = read_dump_mesh(str(tmp_path / "dump.bin"))
    assert len(mesh_in) == 10
    n0 = bodies.shape[0]
    B = np.zeros((n_max, 8))
    B[:n0] = bodies
    E = np.zeros(n_max, np.int32)
    E[:n0] = ex
    DC = np.zeros(n_max, np.int32)
    DC[:n0] = dc
    bt, et, dct = (torch.from_numpy(a).cuda() for a in (B, E, DC))
    ct = torch.from_numpy(np.concatenate([cm[0:7], [1.0]])).cuda()
    dt = 2.0 * ts  # dumped inside the half kick (timestep halved)
    params = dict(gam=gam, gam_min=2.56 / 256, how_many_neighbours=100.0, cells=256)
    state = dict(largest_id=n0 - 1, how_many_free=0, free_indexes=np.zeros(n_max, np.int32),
                 need_to_reduce_com=0)
    model = _abi.outcome_model(kind, G=1.0)
    kw = dict(G=1.0, softening=0.0, precision="exact", did_collide_t=dct, r_disruption=rdis,
              r_high_dumping=rhd)
    ctx.mesh_leapfrog_model_dev(bt, ct, et, 0, dt, params, state, model, half_kick_start=True,
                                **kw)
    total = 0
    for s, h in enumerate(mesh_in):
        ctx.synchronize()
        n = state["largest_id"] + 1
        st = bt.cpu().numpy()[:n].copy()
        e = et.cpu().numpy()[:n].copy()
        three_largest_first(st, e)
        assert orc.state_hash(st, e) == h, "step %d" % s
        nc = ctx.mesh_leapfrog_model_dev(bt, ct, et, 1, dt, params, state, model, **kw)
        assert nc <= coll[s], "step %d" % s
        total += nc
    print("%s: %d collisions (%d treatment calls), mergers %d, super-catastrophic %d, "
          "partial %d, full %d, removed %d, tidy_ups %d, largest_id %d, host passes %d" % (
              cfg, total, sum(coll), state["mergers"], state["super_catastrophic"],
              state["partial"], state["full"], state["removed"], state["tidy_ups"],
              state["largest_id"], state["host_passes"]))
    assert total > 100
    if cfg == "run_mesh_rm":
        assert state["removed"] > 100 and state["tidy_ups"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fast64", "mixed"])
def test_leapfrog_model_tiled_field_one_call_equals_steps(ctx, tmp_path, prec):
    """With a tiled field (whose spatial order is rebuilt every step over the
    slots in use) and the fragmentation model with removals, one 39-step
    call replayed from graphs ends in the same state, bit for bit, as 39
    calls of one step: windows, re-runs and recaptures after the slot count
    changes do not alter the run."""
    runs = []
    for one_call in (True, False):
        d = tmp_path / ("one_call" if one_call else "steps")
        d.mkdir()
        t, state, model, dt, vf, ncoll = _model_setup(ctx, d, "run_c1_rm", "fragmentation", 4000)
        kw = dict(G=1.0, softening=0.0, precision=prec, did_collide_t=t["dc"],
                  r_disruption=0.93, r_high_dumping=1.07)
        if one_call:
            nc = ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], 39, dt, state, model, graph=True,
                                        half_drift_start=True, **kw)
        else:
            nc = ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], 0, dt, state, model,
                                        half_drift_start=True, **kw)
            for _ in range(39):
                nc += ctx.leapfrog_model_dev(t["b"], t["c"], t["e"], 1, dt, state, model, **kw)
        ctx.synchronize()
        n = state["largest_id"] + 1
        runs.append((nc, n, t["b"].cpu().numpy()[:n].copy(), t["e"].cpu().numpy()[:n].copy(),
                     t["c"].cpu().numpy().copy(), state["removed"], state["full"]))
    (n1, s1, b1, e1, c1, r1, f1), (n2, s2, b2, e2, c2, r2, f2) = runs
    assert n1 == n2 > 100 and s1 == s2 and r1 == r2 > 100 and f1 == f2 > 0
    assert np.array_equal(e1, e2)
    assert np.array_equal(b1[e1 != 0].view(np.uint64), b2[e2 != 0].view(np.uint64))
    assert np.array_equal(c1.view(np.uint64), c2.view(np.uint64))


@pytest.mark.gpu
def test_mesh_loop_refuses_a_threshold_crossing(ctx):
    """Without an outcome model the mesh loop applies no removal: a body that
    crosses high_dumping_threshold during the run makes the call fail (the
    check is read with the next pass's counters or at the end), while the
    same run without the crossing succeeds."""
    torch = tt()
    from ncorpion_amd import NcgError
    b = ic.disk(4000, seed=3, mass=1e-10, radius=1e-4)
    c8 = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    for crossing in (False, True):
        bb = b.copy()
        if crossing:
            bb[10, 0:3] = [1.199, 0.0, 0.0]
            bb[10, 3:6] = [2.0, 0.0, 0.0]
        bt = torch.from_numpy(bb).cuda()
        ct = torch.from_numpy(c8).cuda()
        et = torch.ones(bb.shape[0], dtype=torch.int32, device="cuda")
        dct = torch.zeros(bb.shape[0], dtype=torch.int32, device="cuda")
        params = dict(gam=0.05, gam_min=0.01, how_many_neighbours=30.0,
                      disruption_threshold=0.05, high_dumping_threshold=1.2, cells=64)
        run = lambda: ctx.mesh_leapfrog_dev(bt, ct, 3, 0.01, params, 1.0, 0.0, "fast64", 1.5, 0,  # noqa: E731
                                            exists_t=et, did_collide_t=dct)
        if crossing:
            with pytest.raises(NcgError, match="high_dumping_threshold"):
                run()
        else:
            run()
        ctx.synchronize()
