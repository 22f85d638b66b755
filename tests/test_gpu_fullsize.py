"""Parity at the benchmark workloads' own sizes.

* Brute force, N = 262,144 (the disk of the all-pairs benchmark: masses
  log-uniform in [1e-11, 1e-9], radius 1e-4 (m/1e-10)^(1/3), relative
  softening 1e-6).  The full O(N^2) oracle is too slow here, so the oracle's
  row form (each sampled body summed over all N, pinned bit for bit to the
  reference's own accelerations) checks 1024 seeded bodies: the exact form bit
  for bit, FP64-fast to FP64 rounding, and the mixed form -- under the tuning
  table's row for this size -- within 1e-6 max relative error of the mutual
  acceleration.
* Mesh module, 2^20 bodies (m = 1e-10, R = 5e-5), cells of side 0.01 (256 per
  axis): NcorpiON's own get_neighbours_mesh + vector_field and mesh()
  (compiled from source, oracle/_ref/probe_c2) against the library: the exact
  mesh field bit for bit, mesh()'s collision sequence, counters and final
  state bit for bit (host replay and device replay).
* The GPU-resident mesh loop with the mixed field against the FP64-fast one.
"""
import math
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE_C2 = os.path.join(ROOT, "oracle", "_ref", "probe_c2", "ref_probe")

C3_N, C3_SOFT, C3_SAMPLE = 262144, 1e-6, 1024
C2_N, C2_GAM, C2_CELLS, C2_CUBE = 1 << 20, 0.01, 256, 2.56
C2_DT = 2.0 * math.pi / 512.0


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def c3():
    m, R = ic.disk_c3_masses(C3_N, 0)
    b = ic.disk(C3_N, seed=0, masses=m, radii=R)
    rows = np.sort(np.random.default_rng(2024).choice(C3_N, C3_SAMPLE, replace=False))
    print("c3: oracle rows (%d of %d bodies)" % (C3_SAMPLE, C3_N), flush=True)
    mutual = orc.vector_field_rows(b, rows, None, None, 1.0, C3_SOFT)
    full = orc.vector_field_rows(b, rows, None, (0.0, 0.0, 0.0, 1.0), 1.0, C3_SOFT)
    return b, rows, mutual, full


def test_c3_exact_rows_bit_identical(ctx, c3):
    b, rows, _, full = c3
    print("c3: exact form, n=%d" % C3_N, flush=True)
    acc, _ = ctx.gravity_allpairs(b, None, (0.0, 0.0, 0.0, 1.0), 1.0, C3_SOFT, "exact")
    assert np.array_equal(bits(acc[rows]), bits(full))


@pytest.mark.parametrize("prec,bound", [("fast64", 1e-11), ("mixed", 1e-6)])
def test_c3_fast_forms_on_sampled_rows(ctx, c3, prec, bound):
    b, rows, mutual, _ = c3
    print("c3: %s form, n=%d" % (prec, C3_N), flush=True)
    acc, _ = ctx.gravity_allpairs(b, None, None, 1.0, C3_SOFT, prec)
    d = np.linalg.norm(acc[rows] - mutual, axis=1)
    r = np.linalg.norm(mutual, axis=1)
    rel = d / r
    print("c3 %s: max rel %.3e, p99 %.3e, normwise %.3e" %
          (prec, rel.max(), np.quantile(rel, 0.99), d.max() / np.sqrt(np.mean(r * r))))
    assert rel.max() <= bound, rel.max()


def test_c3_mixed_every_body_against_fast64(ctx, c3):
    """The mixed form under the tuning table's row for this size, on all
    262,144 bodies: within 1e-6 max relative error of the FP64-fast form
    (itself held to the oracle's rows at 1e-11 above)."""
    b, _, _, _ = c3
    f64, _ = ctx.gravity_allpairs(b, None, None, 1.0, C3_SOFT, "fast64")
    mix, _ = ctx.gravity_allpairs(b, None, None, 1.0, C3_SOFT, "mixed")
    rel = np.linalg.norm(mix - f64, axis=1) / np.linalg.norm(f64, axis=1)
    print("c3 mixed vs fast64, all bodies: max rel %.3e, p99.9 %.3e" %
          (rel.max(), np.quantile(rel, 0.999)))
    assert rel.max() <= 1e-6, rel.max()


def c2_workload():
    return ic.disk(C2_N, seed=0, mass=1e-10, radius=5e-5)


def probe_c2(op, b):
    """NcorpiON's mesh configuration of the 2^20-body workload: op 5
    (get_neighbours_mesh then vector_field), op 2 | 0x200 (mesh(), inelastic
    bounces f = 1.5, without the pair list)."""
    n = b.shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, op, n, 0))
            f.write(np.array([C2_DT, C2_GAM, C2_CUBE, 0, 0, 0, 0, 0]).tobytes())
            f.write(np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]).tobytes())
            f.write(b.tobytes())
            f.write(np.ones(n, np.int32).tobytes())
            f.write(np.zeros(n, np.int32).tobytes())
        r = subprocess.run([PROBE_C2, fin, fout], capture_output=True, text=True, timeout=110)
        assert r.returncode == 0, r.stderr
        hdr = np.fromfile(fout, dtype=np.int32, count=8)
        cm = np.fromfile(fout, dtype=np.float64, count=3, offset=32)
        out = np.fromfile(fout, dtype=np.float64, count=8 * n, offset=64).reshape(n, 8)
        coll = np.fromfile(fout, dtype=np.int32, count=2 * int(hdr[2]),
                           offset=64 + 64 * n).reshape(-1, 2)
    return hdr, cm, out, coll


needs_probe = pytest.mark.skipif(not os.path.exists(PROBE_C2),
                                 reason="oracle/_ref/probe_c2 not built (oracle/build_ref.py)")


@needs_probe
def test_c2_exact_mesh_field_equals_reference(ctx):
    b = c2_workload()
    print("c2: reference get_neighbours_mesh + vector_field, n=%d" % C2_N, flush=True)
    hdr, cm, out, _ = probe_c2(5, b)
    ex = np.ones(C2_N, np.int32)
    st = ctx.mesh_pass(b.copy(), ex, 0.0, C2_GAM, C2_CUBE, C2_CELLS, 1)
    assert st["total_neighbours"] == int(hdr[4])
    acc, gcm = ctx.mesh_gravity(b, ex, (0.0, 0.0, 0.0, 1.0), 1.0, 0.0, "exact")
    assert np.array_equal(bits(acc), bits(out[:, 3:6]))
    assert np.array_equal(bits(gcm), bits(cm))


@needs_probe
def test_c2_mesh_pass_collisions_equal_reference(ctx):
    import torch
    b = c2_workload()
    print("c2: reference mesh(), n=%d" % C2_N, flush=True)
    hdr, _, out, coll = probe_c2(2 | 0x200, b)
    assert len(coll) > 100
    # Host replay with the oracle's decisions and outcome (what NcorpiON's
    # glue does with its own closest_approach and collision_treatment).
    r = orc.Replay(b, C2_DT, 1, 1.5, 0)
    st = ctx.mesh_pass(r.b, r.e, C2_DT, C2_GAM, C2_CUBE, C2_CELLS, 0, r.test, r.resolve)
    assert np.array_equal(np.array(r.log, np.int32).reshape(-1, 2), coll)
    assert st["total_neighbours"] == int(hdr[4])
    assert st["how_many_small"] == int(hdr[5]) and st["how_many_big"] == int(hdr[6])
    assert np.array_equal(bits(r.b), bits(out))
    # Device replay with the library's bounce.
    bt = torch.from_numpy(b.copy()).cuda()
    torch.cuda.synchronize()
    sd = ctx.mesh_pass_dev(bt, C2_DT, C2_GAM, C2_CUBE, C2_CELLS, 0, 1.5, 0)
    ctx.synchronize()
    assert sd["collisions"] == len(coll)
    assert sd["total_neighbours"] == int(hdr[4])
    assert np.array_equal(bits(bt.cpu().numpy()), bits(out))


def test_c2_mesh_loop_mixed_field_close_to_fp64(ctx):
    """Three steps of the GPU-resident mesh integrator (2^20 bodies,
    collisions) with the mixed field against the FP64-fast one: on the mesh
    path the mixed form is the FP64-fast one (every neighbour term is a
    near one), so the runs agree bit for bit."""
    import torch
    b = c2_workload()
    res = {}
    for prec in ("fast64", "mixed"):
        print("c2: mesh loop, %s field" % prec, flush=True)
        bt = torch.from_numpy(b.copy()).cuda()
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
        et = torch.ones(C2_N, dtype=torch.int32, device="cuda")
        params = {"gam": C2_GAM, "gam_min": C2_GAM, "how_many_neighbours": 100.0,
                  "disruption_threshold": 0.05, "high_dumping_threshold": 100.0,
                  "cells": C2_CELLS}
        torch.cuda.synchronize()
        nc = ctx.mesh_leapfrog_dev(bt, ct, 3, C2_DT, params, 1.0, 0.0, prec, collision_f=1.5,
                                   exists_t=et)
        ctx.synchronize()
        res[prec] = (bt.cpu().numpy(), nc)
    (bf, nf), (bm, nm) = res["fast64"], res["mixed"]
    assert nf > 100 and nm == nf
    assert np.array_equal(bm.view(np.uint64), bf.view(np.uint64))
