#!/usr/bin/env python3
"""Benchmark of the mutual-interaction step on one MI355X.

Workload: brute-force all-pairs gravity of a thin disk of N = 262,144 bodies
around a central mass 1 (G = 1; semi-major axes uniform in [0.9, 1.1],
Rayleigh eccentricity 0.01 and inclination 0.005 rad; masses log-uniform in
[1e-11, 1e-9], radius 1e-4 (m/1e-10)^(1/3); relative softening 1e-6),
integrated with the device RK4 driver (4 force evaluations per step,
dt = 2 pi / 256, hipGraph-captured stages).  Synthetic data, seeded.

A step is one RK4 step.  value = pair-interactions/s of the mixed form,
counting N(N-1) ordered interactions per force evaluation.  Printed as one
JSON line; extra fields give the FP64 form, the accuracy of the mixed form
against it, the roofline of the pair kernel and the reference CPU baseline.
"""
import argparse
import json
import math
import os
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-interactions/s & RK4 steps/s at N=262k (mixed vs FP64); adjoint/fwd time"
FLOP_PER_INTERACTION = 20.0        # convention for all-pairs gravity
PEAK_FP32_VECTOR_TFLOPS = 157.3    # MI355X published FP32 vector peak
PEAK_FP64_VECTOR_TFLOPS = 78.6


COPY_MIB = 256  # plain copy that calibrates the HBM byte counters


def traffic_child(n):
    """Run under rocprofv3 --pmc by traffic_run: one mixed evaluation of the
    workload (its pair kernel is the launch measured) and three passes of a
    plain elementwise copy of COPY_MIB MiB (reads and writes that many bytes
    once each), the reference point of the counters."""
    import torch
    from ncorpion_amd import Context
    b = make_workload(n)
    ctx = Context(0)
    bt = torch.from_numpy(b).cuda()
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    ctx.gravity_allpairs_dev(bt, acc, None, None, 1.0, 1e-6, "mixed")
    ctx.synchronize()
    x = torch.zeros(COPY_MIB * (1 << 18), dtype=torch.float32, device="cuda")
    for _ in range(3):
        y = x * 1.0
    torch.cuda.synchronize()
    del y
    ctx.close()
    return 0


def _pmc_pass(counter, n, timeout):
    """Mean value of `counter` per dispatch of the pair kernel and of the copy
    kernel, from one rocprofv3 pass over traffic_child."""
    import csv
    import glob
    import shutil
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", td, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-child", "--n", str(n)]
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, cwd=td, env=env, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return None
        files = glob.glob(os.path.join(td, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None
        per = {}  # dispatch -> (kernel name, summed value)
        for fn in files:
            with open(fn, newline="") as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    key = (fn, row.get("Dispatch_Id"))
                    name = row.get("Kernel_Name", "")
                    v = float(row.get("Counter_Value", 0.0))
                    per[key] = (name, per.get(key, (name, 0.0))[1] + v)
    pairs = [v for name, v in per.values() if "k_pairs" in name]
    copies = [v for name, v in per.values() if "elementwise" in name or "vectorized" in name]
    if not pairs or not copies:
        return None
    copies = sorted(copies)[-3:]  # the three copy passes (the largest elementwise launches)
    return sum(pairs) / len(pairs), sum(copies) / len(copies), r.returncode


def traffic_run(n, timeout=150):
    """HBM bytes per launch of the mixed pair kernel: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch, each from its own rocprofv3 --pmc pass),
    each scaled by what the same counter reads for a copy of known size."""
    out = {}
    copy_kib = COPY_MIB * 1024.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        r = _pmc_pass(counter, n, timeout)
        if r is None:
            return None
        kib, copy, _ = r
        scale = copy_kib / copy if copy > 0 else float("nan")
        out[counter] = {"kib_per_dispatch": kib, "copy_kib": copy, "scale": scale,
                        "bytes": kib * scale * 1024.0}
    out["bytes"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return out


def pairs_algorithmic_bytes(n):
    """Bytes the pair kernel must move per launch: the sorted FP64 records
    (32 B) and softening lengths (8 B) and the FP32 far-form records (16 B)
    read once, 24 B of acceleration partials written per body."""
    return n * (32 + 8 + 16 + 24)


def make_workload(n, seed=0):
    from ncorpion_amd import ic
    m, R = ic.disk_c3_masses(n, seed)
    return ic.disk(n, seed=seed, masses=m, radii=R)


def cpu_baseline(bodies, n_sample, timeout=300):
    """NcorpiON's own vector_field (compiled from source, 1 thread) on the
    first n_sample bodies of the workload, through the function-level probe."""
    import numpy as np
    probe = os.path.join(ROOT, "oracle", "_ref", "probe_c3", "ref_probe")
    if not os.path.exists(probe):
        return None
    b = np.ascontiguousarray(bodies[:n_sample], dtype=np.float64)
    n = b.shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, 0, n, 0))
            f.write(np.zeros(8).tobytes())
            f.write(np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]).tobytes())
            f.write(b.tobytes())
            f.write(np.ones(n, np.int32).tobytes())
            f.write(np.zeros(n, np.int32).tobytes())
        r = subprocess.run([probe, fin, fout], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=timeout)
        if r.returncode != 0:
            return None
        secs = float(r.stderr.strip().split("op_seconds=")[-1].split()[0])
    return {"value": n * (n - 1) / secs, "unit": "pair-interactions/s", "cores": 1,
            "kind": "reference",
            "sample": "NcorpiON vector_field (brute force, gcc -O2, 1 thread) on the first %d bodies "
                      "of the workload, one evaluation of %.1f s" % (n, secs)}


MESH_N = 1 << 20
MESH_GAM, MESH_CELLS, MESH_CUBE = 0.01, 256, 2.56
# Algorithmic bytes per body of one binning pass: read the 64-byte record,
# write a 64-byte sorted copy, write and read the 4-byte key, write the
# 4-byte permutation.  (The dense cell table is not swept: only the cells of
# the previous and current pass are touched.)
MESH_BIN_BYTES_PER_BODY = 64 + 64 + 4 + 4 + 4


def mesh_workload(n=MESH_N, seed=0):
    from ncorpion_amd import ic
    return ic.disk(n, seed=seed, mass=1e-10, radius=5e-5)


def mesh_run(ctx, torch, passes=3):
    """Mesh module on a disk of 2^20 bodies (m = 1e-10, R = 5e-5) with cells
    of side 0.01 (256 per axis, cube 2.56), dt = 2 pi / 512:
    get_neighbours_mesh passes (mode 1), mesh() passes with the library's
    built-in inelastic bounce (mode 0, f = 1.5), and the neighbour-cell
    gravity in its three forms.  Kernel times from HIP events."""
    import numpy as np
    b0 = mesh_workload()
    n = b0.shape[0]
    dt = 2.0 * math.pi / 512.0
    ex = np.ones(n, np.int32)
    out = {"n": n, "cells": MESH_CELLS, "gam": MESH_GAM}
    # Warm-up: buffers, dense cell table.
    ctx.mesh_pass(b0.copy(), ex, dt, MESH_GAM, MESH_CUBE, MESH_CELLS, 1)
    ctx.profile(True)
    for k in ("mesh_bin", "mesh_count", "mesh_gravity", "mesh_lists"):
        ctx.profile_read(k, reset=True)
    b = b0.copy()
    t0 = time.perf_counter()
    for _ in range(passes):
        st = ctx.mesh_pass(b, ex, dt, MESH_GAM, MESH_CUBE, MESH_CELLS, 1)
    t_gnm = (time.perf_counter() - t0) / passes
    bin_ms, nb = ctx.profile_read("mesh_bin", reset=True)
    cnt_ms, nc = ctx.profile_read("mesh_count", reset=True)
    P = st["total_neighbours"]
    out["get_neighbours_mesh"] = {
        "seconds_per_pass": t_gnm, "total_neighbours": P,
        "bin_ms": bin_ms / nb, "count_ms": cnt_ms / nc,
        "bin_GBps": MESH_BIN_BYTES_PER_BODY * n / (bin_ms / nb * 1e-3) / 1e9,
        "note": "pass time includes the 64 MiB host-to-device upload"}
    # mesh() with collisions.
    coll = 0
    t_mesh = 0.0
    for _ in range(passes):
        b = b0.copy()
        dc = np.zeros(n, np.int32)
        ops = ctx.builtin_ops(b, dt, exists=ex, did_collide=dc, f=1.5)
        t0 = time.perf_counter()
        st = ctx.mesh_pass(b, ex, dt, MESH_GAM, MESH_CUBE, MESH_CELLS, 0, ops=ops)
        t_mesh += time.perf_counter() - t0
        coll = ops.collisions
    bin_ms, nb = ctx.profile_read("mesh_bin", reset=True)
    cnt_ms, nc = ctx.profile_read("mesh_count", reset=True)
    out["mesh"] = {"seconds_per_pass": t_mesh / passes, "collisions": coll,
                   "candidates": st["candidates"], "host_tests": st["host_tests"],
                   "total_neighbours": st["total_neighbours"],
                   "bin_ms": bin_ms / nb, "count_ms": cnt_ms / nc,
                   "predicate_tests_per_s": st["total_neighbours"] / (cnt_ms / nc * 1e-3)}
    # Neighbour-cell gravity at the post-pass state (every pair both ways).
    grav = {}
    ref = None
    for prec in ("exact", "fast64", "mixed"):
        acc, _ = ctx.mesh_gravity(b, ex, None, 1.0, 0.0, prec)
        ctx.profile_read("mesh_gravity", reset=True)  # the first call of a form is not timed
        acc, _ = ctx.mesh_gravity(b, ex, None, 1.0, 0.0, prec)
        ms, k = ctx.profile_read("mesh_gravity", reset=True)
        ms /= k
        g = {"kernel_ms": ms,
             "pair_terms_per_s": 2.0 * st["total_neighbours"] / (ms * 1e-3)}
        if prec == "exact":
            ref = acc
            lms, lk = ctx.profile_read("mesh_lists", reset=True)
            g["order"] = "reference pair-list order (bit for bit)"
            g["lists_ms_per_pass"] = lms / max(lk, 1)
        else:
            d = np.linalg.norm(acc - ref, axis=1)
            r = np.linalg.norm(ref, axis=1)
            g["max_rel_err_vs_exact"] = float((d / r).max())
            g["normwise_err_vs_exact"] = float(d.max() / np.sqrt(np.mean(r * r)))
        if prec == "mixed":
            g["form"] = ("the FP64-fast kernel: every neighbour term is a near one, where FP32 "
                         "terms lose ~1e-5 to cancellation")
        grav[prec] = g
    out["gravity"] = grav
    ctx.profile(False)
    return out


def mesh_loop_run(ctx, torch, steps=10):
    """NcorpiON's mesh integrator resident on the GPU (ncg_mesh_leapfrog_dev,
    SBAB1) on the 2^20-body mesh workload: dt = 2 pi / 512, inelastic bounces
    f = 1.5, 256 cells per axis, mesh size 0.01 (its minimum, where the
    reference keeps it on this disk), target 100 neighbours; after the half
    kick, `steps` steps timed with the exact field and with the FP64-fast
    one.  Each step: top-3 swap, binning, device mesh() replay, drift, field,
    kick, mesh-size update (one host synchronisation)."""
    b0 = mesh_workload()
    dt = 2.0 * math.pi / 512.0
    out = {"n": int(b0.shape[0]), "steps": steps}
    for prec in ("exact", "fast64"):
        bt = torch.from_numpy(b0.copy()).cuda()
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
        params = dict(gam=MESH_GAM, gam_min=MESH_GAM, how_many_neighbours=100.0,
                      disruption_threshold=0.05, high_dumping_threshold=100.0, cells=MESH_CELLS)
        ctx.mesh_leapfrog_dev(bt, ct, 1, dt, params, 1.0, 0.0, prec, 1.5, 0,
                              half_kick_start=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nc = ctx.mesh_leapfrog_dev(bt, ct, steps, dt, params, 1.0, 0.0, prec, 1.5, 0)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out[prec] = {"seconds_per_step": t / steps, "steps_per_s": steps / t, "collisions": nc,
                     "mesh_size": params["gam"]}
    return out


def _probe_mesh(op, b, timeout):
    """Runs the function-level probe (mesh configuration of the 2^20-body
    workload) with `op` on bodies b; returns (seconds, header, CM_acc, bodies)."""
    import numpy as np
    probe = os.path.join(ROOT, "oracle", "_ref", "probe_c2", "ref_probe")
    if not os.path.exists(probe):
        return None
    n = b.shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, op, n, 0))
            f.write(np.array([2.0 * math.pi / 512.0, MESH_GAM, MESH_CUBE, 0, 0, 0, 0, 0]).tobytes())
            f.write(np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]).tobytes())
            f.write(b.tobytes())
            f.write(np.ones(n, np.int32).tobytes())
            f.write(np.zeros(n, np.int32).tobytes())
        r = subprocess.run([probe, fin, fout], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=timeout)
        if r.returncode != 0:
            return None
        secs = float(r.stderr.strip().split("op_seconds=")[-1].split()[0])
        if "gnm_seconds=" in r.stderr:
            secs = (float(r.stderr.split("gnm_seconds=")[-1].split()[0]), secs)
        hdr = np.fromfile(fout, dtype=np.int32, count=8)
        cm = np.fromfile(fout, dtype=np.float64, count=3, offset=32)
        bodies = None
        if not op & 0x100:
            bodies = np.fromfile(fout, dtype=np.float64, count=8 * n, offset=64).reshape(n, 8)
    return secs, hdr, cm, bodies


def cpu_mesh_baseline(ctx=None, timeout=300):
    """NcorpiON's own mesh module, compiled from source, 1 thread, on the full
    2^20-body mesh workload through the function-level probe: mesh() with
    inelastic bounces (f = 1.5), then get_neighbours_mesh and vector_field
    (mesh branch over the pairs it registered).  With a context, the device's
    exact mesh field on the same state is compared with the reference's bit
    for bit."""
    import numpy as np
    b = mesh_workload()
    n = b.shape[0]
    v = _probe_mesh(5, b, timeout)
    if v is None:
        return None
    (gsecs, vsecs), hdr, cm, ob = v
    out = {"value": 1.0 / gsecs, "unit": "get_neighbours_mesh passes/s", "cores": 1,
           "kind": "reference", "seconds": gsecs, "total_neighbours": int(hdr[4]),
           "sample": "NcorpiON get_neighbours_mesh (gcc -O2, 1 thread), one call on the full "
                     "2^20-body mesh workload"}
    out["vector_field"] = {
        "seconds": vsecs, "pair_terms_per_s": 2.0 * int(hdr[4]) / vsecs,
        "sample": "NcorpiON vector_field, mesh branch over the pairs of its own "
                  "get_neighbours_mesh, central mass 1 (gcc -O2, 1 thread)"}
    if ctx is not None:
        ex = np.ones(n, np.int32)
        ctx.mesh_pass(b.copy(), ex, 0.0, MESH_GAM, MESH_CUBE, MESH_CELLS, 1)
        acc, gcm = ctx.mesh_gravity(b, ex, (0.0, 0.0, 0.0, 1.0), 1.0, 0.0, "exact")
        bad = np.any(acc.view(np.uint64) != np.ascontiguousarray(ob[:, 3:6]).view(np.uint64), axis=1)
        out["vector_field"]["gpu_exact_bodies_differing"] = int(bad.sum())
        out["vector_field"]["gpu_exact_cm_acc_identical"] = bool(
            np.array_equal(gcm.view(np.uint64), cm.view(np.uint64)))
    m = _probe_mesh(2 | 0x100, b, timeout)
    if m is not None:
        out["mesh"] = {"seconds": m[0], "collisions": int(m[1][2]),
                       "sample": "NcorpiON mesh() with inelastic bounces f = 1.5, dt = 2 pi/512 "
                                 "(gcc -O2, 1 thread), one call"}
        out["seconds_per_step"] = m[0] + vsecs
        out["step_note"] = ("one SBAB1 step of the mesh integrator = mesh() + vector_field "
                            "(+ O(N) kick and drift, not counted)")
    return out


C1_N, C1_STEPS = 2000, 200


OUT_COLUMNS = ("x", "y", "z", "vx", "vy", "vz", "a", "lambda", "k", "h", "q", "p", "radius", "mass")


def output_run(ctx, passes=3):
    """Output path of display() on the 2^20-body disk (central mass 1 at
    rest at the origin, G = 1): elliptic elements (ncg_cart2ell) and the
    "%.14lf " text of the 14 per-body columns (ncg_format_fixed), written to
    files in a temporary directory.  Seconds per display, end to end."""
    import numpy as np
    b = mesh_workload()
    n = b.shape[0]
    cen = np.zeros(6)
    ctx.cart2ell(b[:1024], None, central=cen, mu=1.0, G=1.0)  # warm-up
    ctx.format_fixed(b[:1024, 0])
    t_lib = t_all = 0.0
    nbytes = 0
    with tempfile.TemporaryDirectory() as td:
        for _ in range(passes):
            t0 = time.perf_counter()
            el, _ = ctx.cart2ell(b, None, central=cen, mu=1.0, G=1.0)
            cols = [b[:, 0], b[:, 1], b[:, 2], b[:, 3], b[:, 4], b[:, 5], el[:, 0], el[:, 1],
                    el[:, 2], el[:, 3], el[:, 4], el[:, 5], b[:, 7], b[:, 6]]
            texts = [ctx.format_fixed(c) for c in cols]
            t1 = time.perf_counter()
            for name, t in zip(OUT_COLUMNS, texts):
                with open(os.path.join(td, name + ".txt"), "ab") as f:
                    f.write(t)
                    f.write(b"\n")
            t2 = time.perf_counter()
            t_lib += t1 - t0
            t_all += t2 - t0
            nbytes = sum(len(t) for t in texts)
    return {"n": n, "seconds_per_display": t_all / passes, "library_seconds": t_lib / passes,
            "text_bytes": nbytes, "columns": len(OUT_COLUMNS)}


def cpu_output_baseline(timeout=300):
    """NcorpiON's display() (gcc -O2, 1 thread) on the same bodies, through
    the function-level probe (op 7: display then file_closing, files in a
    temporary directory)."""
    import numpy as np
    probe = os.path.join(ROOT, "oracle", "_ref", "probe_out", "ref_probe")
    if not os.path.exists(probe):
        return None
    b = mesh_workload()
    n = b.shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, 7 | 0x100, n, 0))
            f.write(np.zeros(8).tobytes())
            f.write(np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]).tobytes())
            f.write(b.tobytes())
            f.write(np.ones(n, np.int32).tobytes())
            f.write(np.zeros(n, np.int32).tobytes())
        r = subprocess.run([probe, fin, fout], cwd=td, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=timeout)
        if r.returncode != 0:
            return None
        secs = float(r.stderr.strip().split("op_seconds=")[-1].split()[0])
    return {"seconds_per_display": secs, "cores": 1, "kind": "reference",
            "sample": "NcorpiON display() of the %d-body disk: cartesian and elliptic elements, "
                      "radii, masses (one call, files in a temporary directory)" % n}


def leapfrog_run(ctx, torch):
    """NcorpiON's brute-force leapfrog on the device (ncg_leapfrog_dev): disk of
    2000 bodies, m = 1e-8, R = 2e-4, dt = 2 pi / 256, 200 steps after the half
    drift, inelastic bounces (f = 1.5); exact field (the reference's
    arithmetic), and the FP64-fast and mixed fields."""
    import numpy as np
    from ncorpion_amd import ic
    b = ic.disk(C1_N, seed=0, mass=1e-8, radius=2e-4)
    dt = 2.0 * math.pi / 256.0
    out = {"n": C1_N, "steps": C1_STEPS}
    for prec in ("exact", "fast64", "mixed"):
        bt = torch.from_numpy(b).cuda()
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
        wb, wc = bt.clone(), ct.clone()
        ctx.leapfrog_dev(wb, wc, 2, dt, 1.0, 0.0, prec, 1.5, 0, graph=True)  # warm-up
        ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nc = ctx.leapfrog_dev(bt, ct, C1_STEPS, dt, 1.0, 0.0, prec, 1.5, 0, graph=True,
                              half_drift_start=True)
        ctx.synchronize()
        t = time.perf_counter() - t0
        out[prec] = {"seconds": t, "steps_per_s": C1_STEPS / t, "collisions": nc}
    return out


def cpu_c1_baseline(timeout=300):
    """NcorpiON's own brute-force run of the same workload (compiled from source,
    1 thread): time inside vector_field and brute_force, from the traced build."""
    import numpy as np
    from ncorpion_amd import ic
    exe = os.path.join(ROOT, "oracle", "_ref", "run_c1", "ncorpion_trace")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        ic.write_init_txt(os.path.join(td, "init.txt"),
                          ic.disk(C1_N, seed=0, mass=1e-8, radius=2e-4))
        env = dict(os.environ, NCG_TRACE=os.path.join(td, "trace.txt"))
        r = subprocess.run([exe], cwd=td, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=timeout)
        if r.returncode != 0:
            return None
        t = {}
        for line in open(os.path.join(td, "trace.txt")):
            if line.startswith("TIME"):
                f = line.split()
                t[f[1]] = float(f[3].split("=")[1])
    hot = t.get("vector_field", 0.0) + t.get("brute_force", 0.0)
    return {"value": C1_STEPS / hot, "unit": "steps/s", "cores": 1, "kind": "reference",
            "seconds_in_vector_field_and_brute_force": hot,
            "sample": "NcorpiON brute-force run, 2000 bodies, 200 steps (gcc -O2, 1 thread): "
                      "wall time inside vector_field and brute_force"}


DROPIN = os.path.join(ROOT, "ncorpion_amd", "glue", "_build")


def _timing_lines(text):
    """{("call" | "kernel", name): {field: value}} from the glue's NCG_TIMING lines."""
    out = {}
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 4 and f[0] == "ncg_timing":
            out[(f[1], f[2])] = {k: float(v) for k, v in (x.split("=") for x in f[3:])}
    return out


def _run_dropin(cfg, bodies, precision, timeout=600):
    """Runs the drop-in NcorpiON of configuration `cfg` (NcorpiON's own main
    and integrator, its hot-path calls redirected to libncg through the glue)
    on `bodies` with NCG_TIMING; returns (wall seconds, timing lines)."""
    from ncorpion_amd import ic
    exe = os.path.join(DROPIN, cfg, "ncorpion_gpu")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        ic.write_init_txt(os.path.join(td, "init.txt"), bodies)
        env = dict(os.environ, NCG_TIMING="1", NCG_PRECISION=precision)
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=td, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=timeout)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError("drop-in %s failed: %s" % (cfg, r.stderr[-2000:]))
    return wall, _timing_lines(r.stderr)


def dropin_run():
    """The drop-in NcorpiON a user builds (ncorpion_amd/glue/build_dropin.py),
    whole runs with per-call uploads: the brute-force workload (2000 bodies,
    200 steps, inelastic bounces), exact and mixed fields, steps/s counted
    over the time inside vector_field and brute_force (as the reference's
    baseline is); and the mesh workload (2^20 bodies, 3 steps), with the time
    inside mesh(), vector_field and get_neighbours_mesh split into device
    kernel time and the rest (uploads, host replay of mesh()'s visiting
    order, glue)."""
    from ncorpion_amd import ic
    out = {}
    b1 = ic.disk(C1_N, seed=0, mass=1e-8, radius=2e-4)
    c1 = {"n": C1_N, "steps": C1_STEPS}
    for prec in ("exact", "mixed"):
        r = _run_dropin("c1", b1, prec)
        if r is None:
            return None
        wall, t = r
        calls = [t.get(("call", k), {}) for k in ("vector_field", "brute_force")]
        hot = sum(c.get("seconds", 0.0) for c in calls)
        first = sum(c.get("first", 0.0) for c in calls)
        # Steady rate: the first call of each function (GPU context creation,
        # kernel loading) is counted apart.
        c1[prec] = {"steps_per_s": (C1_STEPS - 1) / (hot - first),
                    "seconds_in_vector_field_and_brute_force": hot,
                    "first_calls_seconds": first, "process_wall_seconds": wall,
                    "vector_field_calls": int(calls[0].get("calls", 0))}
    out["c1"] = c1
    r = _run_dropin("c2", mesh_workload(), "exact")
    if r is not None:
        wall, t = r
        calls = {k: t[("call", k)] for k in ("mesh", "vector_field", "get_neighbours_mesh")
                 if ("call", k) in t}
        kern = {k[1]: v for k, v in t.items() if k[0] == "kernel"}
        hot = sum(v["seconds"] for v in calls.values())
        first = sum(v.get("first", 0.0) for v in calls.values())
        dev = sum(v["ms"] for v in kern.values()) * 1e-3
        out["c2"] = {"n": MESH_N, "steps": 3, "precision": "exact",
                     "seconds_per_step": hot / 3, "device_kernel_seconds_per_step": dev / 3,
                     "host_and_transfer_seconds_per_step": (hot - dev) / 3,
                     # Each function's first call also creates the GPU context
                     # and loads its kernels; without them, about two steps.
                     "first_calls_seconds": first,
                     "seconds_per_step_without_first_calls": (hot - first) / 2,
                     "calls": {k: {"calls": int(v["calls"]), "seconds": v["seconds"],
                                   "first": v.get("first")}
                               for k, v in calls.items()},
                     "kernels_ms": {k: v["ms"] for k, v in kern.items()},
                     "process_wall_seconds": wall}
    return out


def ensemble_run(ctx, torch, nsys=512, steps=100):
    """512 independent disks (seeds 0..511), N log-uniform in [256, 8192],
    m = 1e-8, R = 2e-4, brute-force gravity + collision detection, mixed form,
    RK4 dt = 2 pi / 256 with hipGraph-captured stages."""
    import numpy as np
    from ncorpion_amd import ic
    sizes = ic.seeded_ensemble_sizes(nsys)
    systems = [ic.disk(int(nn), seed=s, mass=1e-8, radius=2e-4) for s, nn in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cen = np.zeros((nsys, 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    bt = torch.from_numpy(np.concatenate(systems)).cuda()
    ct = torch.from_numpy(cen).cuda()
    dt = 2.0 * math.pi / 256.0
    torch.cuda.synchronize()
    ctx.ensemble_rk4_dev(bt, off, ct, 1, dt, 1.0, 0.0, "mixed", graph=True, collisions=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    nc = ctx.ensemble_rk4_dev(bt, off, ct, steps, dt, 1.0, 0.0, "mixed", graph=True,
                              collisions=True)
    ctx.synchronize()
    t = time.perf_counter() - t0
    inter = float(np.sum(sizes.astype(np.float64) * (sizes - 1))) * 4 * steps
    # Accuracy of one mixed evaluation of every system against the FP64-fast
    # ensemble form, on the mutual accelerations (central bodies absent).
    b0 = torch.from_numpy(np.concatenate(systems)).cuda()
    c0 = torch.from_numpy(np.where(np.arange(8) == 7, 0.0, cen)).cuda()
    acc = {}
    for prec in ("mixed", "fast64"):
        a = torch.zeros((b0.shape[0], 3), dtype=torch.float64, device="cuda")
        cm = torch.zeros((nsys, 3), dtype=torch.float64, device="cuda")
        ctx.ensemble_gravity_dev(b0, off, c0, a, cm, 1.0, 0.0, prec)
        ctx.synchronize()
        acc[prec] = a
    d = torch.linalg.norm(acc["mixed"] - acc["fast64"], dim=1)
    r = torch.linalg.norm(acc["fast64"], dim=1)
    rel = (d / r).cpu().numpy()
    return {"systems": nsys, "bodies": int(sizes.sum()), "steps": steps, "seconds": t,
            "pair_interactions_per_s": inter / t, "system_steps_per_s": nsys * steps / t,
            "collisions": nc, "precision": "mixed", "graph": True,
            "accuracy": {"max_rel_acc_err": float(rel.max()),
                         "p999_rel_acc_err": float(np.quantile(rel, 0.999)),
                         "reference": "FP64-fast ensemble form, mutual acceleration, "
                                      "one evaluation of every system"}}


def collision_filter_run(ctx, b, dt):
    """closest_approach predicate on every pair i < j of the workload (the
    GPU filter of brute_force), timed with HIP events."""
    n = b.shape[0]
    ctx.collide_candidates(b, None, dt)  # warm-up
    ctx.profile(True)
    ctx.profile_read("candidates", reset=True)
    for _ in range(2):
        cand = ctx.collide_candidates(b, None, dt)
    ms, k = ctx.profile_read("candidates", reset=True)
    ctx.profile(False)
    ms /= k
    tests = n * (n - 1) / 2.0
    return {"kernel_ms": ms, "pair_tests_per_s": tests / (ms * 1e-3),
            "candidates": int(cand.shape[0]), "n": n,
            "flop_per_test": 27, "achieved_tflops": 27 * tests / (ms * 1e-3) / 1e12}


def precision_run(ctx, torch, b, soft, dt, steps=20):
    """Energy drift over `steps` RK4 steps of the workload in the mixed and
    the FP64 form (the FP64 energy kernel measures both), and how far the two
    final states are apart."""
    dev = torch.device("cuda:0")
    out, finals = {}, {}
    for prec in ("mixed", "fast64"):
        bt = torch.from_numpy(b).to(dev)
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        e0 = ctx.energy_dev(bt, ct, None, 1.0, soft).clone()
        ctx.rk4_steps_dev(bt, ct, steps, dt, 1.0, soft, prec, graph=True)
        e1 = ctx.energy_dev(bt, ct, None, 1.0, soft).clone()
        ctx.synchronize()
        e0v, e1v = e0.cpu().numpy(), e1.cpu().numpy()
        out["energy_drift_" + prec] = float((e1v[3] - e0v[3]) / abs(e0v[3]))
        finals[prec] = bt
    out["energy_drift_difference"] = out["energy_drift_mixed"] - out["energy_drift_fast64"]
    out["max_position_difference"] = float((finals["mixed"][:, 0:3] - finals["fast64"][:, 0:3])
                                           .abs().max())
    out["steps"] = steps
    return out


ADJ_SOFT = 1.0  # softening length R_i + R_j: smooth field without collisions


def adjoint_run(ctx, torch, n=16384, steps=200, K=20, fd_dirs=4):
    """Collision-free disk of 16384 bodies (m = 1e-8, R = 2e-4, softening
    length R_i + R_j), 200 FP64 RK4 steps dt = 2 pi/256; loss = mean squared
    distance of the final positions to a seeded target (sigma 1e-3); adjoint
    w.r.t. initial positions, velocities and masses with a checkpoint every
    20 steps.  Times: the plain forward run (RK4 driver, graph replay), the
    forward run that keeps the checkpoints, the backward pass from them, and
    the all-in-one entry point that runs its own forward pass."""
    import numpy as np
    from ncorpion_amd import ic
    b = ic.disk(n, seed=0, mass=1e-8, radius=2e-4)
    dt = 2.0 * math.pi / 256.0
    bt = torch.from_numpy(b).cuda()
    c0 = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
    g0 = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    gm = torch.zeros(n, dtype=torch.float64, device="cuda")
    gc = torch.zeros(7, dtype=torch.float64, device="cuda")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        ctx.synchronize()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    # Warm-up (buffers, graph capture, adjoint work arrays), untimed.
    wb, wc = bt.clone(), c0.clone()
    ck, ckc = ctx.rk4_forward_ckpt_dev(wb, wc, K, dt, K, 1.0, ADJ_SOFT)
    ctx.rk4_adjoint_ckpt_dev(ck, ckc, g0, g0.clone(), gm, gc, K, dt, K, 1.0, ADJ_SOFT)
    ctx.rk4_adjoint_dev(bt, c0, g0, g0.clone(), gm, gc, K, dt, K, 1.0, ADJ_SOFT)
    ctx.synchronize()
    fb, fc = bt.clone(), c0.clone()
    t_fwd, _ = timed(lambda: ctx.rk4_steps_dev(fb, fc, steps, dt, 1.0, ADJ_SOFT, "fast64",
                                               graph=True))
    target = fb[:, 0:3] + 1e-3 * torch.tensor(np.random.default_rng(0).normal(size=(n, 3)),
                                              device="cuda")
    gfin = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    gfin[:, 0:3] = 2.0 / n * (fb[:, 0:3] - target)
    kb, kc = bt.clone(), c0.clone()
    t_fck, (ck, ckc) = timed(lambda: ctx.rk4_forward_ckpt_dev(kb, kc, steps, dt, K, 1.0, ADJ_SOFT))
    t_bwd, _ = timed(lambda: ctx.rk4_adjoint_ckpt_dev(ck, ckc, gfin, g0, gm, gc, steps, dt, K, 1.0,
                                                      ADJ_SOFT))
    g_ck = g0.clone()
    t_all, _ = timed(lambda: ctx.rk4_adjoint_dev(bt, c0, gfin, g0, gm, gc, steps, dt, K, 1.0,
                                                 ADJ_SOFT))
    same = bool(torch.equal(g_ck, g0))
    # Directional derivatives of the loss along seeded directions in
    # (x0, v0, m) against Richardson-extrapolated central differences,
    # (4 D(h) - D(2h)) / 3 with h = 2e-8 (the recipe of the 64-direction
    # acceptance test in tests/test_gpu_adjoint.py).
    def loss(bb0):
        bb, cc = bb0.clone(), c0.clone()
        torch.cuda.synchronize()
        ctx.rk4_steps_dev(bb, cc, steps, dt, 1.0, ADJ_SOFT, "fast64", graph=True)
        ctx.synchronize()
        return float(((bb[:, 0:3] - target) ** 2).sum() / n)

    h, fd_err = 2e-8, []
    for q in range(fd_dirs):
        r = np.random.default_rng(100 + q)
        d = np.zeros((n, 8))
        d[:, 0:6] = r.normal(size=(n, 6))
        d[:, 6] = r.normal(size=n) * b[:, 6]
        d = torch.from_numpy(d).cuda()
        ad = float((g0 * d[:, 0:6]).sum() + (gm * d[:, 6]).sum())
        d1 = (loss(bt + h * d) - loss(bt - h * d)) / (2 * h)
        d2 = (loss(bt + 2 * h * d) - loss(bt - 2 * h * d)) / (4 * h)
        fd_err.append(abs((4.0 * d1 - d2) / 3.0 - ad) / abs(ad))
    return {"n": n, "steps": steps, "checkpoint_every": K, "forward_s": t_fwd,
            "forward_with_checkpoints_s": t_fck, "backward_s": t_bwd,
            "adjoint_over_forward": t_bwd / t_fwd,
            "forward_and_backward_over_forward": (t_fck + t_bwd) / t_fwd,
            "all_in_one_s": t_all, "all_in_one_over_forward": t_all / t_fwd,
            "gradients_identical_both_ways": same,
            "fd_check": {"directions": fd_dirs, "h": h,
                         "max_rel_err_richardson": max(fd_err) if fd_err else None,
                         "reference": "Richardson-extrapolated central differences of the "
                                      "loss along seeded directions in (x0, v0, m)"},
            "note": "backward = segment recomputation from the checkpoints + 4 VJPs per step"}



def progress(section):
    """One line on stderr per finished section (the JSON line stays alone on stdout)."""
    print("bench.py: %s done" % section, file=sys.stderr, flush=True)

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--cpu-sample", type=int, default=45000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--no-collisions", action="store_true")
    ap.add_argument("--no-leapfrog", action="store_true")
    ap.add_argument("--no-output", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.traffic_child:
        return traffic_child(args.n)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path is measured at one GPU's "
              "share); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from ncorpion_amd import Context

    n = args.n
    dt = 2.0 * math.pi / 256.0
    soft = 1e-6
    b = make_workload(n)
    ctx = Context(0)
    dev = torch.device("cuda:0")
    interactions = float(n) * (n - 1)

    def rk4_run(precision, steps, warmup, graph):
        """Times `steps` RK4 steps after `warmup` untimed ones.  Without the
        graph, the pair kernel is bracketed by HIP events on its stream."""
        bt = torch.from_numpy(b).to(dev)
        ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.rk4_steps_dev(bt, ct, warmup, dt, 1.0, soft, precision, graph=graph)
        ctx.synchronize()
        kname = "pairs_" + precision
        ctx.profile(not graph)
        ctx.profile_read(kname, reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.rk4_steps_dev(bt, ct, steps, dt, 1.0, soft, precision, graph=graph)
        ctx.synchronize()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        kms, launches = ctx.profile_read(kname, reset=True)
        ctx.profile(False)
        return t, kms, launches, bt, ct

    # Timed region: eager launches, pair kernel timed with HIP events.
    t_mix, kms_mix, nl_mix, bt_mix, _ = rk4_run("mixed", args.steps, args.warmup, False)
    t_graph = None
    if not args.no_graph:
        t_graph = rk4_run("mixed", args.steps, 1, True)[0]
    evals = 4 * args.steps
    value = interactions * evals / t_mix
    if nl_mix == 0:
        raise RuntimeError("no timed launches of the pair kernel were recorded")
    kavg_ms = kms_mix / nl_mix
    achieved_tflops = FLOP_PER_INTERACTION * interactions / (kavg_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mix / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 positions/accumulation, f32 far-field forces",
        "data": "synthetic (seeded disk)",
        "config": {"workload": "brute-force all-pairs gravity, RK4 (4 evaluations/step), disk of "
                               "N=262144 bodies around a central mass, softening 1e-6, dt=2pi/256",
                   "n": n, "precision": "mixed", "timed_launches": "eager (events)",
                   "parallelism": "single GPU"},
        "rk4_steps_per_s": args.steps / t_mix,
        "rk4_steps_per_s_graph": None if t_graph is None else args.steps / t_graph,
        "pair_interactions_per_s_graph": None if t_graph is None else interactions * evals / t_graph,
        "roofline": {"bound": "valu", "achieved": achieved_tflops,
                     "peak": PEAK_FP32_VECTOR_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_VECTOR_TFLOPS,
                     "traffic": None,
                     "kernel": "k_pairs<mixed>", "kernel_avg_ms": kavg_ms,
                     "kernel_launches": nl_mix,
                     "flop_per_interaction": FLOP_PER_INTERACTION},
    }

    if not args.no_traffic:
        tr = traffic_run(n)
        if tr is not None:
            out["roofline"]["traffic"] = tr["bytes"]
            out["roofline"]["traffic_counters"] = tr
            out["roofline"]["algorithmic_bytes"] = pairs_algorithmic_bytes(n)

    if not args.no_fp64:
        fsteps = max(1, args.steps // 4)
        t64, kms64, nl64, bt64, _ = rk4_run("fast64", fsteps, 1, False)
        v64 = interactions * 4 * fsteps / t64
        if nl64 == 0:
            raise RuntimeError("no timed launches of the FP64 pair kernel were recorded")
        k64 = kms64 / nl64
        out["fp64"] = {"pair_interactions_per_s": v64, "rk4_steps_per_s": fsteps / t64,
                       "kernel_avg_ms": k64,
                       "achieved_tflops": FLOP_PER_INTERACTION * interactions / (k64 * 1e-3) / 1e12,
                       "frac_of_fp64_peak": FLOP_PER_INTERACTION * interactions / (k64 * 1e-3)
                       / 1e12 / PEAK_FP64_VECTOR_TFLOPS}
        out["mixed_over_fp64"] = value / v64
        # Accuracy of one mixed evaluation against the FP64 form, mutual part only.
        bt = torch.from_numpy(b).to(dev)
        am = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        af = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        ctx.gravity_allpairs_dev(bt, am, None, None, 1.0, soft, "mixed")
        ctx.gravity_allpairs_dev(bt, af, None, None, 1.0, soft, "fast64")
        ctx.synchronize()
        d = torch.linalg.norm(am - af, dim=1)
        r = torch.linalg.norm(af, dim=1)
        rel = (d / r).cpu().numpy()
        out["accuracy"] = {"max_rel_acc_err": float(rel.max()),
                           "p999_rel_acc_err": float(np.quantile(rel, 0.999)),
                           "normwise_max_err": float(d.max() / torch.sqrt(torch.mean(r * r))),
                           "reference": "FP64-fast form (mutual acceleration only)"}

    progress("headline")
    if not args.no_collisions:
        out["collision_filter"] = collision_filter_run(ctx, b, dt)
        progress("collision_filter")
    if not args.no_precision:
        out["precision_study"] = precision_run(ctx, torch, b, soft, dt)
        progress("precision_study")
    if not args.no_leapfrog:
        out["leapfrog_c1"] = leapfrog_run(ctx, torch)
        progress("leapfrog_c1")
    if not args.no_mesh:
        out["mesh_module"] = mesh_run(ctx, torch)
        progress("mesh_module")
        out["mesh_loop"] = mesh_loop_run(ctx, torch)
        progress("mesh_loop")
    if not args.no_output:
        out["output"] = output_run(ctx)
        progress("output")
    if not args.no_dropin:
        out["dropin"] = dropin_run()
        progress("dropin")
    if not args.no_ensemble:
        out["ensemble"] = ensemble_run(ctx, torch)
        progress("ensemble")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_run(ctx, torch)
        progress("adjoint")
    if args.copy_calibration_mib > 0:
        # Plain elementwise copy (reads and writes the stated bytes once):
        # the reference point for the HBM byte counters.
        x = torch.zeros(args.copy_calibration_mib * (1 << 18), dtype=torch.float32, device=dev)
        for _ in range(3):
            y = x * 1.0
        torch.cuda.synchronize()
        del y
        out["copy_calibration_bytes"] = args.copy_calibration_mib * (1 << 20)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(b, args.cpu_sample)
        if out["cpu_baseline"] is not None and not args.no_mesh:
            out["cpu_baseline"]["mesh"] = cpu_mesh_baseline(ctx)
        if out["cpu_baseline"] is not None and not args.no_leapfrog:
            out["cpu_baseline"]["c1_run"] = cpu_c1_baseline()
        if out["cpu_baseline"] is not None and not args.no_output:
            out["cpu_baseline"]["output"] = cpu_output_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
