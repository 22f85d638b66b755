#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs the function-level probe (oracle/ref_probe.c linked with NcorpiON's own
objects by oracle/build_ref.py) on seeded inputs and stores inputs and
outputs as .npz files.  Needs /root/reference (to build the probes) and is run
by hand when the fixture set changes:

    python3 oracle/build_ref.py && python3 tests/golden/make_golden.py
"""
import math
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from ncorpion_amd import ic  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
G_DEFAULT = 39.47841760435743447533796  # parameters.h:130

OPS = {"vf": 0, "bf": 1, "mesh": 2, "gnm": 3, "ca": 4}


def run_probe(cfg, op, bodies, exists=None, did_collide=None, timestep=0.0, gam=0.0, cube=0.0,
              cm=None, pairs=None):
    n = bodies.shape[0]
    exists = np.ones(n, np.int32) if exists is None else exists.astype(np.int32)
    did_collide = np.zeros(n, np.int32) if did_collide is None else did_collide.astype(np.int32)
    cm = np.array([0, 0, 0, 0, 0, 0, 1.0, 1.0]) if cm is None else np.asarray(cm, np.float64)
    pairs = np.zeros((0, 2), np.int32) if pairs is None else np.asarray(pairs, np.int32)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, OPS[op], n, pairs.shape[0]))
            f.write(np.array([timestep, gam, cube, 0, 0, 0, 0, 0], np.float64).tobytes())
            f.write(cm.astype(np.float64).tobytes())
            f.write(np.ascontiguousarray(bodies, np.float64).tobytes())
            f.write(exists.tobytes())
            f.write(did_collide.tobytes())
            f.write(pairs.tobytes())
        subprocess.run([os.path.join(REF, cfg, "ref_probe"), fin, fout], check=True,
                       stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    hdr = struct.unpack_from("<8i", raw, 0)
    off = 32
    cm_acc = np.frombuffer(raw, np.float64, 4, off)[:3].copy()
    off += 32
    out_b = np.frombuffer(raw, np.float64, 8 * n, off).reshape(n, 8).copy()
    off += 64 * n
    ncoll, npairs, n_ca = hdr[2], hdr[3], hdr[7]
    coll = np.frombuffer(raw, np.int32, 2 * ncoll, off).reshape(ncoll, 2).copy()
    off += 8 * ncoll
    prs = np.frombuffer(raw, np.int32, 2 * npairs, off).reshape(npairs, 2).copy()
    off += 8 * npairs
    ca = None
    if n_ca:
        rec = np.dtype([("ab", "<i4", 2), ("hit", "<i4"), ("pad", "<i4"), ("tc", "<f8"),
                        ("ap", "<f8", 6)])
        ca = np.frombuffer(raw, rec, n_ca, off).copy()
    return dict(bodies=out_b, cm_acc=cm_acc, coll=coll, pairs=prs, total_neighbours=hdr[4],
                how_many_small=hdr[5], how_many_big=hdr[6], ca=ca)


def cluster(n, seed, box=0.05, radius=2e-3, vscale=1.0, mass=(1e-9, 1e-8)):
    """Bodies packed in a small box with random velocities (many collisions)."""
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-box / 2, box / 2, (n, 3)) + np.array([1.0, 0.0, 0.0])
    b[:, 3:6] = rng.normal(0.0, vscale, (n, 3))
    b[:, 6] = rng.uniform(mass[0], mass[1], n)
    b[:, 7] = radius * rng.uniform(0.5, 1.5, n)
    return b


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", os.path.relpath(path, ROOT), {k: getattr(v, "shape", v) for k, v in arrays.items()})


def main():
    cm8 = np.array([1e-3, -2e-3, 5e-4, 0, 0, 0, 1.0, 0.05])
    # ---- vector_field, brute force ------------------------------------------------
    b3 = np.array([[1.0, 0.0, 0.0, 0, 0, 0, 1e-3, 1e-3],
                   [0.0, 2.0, 0.0, 0, 0, 0, 2e-3, 1e-3],
                   [0.0, 0.0, -0.5, 0, 0, 0, 3e-3, 1e-3]])
    r = run_probe("probe_bf_g1", "vf", b3, cm=cm8)
    save("vf_bf_g1_n3", bodies=b3, central=cm8, G=1.0, soft=0.0, acc=r["bodies"][:, 3:6],
         cm_acc=r["cm_acc"])

    b = ic.disk(257, seed=1, mass=1e-6, radius=1e-3)
    ex = (np.random.default_rng(2).uniform(size=257) > 0.1).astype(np.int32)
    r = run_probe("probe_bf_g1", "vf", b, exists=ex, cm=cm8)
    save("vf_bf_g1_n257_holes", bodies=b, exists=ex, central=cm8, G=1.0, soft=0.0,
         acc=r["bodies"][:, 3:6], cm_acc=r["cm_acc"])

    b = cluster(64, seed=3, box=0.02, radius=3e-3)
    r = run_probe("probe_bf_soft", "vf", b, cm=cm8)
    save("vf_bf_soft_n64", bodies=b, central=cm8, G=G_DEFAULT, soft=0.5, acc=r["bodies"][:, 3:6],
         cm_acc=r["cm_acc"])

    b = ic.disk(200, seed=4, mass=1e-5, radius=2e-3)
    r = run_probe("probe_bf_nocm", "vf", b, cm=cm8)
    save("vf_bf_nocm_n200", bodies=b, G=1.0, soft=0.1, acc=r["bodies"][:, 3:6])

    m, R = ic.disk_c3_masses(4096, seed=5)
    b = ic.disk(4096, seed=5, masses=m, radii=R)
    r = run_probe("probe_c3", "vf", b, cm=np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]))
    save("vf_c3_n4096", bodies=b, central=np.array([0, 0, 0, 0, 0, 0, 1.0, 0.05]), G=1.0,
         soft=1e-6, acc=r["bodies"][:, 3:6], cm_acc=r["cm_acc"])

    # ---- closest_approach on every pair -------------------------------------------
    b = cluster(48, seed=6, box=0.02, radius=2e-3, vscale=1.0)
    r = run_probe("probe_bf_g1", "ca", b, timestep=0.01)
    ca = r["ca"]
    save("ca_n48", bodies=b, dt=0.01, ab=ca["ab"], hit=ca["hit"], tc=ca["tc"], approach=ca["ap"])
    print("   closest_approach hits:", int(ca["hit"].sum()))

    # ---- brute_force passes -------------------------------------------------------
    b = cluster(300, seed=7, box=0.06, radius=2e-3, vscale=1.0)
    r = run_probe("probe_bf_g1", "bf", b, timestep=0.01)
    save("bf_inelastic_n300", bodies=b, dt=0.01, f=1.5, ctype=1, one_coll=0,
         out_bodies=r["bodies"], coll=r["coll"])
    print("   collisions:", r["coll"].shape[0])

    b = cluster(200, seed=8, box=0.05, radius=2.5e-3, vscale=1.0)
    ex = (np.random.default_rng(9).uniform(size=200) > 0.05).astype(np.int32)
    r = run_probe("probe_bf_soft", "bf", b, exists=ex, timestep=0.01)
    save("bf_elastic_once_n200", bodies=b, exists=ex, dt=0.01, f=2.0, ctype=0, one_coll=1,
         out_bodies=r["bodies"], coll=r["coll"])
    print("   collisions:", r["coll"].shape[0])

    b = cluster(250, seed=10, box=0.06, radius=2e-3, vscale=1.0)
    r = run_probe("probe_bf_nocm", "bf", b, timestep=0.01)
    save("bf_inelastic125_n250", bodies=b, dt=0.01, f=1.25, ctype=1, one_coll=0,
         out_bodies=r["bodies"], coll=r["coll"])
    print("   collisions:", r["coll"].shape[0])

    # ---- mesh (cells 16) ----------------------------------------------------------
    cells, gam = 16, 0.01
    cube = gam * cells
    rng = np.random.default_rng(11)
    n = 1500
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-0.085, 0.085, (n, 3))     # some bodies outside the cube (|x| > 0.08)
    b[:40, 0] = rng.uniform(0.07, 0.08, 40)            # near the +x edge: key aliasing
    b[40:80, 1] = rng.uniform(0.07, 0.08, 40)          # near the +y edge
    b[:, 3:6] = rng.normal(0, 0.3, (n, 3))
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = rng.uniform(1e-3, 4e-3, n)
    b[[5, 300, 900], 7] = [0.012, 0.02, 0.015]          # three big bodies (R > gam)
    ex = (rng.uniform(size=n) > 0.03).astype(np.int32)
    r = run_probe("probe_mesh", "mesh", b, exists=ex, timestep=0.01, gam=gam, cube=cube)
    save("mesh_inelastic_n1500", bodies=b, exists=ex, dt=0.01, gam=gam, cube=cube, cells=cells,
         f=1.5, ctype=1, one_coll=0, out_bodies=r["bodies"], coll=r["coll"], pairs=r["pairs"],
         total_neighbours=r["total_neighbours"], how_many_small=r["how_many_small"],
         how_many_big=r["how_many_big"])
    print("   mesh collisions:", r["coll"].shape[0], "pairs:", r["pairs"].shape[0],
          "tn:", r["total_neighbours"])

    r = run_probe("probe_mesh", "gnm", b, exists=ex, timestep=0.01, gam=gam, cube=cube)
    save("gnm_n1500", bodies=b, exists=ex, gam=gam, cube=cube, cells=cells, pairs=r["pairs"],
         total_neighbours=r["total_neighbours"])
    gnm_pairs = r["pairs"]

    # vector_field, mesh branch, with the pair list of get_neighbours_mesh
    b2 = b.copy()
    b2[:, 0:3] += 1e-3 * rng.normal(size=(n, 3))        # positions moved after the binning
    r = run_probe("probe_mesh", "vf", b2, exists=ex, gam=gam, cube=cube, cm=cm8, pairs=gnm_pairs)
    save("vf_mesh_n1500", bodies=b2, exists=ex, central=cm8, G=1.0, soft=0.0, pairs=gnm_pairs,
         acc=r["bodies"][:, 3:6], cm_acc=r["cm_acc"])

    cells, gam = 32, 0.005
    cube = gam * cells
    rng = np.random.default_rng(12)
    n = 1200
    b = np.zeros((n, 8))
    b[:, 0:3] = rng.uniform(-0.07, 0.07, (n, 3))
    b[:, 3:6] = rng.normal(0, 0.3, (n, 3))
    b[:, 6] = rng.uniform(1e-9, 1e-8, n)
    b[:, 7] = rng.uniform(1e-3, 2.4e-3, n)
    r = run_probe("probe_mesh_el", "mesh", b, timestep=0.01, gam=gam, cube=cube)
    save("mesh_elastic_once_n1200", bodies=b, dt=0.01, gam=gam, cube=cube, cells=cells, f=2.0,
         ctype=0, one_coll=1, out_bodies=r["bodies"], coll=r["coll"], pairs=r["pairs"],
         total_neighbours=r["total_neighbours"], how_many_small=r["how_many_small"],
         how_many_big=r["how_many_big"])
    print("   mesh collisions:", r["coll"].shape[0], "pairs:", r["pairs"].shape[0])


if __name__ == "__main__":
    main()
