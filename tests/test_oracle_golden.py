"""The CPU oracle against golden vectors of the compiled reference (bit-exact).

The fixtures in tests/golden/ were written by tests/golden/make_golden.py from
NcorpiON's own objects (oracle/build_ref.py).  These tests pin the oracle
before the oracle is used to check the GPU library.
"""
import glob
import os

import numpy as np
import pytest

import oracle_lib as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def central4(c8):
    return (c8[0], c8[1], c8[2], c8[6])


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", ["vf_bf_g1_n3", "vf_bf_g1_n257_holes", "vf_bf_soft_n64",
                                  "vf_bf_nocm_n200", "vf_c3_n4096"])
def test_vector_field_brute_bitwise(name):
    g = load(name)
    cen = central4(g["central"]) if "central" in g else None
    acc, cm = orc.vector_field_brute(g["bodies"], g.get("exists"), cen, float(g["G"]),
                                     float(g["soft"]))
    ex = g["exists"].astype(bool) if "exists" in g else np.ones(len(acc), bool)
    assert np.array_equal(bits(acc[ex]), bits(g["acc"][ex]))
    if cen is not None:
        assert np.array_equal(bits(cm), bits(g["cm_acc"]))


def test_vector_field_mesh_bitwise():
    g = load("vf_mesh_n1500")
    acc, cm = orc.vector_field_mesh(g["bodies"], g["pairs"], g["exists"], central4(g["central"]),
                                    float(g["G"]), float(g["soft"]))
    ex = g["exists"].astype(bool)
    assert np.array_equal(bits(acc[ex]), bits(g["acc"][ex]))
    assert np.array_equal(bits(cm), bits(g["cm_acc"]))


def test_closest_approach_every_pair():
    g = load("ca_n48")
    for (a, b), hit, tc, ap in zip(g["ab"], g["hit"], g["tc"], g["approach"]):
        h, t, app = orc.closest_approach(g["bodies"], int(a), int(b), float(g["dt"]))
        assert h == bool(hit)
        if hit:
            assert t == tc
            assert np.array_equal(bits(app), bits(ap))
    assert g["hit"].sum() > 10


@pytest.mark.parametrize("name", ["vf_bf_g1_n257_holes", "vf_bf_soft_n64", "vf_bf_nocm_n200",
                                  "vf_c3_n4096"])
def test_vector_field_rows_equal_reference_rows(name):
    # The row-subset form (the checker of fields too large for the full
    # O(N^2) oracle) against the reference's own accelerations of those rows.
    g = load(name)
    cen = central4(g["central"]) if "central" in g else None
    n = len(g["bodies"])
    ex = g.get("exists")
    keep = ex.astype(bool) if ex is not None else np.ones(n, bool)
    rows = np.flatnonzero(keep)[np.random.default_rng(n).permutation(int(keep.sum()))[:97]]
    acc = orc.vector_field_rows(g["bodies"], rows, ex, cen, float(g["G"]), float(g["soft"]))
    assert np.array_equal(bits(acc), bits(g["acc"][rows]))


@pytest.mark.parametrize("name", ["bf_inelastic_n300", "bf_elastic_once_n200",
                                  "bf_inelastic125_n250"])
def test_brute_force_pass(name):
    g = load(name)
    out, dc, pairs = orc.brute_force(g["bodies"], float(g["dt"]), int(g["ctype"]), float(g["f"]),
                                     int(g["one_coll"]), g.get("exists"))
    assert np.array_equal(pairs, g["coll"])
    assert np.array_equal(bits(out), bits(g["out_bodies"]))
    assert len(pairs) > 20


@pytest.mark.parametrize("name", ["mesh_inelastic_n1500", "mesh_elastic_once_n1200"])
def test_mesh_pass(name):
    g = load(name)
    r = orc.mesh(g["bodies"], float(g["dt"]), float(g["gam"]), float(g["cube"]), int(g["cells"]),
                 int(g["ctype"]), float(g["f"]), int(g["one_coll"]), 1, g.get("exists"))
    assert np.array_equal(r["collisions"], g["coll"])
    assert np.array_equal(r["pairs"], g["pairs"])
    assert r["total_neighbours"] == int(g["total_neighbours"])
    assert r["how_many_small"] == int(g["how_many_small"])
    assert r["how_many_big"] == int(g["how_many_big"])
    assert np.array_equal(bits(r["bodies"]), bits(g["out_bodies"]))


def test_get_neighbours_mesh():
    g = load("gnm_n1500")
    pairs, tn = orc.get_neighbours_mesh(g["bodies"], float(g["gam"]), float(g["cube"]),
                                        int(g["cells"]), g["exists"])
    assert np.array_equal(pairs, g["pairs"])
    assert tn == int(g["total_neighbours"])


def test_fixture_set_complete():
    names = {os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz"))}
    assert len(names) >= 13
