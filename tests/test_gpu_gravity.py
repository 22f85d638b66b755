"""GPU all-pairs gravity (libncg.so) against the oracle.

NCG_F64_EXACT must be bit-identical to the reference's vector_field; the FP64-
fast form must agree to FP64 rounding; the mixed form must stay within the
north-star tolerance of 1e-6 max relative error on the mutual acceleration.
"""
import os

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def rel_err(a, ref):
    """Per-body |a - ref| / |ref| (max over bodies) and normwise max |da| / rms|ref|."""
    d = np.linalg.norm(a - ref, axis=1)
    r = np.linalg.norm(ref, axis=1)
    return float(np.max(d / r)), float(np.max(d) / np.sqrt(np.mean(r * r)))


@pytest.mark.parametrize("name", ["vf_bf_g1_n3", "vf_bf_g1_n257_holes", "vf_bf_soft_n64",
                                  "vf_bf_nocm_n200", "vf_c3_n4096"])
def test_exact_matches_reference_golden(ctx, name):
    g = dict(np.load(os.path.join(GOLD, name + ".npz")))
    c8 = g.get("central")
    cen = (c8[0], c8[1], c8[2], c8[6]) if c8 is not None else None
    ex = g.get("exists")
    acc, cm = ctx.gravity_allpairs(g["bodies"], ex, cen, float(g["G"]), float(g["soft"]), "exact")
    keep = ex.astype(bool) if ex is not None else np.ones(len(acc), bool)
    assert np.array_equal(bits(acc[keep]), bits(g["acc"][keep]))
    if cen is not None:
        assert np.array_equal(bits(cm), bits(g["cm_acc"]))


@pytest.mark.parametrize("n,seed,soft,central,holes", [
    (1, 0, 0.0, True, False), (2, 1, 0.0, False, False), (255, 2, 0.0, True, True),
    (256, 3, 0.3, True, False), (257, 4, 1e-6, False, True), (1000, 5, 0.0, True, True),
    (3001, 6, 0.05, True, False)])
def test_exact_bitwise_random(ctx, n, seed, soft, central, holes):
    b = ic.disk(n, seed=seed, mass=1e-6, radius=1e-3)
    b[:, 6] *= np.random.default_rng(seed).uniform(0.5, 2.0, n)
    ex = None
    if holes:
        ex = (np.random.default_rng(seed + 9).uniform(size=n) > 0.15).astype(np.int32)
    cen = (1e-3, 2e-3, -1e-3, 1.0) if central else None
    ref, rcm = orc.vector_field_brute(b, ex, cen, 1.0, soft)
    acc, cm = ctx.gravity_allpairs(b, ex, cen, 1.0, soft, "exact")
    keep = ex.astype(bool) if ex is not None else np.ones(n, bool)
    assert np.array_equal(bits(acc[keep]), bits(ref[keep]))
    if central:
        assert np.array_equal(bits(cm), bits(rcm))


def test_holes_left_untouched(ctx):
    b = ic.disk(100, seed=7)
    ex = np.ones(100, np.int32)
    ex[::7] = 0
    acc = np.full((100, 3), 123.0)
    ctx.gravity_allpairs(b, ex, (0, 0, 0, 1.0), 1.0, 0.0, "exact", acc=acc)
    assert np.all(acc[~ex.astype(bool)] == 123.0)


def c3_disk(n, seed=0):
    m, R = ic.disk_c3_masses(n, seed)
    return ic.disk(n, seed=seed, masses=m, radii=R)


# 1000 and 4096 run the direct form (up to 4096 bodies), 4097 and 8192 the
# tiled one.
@pytest.mark.parametrize("n", [1000, 4096, 4097, 8192])
def test_fast64_matches_oracle(ctx, n):
    b = c3_disk(n, 11)
    ref, rcm = orc.vector_field_brute(b, None, None, 1.0, 1e-6)
    acc, _ = ctx.gravity_allpairs(b, None, None, 1.0, 1e-6, "fast64")
    mx, nw = rel_err(acc, ref)
    assert mx < 1e-12, mx
    # With the central body the reference adds n mutual terms one by one onto
    # an O(1) central term, so its own rounding is ~sqrt(n) * 1e-16 of |a|.
    ref2, rcm2 = orc.vector_field_brute(b, None, (0, 0, 0, 1.0), 1.0, 1e-6)
    acc2, cm2 = ctx.gravity_allpairs(b, None, (0, 0, 0, 1.0), 1.0, 1e-6, "fast64")
    assert np.max(np.abs(acc2 - ref2) / np.abs(ref2).max()) < 4e-16 * np.sqrt(n) + 1e-15
    assert np.allclose(cm2, rcm2, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [1000, 4096, 4097, 16384])
def test_mixed_within_tolerance(ctx, n):
    # Mutual part only (the central term is FP64 in every form).
    b = c3_disk(n, 12)
    ref, _ = orc.vector_field_brute(b, None, None, 1.0, 1e-6)
    acc, _ = ctx.gravity_allpairs(b, None, None, 1.0, 1e-6, "mixed")
    mx, nw = rel_err(acc, ref)
    assert mx <= 1e-6, (mx, nw)


def test_mixed_with_holes_and_softening(ctx):
    b = ic.disk(3000, seed=13, mass=1e-7, radius=5e-3)
    ex = (np.random.default_rng(1).uniform(size=3000) > 0.2).astype(np.int32)
    ref, rcm = orc.vector_field_brute(b, ex, (0, 0, 0, 1.0), 1.0, 0.5)
    acc, cm = ctx.gravity_allpairs(b, ex, (0, 0, 0, 1.0), 1.0, 0.5, "mixed")
    keep = ex.astype(bool)
    refm, _ = orc.vector_field_brute(b, ex, None, 1.0, 0.5)
    accm, _ = ctx.gravity_allpairs(b, ex, None, 1.0, 0.5, "mixed")
    mx, _ = rel_err(accm[keep], refm[keep])
    assert mx <= 1e-6, mx
    assert np.allclose(cm, rcm, rtol=1e-12, atol=0)


def test_device_entry_point_matches_host(ctx):
    import torch
    b = c3_disk(5000, 14)
    acc_h, cm_h = ctx.gravity_allpairs(b, None, (0, 0, 0, 1.0), 1.0, 1e-6, "mixed")
    bt = torch.from_numpy(b).cuda()
    at = torch.zeros((5000, 3), dtype=torch.float64, device="cuda")
    ct = torch.zeros(3, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.gravity_allpairs_dev(bt, at, None, (0, 0, 0, 1.0), 1.0, 1e-6, "mixed", ct)
    ctx.synchronize()
    assert np.array_equal(at.cpu().numpy(), acc_h)
    assert np.array_equal(ct.cpu().numpy(), cm_h)


@pytest.mark.parametrize("n", [1000, 4096, 4097, 8192])
@pytest.mark.parametrize("prec,tol", [("fast64", 1e-12), ("mixed", 1e-6)])
def test_no_softening_with_large_gm(ctx, n, prec, tol):
    # Without softening (NcorpiON's default) a body's pair with itself has
    # r = 0.  With G m far above 1 (SI or cgs units: here G m spans 1e7..1e9)
    # that term must still add exactly nothing rather than inf * 0.  Both
    # sides of the direct form's 4096-body limit.
    b = c3_disk(n, 15)
    G = 1e18
    ref, _ = orc.vector_field_brute(b, None, None, G, 0.0)
    acc, _ = ctx.gravity_allpairs(b, None, None, G, 0.0, prec)
    assert np.all(np.isfinite(acc))
    mx, _ = rel_err(acc, ref)
    assert mx <= tol, mx


def test_exact_device_entry_point_cm_acc_in_reference_order(ctx):
    # The device entry point sums CM_acc body by body as vector_field does
    # (physics.c:141-146): bit-identical to the oracle, holes included.
    import torch
    n = 3000
    b = ic.disk(n, seed=16, mass=1e-6, radius=1e-3)
    b[:, 6] *= np.random.default_rng(16).uniform(0.5, 2.0, n)
    ex = (np.random.default_rng(17).uniform(size=n) > 0.1).astype(np.int32)
    cen = (1e-3, 2e-3, -1e-3, 1.0)
    ref, rcm = orc.vector_field_brute(b, ex, cen, 1.0, 0.01)
    bt = torch.from_numpy(b).cuda()
    et = torch.from_numpy(ex).cuda()
    at = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    ct = torch.zeros(3, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.gravity_allpairs_dev(bt, at, et, cen, 1.0, 0.01, "exact", ct)
    ctx.synchronize()
    keep = ex.astype(bool)
    assert np.array_equal(bits(at.cpu().numpy()[keep]), bits(ref[keep]))
    assert np.array_equal(bits(ct.cpu().numpy()), bits(rcm))


def test_empty_and_single(ctx):
    acc, cm = ctx.gravity_allpairs(np.zeros((0, 8)), None, (0, 0, 0, 1.0), 1.0, 0.0, "mixed")
    assert acc.shape == (0, 3)
    b = np.array([[1.0, 0, 0, 0, 0, 0, 1e-3, 1e-3]])
    for p in ("exact", "fast64", "mixed"):
        acc, _ = ctx.gravity_allpairs(b, None, (0, 0, 0, 2.0), 1.0, 0.0, p)
        assert acc[0, 0] == -2.0 and acc[0, 1] == 0.0


def central_start(b, ex, cen, G):
    """The central term of vector_field (physics.c:121-136) in numpy, which
    rounds each operation as the reference does: ((-mu) DX) / r^3."""
    cx, cy, cz, M = cen
    mu = G * M
    DX, DY, DZ = b[:, 0] - cx, b[:, 1] - cy, b[:, 2] - cz
    rk = np.sqrt(DX * DX + DY * DY + DZ * DZ)
    r3 = rk * rk * rk
    st = np.stack([-mu * DX / r3, -mu * DY / r3, -mu * DZ / r3], axis=1)
    st[~ex] = 0.0
    return st


@pytest.mark.parametrize("n,prec", [(1000, "exact"), (40000, "exact"), (1000, "fast64"),
                                    (8192, "fast64"), (8192, "mixed")])
def test_start_values_equal_central_term(ctx, n, prec):
    """ncg_gravity_allpairs_from with the central term as the start values
    equals ncg_gravity_allpairs with the central body: bit for bit in the
    exact form (both paths of the exact kernel, n below and above 32,768),
    within the form's tolerance in the others (which add the start values
    after their sums)."""
    rng = np.random.default_rng(n)
    b = ic.disk(n, seed=n, mass=1e-8, radius=2e-4)
    ex = np.ones(n, np.int32)
    ex[rng.choice(n, n // 50, replace=False)] = 0
    cen = (1e-3, -2e-3, 0.0, 1.0)
    G, soft = 1.0, 0.25
    ref, _ = ctx.gravity_allpairs(b, ex, cen, G, soft, prec)
    got = ctx.gravity_allpairs_from(b, central_start(b, ex != 0, cen, G), ex, G, soft, prec)
    live = ex != 0
    if prec == "exact":
        assert np.array_equal(bits(got[live]), bits(ref[live]))
    else:
        tol = 1e-12 if prec == "fast64" else 1e-6
        assert rel_err(got[live], ref[live])[0] < tol
