"""ncg_gravity_allpairs_static_dev (the allocation-free entry point) against
ncg_gravity_allpairs_dev on the same inputs, through the C ABI.

Two calls in a row, the second with fewer bodies (the static scratch keeps
the first call's contents): NCG_F64_EXACT bit for bit; the fast forms to
FP64 rounding of the j-split sums (the split count may be lowered to fit the
scratch), far inside the mixed form's 1e-6 bar.  Sizes on both sides of the
direct/tiled boundary (4096 bodies), holes, and the central body's CM_acc.
"""
import numpy as np
import pytest

from ncorpion_amd import ic

pytestmark = pytest.mark.gpu


def _case(n, seed):
    m, R = ic.disk_c3_masses(n, seed)
    b = ic.disk(n, seed=seed, masses=m, radii=R)
    ex = np.ones(n, np.int32)
    ex[np.random.default_rng(seed).choice(n, n // 50, replace=False)] = 0
    return b, ex


def _both(ctx, torch, b, ex, prec, soft):
    n = b.shape[0]
    bt = torch.from_numpy(b).cuda()
    et = torch.from_numpy(ex).cuda()
    central = (0.0, 0.0, 0.0, 1.0)
    out = []
    for fn in (ctx.gravity_allpairs_dev, ctx.gravity_allpairs_static_dev):
        acc = torch.full((n, 3), 7.0, dtype=torch.float64, device="cuda")
        cm = torch.zeros(3, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        fn(bt, acc, et, central, 1.0, soft, prec, cm)
        ctx.synchronize()
        out.append((acc.cpu().numpy(), cm.cpu().numpy()))
    return out


@pytest.mark.parametrize("prec", ["exact", "fast64", "mixed"])
@pytest.mark.parametrize("sizes", [(20000, 3001), (32768, 9000)])
def test_static_entry_point_matches_dev(ctx, prec, sizes):
    import torch
    soft = 1e-6
    for call, n in enumerate(sizes):
        print("static entry point: %s, call %d, n=%d" % (prec, call, n), flush=True)
        b, ex = _case(n, 11 + call)
        (a_ref, cm_ref), (a_st, cm_st) = _both(ctx, torch, b, ex, prec, soft)
        on = ex != 0
        # Holes keep what the caller had (the 7.0 fill) on both paths.
        assert np.array_equal(a_ref[~on], a_st[~on])
        if prec == "exact":
            assert np.array_equal(a_ref.view(np.uint64), a_st.view(np.uint64))
            assert np.array_equal(cm_ref.view(np.uint64), cm_st.view(np.uint64))
        else:
            d = np.linalg.norm(a_ref[on] - a_st[on], axis=1)
            r = np.linalg.norm(a_ref[on], axis=1)
            assert (d / r).max() <= 1e-12, (d / r).max()
            assert np.linalg.norm(cm_ref - cm_st) <= 1e-12 * np.linalg.norm(cm_ref)


def test_static_entry_point_refuses_oversize(ctx):
    import torch
    from ncorpion_amd._abi import NcgError
    n = 32769
    bt = torch.zeros((n, 8), dtype=torch.float64, device="cuda")
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(NcgError, match="static scratch"):
        ctx.gravity_allpairs_static_dev(bt, acc, None, None, 1.0, 0.0, "exact")
