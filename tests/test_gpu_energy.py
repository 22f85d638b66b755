"""Energy kernel (ncg_energy_dev) against the oracle's compensated FP64 sum,
and the energy drift of the device RK4 (mixed and FP64) against the oracle's
RK4 on the same initial state."""
import math

import numpy as np
import pytest

import oracle_lib as orc
from ncorpion_amd import ic

pytestmark = pytest.mark.gpu


def tt():
    import torch
    return torch


@pytest.mark.parametrize("central,holes,soft", [(True, False, 0.0), (True, True, 0.05),
                                                (False, True, 1e-3)])
def test_energy_matches_oracle(ctx, central, holes, soft):
    torch = tt()
    n = 1500
    b = ic.disk(n, seed=11, mass=1e-6, radius=1e-3)
    rng = np.random.default_rng(12)
    b[:, 6] *= rng.uniform(0.1, 10.0, n)
    ex = np.ones(n, np.int32)
    if holes:
        ex[rng.choice(n, 200, replace=False)] = 0
        b[ex == 0, 0:3] = b[0, 0:3]  # stale records may sit anywhere, even on a live body
    c7 = np.array([0.01, -0.02, 0.001, 1e-4, 2e-4, -3e-4, 1.3])
    ref = orc.energy(b, c7 if central else None, 1.3, soft, exists=ex)
    bt = torch.from_numpy(b).cuda()
    et = torch.from_numpy(ex).cuda()
    ct = torch.tensor(np.concatenate([c7, [1.0 if central else 0.0]]), device="cuda")
    torch.cuda.synchronize()
    out = ctx.energy_dev(bt, ct, et, 1.3, soft)
    ctx.synchronize()
    got = out.cpu().numpy()
    assert abs(got[3] - ref) <= 1e-13 * abs(ref), (got, ref)
    assert got[3] == got[0] + got[1] + got[2]
    if not central:
        assert got[2] == 0.0


def test_energy_two_bodies_closed_form(ctx):
    torch = tt()
    b = np.zeros((2, 8))
    b[0, :] = [0.3, 0.0, 0.0, 0.0, 0.5, 0.0, 2e-3, 1e-3]
    b[1, :] = [-0.1, 0.2, 0.0, 0.1, 0.0, -0.2, 5e-3, 2e-3]
    G, s = 0.7, 0.5
    r2 = 0.4 ** 2 + 0.2 ** 2
    eps = s * (1e-3 + 2e-3)
    expect_pot = -G * 2e-3 * 5e-3 / math.sqrt(r2 + eps * eps)
    expect_kin = 0.5 * 2e-3 * 0.25 + 0.5 * 5e-3 * (0.01 + 0.04)
    out = ctx.energy_dev(torch.from_numpy(b).cuda(), None, None, G, s)
    ctx.synchronize()
    got = out.cpu().numpy()
    assert math.isclose(got[1], expect_pot, rel_tol=1e-14)
    assert math.isclose(got[0], expect_kin, rel_tol=1e-14)


@pytest.mark.parametrize("prec,tol", [("fast64", 1e-13), ("mixed", 1e-9)])
def test_rk4_energy_drift_matches_oracle(ctx, prec, tol):
    """20 RK4 steps of a disk with the C3 mass spectrum (softening 1e-6):
    the drift (E_20 - E_0) / |E_0| of the device run against the oracle's."""
    torch = tt()
    n, dt, steps, soft = 1024, 2 * math.pi / 256, 20, 1e-6
    m, R = ic.disk_c3_masses(n, seed=21)
    b = ic.disk(n, seed=21, masses=m, radii=R)
    c7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
    e0 = orc.energy(b, c7, 1.0, soft)
    ref, rc = b.copy(), c7.copy()
    for _ in range(steps):
        ref, rc = orc.rk4_step(ref, rc, 1.0, soft, dt)
    drift_ref = (orc.energy(ref, rc, 1.0, soft) - e0) / abs(e0)
    bt = torch.from_numpy(b).cuda()
    ct = torch.tensor([0, 0, 0, 0, 0, 0, 1.0, 1.0], dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    e0g = ctx.energy_dev(bt, ct, None, 1.0, soft).clone()
    ctx.rk4_steps_dev(bt, ct, steps, dt, 1.0, soft, prec, graph=False)
    e1g = ctx.energy_dev(bt, ct, None, 1.0, soft)
    ctx.synchronize()
    e0g, e1g = float(e0g[3]), float(e1g[3])
    assert abs(e0g - e0) <= 1e-14 * abs(e0)
    drift = (e1g - e0g) / abs(e0g)
    print("drift %s %.3e oracle %.3e" % (prec, drift, drift_ref))
    assert abs(drift - drift_ref) <= tol, (drift, drift_ref)
