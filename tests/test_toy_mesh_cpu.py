"""The Python restatement of mesh() used by the merger/fragmentation GPU tests
(toy_mesh in test_gpu_mesh.py) walks the reference's traversal: with no
collision it registers exactly get_neighbours_mesh's pair list and count,
and with bodies larger than the mesh size and bounces it gives exactly the
oracle's mesh() (both oracle functions pinned on the reference's golden pair
lists and collision sequences, big bodies included)."""
import numpy as np

import oracle_lib as orc
from test_gpu_mesh import MeshToyWorld, blob, toy_mesh


def test_toy_mesh_traversal_is_the_reference_order():
    n0, cap, dt, cells, gam = 3000, 6000, 0.01, 16, 0.01
    b = blob(n0, 31, half=0.06, rmin=1.5e-3, rmax=3e-3)
    w = MeshToyWorld(b, cap, dt)
    w.test = lambda a, c: False
    total, small = toy_mesh(w, n0, gam, cells * gam, cells)
    pairs, tn = orc.get_neighbours_mesh(b, gam, cells * gam, cells)
    assert total == tn and small == n0
    assert np.array_equal(np.array(w.pairs, np.int32).reshape(-1, 2), pairs)


class BounceWorld:
    """Decision and outcome of the oracle (closest_approach, inelastic
    bounce f = 1.5) on a live array, with the toy world's fields."""

    def __init__(self, b, dt):
        self.r = orc.Replay(b, dt, 1, 1.5, 0)
        self.b, self.e = self.r.b, self.r.e
        self.largest = b.shape[0] - 1
        self.pairs = []

    def test(self, a, c):
        return self.r.test(a, c)

    def resolve(self, a, c):
        self.r.resolve(a, c)
        self.pairs.append((a, c))


def test_toy_mesh_with_big_bodies_is_the_oracles_mesh():
    n0, dt, cells, gam = 3000, 0.01, 16, 0.01
    b = blob(n0, 33, half=0.06, rmin=1.5e-3, rmax=3e-3)
    rng = np.random.default_rng(33)
    bigs = rng.choice(np.arange(n0), 12, replace=False)
    b[bigs, 7] = rng.uniform(1.2 * gam, 2.5 * gam, 12)
    b[bigs[:3], 7] = gam  # equal to the mesh size: still small (collision.c:344)
    w = BounceWorld(b, dt)
    total, small, big = toy_mesh(w, n0, gam, cells * gam, cells, with_big=True)
    ref = orc.mesh(b, dt, gam, cells * gam, cells, 1, 1.5, 0, 1)
    assert np.array_equal(np.array(w.r.log, np.int32).reshape(-1, 2), ref["collisions"])
    assert len(ref["collisions"]) > 20
    assert (total, small, big) == (ref["total_neighbours"], ref["how_many_small"],
                                   ref["how_many_big"])
    assert big == 9
    assert np.array_equal(np.array(w.pairs, np.int32).reshape(-1, 2), ref["pairs"])
    assert np.array_equal(w.b.view(np.uint64), ref["bodies"].view(np.uint64))
