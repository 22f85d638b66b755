"""The Python restatement of mesh() used by the merger/fragmentation GPU test
(toy_mesh in test_gpu_mesh.py) walks the reference's traversal: with no
collision it registers exactly get_neighbours_mesh's pair list and count
(the oracle's, itself pinned on the reference's golden pair lists)."""
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
