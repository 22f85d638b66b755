"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

The oracle restates NcorpiON's hot path in plain C (oracle/ncg_oracle.c); it
is the checker of the HIP library and never part of the product.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
L = ctypes.c_long
D = ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_vector_field_brute.argtypes = [P, I, P, I, P, D, D, P, P]
        _lib.oracle_vector_field_mesh.argtypes = [P, I, P, I, P, D, D, P, L, P, P]
        _lib.oracle_vector_field_rows.argtypes = [P, I, P, I, P, D, D, P, I, P]
        _lib.oracle_closest_approach.argtypes = [P, P, I, I, D, P, P]
        _lib.oracle_closest_approach.restype = I
        _lib.oracle_brute_force.argtypes = [P, I, P, P, D, I, D, I, P, L]
        _lib.oracle_brute_force.restype = L
        _lib.oracle_cell_indexes.argtypes = [D, D, D, D, D, I, P]
        _lib.oracle_cell_indexes.restype = I
        _lib.oracle_mesh.argtypes = [P, I, P, P, D, D, D, I, I, D, I, I, P, L, P, P, L, P, P, P, P]
        _lib.oracle_mesh.restype = I
        _lib.oracle_get_neighbours_mesh.argtypes = [P, I, P, D, D, I, P, L, P, P]
        _lib.oracle_get_neighbours_mesh.restype = I
        _lib.oracle_rk4_step.argtypes = [P, I, P, I, P, D, D, D]
        _lib.oracle_energy.argtypes = [P, I, P, I, P, D, D]
        _lib.oracle_energy.restype = D
        _lib.oracle_state_hash.argtypes = [P, I, P]
        _lib.oracle_state_hash.restype = ctypes.c_uint64
        _lib.oracle_resolve.argtypes = [P, P, P, I, I, I, D, I, P, D]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ex(exists, n):
    if exists is None:
        return np.ones(n, dtype=np.int32)
    return np.ascontiguousarray(exists, dtype=np.int32)


def vector_field_brute(bodies, exists=None, central=None, G=1.0, soft=0.0):
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = _ex(exists, n)
    acc = np.zeros((n, 3))
    cm = np.zeros(3)
    c = np.array(central if central is not None else (0, 0, 0, 0), dtype=np.float64)
    lib().oracle_vector_field_brute(_p(b), n, _p(e), int(central is not None), _p(c), G, soft,
                                    _p(acc), _p(cm))
    return acc, cm


def vector_field_rows(bodies, rows, exists=None, central=None, G=1.0, soft=0.0):
    """Field of the bodies `rows` only (each summed over all bodies), as
    vector_field_brute gives it for those rows."""
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = _ex(exists, n)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    acc = np.zeros((len(r), 3))
    c = np.array(central if central is not None else (0, 0, 0, 0), dtype=np.float64)
    lib().oracle_vector_field_rows(_p(b), n, _p(e), int(central is not None), _p(c), G, soft,
                                   _p(r), len(r), _p(acc))
    return acc


def vector_field_mesh(bodies, pairs, exists=None, central=None, G=1.0, soft=0.0):
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = _ex(exists, n)
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    acc = np.zeros((n, 3))
    cm = np.zeros(3)
    c = np.array(central if central is not None else (0, 0, 0, 0), dtype=np.float64)
    lib().oracle_vector_field_mesh(_p(b), n, _p(e), int(central is not None), _p(c), G, soft,
                                   _p(pr), pr.shape[0], _p(acc), _p(cm))
    return acc, cm


def closest_approach(bodies, a, b, dt, exists=None):
    bb = np.ascontiguousarray(bodies, dtype=np.float64)
    e = _ex(exists, bb.shape[0])
    ap = np.zeros(6)
    tc = ctypes.c_double(0.0)
    hit = lib().oracle_closest_approach(_p(bb), _p(e), a, b, dt, _p(ap), ctypes.byref(tc))
    return bool(hit), tc.value, ap


def brute_force(bodies, dt, ctype=1, f=1.5, one_coll=0, exists=None, did_collide=None):
    """Returns (new bodies, did_collide, pairs)."""
    b = np.array(bodies, dtype=np.float64, order="C")
    n = b.shape[0]
    e = _ex(exists, n)
    dc = np.zeros(n, np.int32) if did_collide is None else np.array(did_collide, np.int32)
    cap = 1 << 20
    out = np.zeros((cap, 2), np.int32)
    cnt = lib().oracle_brute_force(_p(b), n, _p(e), _p(dc), dt, ctype, f, one_coll, _p(out), cap)
    return b, dc, out[: min(cnt, cap)].copy()


def mesh(bodies, dt, gam, cube, cells, ctype=1, f=1.5, one_coll=0, mutual=1, exists=None,
         did_collide=None, max_pairs=None):
    b = np.array(bodies, dtype=np.float64, order="C")
    n = b.shape[0]
    e = _ex(exists, n)
    dc = np.zeros(n, np.int32) if did_collide is None else np.array(did_collide, np.int32)
    mp = max_pairs or max(1 << 16, 64 * n)
    pairs = np.zeros((mp, 2), np.int32)
    coll = np.zeros((1 << 16, 2), np.int32)
    np_, nc, tn, hs, hb = (ctypes.c_long() for _ in range(5))
    r = lib().oracle_mesh(_p(b), n, _p(e), _p(dc), dt, gam, cube, cells, ctype, f, one_coll, mutual,
                          _p(pairs), mp, ctypes.byref(np_), _p(coll), coll.shape[0],
                          ctypes.byref(nc), ctypes.byref(tn), ctypes.byref(hs), ctypes.byref(hb))
    if r != 0:
        raise RuntimeError("oracle_mesh failed (pair buffer too small?)")
    return dict(bodies=b, did_collide=dc, pairs=pairs[: np_.value].copy(),
                collisions=coll[: nc.value].copy(), total_neighbours=tn.value,
                how_many_small=hs.value, how_many_big=hb.value)


def get_neighbours_mesh(bodies, gam, cube, cells, exists=None, max_pairs=None):
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = _ex(exists, n)
    mp = max_pairs or max(1 << 16, 64 * n)
    pairs = np.zeros((mp, 2), np.int32)
    np_, tn = ctypes.c_long(), ctypes.c_long()
    r = lib().oracle_get_neighbours_mesh(_p(b), n, _p(e), gam, cube, cells, _p(pairs), mp,
                                         ctypes.byref(np_), ctypes.byref(tn))
    if r != 0:
        raise RuntimeError("oracle_get_neighbours_mesh failed")
    return pairs[: np_.value].copy(), tn.value


def cell_indexes(x, y, z, gam, cube, cells):
    idx = np.zeros(27, dtype=np.int64)
    ok = lib().oracle_cell_indexes(x, y, z, gam, cube, cells, _p(idx))
    return idx if ok else None


def rk4_step(bodies, central7, G, soft, dt, exists=None):
    """central7: (x, y, z, vx, vy, vz, M) or None. Returns new (bodies, central7)."""
    b = np.array(bodies, dtype=np.float64, order="C")
    n = b.shape[0]
    e = _ex(exists, n)
    c = np.array(central7 if central7 is not None else np.zeros(7), dtype=np.float64)
    lib().oracle_rk4_step(_p(b), n, _p(e), int(central7 is not None), _p(c), G, soft, dt)
    return b, c


def energy(bodies, central7, G, soft, exists=None):
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = _ex(exists, n)
    c = np.array(central7 if central7 is not None else np.zeros(7), dtype=np.float64)
    return lib().oracle_energy(_p(b), n, _p(e), int(central7 is not None), _p(c), G, soft)


class Replay:
    """Host decision/outcome callbacks backed by the oracle, on a live array
    (what the NcorpiON glue does with closest_approach/collision_treatment)."""

    def __init__(self, bodies, dt, ctype=1, f=1.5, one_coll=0, exists=None, did_collide=None):
        self.b = np.array(bodies, dtype=np.float64, order="C")
        n = self.b.shape[0]
        self.e = _ex(exists, n)
        self.dc = np.zeros(n, np.int32) if did_collide is None else np.array(did_collide, np.int32)
        self.dt, self.ctype, self.f, self.one = dt, ctype, f, one_coll
        self.ap = np.zeros(6)
        self.tc = ctypes.c_double(0.0)
        self.log = []
        self.tests = 0

    def test(self, a, b):
        self.tests += 1
        hit = lib().oracle_closest_approach(_p(self.b), _p(self.e), a, b, self.dt, _p(self.ap),
                                            ctypes.byref(self.tc))
        return bool(hit) and not self.dc[a] and not self.dc[b]

    def resolve(self, a, b):
        lib().oracle_resolve(_p(self.b), _p(self.e), _p(self.dc), a, b, self.ctype, self.f,
                             self.one, _p(self.ap), self.tc.value)
        self.log.append((a, b))


def state_hash(bodies, exists=None):
    """FNV-1a of the existing bodies as the reference trace logs it."""
    b = np.ascontiguousarray(bodies, dtype=np.float64)
    n = b.shape[0]
    e = None if exists is None else np.ascontiguousarray(exists, dtype=np.int32)
    return int(lib().oracle_state_hash(_p(b), n, None if e is None else _p(e)))
