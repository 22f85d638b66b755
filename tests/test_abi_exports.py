"""libncg.so loads and exports every function include/ncg.h declares (no GPU needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "ncg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ncg_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("ncg_create", "ncg_gravity_allpairs", "ncg_brute_force_pass", "ncg_mesh_pass",
                 "ncg_mesh_gravity"):
        assert must in names


def test_library_exports_every_symbol():
    from ncorpion_amd import LIB_PATH
    assert os.path.exists(LIB_PATH), "build libncg.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.ncg_abi_version.restype = ctypes.c_int
    assert lib.ncg_abi_version() == 4


def test_python_binding_covers_header():
    from ncorpion_amd import _abi
    bound = set(_abi._SIGNATURES)
    missing = [n for n in declared() if n not in bound]
    assert not missing, missing
