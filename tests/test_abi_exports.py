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
    assert lib.ncg_abi_version() == 5


def test_python_binding_covers_header():
    from ncorpion_amd import _abi
    bound = set(_abi._SIGNATURES)
    missing = [n for n in declared() if n not in bound]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the C sizes and field offsets (gcc on the header)."""
    import shutil
    import subprocess
    import pytest
    from ncorpion_amd import _abi
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    mirrors = {"ncg_outcome_model": _abi.OutcomeModel,
               "ncg_leapfrog_model_params": _abi.LeapfrogModelParams,
               "ncg_mesh_loop_params": _abi.MeshLoopParams,
               "ncg_builtin_collider": _abi.BuiltinCollider,
               "ncg_mesh_stats": _abi.MeshStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "ncg.h"', "int main(void) {"]
    for c, py in mirrors.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (c, c))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for c, py in mirrors.items():
        assert int(got[c]) == ctypes.sizeof(py), c
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (c, f)]) == getattr(py, f).offset, (c, f)
