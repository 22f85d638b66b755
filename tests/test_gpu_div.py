"""FP64 division with a shared reciprocal (csrc/exact_div.hpp) against `/`.

The exact gravity kernels form the reference's three quotients
G m dx / D^3, G m dy / D^3, G m dz / D^3 (physics.c:186-200) with one
reciprocal of D^3; tests/native/div_check (built by tests/native/Makefile with
the library's device flags) compares that against the division operator bit for bit on 2^24
triples covering every bit pattern, the edges of the fast window and the
magnitudes of a disk.
"""
import pathlib
import subprocess

import pytest

TOOL = pathlib.Path(__file__).resolve().parent / "native" / "div_check"


@pytest.mark.gpu
def test_shared_reciprocal_division_is_bit_identical():
    assert TOOL.exists(), f"{TOOL} is not built (make -C tests/native)"
    print("div_check: 2^24 triples", flush=True)
    r = subprocess.run([str(TOOL)], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
