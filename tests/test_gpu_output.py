"""Output path (display(), display.c:239-451) on the GPU.

ncg_format_fixed against the exact decimal conversion of printf("%.<D>f")
(Python's formatting: the exact binary value, ties to even, as glibc's), and
ncg_cart2ell against NcorpiON's own cart2ell (structure.c:172-284) compiled
into the function-level probe, bit for bit.  Whole runs with output files are
compared byte for byte in test_gpu_dropin.py.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "oracle", "_ref", "probe_bf_g1", "ref_probe")


def printf_fixed(v, d):
    if math.isnan(v):
        return ("-nan " if math.copysign(1.0, v) < 0 else "nan ")
    return "%.*f " % (d, v)


def special_values():
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 1.5, 2.5, 2.0 ** -15, 3 * 2.0 ** -15, -(2.0 ** -15), 1e-20,
         -1e-20, 1e300, -1e300, 1.7976931348623157e308, 5e-324, 2.2250738585072014e-308,
         123456789012345678.0, 0.1, 0.7, 1e17, 9.999999999999999e-15, 5e-15, 2.0 / 3.0,
         float("nan"), -float("nan"), float("inf"), -float("inf")]
    return np.array(v)


@pytest.mark.parametrize("decimals", [14, 0, 1, 17])
def test_format_fixed_matches_printf(ctx, decimals):
    rng = np.random.default_rng(decimals)
    vals = [special_values(),
            rng.integers(0, 2 ** 63, 20000, dtype=np.int64).astype(np.uint64).view(np.float64),
            -rng.integers(0, 2 ** 63, 20000, dtype=np.int64).astype(np.uint64).view(np.float64),
            rng.normal(0, 1, 20000) * 10.0 ** rng.integers(-20, 20, 20000),
            # Exact ties at the last printed digit: odd multiples of 2^-(D+1)
            # (v 10^D = odd 5^D / 2); and odd multiples of other powers of two.
            np.ldexp(rng.integers(0, 2 ** 40, 20000) * 2 + 1.0, -(decimals + 1)),
            -np.ldexp(rng.integers(0, 2 ** 20, 20000) * 2 + 1.0, -(decimals + 1)),
            np.ldexp(rng.integers(0, 2 ** 20, 20000) * 2 + 1.0, -rng.integers(1, 60, 20000))]
    v = np.concatenate(vals)
    got = ctx.format_fixed(v, decimals).decode()
    want = "".join(printf_fixed(float(x), decimals) for x in v)
    assert got == want


def test_format_fixed_empty_and_capacity(ctx):
    assert ctx.format_fixed(np.zeros(0)) == b""
    assert ctx.format_fixed(np.array([-0.0])) == b"-0.00000000000000 "


def probe_cart2ell(b, ex, cm, mu):
    import tempfile
    n = b.shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", 0x4E434750, 6, n, 0))
            f.write(np.array([0.0, 0.0, 0.0, mu, 0, 0, 0, 0]).tobytes())
            f.write(np.asarray(cm, dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(b, dtype=np.float64).tobytes())
            f.write(ex.astype(np.int32).tobytes())
            f.write(np.zeros(n, np.int32).tobytes())
        subprocess.run([PROBE, fin, fout], check=True, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=120)
        raw = open(fout, "rb").read()
    return np.frombuffer(raw, dtype=np.float64, count=8 * n, offset=64).reshape(n, 8)[:, :6]


def test_cart2ell_matches_reference(ctx):
    if not os.path.exists(PROBE):
        pytest.skip("reference probe missing (oracle/build_ref.py)")
    from ncorpion_amd import ic
    rng = np.random.default_rng(5)
    b = ic.disk(4000, seed=5, mass=1e-8, radius=2e-4)
    b[:200, 3:6] *= 1.6                     # hyperbolic
    b[200:300, 3:6] *= -1.0                 # retrograde
    b[300:400, 2] += rng.normal(0, 0.3, 100)  # inclined
    b[400:450, 3:6] = rng.normal(0, 0.5, (50, 3))
    cm = np.array([1e-3, -2e-3, 5e-4, 1e-4, 2e-4, -3e-4, 1.0, 0.05])
    b[450, 0:6] = cm[0:6]                   # at the central body, at rest: R = 0, V2 = 0
    b[451, 3:6] = cm[3:6]                   # V2 = 0
    ex = (rng.uniform(size=4000) > 0.05).astype(np.int32)
    ex[450] = ex[451] = 1
    mu = 1.0 * cm[6]
    got, warn = ctx.cart2ell(b, ex, central=cm[:6], mu=mu, G=1.0)
    want = probe_cart2ell(b, ex, cm, mu)
    live = ex != 0
    g = np.ascontiguousarray(got[live]).view(np.uint64)
    w = np.ascontiguousarray(want[live]).view(np.uint64)
    # NaN results (R = 0) compare by bits as well: the same operations give the same NaN.
    assert np.array_equal(g, w), np.argwhere(g != w)[:5]
