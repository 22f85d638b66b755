"""The decimal formatter of the output path (fmt_fixed in
ncorpion_amd/csrc/output.hip) compiled for the host and compared with the C
library's printf("%.<D>f ") over random bit patterns (every exponent,
subnormals, NaN payloads), exact decimal ties and special values, D = 0..17.
The GPU kernels run the same function; tests/test_gpu_output.py checks them
on the device."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HEADER = r'''
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#define __device__
static inline long long __double_as_longlong(double v) { long long r; memcpy(&r, &v, 8); return r; }
static inline unsigned long long __umul64hi(unsigned long long a, unsigned long long b)
{ return (unsigned long long)(((unsigned __int128)a * b) >> 64); }
'''

MAIN = r'''
static long bad = 0, tot = 0;
static void check(double v, int D)
{
    char a[2048], b[2048];
    const int n = fmt_fixed(v, D, a), n2 = fmt_fixed(v, D, nullptr);
    snprintf(b, sizeof b, "%.*f ", D, v);
    ++tot;
    if (n != n2 || (int)strlen(b) != n || memcmp(a, b, n)) {
        if (bad < 5) { a[n] = 0; printf("D=%d v=%a got '%s' want '%s'\n", D, v, a, b); }
        ++bad;
    }
}
int main()
{
    std::mt19937_64 r(7);
    const double sp[] = {0., -0., 1., -1., 0.5, 1.5, 2.5, std::ldexp(1, -15), 3 * std::ldexp(1, -15),
                         1e-20, -1e-20, 1e300, -1.7976931348623157e308, std::ldexp(1, -1074),
                         std::ldexp(1, -1022), 123456789012345678.0, 0.1, 9.999999999999999e-15,
                         5e-15, NAN, -NAN, INFINITY, -INFINITY};
    for (double v : sp)
        for (int D = 0; D <= 17; ++D) check(v, D);
    for (long i = 0; i < 400000; ++i) {
        const uint64_t bits = r();
        double v;
        memcpy(&v, &bits, 8);
        check(v, (int)(r() % 18));
        const int D = (int)(r() % 18);
        check(std::ldexp((double)((r() >> 24) | 1), -(D + 1)), D);            // exact ties
        check(std::ldexp((double)(r() >> 11), (int)(r() % 140) - 110), 14);  // display's range
    }
    printf("checked %ld mismatches %ld\n", tot, bad);
    return bad != 0;
}
'''


def test_formatter_matches_printf(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = open(os.path.join(ROOT, "ncorpion_amd", "csrc", "output.hip")).read()
    consts = "\n".join(re.findall(r"^constexpr int (?:FMT_MAXD|LIMBS) = \d+;$", src, re.M))
    a = src.index("__device__ int fmt_fixed")
    b = src.index("__global__", a)
    prog = tmp_path / "fmt.cpp"
    prog.write_text(HEADER + consts + "\n" + src[a:b] + MAIN)
    exe = tmp_path / "fmt"
    subprocess.run([cxx, "-O2", "-std=c++17", str(prog), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
