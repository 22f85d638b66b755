// Checks div3 (exact_div.hpp) against the `/` operator bit for bit on the
// GPU: 2^24 triples of numerators over one denominator, drawn from every bit
// pattern (zeros, denormals, infinities, NaNs), from exponents around and on
// the edges of div3's window, and from the magnitudes of G m dx and D^3 in a
// disk of bodies.  Prints the counts and exits 1 on any mismatch.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "exact_div.hpp"

#pragma clang fp contract(off)

namespace {

__global__ void k_div_check(const double *n, const double *d, double *q_div3, double *q_op, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double a = n[3 * i], b = n[3 * i + 1], c = n[3 * i + 2], D = d[i];
    double q0, q1, q2;
    div3(a, b, c, D, q0, q1, q2);
    q_div3[3 * i] = q0;
    q_div3[3 * i + 1] = q1;
    q_div3[3 * i + 2] = q2;
    q_op[3 * i] = a / D;
    q_op[3 * i + 1] = b / D;
    q_op[3 * i + 2] = c / D;
}

double from_bits(uint64_t u)
{
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

double draw(std::mt19937_64 &g, int mode)
{
    const uint64_t u = g();
    const uint64_t sm = u & 0x800FFFFFFFFFFFFFull;  // sign and significand
    switch (mode) {
    case 0: return from_bits(u);                                            // any pattern
    case 1: return from_bits(sm | ((700 + (u >> 53) % 650) << 52));         // around the window
    case 2: {                                                               // window edges
        const uint64_t e = ((u >> 53) & 1) ? 722 + (u >> 40) % 3 : 1322 + (u >> 40) % 3;
        return from_bits(sm | (e << 52));
    }
    case 3: {                                                               // G m dx of a disk
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        return U(g) * 1e-14 * (double)(1 + u % 7);
    }
    default: return from_bits(u & 0x8000000000000000ull);                   // signed zeros
    }
}

}  // namespace

int main()
{
    const int N = 1 << 24;
    std::vector<double> n(3 * (size_t)N), d(N), q1(3 * (size_t)N), q2(3 * (size_t)N);
    std::mt19937_64 g(12345);
    for (int i = 0; i < N; ++i) {
        const int md = i % 5 < 3 ? i % 5 : 3;
        d[i] = md == 3 ? std::fabs(draw(g, 3)) * 1e8 : draw(g, md);
        for (int c = 0; c < 3; ++c) n[3 * (size_t)i + c] = draw(g, g() % 10 == 0 ? 4 : (int)(g() % 4));
    }
    double *dn = nullptr, *dd = nullptr, *d1 = nullptr, *d2 = nullptr;
    if (hipMalloc(&dn, 24 * (size_t)N) != hipSuccess || hipMalloc(&dd, 8 * (size_t)N) != hipSuccess ||
        hipMalloc(&d1, 24 * (size_t)N) != hipSuccess || hipMalloc(&d2, 24 * (size_t)N) != hipSuccess) {
        std::printf("device allocation failed\n");
        return 2;
    }
    if (hipMemcpy(dn, n.data(), 24 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd, d.data(), 8 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess)
        return 2;
    hipLaunchKernelGGL(k_div_check, dim3((N + 255) / 256), dim3(256), 0, 0, dn, dd, d1, d2, N);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::printf("kernel failed\n");
        return 2;
    }
    if (hipMemcpy(q1.data(), d1, 24 * (size_t)N, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(q2.data(), d2, 24 * (size_t)N, hipMemcpyDeviceToHost) != hipSuccess)
        return 2;
    long bad = 0, in_window = 0;
    for (size_t i = 0; i < 3 * (size_t)N; ++i) {
        if (std::memcmp(&q1[i], &q2[i], 8)) {
            if (bad < 10)
                std::printf("mismatch n=%a d=%a div3=%a op=%a\n", n[i], d[i / 3], q1[i], q2[i]);
            ++bad;
        }
        uint64_t bits;
        std::memcpy(&bits, &d[i / 3], 8);
        in_window += ((unsigned)(bits >> 52) & 0x7ffu) - 723u <= 600u;
    }
    std::printf("quotients %zu, denominator in the window %ld, mismatches %ld\n", 3 * (size_t)N,
                in_window, bad);
    for (double *p : {dn, dd, d1, d2}) (void)hipFree(p);
    return bad ? 1 : 0;
}
