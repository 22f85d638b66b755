// Three FP64 quotients over one denominator, each bit-identical to the `/`
// operator (the correctly rounded IEEE quotient).
//
// gfx950's FP64 division is the sequence s0 = div_scale(d, d, n), y0 =
// rcp(s0), y1 = fma(y0, fma(-s0, y0, 1), y0), y = fma(y1, fma(-s0, y1, 1), y1),
// s1 = div_scale(n, d, n), q = s1 * y, r = fma(-s0, q, s1), then
// div_fmas(r, y, q) and div_fixup(., d, n).  div_scale leaves its operand as
// it is (and div_fmas is a plain fma) unless an operand is zero, denormal,
// infinite or NaN, the exponent of n exceeds that of d by 768 or more, 1/d or
// n/d is denormal, or the biased exponent of n is 53 or less; div_fixup of a
// finite, nonzero, normal quotient only sets the sign, which q already has.
// With both biased exponents in [723, 1323] (magnitudes in [2^-300, 2^301))
// none of those cases can arise, so the reciprocal y depends on d alone and
// is formed once; each numerator then costs a multiply and two FMAs instead
// of the whole sequence.  A zero numerator gives n * y, the zero of sign
// sign(n) xor sign(d) that div_fixup returns; any other operand outside the
// window takes the `/` operator.
#pragma once

namespace {

__device__ inline bool div_window(double v)
{
    const unsigned e = (unsigned)((unsigned long long)__double_as_longlong(v) >> 52) & 0x7ffu;
    return e - 723u <= 600u;
}

__device__ inline double div_by(double n, double d, double y)
{
    if (n == 0.0) return n * y;
    if (!div_window(n)) return n / d;
    const double q = n * y;
    const double r = __builtin_fma(-d, q, n);
    return __builtin_fma(r, y, q);
}

__device__ inline void div3(double n0, double n1, double n2, double d, double &q0, double &q1,
                            double &q2)
{
    if (!div_window(d)) {
        q0 = n0 / d;
        q1 = n1 / d;
        q2 = n2 / d;
        return;
    }
    const double y0 = __builtin_amdgcn_rcp(d);
    const double y1 = __builtin_fma(y0, __builtin_fma(-d, y0, 1.0), y0);
    const double y = __builtin_fma(y1, __builtin_fma(-d, y1, 1.0), y1);
    q0 = div_by(n0, d, y);
    q1 = div_by(n1, d, y);
    q2 = div_by(n2, d, y);
}

}  // namespace
