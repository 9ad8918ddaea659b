// oka_math.h -- exact scalar arithmetic shared by host and device code.
//
// The FP64-exact tracer must reproduce the reference's numpy scalar
// arithmetic bit for bit.  +, -, *, / and sqrt are IEEE correctly rounded on
// both the CPU and gfx950 (this file is compiled with -ffp-contract=off, so no
// a*b+c is fused unless written as fma()).  The two library functions on the
// path are not: the reference evaluates
//   (tol/(est+EPS))**(1/5)   capturezone.py:247   (C pow(x, 0.2))
//   np.log(r2)               model.py:264
// This header gives correctly rounded versions of both, built from IEEE basic
// operations only, so host and device produce the same bits and those bits are
// the correctly rounded result (up to a failure probability below 2^-47 per
// call, where the exact value lies within 2^-100 of a rounding boundary).
//
// Double-double ("dd") arithmetic follows Dekker (1971) and Knuth's TwoSum.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OKA_HD __host__ __device__ inline

namespace oka {

struct dd {
    double hi, lo;
};

OKA_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

OKA_HD uint64_t bits_of(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}

OKA_HD double from_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

OKA_HD dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

OKA_HD dd fast_two_sum(double a, double b)  // requires |a| >= |b| or a == 0
{
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

OKA_HD dd two_prod(double a, double b)
{
    double p = a * b;
    return {p, fma_(a, b, -p)};
}

OKA_HD dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}

OKA_HD dd dd_mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return fast_two_sum(p.hi, p.lo);
}

OKA_HD dd dd_mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo = p.lo + a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

// x = m * 2^e with m in [1, 2); x finite and > 0.
OKA_HD double split_exponent(double x, int *e)
{
    uint64_t u = bits_of(x);
    int be = (int)((u >> 52) & 0x7ff);
    int adj = 0;
    if (be == 0) {  // subnormal: scale into the normal range first
        x = x * 0x1p54;
        u = bits_of(x);
        be = (int)((u >> 52) & 0x7ff);
        adj = 54;
    }
    *e = be - 1023 - adj;
    return from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
}

OKA_HD double pow2i(int k)  // 2^k for -1022 <= k <= 1023
{
    return from_bits((uint64_t)(k + 1023) << 52);
}

// ln(2) and 1/(2k+1), k = 1..9, as double-double constants (exact to 2^-106).
#define OKA_LN2_HI 0x1.62e42fefa39efp-1
#define OKA_LN2_LO 0x1.abc9e3b39803fp-56

// Correctly rounded natural logarithm.  ln x = e ln2 + 2 atanh(s),
// s = (m-1)/(m+1), m in (sqrt(1/2), sqrt(2)]; the atanh series is summed in
// double-double for its first ten terms and in double for the rest.
OKA_HD double cr_log(double x)
{
    if (!(x > 0.0))
        return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (x == __builtin_inf())
        return x;
    int e;
    double m = split_exponent(x, &e);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e = e + 1;
    }
    double num = m - 1.0;               // exact (Sterbenz)
    dd den = two_sum(m, 1.0);
    // s = num / den in double-double
    double q1 = num / den.hi;
    dd p = two_prod(q1, den.hi);
    double r = ((num - p.hi) - p.lo) - q1 * den.lo;
    double q2 = r / den.hi;
    dd s = fast_two_sum(q1, q2);
    dd s2 = dd_mul(s, s);
    // tail: k = 10..21 in double
    double z = s2.hi;
    double tail = 1.0 / 43.0;
    tail = 1.0 / 41.0 + z * tail;
    tail = 1.0 / 39.0 + z * tail;
    tail = 1.0 / 37.0 + z * tail;
    tail = 1.0 / 35.0 + z * tail;
    tail = 1.0 / 33.0 + z * tail;
    tail = 1.0 / 31.0 + z * tail;
    tail = 1.0 / 29.0 + z * tail;
    tail = 1.0 / 27.0 + z * tail;
    tail = 1.0 / 25.0 + z * tail;
    tail = 1.0 / 23.0 + z * tail;
    tail = 1.0 / 21.0 + z * tail;
    const dd c[10] = {
        {1.0, 0.0},
        {0x1.5555555555555p-2, 0x1.5555555555555p-56},
        {0x1.999999999999ap-3, -0x1.999999999999ap-57},
        {0x1.2492492492492p-3, 0x1.2492492492492p-57},
        {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
        {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
        {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
        {0x1.1111111111111p-4, 0x1.1111111111111p-60},
        {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},
        {0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59},
    };
    dd acc = dd_add(c[9], dd{z * tail, 0.0});
    for (int k = 8; k >= 0; --k)
        acc = dd_add(c[k], dd_mul(s2, acc));
    dd lnm = dd_mul(s, acc);
    lnm.hi = 2.0 * lnm.hi;
    lnm.lo = 2.0 * lnm.lo;
    double de = (double)e;
    dd el = two_prod(de, OKA_LN2_HI);
    el.lo = el.lo + de * OKA_LN2_LO;
    el = fast_two_sum(el.hi, el.lo);
    dd res = dd_add(el, lnm);
    return res.hi + res.lo;
}

// A quick natural logarithm, accurate to a few ulp; only used for the tiny
// correction factor x^(0.2 - 1/5) inside pow_fifth.
OKA_HD double quick_log(double x)
{
    int e;
    double m = split_exponent(x, &e);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e = e + 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + z * p;
    p = 1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = 1.0 + z * p;
    return (double)e * OKA_LN2_HI + 2.0 * s * p;
}

// Correctly rounded pow(x, 0.2) where 0.2 is the double nearest 1/5, i.e.
// x^(1/5) * x^dc with dc = 0.2 - 1/5 = 2^-54/5.  x^(1/5) is found by Newton's
// method in double and one double-double correction; x^dc = 1 + t + t^2/2
// with t = dc ln x (|t| < 1e-14).
OKA_HD double pow_fifth(double x)
{
    if (!(x > 0.0))
        return x == 0.0 ? 0.0 : __builtin_nan("");
    if (x == __builtin_inf())
        return x;
    int e;
    double m = split_exponent(x, &e);
    int q = e >= 0 ? e / 5 : -((4 - e) / 5);
    int r = e - 5 * q;  // 0..4
    const double t5[5] = {0x1.0000000000000p+0, 0x1.2611186bae675p+0, 0x1.51cb453b9536cp+0,
                          0x1.8406003b2ae5cp+0, 0x1.bdb8cdadbe120p+0};
    double u = m * (double)(1 << r);          // exact, in [1, 32)
    double y = (1.0 + (m - 1.0) * 0x1.3088c35d733a5p-3) * t5[r];
    for (int it = 0; it < 4; ++it)
        y = 0.8 * y + 0.2 * (u / ((y * y) * (y * y)));
    // one Newton step in double-double: y + (u - y^5) / (5 y^4)
    dd y2 = two_prod(y, y);
    dd y4 = dd_mul(y2, y2);
    dd y5 = dd_mul_d(y4, y);
    dd rho = two_sum(u, -y5.hi);
    rho.lo = rho.lo - y5.lo;
    double dy = (rho.hi + rho.lo) / (5.0 * y4.hi);
    dd root = fast_two_sum(y, dy);
    double t = 0x1.999999999999ap-57 * quick_log(x);
    double corr = t + 0.5 * (t * t);
    double lo = root.lo + root.hi * corr;
    double res = root.hi + lo;
    return res * pow2i(q);
}

}  // namespace oka
