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
// call, where the exact value lies within 2^-100 of a rounding boundary):
// pow_fifth, and log_rn (a table-driven fast path that defers to cr_log in
// the rare calls whose rounding its accuracy cannot decide).
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

// ln x correctly rounded, quickly: a table-driven evaluation accurate to a
// relative 2^-65, rounded only when that accuracy decides the rounding
// (Ziv's strategy), else cr_log.  With x = m 2^e, m in [1, 2), the top 7
// fraction bits of m pick c ~ 1/m (at most 8 significant bits, so that
// r = m c - 1 is exact under fma, |r| < 2^-7) and T = -ln c as a
// double-double (in the upper half, m >= 1 + 53/128, T = -ln(2 c) and e + 1,
// and the first and last rows have c = 1 and 1/2, T = 0, so that x near 1
// from either side reduces to ln(1 + r) without cancellation):
//   ln x = e ln2 + T + ln(1 + r),
//   ln(1 + r) = r - r^2 / 2 + r^3 (1/3 - r/4 + ... - r^7/10) + O(r^11).
// e ln2_hi is exact (ln2_hi has 42 significant bits, |e| < 2^11) and the
// leading sums are error-free (two_sum, two_prod); the remaining terms are
// summed in double.  Error budget, relative to ln x: table 2^-106, series
// truncation 2^-73, the low-order sums 2^-66 at worst (x within 2^-7 of a
// table node with e + the row's shift = 0).  The result rounds to RN(hi + lo)
// when hi + (lo -+ 2^-62 |hi|) round alike; otherwise (about 1 call in 256)
// cr_log decides.
struct LogRow {
    double c, t_hi, t_lo;
};

constexpr LogRow kLogRows[128] = {
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
    {0x1.fa00000000000p-1, 0x1.82448a388a2aap-7, 0x1.04b16137f09a0p-62},
    {0x1.f600000000000p-1, 0x1.432a925980cc1p-6, -0x1.8cdaf39004192p-60},
    {0x1.f200000000000p-1, 0x1.c63d2ec14aaf2p-6, -0x1.ce030a686bd86p-60},
    {0x1.ee00000000000p-1, 0x1.252f32f8d183fp-5, -0x1.947f792615916p-59},
    {0x1.ea00000000000p-1, 0x1.67c94f2d4bb58p-5, 0x1.0413e6505e603p-59},
    {0x1.e800000000000p-1, 0x1.894aa149fb343p-5, 0x1.a8be97660a23dp-60},
    {0x1.e400000000000p-1, 0x1.ccb73cdddb2ccp-5, -0x1.e48fb0500efd4p-59},
    {0x1.e000000000000p-1, 0x1.08598b59e3a07p-4, -0x1.dd7009902bf32p-58},
    {0x1.dc00000000000p-1, 0x1.2aa04a44717a5p-4, -0x1.d15d38d2fa3f7p-58},
    {0x1.da00000000000p-1, 0x1.3bdf5a7d1ee64p-4, 0x1.7a976d3b5b45fp-59},
    {0x1.d600000000000p-1, 0x1.5e95a4d9791cbp-4, 0x1.f38745c5c450ap-58},
    {0x1.d200000000000p-1, 0x1.8197e2f40e3f0p-4, 0x1.b9f2dffbeed43p-60},
    {0x1.d000000000000p-1, 0x1.9335e5d594989p-4, -0x1.478a85704ccb7p-58},
    {0x1.cc00000000000p-1, 0x1.b6ac88dad5b1cp-4, -0x1.0057eed1ca59fp-59},
    {0x1.c800000000000p-1, 0x1.da727638446a2p-4, 0x1.401fa71733019p-58},
    {0x1.c600000000000p-1, 0x1.ec739830a1120p-4, -0x1.a2bf991780d3fp-59},
    {0x1.c200000000000p-1, 0x1.08598b59e3a07p-3, -0x1.dd7009902bf32p-57},
    {0x1.c000000000000p-1, 0x1.1178e8227e47cp-3, -0x1.0e63a5f01c691p-58},
    {0x1.bc00000000000p-1, 0x1.23d712a49c202p-3, -0x1.6e38161051d69p-57},
    {0x1.ba00000000000p-1, 0x1.2d1610c86813ap-3, -0x1.499a3f25af95fp-58},
    {0x1.b600000000000p-1, 0x1.3fb45a59928ccp-3, -0x1.d87e6a354d056p-57},
    {0x1.b400000000000p-1, 0x1.4913d8333b561p-3, -0x1.0d5604930f135p-58},
    {0x1.b000000000000p-1, 0x1.5bf406b543db2p-3, -0x1.1f5b44c0df7e7p-61},
    {0x1.ae00000000000p-1, 0x1.6574ebe8c133ap-3, -0x1.d34f0f4621bedp-60},
    {0x1.aa00000000000p-1, 0x1.7898d85444c73p-3, 0x1.ef8f6ebcfb201p-58},
    {0x1.a800000000000p-1, 0x1.823c16551a3c2p-3, -0x1.1232ce70be781p-57},
    {0x1.a600000000000p-1, 0x1.8beafeb38fe8cp-3, 0x1.55aa8b6997a40p-58},
    {0x1.a200000000000p-1, 0x1.9f6c407089664p-3, 0x1.35a19605e67efp-59},
    {0x1.a000000000000p-1, 0x1.a93ed3c8ad9e3p-3, 0x1.bcafa9de97203p-57},
    {0x1.9e00000000000p-1, 0x1.b31d8575bce3dp-3, -0x1.6353ab386a94dp-57},
    {0x1.9a00000000000p-1, 0x1.c6ffbc6f00f71p-3, -0x1.8e58b2c57a4a5p-57},
    {0x1.9800000000000p-1, 0x1.d1037f2655e7bp-3, 0x1.60629242471a2p-57},
    {0x1.9600000000000p-1, 0x1.db13db0d48940p-3, 0x1.aa11d49f96cb9p-58},
    {0x1.9400000000000p-1, 0x1.e530effe71012p-3, 0x1.2276041f43042p-59},
    {0x1.9000000000000p-1, 0x1.f991c6cb3b379p-3, 0x1.f665066f980a2p-57},
    {0x1.8e00000000000p-1, 0x1.01eae5626c691p-2, -0x1.18290bd2932e2p-59},
    {0x1.8c00000000000p-1, 0x1.07138604d5862p-2, 0x1.cdb16ed4e9138p-56},
    {0x1.8a00000000000p-1, 0x1.0c42d676162e3p-2, 0x1.162c79d5d11eep-58},
    {0x1.8800000000000p-1, 0x1.1178e8227e47cp-2, -0x1.0e63a5f01c691p-57},
    {0x1.8400000000000p-1, 0x1.1bf99635a6b95p-2, -0x1.12aeb84249223p-57},
    {0x1.8200000000000p-1, 0x1.214456d0eb8d4p-2, 0x1.f7ae91aeba60ap-57},
    {0x1.8000000000000p-1, 0x1.269621134db92p-2, 0x1.e0efadd9db02bp-56},
    {0x1.7e00000000000p-1, 0x1.2bef07cdc9354p-2, -0x1.82dad7fd86088p-56},
    {0x1.7c00000000000p-1, 0x1.314f1e1d35ce4p-2, -0x1.3d69909e5c3dcp-56},
    {0x1.7a00000000000p-1, 0x1.36b6776be1117p-2, -0x1.324f0e883858ep-58},
    {0x1.7800000000000p-1, 0x1.3c25277333184p-2, -0x1.2ad27e50a8ec6p-56},
    {0x1.7600000000000p-1, 0x1.419b423d5e8c7p-2, 0x1.0dbb243827392p-57},
    {0x1.7400000000000p-1, 0x1.4718dc271c41bp-2, 0x1.8fb4c14c56eefp-60},
    {0x1.7200000000000p-1, 0x1.4c9e09e172c3cp-2, -0x1.123615b147a5dp-58},
    {0x1.7000000000000p-1, 0x1.522ae0738a3d8p-2, -0x1.8f7e9b38a6979p-57},
    {0x1.6e00000000000p-1, 0x1.57bf753c8d1fbp-2, -0x1.0908d15f88b63p-57},
    {0x1.6c00000000000p-1, 0x1.5d5bddf595f30p-2, -0x1.6541148cbb8a2p-56},
    {0x1.6a00000000000p-1, -0x1.62c82f2b9c795p-2, -0x1.7b7af915300e5p-57},
    {0x1.6800000000000p-1, -0x1.5d1bdbf5809cap-2, -0x1.4236383dc7fe1p-56},
    {0x1.6600000000000p-1, -0x1.5767717455a6cp-2, -0x1.526adb283660cp-56},
    {0x1.6400000000000p-1, -0x1.51aad872df82dp-2, -0x1.3927ac19f55e3p-59},
    {0x1.6200000000000p-1, -0x1.4be5f957778a1p-2, 0x1.259b35b04813dp-57},
    {0x1.6000000000000p-1, -0x1.4618bc21c5ec2p-2, -0x1.f42decdeccf1dp-56},
    {0x1.5e00000000000p-1, -0x1.404308686a7e4p-2, 0x1.0bcfb6082ce6dp-56},
    {0x1.5c00000000000p-1, -0x1.3a64c556945eap-2, 0x1.c68651945f97cp-57},
    {0x1.5a00000000000p-1, -0x1.347dd9a987d55p-2, 0x1.4dd4c580919f8p-57},
    {0x1.5800000000000p-1, -0x1.2e8e2bae11d31p-2, 0x1.8f4cdb95ebdf9p-56},
    {0x1.5600000000000p-1, -0x1.2895a13de86a3p-2, -0x1.7ad24c13f040ep-56},
    {0x1.5400000000000p-1, -0x1.22941fbcf7966p-2, 0x1.76f5eb09628afp-56},
    {0x1.5200000000000p-1, -0x1.1c898c16999fbp-2, 0x1.0e5c62aff1c44p-60},
    {0x1.5000000000000p-1, -0x1.1675cababa60ep-2, -0x1.ce63eab883717p-61},
    {0x1.5000000000000p-1, -0x1.1675cababa60ep-2, -0x1.ce63eab883717p-61},
    {0x1.4e00000000000p-1, -0x1.1058bf9ae4ad5p-2, -0x1.89fa0ab4cb31dp-58},
    {0x1.4c00000000000p-1, -0x1.0a324e27390e3p-2, -0x1.7dcfde8061c03p-56},
    {0x1.4a00000000000p-1, -0x1.0402594b4d041p-2, 0x1.28ec217a5022dp-57},
    {0x1.4800000000000p-1, -0x1.fb9186d5e3e2bp-3, 0x1.caaae64f21acbp-57},
    {0x1.4600000000000p-1, -0x1.ef0adcbdc5936p-3, -0x1.48637950dc20dp-57},
    {0x1.4600000000000p-1, -0x1.ef0adcbdc5936p-3, -0x1.48637950dc20dp-57},
    {0x1.4400000000000p-1, -0x1.e27076e2af2e6p-3, 0x1.61578001e0162p-59},
    {0x1.4200000000000p-1, -0x1.d5c216b4fbb91p-3, -0x1.6e443597e4d40p-57},
    {0x1.4000000000000p-1, -0x1.c8ff7c79a9a22p-3, 0x1.4f689f8434012p-57},
    {0x1.3e00000000000p-1, -0x1.bc286742d8cd6p-3, -0x1.4fce744870f55p-58},
    {0x1.3e00000000000p-1, -0x1.bc286742d8cd6p-3, -0x1.4fce744870f55p-58},
    {0x1.3c00000000000p-1, -0x1.af3c94e80bff3p-3, 0x1.398cff3641985p-58},
    {0x1.3a00000000000p-1, -0x1.a23bc1fe2b563p-3, -0x1.93711b07a998cp-59},
    {0x1.3800000000000p-1, -0x1.9525a9cf456b4p-3, -0x1.d904c1d4e2e26p-57},
    {0x1.3800000000000p-1, -0x1.9525a9cf456b4p-3, -0x1.d904c1d4e2e26p-57},
    {0x1.3600000000000p-1, -0x1.87fa06520c911p-3, 0x1.bf7fdbfa08d9ap-57},
    {0x1.3400000000000p-1, -0x1.7ab890210d909p-3, -0x1.be36b2d6a0608p-59},
    {0x1.3200000000000p-1, -0x1.6d60fe719d21dp-3, 0x1.caae268ecd179p-57},
    {0x1.3200000000000p-1, -0x1.6d60fe719d21dp-3, 0x1.caae268ecd179p-57},
    {0x1.3000000000000p-1, -0x1.5ff3070a793d4p-3, 0x1.bc60efafc6f6ep-58},
    {0x1.2e00000000000p-1, -0x1.526e5e3a1b438p-3, 0x1.746ff8a470d3ap-57},
    {0x1.2e00000000000p-1, -0x1.526e5e3a1b438p-3, 0x1.746ff8a470d3ap-57},
    {0x1.2c00000000000p-1, -0x1.44d2b6ccb7d1ep-3, -0x1.9f4f6543e1f88p-57},
    {0x1.2a00000000000p-1, -0x1.371fc201e8f74p-3, -0x1.de6cb62af18a0p-58},
    {0x1.2a00000000000p-1, -0x1.371fc201e8f74p-3, -0x1.de6cb62af18a0p-58},
    {0x1.2800000000000p-1, -0x1.29552f81ff523p-3, -0x1.301771c407dbfp-57},
    {0x1.2600000000000p-1, -0x1.1b72ad52f67a0p-3, -0x1.483023472cd74p-58},
    {0x1.2600000000000p-1, -0x1.1b72ad52f67a0p-3, -0x1.483023472cd74p-58},
    {0x1.2400000000000p-1, -0x1.0d77e7cd08e59p-3, -0x1.9a5dc5e9030acp-57},
    {0x1.2200000000000p-1, -0x1.fec9131dbeabbp-4, 0x1.5746b9981b36cp-58},
    {0x1.2200000000000p-1, -0x1.fec9131dbeabbp-4, 0x1.5746b9981b36cp-58},
    {0x1.2000000000000p-1, -0x1.e27076e2af2e6p-4, 0x1.61578001e0162p-60},
    {0x1.1e00000000000p-1, -0x1.c5e548f5bc743p-4, -0x1.5d617ef8161b1p-60},
    {0x1.1e00000000000p-1, -0x1.c5e548f5bc743p-4, -0x1.5d617ef8161b1p-60},
    {0x1.1c00000000000p-1, -0x1.a926d3a4ad563p-4, -0x1.942f48aa70ea9p-58},
    {0x1.1c00000000000p-1, -0x1.a926d3a4ad563p-4, -0x1.942f48aa70ea9p-58},
    {0x1.1a00000000000p-1, -0x1.8c345d6319b21p-4, 0x1.4a697ab3424a9p-61},
    {0x1.1800000000000p-1, -0x1.6f0d28ae56b4cp-4, 0x1.906d99184b992p-58},
    {0x1.1800000000000p-1, -0x1.6f0d28ae56b4cp-4, 0x1.906d99184b992p-58},
    {0x1.1600000000000p-1, -0x1.51b073f06183fp-4, -0x1.a49e39a1a8be4p-58},
    {0x1.1600000000000p-1, -0x1.51b073f06183fp-4, -0x1.a49e39a1a8be4p-58},
    {0x1.1400000000000p-1, -0x1.341d7961bd1d1p-4, 0x1.b599f227becbbp-58},
    {0x1.1200000000000p-1, -0x1.16536eea37ae1p-4, 0x1.79da3e8c22cdap-60},
    {0x1.1200000000000p-1, -0x1.16536eea37ae1p-4, 0x1.79da3e8c22cdap-60},
    {0x1.1000000000000p-1, -0x1.f0a30c01162a6p-5, -0x1.85f325c5bbacdp-59},
    {0x1.1000000000000p-1, -0x1.f0a30c01162a6p-5, -0x1.85f325c5bbacdp-59},
    {0x1.0e00000000000p-1, -0x1.b42dd711971bfp-5, 0x1.eb9759c130499p-60},
    {0x1.0e00000000000p-1, -0x1.b42dd711971bfp-5, 0x1.eb9759c130499p-60},
    {0x1.0c00000000000p-1, -0x1.77458f632dcfcp-5, -0x1.18d3ca87b9296p-59},
    {0x1.0a00000000000p-1, -0x1.39e87b9febd60p-5, 0x1.5bfa937f551bbp-59},
    {0x1.0a00000000000p-1, -0x1.39e87b9febd60p-5, 0x1.5bfa937f551bbp-59},
    {0x1.0800000000000p-1, -0x1.f829b0e783300p-6, -0x1.33e3f04f1ef23p-60},
    {0x1.0800000000000p-1, -0x1.f829b0e783300p-6, -0x1.33e3f04f1ef23p-60},
    {0x1.0600000000000p-1, -0x1.7b91b07d5b11bp-6, 0x1.5b602ace3a510p-60},
    {0x1.0600000000000p-1, -0x1.7b91b07d5b11bp-6, 0x1.5b602ace3a510p-60},
    {0x1.0400000000000p-1, -0x1.fc0a8b0fc03e4p-7, 0x1.83092c59642a1p-62},
    {0x1.0400000000000p-1, -0x1.fc0a8b0fc03e4p-7, 0x1.83092c59642a1p-62},
    {0x1.0200000000000p-1, -0x1.fe02a6b106789p-8, 0x1.e44b7e3711ebfp-67},
    {0x1.0200000000000p-1, -0x1.fe02a6b106789p-8, 0x1.e44b7e3711ebfp-67},
    {0x1.0000000000000p-1, 0x0.0p+0, 0x0.0p+0}
};

#define OKA_LN2_HI42 0x1.62e42fefa3800p-1
#define OKA_LN2_LO42 0x1.ef35793c76730p-45

// cr_log out of line: log_rn's rare undecided calls, kept away from the
// registers of the loops that call log_rn.
__host__ __device__ __attribute__((noinline)) static double cr_log_call(double x)
{
    return cr_log(x);
}

OKA_HD double log_rn(double x)
{
    if (!(x > 0.0) || x == __builtin_inf())
        return cr_log_call(x);
    int e;
    const double m = split_exponent(x, &e);
    const int i = (int)((bits_of(m) >> 45) & 127u);
    const LogRow t = kLogRows[i];
    if (i >= 53)
        e += 1;
    const double r = fma_(m, t.c, -1.0);   // exact
    const double de = (double)e;
    const dd s = two_sum(de * OKA_LN2_HI42, t.t_hi);
    const dd u = two_sum(s.hi, r);
    const dd p = two_prod(r, r);
    const dd v = two_sum(u.hi, -0.5 * p.hi);
    double q = -1.0 / 10.0;
    q = fma_(q, r, 1.0 / 9.0);
    q = fma_(q, r, -1.0 / 8.0);
    q = fma_(q, r, 1.0 / 7.0);
    q = fma_(q, r, -1.0 / 6.0);
    q = fma_(q, r, 1.0 / 5.0);
    q = fma_(q, r, -1.0 / 4.0);
    q = fma_(q, r, 1.0 / 3.0);
    double lo = s.lo + u.lo;
    lo = lo + v.lo;
    lo = lo + (-0.5 * p.lo);
    lo = lo + de * OKA_LN2_LO42;
    lo = lo + t.t_lo;
    lo = lo + (p.hi * r) * q;
    const double err = fabs(v.hi) * 0x1p-62;
    const double a = v.hi + (lo - err);
    const double b = v.hi + (lo + err);
    if (a == b)
        return a;
    return cr_log_call(x);
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
