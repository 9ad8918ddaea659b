"""Sobol' points with a random digital shift, in plain Python -- TEST
INFRASTRUCTURE ONLY: the checker of the library's QMC draws (OKA_QMC).

The construction is the textbook one (Bratley & Fox, ACM TOMS 14, 1988; the
recurrence as in Joe & Kuo, ACM TOMS 29, 2003): dimension 0 is the van der
Corput sequence; dimension d >= 1 takes the d-th primitive polynomial over
GF(2) (by degree, then by value) p = x^s + a_1 x^(s-1) + ... + a_(s-1) x + 1,
initial direction numbers m_k (odd, below 2^k; here 2 h + 1 with h from
splitmix64 of (d, k)), and v_k = v_(k-s) ^ (v_(k-s) >> s) ^ sum a_i v_(k-i).
Point j (natural order) of dimension d is the XOR of v_(k+1) over the set bits
k of j.  The library scrambles each coordinate with the hash-based nested
(Owen-type) scramble of Burley (JCGT 9(4), 2020, Laine-Karras form) keyed by
word 0 of Philox4x32-10 at counter (d, 2^32 - 1, 0x50B0150, 2) under the seed.
"""
from . import philox

_M64 = (1 << 64) - 1
SHIFT_SLOT = 0x50B0150


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _powmod_x(e, p, deg):
    def mulmod(a, b):
        r = 0
        while b:
            if b & 1:
                r ^= a
            b >>= 1
            a <<= 1
            if (a >> deg) & 1:
                a ^= p
        return r
    r, base = 1, 2
    while e:
        if e & 1:
            r = mulmod(r, base)
        base = mulmod(base, base)
        e >>= 1
    return r


def _prime_factors(n):
    f, out = 2, []
    while f * f <= n:
        if n % f == 0:
            out.append(f)
            while n % f == 0:
                n //= f
        f += 1
    if n > 1:
        out.append(n)
    return out


def is_primitive(p, deg):
    """x has order 2^deg - 1 modulo p (deg >= 2)."""
    n = (1 << deg) - 1
    return _powmod_x(n, p, deg) == 1 and all(_powmod_x(n // f, p, deg) != 1
                                             for f in _prime_factors(n))


def primitive_polynomials(count):
    """The first `count` primitive polynomials as (degree, bits), x + 1 first."""
    out = []
    deg = 1
    while len(out) < count:
        for p in range((1 << deg) | 1, 2 << deg, 2):
            if deg == 1 or is_primitive(p, deg):
                out.append((deg, p))
                if len(out) == count:
                    break
        deg += 1
    return out


def direction_numbers(dims):
    """dims lists of 32 direction numbers (uint32), v_1 first."""
    out = [[1 << (31 - k) for k in range(32)]]
    for d, (deg, p) in enumerate(primitive_polynomials(dims - 1), start=1):
        v = []
        for k in range(1, 33):
            if k <= deg:
                h = splitmix64((d << 32) | k)
                m = 2 * (h & ((1 << (k - 1)) - 1)) + 1
                x = (m << (32 - k)) & 0xFFFFFFFF
            else:
                x = v[k - deg - 1] ^ (v[k - deg - 1] >> deg)
                for i in range(1, deg):
                    if (p >> (deg - i)) & 1:
                        x ^= v[k - i - 1]
            v.append(x)
        out.append(v)
    return out


def point(v, j):
    """Unshifted point j of the dimension with direction numbers v (uint32)."""
    x, k = 0, 0
    while j >> k:
        if (j >> k) & 1:
            x ^= v[k]
        k += 1
    return x


def key(seed, d):
    return philox.philox4x32_10([d, 0xFFFFFFFF, SHIFT_SLOT, 2],
                                [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF])[0]


def _reverse32(x):
    return int(f"{x:032b}"[::-1], 2)


def nested_scramble(x, k):
    """Bit-reversed keyed offset and four x ^= x * c steps (c even), reversed back."""
    x = (_reverse32(x) + k) & 0xFFFFFFFF
    for c in (0x6c50b47c, 0xb82f1e52, 0xc7afe638, 0x8d22f6e6):
        x ^= (x * c) & 0xFFFFFFFF
    return _reverse32(x)


def uniform(v, j, seed, d):
    """The library's QMC uniform of dimension d (direction numbers v) at point j."""
    return (nested_scramble(point(v, j), key(seed, d)) + 0.5) * 2.0 ** -32
