"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011) in plain Python --
TEST INFRASTRUCTURE ONLY: the checker of the library's device draws.

uniforms() restates how the library turns one 128-bit output into two
uniforms in (0, 1) with 53 random bits each.
"""
_M = 0xFFFFFFFF


def philox4x32_10(counter, key):
    c = list(counter)
    k = list(key)
    for _ in range(10):
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & _M, p1 & _M, ((p0 >> 32) ^ c[3] ^ k[1]) & _M, p0 & _M]
        k = [(k[0] + 0x9E3779B9) & _M, (k[1] + 0xBB67AE85) & _M]
    return c


def uniforms(seed, r, slot, sub=0):
    o = philox4x32_10([r & _M, (r >> 32) & _M, slot, sub], [seed & _M, (seed >> 32) & _M])
    mask = (1 << 53) - 1
    a = ((o[0] << 21) ^ (o[1] >> 11)) & mask
    b = ((o[2] << 21) ^ (o[3] >> 11)) & mask
    return (a + 0.5) * 2.0 ** -53, (b + 0.5) * 2.0 ** -53
