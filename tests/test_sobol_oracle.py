"""The Sobol' checker (oracle/sobol.py) against what the mathematics fixes:
primitive-polynomial counts, the van der Corput dimension, the published
second coordinate, stratification of every dimension and the (0, m, 2)-net of
the first two."""
from math import gcd

import pytest

from oracle import sobol


def _totient(n):
    return sum(1 for k in range(1, n + 1) if gcd(k, n) == 1)


def test_primitive_polynomial_counts_and_first_ones():
    polys = sobol.primitive_polynomials(1 + 1 + 2 + 2 + 6 + 6 + 18 + 16 + 48 + 60)
    by_deg = {}
    for deg, p in polys:
        by_deg[deg] = by_deg.get(deg, 0) + 1
    for deg in range(1, 11):
        assert by_deg[deg] == _totient(2 ** deg - 1) // deg
    assert polys[:6] == [(1, 0b11), (2, 0b111), (3, 0b1011), (3, 0b1101), (4, 0b10011),
                         (4, 0b11001)]
    # a reducible polynomial (x^2 + x + 1)^2 and an irreducible non-primitive one
    # (x^4 + x^3 + x^2 + x + 1, order 5) are rejected
    assert not sobol.is_primitive(0b10101, 4) and not sobol.is_primitive(0b11111, 4)


def test_first_two_dimensions_are_the_published_ones():
    v = sobol.direction_numbers(2)
    for j in range(1, 4096):
        assert sobol.point(v[0], j) == int(f"{j:032b}"[::-1], 2)   # van der Corput
    # x + 1 with m_1 = 1: the published m_k = 1, 3, 5, 15, 17, 51, 85, 255 and,
    # in natural order, the points 0, 1/2, 3/4, 1/4, 5/8, 1/8, 3/8, 7/8
    assert [v[1][k] >> (31 - k) for k in range(8)] == [1, 3, 5, 15, 17, 51, 85, 255]
    got = [sobol.point(v[1], j) / 2.0 ** 32 for j in range(8)]
    assert got == [0.0, 0.5, 0.75, 0.25, 0.625, 0.125, 0.375, 0.875]


@pytest.mark.parametrize("k", [1, 4, 7, 10])
def test_every_dimension_is_stratified_before_and_after_scrambling(k):
    v = sobol.direction_numbers(48)
    for d, vd in enumerate(v):
        kd = sobol.key(12345, d)
        top = sorted(sobol.point(vd, j) >> (32 - k) for j in range(2 ** k))
        assert top == list(range(2 ** k)), d
        top = sorted(sobol.nested_scramble(sobol.point(vd, j), kd) >> (32 - k)
                     for j in range(2 ** k))
        assert top == list(range(2 ** k)), d


@pytest.mark.parametrize("scrambled", [False, True])
def test_first_two_dimensions_form_a_0_m_2_net(scrambled):
    v = sobol.direction_numbers(2)
    m = 8
    f = ((lambda x, d: sobol.nested_scramble(x, sobol.key(99, d))) if scrambled
         else (lambda x, d: x))
    pts = [(f(sobol.point(v[0], j), 0), f(sobol.point(v[1], j), 1)) for j in range(2 ** m)]
    for a in range(m + 1):
        b = m - a
        boxes = {(x >> (32 - a) if a else 0, y >> (32 - b) if b else 0) for x, y in pts}
        assert len(boxes) == 2 ** m, a


def test_scramble_is_nested_and_bijective():
    """Bit i of the scrambled value (from the top) depends only on bits 0..i
    of the input: inputs that share their top t bits share the top t bits of
    their images; and the map is one-to-one (checked on 16-bit prefixes)."""
    import random
    rnd = random.Random(5)
    for _ in range(200):
        kk = rnd.getrandbits(32)
        x = rnd.getrandbits(32)
        t = rnd.randrange(1, 32)
        y = (x & ~((1 << (32 - t)) - 1)) | rnd.getrandbits(32 - t)
        assert (sobol.nested_scramble(x, kk) >> (32 - t)) == (sobol.nested_scramble(y, kk) >> (32 - t))
    kk = 0xDEADBEEF
    images = {sobol.nested_scramble(x << 16, kk) >> 16 for x in range(2 ** 16)}
    assert len(images) == 2 ** 16


def test_uniforms_are_inside_the_open_unit_interval():
    v = sobol.direction_numbers(3)
    for j in (0, 1, 2 ** 31, 2 ** 32 - 1):
        for d in range(3):
            u = sobol.uniform(v[d], j, 7, d)
            assert 0.0 < u < 1.0
