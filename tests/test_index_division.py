"""The multiply-high division behind the step kernels' index arithmetic
(`fastdiv_make` / `fastdiv` in snac_amd/csrc/snh_kernels.h): with
l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1, the quotient of
0 <= n < 2^31 by d is (mulhi(m, n) + n) >> l.  Checked here against exact
integer division over every divisor up to 4096 and sampled divisors up to
2^31, at the dividends where an off-by-one would show (multiples of d and
their neighbours, the range ends); on the GPU the same arithmetic is pinned
by the bit-identical parity runs, whose element and node indices come from
it."""
import random


def make(d):
    l = 0
    while (1 << l) < d:
        l += 1
    m = (((1 << l) - d) << 32) // d + 1
    return m, l


def fastdiv(n, ml):
    m, l = ml
    return (((n * m) >> 32) + n) >> l


def _dividends(d, rng):
    top = (1 << 31) - 1
    out = {0, 1, d - 1, d, d + 1, top, top - 1, top - (top % d), top - (top % d) - 1}
    for k in (1, 2, 3, 1000, top // d - 1, top // d):
        for r in (-1, 0, 1, d - 1):
            out.add(k * d + r)
    out.update(rng.randrange(0, 1 << 31) for _ in range(64))
    return [n for n in out if 0 <= n <= top]


def test_magic_number_fits_32_bits():
    for d in list(range(1, 4097)) + [(1 << k) + o for k in range(12, 31) for o in (-1, 0, 1)] + [(1 << 31) - 1]:
        m, l = make(d)
        assert 1 <= m < (1 << 32) and l <= 31, d


def test_quotient_exact_below_2_pow_31():
    rng = random.Random(5)
    divisors = list(range(1, 4097)) + [rng.randrange(4097, 1 << 31) for _ in range(2000)] + \
        [(1 << k) + o for k in range(12, 31) for o in (-1, 0, 1)]
    for d in divisors:
        ml = make(d)
        for n in _dividends(d, rng):
            assert fastdiv(n, ml) == n // d, (n, d)
