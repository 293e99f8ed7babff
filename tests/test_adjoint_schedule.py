"""The binomial checkpointing schedule of the adjoint (snh_adjoint_*, with
snapshots > 0): the closed form of its cost, which the reverse sweep
minimises when it places a snapshot, against a brute-force search over all
recursive schedules."""
from functools import lru_cache
from math import comb


def binomial_cost(n, s):
    """Blocks advanced to reverse n blocks from a kept state with s snapshots
    free: t n - C(s + 1 + t, t - 1), t the least integer with
    C(s + 1 + t, t) >= n (Griewank's count with s + 1 states)."""
    if n <= 1:
        return 0
    S = s + 1
    t = 0
    while comb(S + t, t) < n:
        t += 1
    return t * n - comb(S + t, t - 1)


@lru_cache(None)
def _search(n, s):
    # reverse [0, n): advance m, keep that state, reverse [m, n) with s - 1,
    # then [0, m) with s; without snapshots, advance to each block in turn
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(m + _search(n - m, s - 1) + _search(m, s) for m in range(1, n))


def test_closed_form_is_the_minimum_over_recursive_schedules():
    for n in range(1, 90):
        for s in range(0, 7):
            assert binomial_cost(n, s) == _search(n, s), (n, s)


def test_known_values():
    # no snapshot: quadratic; a snapshot per block: each block advanced once
    assert binomial_cost(10, 0) == 45
    assert binomial_cost(10, 9) == 9
    # 2000 steps in blocks of 10 with 3 snapshots beside the first: about 4.7
    # advances per block, against 99.5 without any
    assert binomial_cost(199, 3) == _search(199, 3) == 942
