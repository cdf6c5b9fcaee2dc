"""The closed form the device hit resolver (k_hit_resolve / chain_rank) uses
for the cheapest unary chain over a hit tree, pinned by enumerating every
chain of '~' / neg nodes: a '~' node at size s maps x to -x - 1 and adds 0
to its child's rank, a neg node maps x to -x and adds T[s][0..3]
(nocache.cu:52-63).  For each parity, offset k, minimum length and size
bound, the shortest chain and its smallest rank must agree."""
import itertools

import pytest

from oracle import oracle as orc


def _negoff(T, s):
    return sum(int(T[s][o]) for o in range(4))


def _closed(T, sb, k, odd, u_min, max_size):
    """chain_rank of simba_kernels.hip, restated."""
    ak = abs(k)
    need = (2 * k + 1 if k >= 0 else 2 * ak - 1) if odd else 2 * ak
    lo = (u_min | 1) if odd else u_min + (u_min & 1)
    u = max(need, lo)
    if sb + u > max_size:
        return None
    d = k - (-1 if odd else 0)
    i = 1 if (d >= 0) == odd else 2
    add = 0
    for _ in range(abs(d)):
        add += _negoff(T, sb + i)
        i += 2
    return u, add


def _brute(T, sb, k, odd, u_min, max_size):
    for u in range(u_min, max_size - sb + 1):
        if u % 2 != int(odd):
            continue
        best = None
        for ops in itertools.product((0, 1), repeat=u):  # innermost first; 1 = neg
            sigma, c, add = 1, 0, 0
            for i, neg in enumerate(ops, start=1):
                sigma, c = -sigma, (-c if neg else -c - 1)
                if neg:
                    add += _negoff(T, sb + i)
            assert sigma == (-1 if odd else 1)
            if c == k and (best is None or add < best):
                best = add
        if best is not None:
            return u, best
    return None


@pytest.mark.parametrize("L", [1, 2, 3, 5])
def test_closed_form_chain_matches_enumeration(L):
    T = orc.count_table(L, 14)
    n = 0
    for sb in (1, 2, 3, 5):
        for max_size in range(sb, min(sb + 7, 14) + 1):
            U = max_size - sb
            for u_min in range(0, U + 2):
                for k in range(-U - 1, U + 2):
                    for odd in (False, True):
                        assert _closed(T, sb, k, odd, u_min, max_size) == \
                            _brute(T, sb, k, odd, u_min, max_size), (sb, max_size, u_min, k, odd)
                        n += 1
    assert n > 1000
