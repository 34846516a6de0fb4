"""A1 outer-lag tables (wz_match.cpp:547-598) rebuilt from their definition by brute force:
place the boundary quads and the next k quads into real length-L arrays, compute the actual
correlation contributions of the newly placed positions at lags L-1-d-j, and collect the
reachable tuples with their first-quad masks. Checks the oracle's tables at two (L, d, k)
geometries per k, which also shows the tables do not depend on L or d."""
import ctypes
import itertools

import numpy as np
import pytest

from oracle_lib import lib

ALL = [tuple(1 if (i >> b) & 1 else -1 for b in (3, 2, 1, 0)) for i in range(16)]
POS = [q for q in ALL if q[0] * q[1] * q[2] * q[3] == 1]
NEG = [q for q in ALL if q[0] * q[1] * q[2] * q[3] == -1]


def brute(L, d, k):
    roots = {int(q[2] > 0): q for q in NEG if q[0] == 1 and q[1] == 1}
    out = {}
    for pat in range(2 * 8 ** (k - 1)):
        A = [0] * L
        B = [0] * L
        rq = roots[pat & 1]
        A[0], B[0], A[L - 1], B[L - 1] = rq
        rest = pat >> 1
        for t in range(1, k):
            q = POS[rest % 8]
            rest //= 8
            A[t], B[t], A[L - 1 - t], B[L - 1 - t] = q
        for t in range(k, d):   # arbitrary filler between the boundary and the open quads
            A[t], B[t], A[L - 1 - t], B[L - 1 - t] = POS[5]
        reach = {}
        for asg in itertools.product(range(8), repeat=k):
            A3, B3 = A[:], B[:]
            for r, qi in enumerate(asg):
                A3[d + r], B3[d + r], A3[L - 1 - d - r], B3[L - 1 - d - r] = POS[qi]
            F = []
            for j in range(k):
                s = L - 1 - d - j
                tot = 0
                for i in range(L - s):
                    new_i = d <= i < d + k or L - d - k <= i < L - d
                    new_j = d <= i + s < d + k or L - d - k <= i + s < L - d
                    if new_i or new_j:
                        tot += A3[i] * A3[i + s] + B3[i] * B3[i + s]
                F.append(tot)
            reach[tuple(F)] = reach.get(tuple(F), 0) | (1 << asg[0])
        out[pat] = reach
    return out


def oracle_tables():
    t2 = np.zeros(16 * 5 * 9, np.uint8)
    t3 = np.zeros(128 * 5 * 9 * 13, np.uint8)
    lib().wzo_outer_tables(t2.ctypes.data, t3.ctypes.data)
    return t2.reshape(16, 5, 9), t3.reshape(128, 5, 9, 13)


@pytest.mark.parametrize('L,d,k', [(11, 3, 2), (13, 4, 2), (13, 3, 3), (17, 5, 3)])
def test_outer_tables_equal_brute_force(L, d, k):
    t2, t3 = oracle_tables()
    want = brute(L, d, k)
    for pat, reach in want.items():
        got = {}
        if k == 2:
            for a, b in zip(*np.nonzero(t2[pat])):
                got[((a - 2) * 2, (b - 4) * 2)] = int(t2[pat][a, b])
        else:
            for a, b, c in zip(*np.nonzero(t3[pat])):
                got[((a - 2) * 2, (b - 4) * 2, (c - 6) * 2)] = int(t3[pat][a, b, c])
        assert got == reach, (L, d, k, pat)
