"""The oracle's full solution sets (oracle_lib.solutions, wzo_solutions) against a brute force
written from the definitions alone.

The set is every (A, B, C, D) of a signed signature with zero summed NPAF at s = 1..n in the
space WZ_JOIN22 searches (src/solver/wz_match.cpp:2894-2904): the Thm 2.2 pair encoding
(:525-530) and X[0] = +1 when X's sum is 0 (:1526-1527). The brute force below takes every
+-1 sequence of each length with the right sum (numpy), forms every A,B and every C,D pair,
keeps the pairs whose mirror quads fit the encoding, and joins on NPAF vectors computed with
np.correlate -- no DFS, no packing, no pruning -- the same ground truth the reference's own
retention test builds (tools/test_orbit_q_retention.py:35-55, restated here). It covers every
class of n = 6..10 and signed variants at n = 6, 7. The banked solutions of n = 7 and 11
(normalized by the sign changes that keep the encoding) must be members of the oracle's set.
"""
import itertools

import numpy as np
import pytest

from golden_data import load
from oracle_lib import solutions
from bs45_quantum_explorer_amd.workload import classes


def seqs_with_sum(L, s):
    """Every +-1 vector of length L with sum s, as rows (element i of row r)."""
    idx = np.arange(1 << L, dtype=np.int64)[:, None]
    X = (1 - 2 * ((idx >> np.arange(L)) & 1)).astype(np.int64)
    X = X[X.sum(1) == s]
    if s == 0:
        X = X[X[:, 0] == 1]
    return X


def npaf_rows(X, n):
    """N_X(s) for s = 1..n of every row, from the definition."""
    L = X.shape[1]
    return np.stack([np.array([int(np.dot(x[:L - s], x[s:])) if s < L else 0 for x in X], dtype=np.int64)
                     for s in range(1, n + 1)], 1) if len(X) else np.zeros((0, n), np.int64)


def pairs_fitting(X, Y, ab):
    """Every (x, y) pair of rows whose mirror quads (x_d, y_d, x_{L-1-d}, y_{L-1-d}) have product +1
    at d >= 1 and, on the A,B side, -1 at d = 0 (any on the C,D side); the odd middle is free."""
    L = X.shape[1]
    out = []
    for i, j in itertools.product(range(len(X)), range(len(Y))):
        x, y = X[i], Y[j]
        ok = True
        for d in range(L // 2):
            p = x[d] * y[d] * x[L - 1 - d] * y[L - 1 - d]
            if d == 0:
                if ab and p != -1:
                    ok = False
                    break
            elif p != 1:
                ok = False
                break
        if ok:
            out.append((i, j))
    return out


def pack(v):
    return sum(1 << i for i, x in enumerate(v) if x < 0)


def brute_force(n, sig):
    a, b, c, d = sig
    A, B = seqs_with_sum(n + 1, a), seqs_with_sum(n + 1, b)
    C, D = seqs_with_sum(n, c), seqs_with_sum(n, d)
    nA, nB, nC, nD = (npaf_rows(X, n) for X in (A, B, C, D))
    cd = {}
    for i, j in pairs_fitting(C, D, ab=False):
        cd.setdefault(tuple(-(nC[i] + nD[j])), []).append((pack(C[i]), pack(D[j])))
    out = []
    for i, j in pairs_fitting(A, B, ab=True):
        for Cw, Dw in cd.get(tuple(nA[i] + nB[j]), []):
            out.append((pack(A[i]), pack(B[j]), Cw, Dw))
    return sorted(out)


def signed(sig):
    out = []
    for mask in range(16):
        v = tuple(-x if (mask >> i) & 1 else x for i, x in enumerate(sig))
        if v not in out:
            out.append(v)
    return out


CASES = [(n, s) for n in range(6, 11) for s in classes(n)] + \
        [(n, v) for n in (6, 7) for s in classes(n) for v in signed(s)[1:]]


@pytest.mark.parametrize('n,sig', CASES, ids=[f'{n}-{s}' for n, s in CASES])
def test_oracle_solution_set_equals_brute_force(n, sig):
    got = solutions(n, sig)
    assert got == brute_force(n, sig)
    assert len(got) == len(set(got))


def test_every_solution_has_zero_npaf():
    for n, sig in [(8, classes(8)[0]), (13, classes(13)[3])]:
        for A, B, C, D in solutions(n, sig):
            unpack = lambda w, L: np.array([-1 if (w >> i) & 1 else 1 for i in range(L)])
            a, b, c, dd = unpack(A, n + 1), unpack(B, n + 1), unpack(C, n), unpack(D, n)
            assert (a.sum(), b.sum(), c.sum(), dd.sum()) == sig
            for s in range(1, n + 1):
                v = int(np.dot(a[:n + 1 - s], a[s:]) + np.dot(b[:n + 1 - s], b[s:]))
                if s < n:
                    v += int(np.dot(c[:n - s], c[s:]) + np.dot(dd[:n - s], dd[s:]))
                assert v == 0


def normalize(A, B, C, D):
    """Negating a sequence keeps NPAF and the pair encoding; a zero-sum sequence starting with -1 is
    negated so it meets the X[0] = +1 pin."""
    out = []
    for X in (A, B, C, D):
        out.append([-x for x in X] if sum(X) == 0 and X[0] < 0 else list(X))
    return out


def test_banked_small_solutions_are_in_the_oracle_set():
    recs = [r for r in load('solutions.json') if not r['quarantined'] and r['n'] <= 11]
    assert recs
    for r in recs:
        n = r['n']
        A, B, C, D = normalize(r['A'], r['B'], r['C'], r['D'])
        sig = (sum(A), sum(B), sum(C), sum(D))
        assert (pack(A), pack(B), pack(C), pack(D)) in set(solutions(n, sig)), r['file']
