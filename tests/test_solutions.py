"""NPAF of the reference's banked solutions (tests/golden/solutions.json): every banked BS(n+1,n)
has zero summed NPAF at s = 1..n through the library's exact check and the oracle's, the
quarantined invalid one does not, and each satisfies the Thm 2.2 pair encoding the search uses."""
import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import load
from oracle_lib import OracleProblem, lib as olib
import numpy as np

SOL = load('solutions.json')


def oracle_npaf_max(n, A, B, C, D):
    a = [np.ascontiguousarray(np.asarray(x, np.int8)) for x in (A, B, C, D)]
    return max(abs(olib().wzo_npaf_at(a[0].ctypes.data, a[1].ctypes.data, n + 1, a[2].ctypes.data,
                                       a[3].ctypes.data, n, s)) for s in range(1, n + 1))


@pytest.mark.parametrize('rec', SOL, ids=[r['file'].split('/')[-1] for r in SOL])
def test_banked_solution_npaf(rec):
    n = rec['n']
    A, B, C, D = rec['A'], rec['B'], rec['C'], rec['D']
    assert len(A) == len(B) == n + 1 and len(C) == len(D) == n
    got = wz.npaf_max(n, A, B, C, D)
    assert got == oracle_npaf_max(n, A, B, C, D)
    if rec['quarantined']:
        assert got != 0
    else:
        assert got == 0
        assert sum(x * x for x in (sum(A), sum(B), sum(C), sum(D))) == 4 * n + 2


def test_solutions_fit_pair_encoding():
    """Thm 2.2: mirror quads (A_d, B_d, A_{L-1-d}, B_{L-1-d}) have product -1 at d = 0 and +1
    inside; on C,D every interior quad is +1 (the stream's P22_POS at d >= 1)."""
    for rec in SOL:
        if rec['quarantined']:
            continue
        n = rec['n']
        A, B, C, D = rec['A'], rec['B'], rec['C'], rec['D']
        L = n + 1
        assert A[0] * B[0] * A[L - 1] * B[L - 1] == -1, rec['file']
        for d in range(1, L // 2):
            assert A[d] * B[d] * A[L - 1 - d] * B[L - 1 - d] == 1, rec['file']
        for d in range(1, n // 2):
            assert C[d] * D[d] * C[n - 1 - d] * D[n - 1 - d] == 1, rec['file']
