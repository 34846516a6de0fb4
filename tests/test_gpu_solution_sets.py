"""Full solution sets on the GPU: the hash join (wzfh_join_build / probe / solutions) lists every
(A, B, C, D) of a signed signature in the space WZ_JOIN22 searches, and the list must equal the
oracle's (wzo_solutions: both sides enumerated from the Thm 2.2 encoding with no profile list
and no hall test, pinned to a brute force at n = 6..10 in tests/test_solution_sets.py).

Covers the exhaustive configurations: BS(7,6), every class and every signing of each, and
BS(20,19), every class. The join's per-profile streams, hall test and exact-vector table all run
on the device; every listed quadruple also passes the exact NPAF check, and the join's FOUND
flag and raw hit count agree with the set."""
import sys

import numpy as np
import pytest

import bs45_quantum_explorer_amd as wz
from oracle_lib import solutions
from bs45_quantum_explorer_amd.workload import classes

pytestmark = pytest.mark.gpu


def signed(sig):
    out = []
    for mask in range(16):
        v = tuple(-x if (mask >> i) & 1 else x for i, x in enumerate(sig))
        if v not in out:
            out.append(v)
    return out


def unpack(w, L):
    return [-1 if (w >> i) & 1 else 1 for i in range(L)]


def gpu_solutions(n, sig, **flags):
    J = wz.Join(n, sig, **flags)
    try:
        J.build()
        r = J.probe()
        sols = J.solutions()
        return sols, r.found, r.hits
    finally:
        J.close()


def check(n, sig, **flags):
    print(f'solution set n={n} sig={sig} {flags}', file=sys.__stderr__, flush=True)
    got, found, hits = gpu_solutions(n, sig, **flags)
    want = solutions(n, sig)
    assert got == want, (n, sig, len(got), len(want))
    assert bool(found) == bool(want)
    assert hits <= len(want) and (hits > 0) == bool(want)
    for A, B, C, D in got[:: max(1, len(got) // 200)]:
        assert wz.npaf_max(n, unpack(A, n + 1), unpack(B, n + 1), unpack(C, n), unpack(D, n)) == 0
    return len(got)


CASES6 = [v for s in classes(6) for v in signed(s)]


@pytest.mark.parametrize('sig', CASES6, ids=[str(s) for s in CASES6])
def test_bs7_6_full_solution_set(sig):
    check(6, sig)


@pytest.mark.parametrize('sig', classes(19), ids=[str(s) for s in classes(19)])
def test_bs20_19_full_solution_set(sig):
    assert check(19, sig) > 0


@pytest.mark.parametrize('n', [8, 11, 13])
def test_small_n_full_solution_sets(n):
    for sig in classes(n):
        check(n, sig)


def test_theorem_filters_keep_a_subset():
    """Thm 2.11b / 2.12 are extra filters on both profile lists: with them on, the set is a subset
    of the unfiltered one (equal when they prune no solution)."""
    n, sig = 13, classes(13)[0]
    on, _, _ = gpu_solutions(n, sig, thm211b=1, thm212=1)
    assert set(on) <= set(solutions(n, sig))
