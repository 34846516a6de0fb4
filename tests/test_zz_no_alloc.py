"""The allocation-free completer (wzfh_complete_noalloc) against wzfh_complete, through the C ABI.

Two calls in a row on one handle (the second with fewer candidates, so stale contents of the
static scratch from the first call must not leak into it), each compared field by field with
wzfh_complete on the same candidates, under the cells' profile rows and a node budget that
makes some completions abort. The device's free memory must not move across the calls (the
entry point allocates nothing), and more candidates than the maximum are refused."""
import ctypes as C

import numpy as np
import pytest

import bs45_quantum_explorer_amd as wz
from bs45_quantum_explorer_amd.workload import classes, sample_cells

pytestmark = pytest.mark.gpu


def _free_bytes():
    """Free device memory as the library's own HIP runtime reports it (hipMemGetInfo)."""
    hip = C.CDLL('libamdhip64.so')
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _candidates(P, cells, node_budget, per_cell):
    st, offs, cd = P.stream(cells, node_budget=node_budget)
    sel, cell_of = [], []
    for j, c in enumerate(cells):
        part = cd[offs[j]:offs[j + 1]]
        order = sorted(range(len(part)), key=lambda i: wz.flat_score(P.n, part[i][0], part[i][1]))[:per_cell]
        sel += [part[i] for i in order]
        cell_of += [c] * len(order)
    return np.array(sel, dtype=np.uint64).reshape(-1, 2), cell_of


def _same(a, b):
    return [(x.verdict, x.nodes, x.A, x.B) for x in a] == [(x.verdict, x.nodes, x.A, x.B) for x in b]


def test_noalloc_completer_matches_the_allocating_one():
    n = 44
    P = wz.Problem(n, classes(n)[3], canon=1, order=1)
    cells = sample_cells(P.live_cells(), 1024, 44 * 1000 + 3)
    cd, cell_of = _candidates(P, cells, 1 << 22, 16)
    assert 24 <= len(cd) <= wz.NOALLOC_MAX_CAND, len(cd)
    print(f'noalloc: {len(cd)} candidates', flush=True)
    P2 = wz.Problem(n, classes(n)[3], canon=1, order=1)   # a handle whose own buffers were never grown
    free0 = _free_bytes()
    first = P2.complete(cd, budget=1 << 14, cell_of=cell_of, noalloc=True)
    second = P2.complete(cd[:len(cd) // 3], budget=1 << 18, cell_of=cell_of[:len(cd) // 3], noalloc=True)
    free1 = _free_bytes()
    assert abs(free1 - free0) < (1 << 20), (free0, free1)
    want1 = P.complete(cd, budget=1 << 14, cell_of=cell_of)
    want2 = P.complete(cd[:len(cd) // 3], budget=1 << 18, cell_of=cell_of[:len(cd) // 3])
    assert _same(first, want1)
    assert _same(second, want2)
    verdicts = {c.verdict for c in want1}
    assert 3 in verdicts and 2 in verdicts, verdicts   # the budget aborts some, others finish clean
    # unconstrained (no rows), exact budget
    third = P2.complete(cd[:40], budget=0, noalloc=True)
    assert _same(third, P.complete(cd[:40], budget=0))
    P.close()
    P2.close()


def test_noalloc_completer_refuses_more_than_its_maximum():
    P = wz.Problem(13, (2, 4, 3, 5), m6=1, canon=1)
    cd = np.zeros((wz.NOALLOC_MAX_CAND + 1, 2), dtype=np.uint64)
    out = (wz.Completion * len(cd))()
    rc = wz.lib().wzfh_complete_noalloc(P.h, cd.ctypes.data, None, len(cd), 0, out)
    assert rc == wz.EINVAL
    assert 'at most' in wz.lib().wzfh_last_error().decode()
    P.close()
