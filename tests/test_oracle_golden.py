"""Pins the CPU oracle (oracle/wz_oracle.cpp) to the reference's golden outputs, so that the
oracle can check the GPU where the fixtures end (random cells, larger n)."""
import hashlib

import numpy as np
import pytest

from golden_data import complete_cases, stream_runs
from oracle_lib import OracleProblem, dump_lines

RUNS = stream_runs()


@pytest.mark.parametrize('n', list(range(6, 14)))
def test_oracle_stream_matches_reference(n):
    for rec in (r for r in RUNS if r['n'] == n):
        P = OracleProblem(n, rec['sig'], canon=rec['canon'], rev=rec['rev'], endpos=rec['endpos'], order=rec['order'])
        assert P.num_cells() == rec['profiles']
        got = []
        for pi in range(P.num_cells()):
            _, _, kept, rows = P.cell(pi)
            if not kept or rows == -1:
                continue
            cnt, cands, st = P.stream(pi, 0, 4096)
            h = hashlib.sha1(''.join(dump_lines(cands)).encode()).hexdigest()
            got.append([pi, cnt, st.leaves, st.hall_ok, st.dfs_nodes, h])
        assert got == rec['cells'], (n, rec['sig'], rec['canon'], rec['rev'])
        P.close()


def test_oracle_stream_n19_one_class():
    rec = next(r for r in RUNS if r['n'] == 19 and r['canon'] == 1 and r['rev'] == 0)
    P = OracleProblem(19, rec['sig'], canon=1, order=rec['order'])
    for row in rec['cells'][::7]:
        cnt, cands, st = P.stream(row[0], 0, 20000)
        h = hashlib.sha1(''.join(dump_lines(cands)).encode()).hexdigest()
        assert [row[0], cnt, st.leaves, st.hall_ok, st.dfs_nodes, h] == row


CASES = complete_cases()


@pytest.mark.parametrize('n', sorted(set(c['n'] for c in CASES)))
def test_oracle_completer_matches_reference(n):
    cache = {}
    for c in (x for x in CASES if x['n'] == n):
        key = (tuple(c['sig']), c['outer_k'])
        if key not in cache:
            cache[key] = OracleProblem(n, c['sig'], outer_k=c['outer_k'], no_cells=1)
        r, nodes, A, B = cache[key].complete(c['C'], c['D'], c['budget'])
        assert (r, nodes) == (c['r'], c['nodes']), c['name']
        if r == 0:
            assert list(A) == c['A'] and list(B) == c['B']


def test_oracle_node_budget_is_a_prefix_of_the_golden_stream():
    """The node budget (a serving policy) truncates the reference recursion at rec() entry:
    the budgeted stream must be a prefix of the golden-pinned full stream, grow monotonically
    with the budget, report min(total nodes, budget) DFS nodes, and equal the full stream
    (with exact leaves) once the budget covers the cell."""
    rec = next(r for r in RUNS if r['n'] == 19 and r['canon'] == 1 and r['rev'] == 0)
    P = OracleProblem(19, rec['sig'], canon=1, order=rec['order'])
    for row in rec['cells'][::11]:
        pi, total_nodes = row[0], row[4]
        cnt, full, st = P.stream(pi, 0, 20000)
        h = hashlib.sha1(''.join(dump_lines(full)).encode()).hexdigest()
        assert [pi, cnt, st.leaves, st.hall_ok, st.dfs_nodes, h] == row   # the reference's stream
        prev = 0
        for budget in sorted({1, 2, 37, 500, 4096, total_nodes // 2, total_nodes - 1, total_nodes, total_nodes + 1}):
            if budget < 1:
                continue
            got, cands, bst = P.stream_budget(pi, 0, budget, 20000)
            assert bst.dfs_nodes == min(total_nodes, budget)
            assert prev <= got <= cnt
            assert np.array_equal(cands, full[:got])
            if budget >= total_nodes:
                assert (got, bst.leaves) == (cnt, st.leaves)
            prev = got
