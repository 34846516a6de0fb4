"""GPU parity: the HIP stream and completer against the reference's golden outputs.

Bit-exact bar: per-cell candidate counts, leaf counts, hall_ok counts, DFS node counts and
the SHA-1 of each cell's stream bytes (WZ_FH_DUMP format) must equal the reference's
(tests/golden/stream_cells.json); completion verdicts, node counts (including the budget
abort unwinding) and first A,B must equal WZ_FH_COMPLETE_C/_D's (complete_only.json).
"""
import hashlib

import numpy as np
import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import complete_cases, load, stream_runs

pytestmark = pytest.mark.gpu


def _problem(rec):
    return wz.Problem(rec['n'], rec['sig'], m6=1, thm211b=1, thm212=1, canon=rec['canon'], rev=rec['rev'],
                      endpos=rec['endpos'], order=rec['order'])


def _cells_of(rec, P, cand_cap=0, split_depth=0):
    cells = [c[0] for c in rec['cells']]
    if not cells:
        return []
    stats, offs, cd = P.stream(cells, cand_cap=cand_cap, split_depth=split_depth)
    rows = []
    for i, pi in enumerate(cells):
        part = cd[offs[i]:offs[i + 1]]
        h = hashlib.sha1(''.join(wz.dump_lines(rec['n'], part)).encode()).hexdigest()
        st = stats[i]
        rows.append([pi, st.cand, st.leaves, st.hall_ok, st.dfs_nodes, h])
    return rows


RUNS = stream_runs()
SMALL = [r for r in RUNS if r['n'] <= 13]
N19 = [r for r in RUNS if r['n'] == 19]


@pytest.mark.parametrize('n', list(range(6, 14)))
def test_stream_small_n_matches_reference(n):
    runs = [r for r in SMALL if r['n'] == n]
    assert runs
    for rec in runs:
        P = _problem(rec)
        try:
            live = P.live_cells()
            assert live == [c[0] for c in rec['cells']], 'cell list differs'
            got = _cells_of(rec, P)
            assert got == rec['cells'], (rec['n'], rec['sig'], rec['canon'], rec['rev'], rec['endpos'])
        finally:
            P.close()


@pytest.mark.parametrize('split', [1, 2, 4])
def test_stream_split_depth_invariance(split):
    rec = next(r for r in SMALL if r['n'] == 13 and r['canon'] == 0 and r['rev'] == 1)
    P = _problem(rec)
    try:
        assert _cells_of(rec, P, split_depth=split) == rec['cells']
    finally:
        P.close()


@pytest.mark.parametrize('idx', range(len(N19)))
def test_stream_n19_matches_reference(idx):
    rec = N19[idx]
    P = _problem(rec)
    try:
        assert _cells_of(rec, P) == rec['cells']
    finally:
        P.close()


def test_stream_candidate_cap_is_exact_prefix():
    """WZ_FH_CELLSIZE semantics: a capped cell streams exactly the first `cap` candidates."""
    rec = next(r for r in N19 if r['canon'] == 1 and r['rev'] == 0)
    P = _problem(rec)
    try:
        cells = [c[0] for c in rec['cells'] if c[1] >= 5][:40]
        full_stats, full_offs, full_cd = P.stream(cells)
        for cap in (1, 3, 17):
            st, offs, cd = P.stream(cells, cand_cap=cap)
            for i in range(len(cells)):
                want = full_cd[full_offs[i]:full_offs[i] + min(cap, full_stats[i].cand)]
                assert np.array_equal(cd[offs[i]:offs[i + 1]], want)
                assert st[i].cand == min(cap, full_stats[i].cand)
                assert st[i].capped == (1 if full_stats[i].cand >= cap else 0)
                assert st[i].dfs_nodes <= full_stats[i].dfs_nodes
    finally:
        P.close()


CASES = complete_cases()


def _complete_group(cases):
    n, sig, k = cases[0]['n'], cases[0]['sig'], cases[0]['outer_k']
    P = wz.Problem(n, sig, m6=0, thm211b=0, thm212=0, outer_k=k)
    try:
        for budget in sorted(set(c['budget'] for c in cases)):
            grp = [c for c in cases if c['budget'] == budget]
            cd = np.array([[wz.pack(c['C']), wz.pack(c['D'])] for c in grp], dtype=np.uint64)
            out = P.complete(cd, budget=budget)
            for c, o in zip(grp, out):
                assert (o.verdict, o.nodes) == (c['r'], c['nodes']), (c['name'], budget, k)
                if c['r'] == 0:
                    assert list(wz.unpack(o.A, n + 1)) == c['A'] and list(wz.unpack(o.B, n + 1)) == c['B']
    finally:
        P.close()


GROUPS = {}
for c in CASES:
    GROUPS.setdefault((c['n'], tuple(c['sig']), c['outer_k']), []).append(c)


@pytest.mark.parametrize('key', sorted(GROUPS), ids=lambda k: f'n{k[0]}-{k[1]}-k{k[2]}')
def test_complete_only_matches_reference(key):
    _complete_group(GROUPS[key])


BUDGETED = sorted(k for k in GROUPS if any(c['budget'] > 0 for c in GROUPS[k]))


@pytest.mark.parametrize('split,cap', [(0, 0), (2, 0), (3, 0), (3, 64)])
@pytest.mark.parametrize('key', BUDGETED, ids=lambda k: f'n{k[0]}-{k[1]}-k{k[2]}')
def test_complete_split_abort_search_matches_reference(key, split, cap, monkeypatch):
    """Budgeted completions with the candidate split into 8^split quad-path items, the item
    holding a budget abort split again down to single items (threshold 1 charge), and the
    last one walked whole for the exact abort count (split 0: every candidate unsplit). With a
    per-item charge cap, an entry whose capped item comes before its decision is run again
    split 2 levels deeper."""
    monkeypatch.setenv('WZFH_COMPLETE_SPLIT', str(split))
    monkeypatch.setenv('WZFH_COMPLETE_ITEM_CAP', str(cap))
    monkeypatch.setenv('WZFH_COMPLETE_EXACT_ITEM', '1')
    monkeypatch.setenv('WZFH_CHECK_REDUCE', '1')   # the device reduce against a host scan of the items
    _complete_group([c for c in GROUPS[key] if c['budget'] > 0])


@pytest.mark.parametrize('a2', [0, 2])
@pytest.mark.parametrize('n,sig,budgets', [(13, (2, 4, 3, 5), (0, 30)), (19, (0, 2, 5, 7), (0, 500)),
                                           (12, (1, 7, 0, 0), (0, 7)), (29, (0, 6, 1, 9), (2000, 20000))])
def test_in_cell_completion_with_profile_rows_matches_oracle(n, sig, budgets, a2):
    """The in-stream completer (cell's A,B profile rows applied, wz_match.cpp:2505-2530) against
    the oracle, which is pinned to the reference's full arm runs (firsthit_runs.json; with the
    A2 row prune, WZ_FH_A2=2: a2_runs.json)."""
    from oracle_lib import OracleProblem
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1, a2=a2)
    O = OracleProblem(n, sig, canon=1, order=1, a2=a2)
    try:
        cells = P.live_cells()[:60]
        st, offs, cd = P.stream(cells, cand_cap=12)
        cell_of = [cells[i] for i in range(len(cells)) for _ in range(offs[i], offs[i + 1])]
        for budget in budgets:
            comp = P.complete(cd, budget=budget, cell_of=cell_of)
            for j, (c, pi) in enumerate(zip(comp, cell_of)):
                Cs, Ds = wz.unpack(cd[j][0], n), wz.unpack(cd[j][1], n)
                r, nodes, A, B = O.complete(Cs, Ds, budget, pi=pi)
                assert (c.verdict, c.nodes) == (r, nodes), (n, sig, budget, pi, j)
                if r == 0:
                    assert list(wz.unpack(c.A, n + 1)) == list(A) and list(wz.unpack(c.B, n + 1)) == list(B)
    finally:
        P.close()


@pytest.mark.parametrize('split,window', [(2, 1), (3, 4), (4, 64)])
def test_stream_node_budget_matches_oracle(split, window):
    """Serving policy: a cell's stream is cut at a DFS node budget (nodes with DFS index below
    the budget are visited). The oracle restates the cut (wzo_stream_cell_budget)."""
    from oracle_lib import OracleProblem
    n, sig = 19, (2, 4, 3, 7)
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1)
    O = OracleProblem(n, sig, canon=1, order=1)
    try:
        cells = P.live_cells()[:48]
        for budget in (1, 37, 500, 4096, 100000):
            st, offs, cd = P.stream(cells, node_budget=budget, split_depth=split, window=window)
            for i, pi in enumerate(cells):
                got, cands, ost = O.stream_budget(pi, 0, budget, 4096)
                assert st[i].cand == got and st[i].dfs_nodes == ost.dfs_nodes, (budget, pi)
                if ost.dfs_nodes < budget:
                    assert st[i].leaves == ost.leaves
                assert wz.dump_lines(n, cd[offs[i]:offs[i + 1]]) == [
                    ''.join(f'{int(v)} ' for v in c[0]) + '| ' + ''.join(f'{int(v)} ' for v in c[1]) + '\n'
                    for c in cands], (budget, pi)
    finally:
        P.close()


def test_stream_cap_and_budget_together_match_oracle():
    from oracle_lib import OracleProblem
    n, sig = 13, (2, 4, 3, 5)
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=0, order=1)
    O = OracleProblem(n, sig, canon=0, order=1)
    try:
        cells = P.live_cells()[:80]
        for cap, budget in ((3, 200), (1, 50), (10, 10 ** 6)):
            st, offs, cd = P.stream(cells, cand_cap=cap, node_budget=budget, split_depth=2, window=2)
            for i, pi in enumerate(cells):
                got, cands, ost = O.stream_budget(pi, cap, budget, 64)
                assert (st[i].cand, st[i].dfs_nodes) == (got, ost.dfs_nodes), (cap, budget, pi)
    finally:
        P.close()


@pytest.mark.parametrize('cap,budget,keep', [(0, 0, 0), (0, 37, 0), (0, 4096, 0), (3, 0, 0), (3, 200, 0),
                                             (0, 37, 1), (0, 4096, 1), (0, 4096, 3), (3, 200, 1), (3, 0, 2)])
def test_stream_forced_donation_matches_oracle(cap, budget, keep, monkeypatch):
    """Running items donate their untried subtrees every 2 DFS moves: item lists, offsets,
    tails and the cut must still reproduce the oracle's stream exactly. keep > 0: under a cap a
    donation gives away every level from the shallowest live one down to depth d-1-keep at once
    (deepest first), so items with tails between levels are exercised."""
    from oracle_lib import OracleProblem
    monkeypatch.setenv('WZFH_DONATE_MOVES', '2')
    monkeypatch.setenv('WZFH_DONATE_LOW', '1000000')
    monkeypatch.setenv('WZFH_DONATE_ANY', '1')
    monkeypatch.setenv('WZFH_DONATE_KEEP', str(keep))
    n, sig = 19, (2, 4, 3, 7)
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1)
    O = OracleProblem(n, sig, canon=1, order=1)
    try:
        cells = P.live_cells()[:24]
        st, offs, cd = P.stream(cells, cand_cap=cap, node_budget=budget, split_depth=2, window=2)
        for i, pi in enumerate(cells):
            got, cands, ost = O.stream_budget(pi, cap, budget, 100000)
            assert (st[i].cand, st[i].dfs_nodes) == (got, ost.dfs_nodes), (cap, budget, pi)
            if cap == 0 and (budget == 0 or ost.dfs_nodes < budget):
                assert st[i].leaves == ost.leaves
            assert wz.dump_lines(n, cd[offs[i]:offs[i + 1]]) == [
                ''.join(f'{int(v)} ' for v in c[0]) + '| ' + ''.join(f'{int(v)} ' for v in c[1]) + '\n'
                for c in cands], (cap, budget, pi)
    finally:
        P.close()


@pytest.mark.parametrize('n,sig', [(29, (2, 4, 7, 7)), (44, (1, 7, 8, 8))])
def test_matrix_core_leaf_test_agrees_with_fp64_on_every_leaf(n, sig, monkeypatch):
    """The fp16x2 MFMA leaf test with its rigorous margin decides each leaf certain or
    uncertain; with WZFH_HALL_VERIFY=1 the kernel also runs the reference's FP64 hall_ok
    on every leaf and counts certain decisions that disagree with it (false rejects and
    false accepts). The count must be 0, and the stream must equal the normal-mode stream."""
    P = wz.Problem(n, sig, canon=1, order=1)
    try:
        cells = P.live_cells()[:: max(1, len(P.live_cells()) // 24)][:24]
        budget = 1 << 18
        monkeypatch.setenv('WZFH_TELEMETRY', '1')
        st0, offs0, cd0 = P.stream(cells, node_budget=budget)
        leaves0 = P.last_kernel(5)[1]
        monkeypatch.setenv('WZFH_HALL_VERIFY', '1')
        st1, offs1, cd1 = P.stream(cells, node_budget=budget)
        tested, disagree = P.last_kernel(5)[1], P.last_kernel(7)[1]
        assert leaves0 > 0 and tested > 0
        assert disagree == 0, (n, disagree, tested)
        assert [(s.cand, s.leaves, s.hall_ok, s.dfs_nodes) for s in st0] == \
               [(s.cand, s.leaves, s.hall_ok, s.dfs_nodes) for s in st1]
        assert np.array_equal(offs0, offs1) and np.array_equal(cd0, cd1)
        assert P.last_kernel(10)[0] > 0   # the margin in use
    finally:
        P.close()


QPRUNE = load('qprune_cells.json')


@pytest.mark.parametrize('n', sorted({r['n'] for r in QPRUNE}))
def test_stream_after_qprune_matches_reference(n):
    """The cells the Q-closure prune leaves (WZ_FH_ORBIT_QPRUNE with the 64-element group) stream
    the reference's candidates, leaves, hall_ok and DFS nodes, byte for byte."""
    for rec in (r for r in QPRUNE if r['n'] == n):
        P = wz.Problem(n, rec['sig'], m6=1, thm211b=1, thm212=1, canon=1, orbit_q=1, orbit_qprune=1,
                       endpos=rec['endpos'], order=rec['order'], rev=0)
        try:
            assert P.live_cells() == [c[0] for c in rec['cells']]
            assert _cells_of(dict(rec, rev=0), P) == rec['cells'], (n, rec['sig'])
        finally:
            P.close()
