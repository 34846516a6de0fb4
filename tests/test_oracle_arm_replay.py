"""Pins the oracle's front-only cell search (wzo_cell_search), its flat score and its
row-constrained completer to the reference's own arm runs (firsthit_runs.json: 116 runs of
`WZ_FIRSTHIT=1 wz_match <n> <a> <b> <c> <d>` at n = 6..13 and 19, with CAND lines).

`replay` restates the arm's cell loop and drain (wz_match.cpp:2491-2530 cell skips,
:2656-2714 drain, :2782-2788 buffer push and WZ_FH_MAX_CAND, :2839 final partial batch) on
top of three oracle functions: the C,D stream, wzo_flat_score and wzo_complete under the
cell's profile rows. Every CAND line (idx, sorted rank, cell, verdict, nodes), the FIRSTHIT
line (idx, cell, nodes, score), the printed A,B,C,D and the summary counters must come out
equal. Then, for every cell the arm drained as a single sorted batch, wzo_cell_search (the
unit the GPU search is checked against) must give the same per-rank verdicts, nodes and
scores, the same streamed count and the same hit.

The same replay with WZ_FH_A2=2 (a2_runs.json) pins the oracle's exact profile-row reachability
prune: the CAND lines and node counts of the pruned arm, and the arm's A2SHADOW row counts
(rows tested and rejected by the A2 test, summed over the arm's completions).
"""
import numpy as np
import pytest

from golden_data import firsthit_runs, load
from oracle_lib import OracleProblem, flat_score

RUNS = firsthit_runs()
A2_RUNS = load('a2_runs.json')


def _stream_all(P, pi, cap):
    got, cands, st = P.stream(pi, cap, 4096)
    if got > 4096:
        got, cands, st = P.stream(pi, cap, got)
    return got, cands


def replay(run, check_cell_search=True):
    n, sig, s = run['n'], run['sig'], run['settings']
    P = OracleProblem(n, sig, canon=s.get('WZ_FH_ORBIT_CANON', 0), order=s.get('WZ_FH_PROF_ORDER', 1),
                      rev=s.get('WZ_FH_STREAM_REV', 0), a2=s.get('WZ_FH_A2', 0))
    a2_rows = [0, 0]
    buf_cap = s.get('WZ_FH_BUF_CAP', 500000)
    top = s.get('WZ_FH_DRAIN_TOP', 0)
    nbat = max(1, s.get('WZ_FH_DRAIN_BATCHES', 1))
    max_cand = s.get('WZ_FH_MAX_CAND', 0)
    budget = s['WZ_FH_AB_BUDGET']
    cand, lines, hit = 0, [], None
    tot = {'clean_no_AB': 0, 'budget_aborted': 0, 'pre_filter_rejected': 0, 'total_AB_nodes': 0}
    checked = 0
    for pi in range(P.num_cells()):
        _, _, kept, rows = P.cell(pi)
        if not kept or rows == -1:          # orbit duplicate / proven-dead cell (:2499-2521)
            continue
        limit = buf_cap * nbat if top > 0 else 0
        size, cands = _stream_all(P, pi, limit)
        take = size
        stopped = False
        if max_cand > 0 and cand + size >= max_cand:   # the arm stops at the max_cand-th candidate
            take = max_cand - cand
            stopped = True
        cell_lines = []
        nb = (take + buf_cap - 1) // buf_cap
        for b in range(nb):
            lo, hi = b * buf_cap, min(take, (b + 1) * buf_cap)
            full = hi - lo == buf_cap
            # a partial last batch drains after the DFS returns, unless the arm stopped and the
            # DFS met another candidate (which set the cell's stop flag first)
            if not full and stopped and size > take:
                break
            at = cand + hi
            batch = cands[lo:hi]
            sc = [flat_score(n, c[0], c[1]) for c in batch]
            order = sorted(range(len(batch)), key=lambda i: sc[i])     # stable: ties keep stream order
            stop_at = min(len(batch), top) if top > 0 else len(batch)
            for ci in range(stop_at):
                c = batch[order[ci]]
                r, nodes, A, B = P.complete(c[0], c[1], budget, pi=pi)
                assert r != 4
                cell_lines.append([at, ci, pi, r, nodes])
                tot['total_AB_nodes'] += nodes
                if r == 1:
                    tot['pre_filter_rejected'] += 1
                elif r == 2:
                    tot['clean_no_AB'] += 1
                elif r == 3:
                    tot['budget_aborted'] += 1
                else:
                    hit = {'firsthit': [at, pi, nodes, sc[order[ci]]], 'A': list(map(int, A)), 'B': list(map(int, B)),
                           'C': list(map(int, c[0])), 'D': list(map(int, c[1])), 'rank': ci, 'idx': lo + order[ci]}
                    break
            if hit:   # the hit stops the stream inside this drain
                take = hi
                break
        lines += cell_lines
        cand += take
        t, r = P.a2_counts()
        a2_rows[0] += t
        a2_rows[1] += r
        single_batch = nb <= 1 and not stopped
        if check_cell_search and single_batch and (top > 0 or take <= buf_cap):
            # the same cell through wzo_cell_search: one sorted batch of the first buf_cap
            res, per = P.search(pi, buf_cap if top > 0 else 0, 0, top if top > 0 else 1 << 30, budget)
            assert res.cand == take, (pi, res.cand, take)
            assert [[cand, k, pi, int(per['verdict'][k]), int(per['nodes'][k])] for k in range(res.completed)] == \
                [[cand] + x[1:] for x in cell_lines], pi
            assert res.hit == (hit is not None)
            if hit:
                assert (res.hit_rank, res.hit_idx) == (hit['rank'], hit['idx'])
                assert int(per['score'][res.hit_rank]) == hit['firsthit'][3]
                assert list(res.A[:n + 1]) == hit['A'] and list(res.B[:n + 1]) == hit['B']
                assert list(res.C[:n]) == hit['C'] and list(res.D[:n]) == hit['D']
            assert res.ab_nodes == sum(x[4] for x in cell_lines)
            P.a2_counts()   # the cell search's own row tests are not the arm's
            checked += 1
        if hit or stopped:
            break
    P.close()
    tot['candidates_streamed'] = cand
    replay.a2_rows = tuple(a2_rows)
    return lines, hit, tot, checked


@pytest.mark.parametrize('i', range(len(RUNS)))
def test_oracle_replays_reference_arm_run(i):
    run = RUNS[i]
    lines, hit, tot, checked = replay(run)
    assert lines == run['cand']
    assert hit is not None and hit['firsthit'] == run['firsthit']
    for k in 'ABCD':
        assert hit[k] == run[k]
    for k, v in tot.items():
        assert v == run['summary'][k], k
    if run['settings'].get('WZ_FH_DRAIN_BATCHES', 1) == 1:
        assert checked > 0   # at least the hit cell went through wzo_cell_search


@pytest.mark.parametrize('i', range(len(A2_RUNS)))
def test_oracle_replays_reference_a2_arm_run(i):
    run = A2_RUNS[i]
    lines, hit, tot, checked = replay(run)
    assert lines == run['cand']
    assert hit is not None and hit['firsthit'] == run['firsthit']
    for k in 'ABCD':
        assert hit[k] == run[k]
    for k, v in tot.items():
        assert v == run['summary'][k], k
    mode, tested, rejected, cuts, saved, total, under = run['a2shadow']
    assert (mode, cuts, saved, under) == (2, 0, 0, 0)
    assert replay.a2_rows == (tested, rejected)
    assert total == tot['total_AB_nodes']


def test_a2_prune_cuts_nodes_and_keeps_hits():
    """Against the unpruned arm runs of the same settings: the A2 prune removes rows whose
    residuals no completion can reach, so an uncapped arm finds the same first hit with no more
    completer nodes, and some runs need strictly fewer."""
    base = {(r['n'], tuple(r['sig']), tuple(sorted(r['settings'].items()))): r for r in RUNS}
    pairs, fewer = 0, 0
    for run in A2_RUNS:
        s = {k: v for k, v in run['settings'].items() if k != 'WZ_FH_A2'}
        key = (run['n'], tuple(run['sig']), tuple(sorted(s.items())))
        if key not in base or s.get('WZ_FH_AB_BUDGET') != 0:
            continue
        ref = base[key]
        assert run['firsthit'][:2] == ref['firsthit'][:2] and run['A'] == ref['A'] and run['B'] == ref['B']
        assert run['summary']['total_AB_nodes'] <= ref['summary']['total_AB_nodes']
        fewer += run['summary']['total_AB_nodes'] < ref['summary']['total_AB_nodes']
        pairs += 1
    assert pairs >= 25 and fewer > 0


def test_replay_detects_a_wrong_tie_order():
    """The replay is sensitive to the stable sort: completing the flattest candidates with ties
    broken in reverse stream order does not reproduce every run's CAND lines."""
    differ = 0
    for run in RUNS[:40]:
        n, s = run['n'], run['settings']
        P = OracleProblem(n, run['sig'], canon=s.get('WZ_FH_ORBIT_CANON', 0), order=s.get('WZ_FH_PROF_ORDER', 1),
                          rev=s.get('WZ_FH_STREAM_REV', 0))
        pi = run['firsthit'][1]
        size, cands = _stream_all(P, pi, s.get('WZ_FH_BUF_CAP', 0) if s.get('WZ_FH_DRAIN_TOP', 0) else 0)
        sc = np.array([flat_score(n, c[0], c[1]) for c in cands[:s.get('WZ_FH_BUF_CAP', 500000)]])
        fwd = sorted(range(len(sc)), key=lambda i: sc[i])
        bwd = sorted(range(len(sc)), key=lambda i: (sc[i], -i))
        differ += fwd != bwd
        P.close()
    assert differ > 0
