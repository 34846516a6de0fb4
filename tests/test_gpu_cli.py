"""The wz_match-compatible CLI (bs45_quantum_explorer_amd/wzfh_match) against the reference's
own full WZ_FIRSTHIT runs: the per-completion CAND sequence (idx, sorted rank, cell, verdict,
nodes), the FIRSTHIT line, the solution and the summary counters must be identical
(tests/golden/firsthit_runs.json, firsthit_n29.json), and every FOUND must pass the exact
NPAF check. Mirrors the reference's identity gates (tools/test_firsthit_cellsize.py,
tools/test_outer_tables.py), which diff these same lines."""
import os
import re
import subprocess
import sys

import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import firsthit_runs, load, stream_runs

pytestmark = pytest.mark.gpu

CAND = re.compile(r'^CAND idx=(\d+) ci=(-?\d+) pi=(\d+) r=(\d) nodes=(\d+)', re.M)
FH = re.compile(r'^FIRSTHIT: idx=(\d+) profile_rank=(\d+) nodes_this_cand=(\d+) score=(\d+)', re.M)
KEYS = ('candidates_streamed', 'pre_filter_rejected', 'backtracks_entered', 'clean_no_AB', 'budget_aborted',
        'total_AB_nodes', 'cells_done', 'cells_prof_dead', 'cells_orbit_dup', 'cells_empty', 'cells_capped',
        'pins_dropped', 'resume_pi', 'resume_batch', 'resume_k')
ENV0 = {k: v for k, v in os.environ.items() if not k.startswith(('WZ_', 'FH_'))}


def run_cli(n, sig, settings, timeout=300, echo_progress=False):
    env = {**ENV0, **{k: str(v) for k, v in settings.items()}}
    if not echo_progress:
        p = subprocess.run([str(wz.CLIPATH), str(n), *map(str, sig)], env=env, text=True, capture_output=True,
                           timeout=timeout)
        return p.returncode, p.stdout
    # long runs: pass the arm's own progress lines through as they are printed
    env['WZ_FH_PROG_SEC'] = '20'
    p = subprocess.Popen([str(wz.CLIPATH), str(n), *map(str, sig)], env=env, text=True, stdout=subprocess.PIPE)
    lines = []
    for line in p.stdout:
        lines.append(line)
        if line.startswith('[firsthit progress]'):
            sys.__stderr__.write(line)
            sys.__stderr__.flush()
    p.wait(timeout=timeout)
    return p.returncode, ''.join(lines)


def summary(out):
    tail = out.split('=== FIRSTHIT SUMMARY')[-1]
    d = {}
    for k in KEYS:
        m = re.search(r'\b' + k + r'=(\d+)', tail)
        if m:
            d[k] = int(m[1])
    return d


RUNS = firsthit_runs()


A2_RUNS = load('a2_runs.json')


@pytest.mark.parametrize('n', sorted(set(r['n'] for r in RUNS)))
def test_cli_firsthit_runs_match_reference(n):
    _check_runs([r for r in RUNS if r['n'] == n])


@pytest.mark.parametrize('n', sorted(set(r['n'] for r in A2_RUNS)))
def test_cli_a2_runs_match_reference(n):
    """WZ_FH_A2=2: the arm with the exact profile-row reachability prune (CAND lines, node
    counts, first hit and summary of the reference's pruned arm runs)."""
    _check_runs([r for r in A2_RUNS if r['n'] == n])


def _check_runs(recs):
    for rec in recs:
        n = rec['n']
        rc, out = run_cli(n, rec['sig'], rec['settings'])
        assert rc == rec['rc'], (rec['settings'], out[-500:])
        got = [[int(g) for g in m.groups()] for m in CAND.finditer(out)]
        assert got == rec['cand'], (n, rec['sig'], rec['settings'])
        want_sum = {k: v for k, v in rec['summary'].items() if k in KEYS}
        assert summary(out) == want_sum, (n, rec['sig'], rec['settings'])
        if 'firsthit' in rec:
            m = FH.search(out)
            assert m and [int(g) for g in m.groups()] == rec['firsthit']
            for name in 'ABCD':
                line = re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)[1]
                assert [int(x) for x in line.split(',')] == rec[name]
            assert 'NPAF==0 confirmed' in out


def test_cli_cellsize_matches_reference_stream():
    rec = next(r for r in stream_runs() if r['n'] == 13 and r['canon'] == 1 and r['rev'] == 1)
    settings = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_PROF_ORDER=rec['order'],
                    WZ_FH_ORBIT_CANON=1, WZ_FH_STREAM_REV=1, WZ_FH_CELLSIZE=10 ** 12)
    rc, out = run_cli(13, rec['sig'], settings)
    assert rc == 3
    cells = re.findall(r'^CELLSIZE pi=(\d+) cand=(\d+) .* leaves=(\d+) hall_ok=(\d+)', out, re.M)
    nodes = dict(re.findall(r'^CDSTAT pi=(\d+) dfs_nodes=(\d+)', out, re.M))
    got = [[int(p), int(c), int(lv), int(h), int(nodes[p])] for p, c, lv, h in cells]
    assert got == [row[:5] for row in rec['cells']]


def test_cli_n29_known_answer():
    """The reference's n = 29 first hit (WZ_FH_M6=1, default completer): idx, cell, nodes."""
    for rec in load('firsthit_n29.json'):
        rc, out = run_cli(29, rec['sig'], rec['settings'], timeout=600, echo_progress=True)
        assert rc == rec['rc'] == 0
        m = FH.search(out)
        assert [int(g) for g in m.groups()] == rec['firsthit']
        for name in 'ABCD':
            line = re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)[1]
            assert [int(x) for x in line.split(',')] == rec[name]
        assert summary(out) == {k: v for k, v in rec['summary'].items() if k in KEYS}


def test_cli_exit_codes():
    rc, _ = run_cli(10, (1, 1, 1, 1), {'WZ_FIRSTHIT': 1})
    assert rc == 1                                     # signature norm error
    rc, out = run_cli(12, (1, 7, 0, 0), {'WZ_FIRSTHIT': 1, 'WZ_FH_SHARD': 5, 'WZ_FH_NSHARD': 2})
    assert rc == 2 and 'BAD SHARD' in out              # refusal
    rc, out = run_cli(6, (5, 1, 0, 0), {'WZ_FIRSTHIT': 1, 'WZ_FH_M6': 1, 'WZ_FH_AB_BUDGET': 0,
                                        'WZ_FH_TEST_INTERNAL_ERROR': 1})
    assert rc == 5 and 'RESULT: INTERNAL ERROR' in out  # completer/verifier disagreement
