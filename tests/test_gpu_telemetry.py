"""FH_TELEM v2 of the wz_match-compatible CLI against the reference's own records.

WZ_FH_TELEMETRY=1 or 64 adds one FH_TELEM record after the summary (wz_match.cpp:901-950,
1545-1552): completion and score counts, node sums, the per-depth node histogram of the sampled
completions (:1198, 1244) and per-decile rank counters, plus timers. For the reference's own
telemetry gate cases (tools/test_firsthit_telemetry.py:94-171; tests/golden/telemetry.json) the
record's deterministic fields must equal the reference's, the result lines must be the same as
without telemetry, and every record must pass the consistency rules of the fleet's aggregator
(tools/aggregate_firsthit_telemetry.py:24-71, restated in validate() below). The resumed cases
check the replay accounting (resume_cells, replay timers) the reference's gate checks."""
import json
import os
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import load

pytestmark = pytest.mark.gpu

ENV0 = {k: v for k, v in os.environ.items() if not k.startswith(('WZ_', 'FH_'))}
SEMANTIC = ('RESULT:', 'candidates_streamed=', 'backtracks_entered=', 'A =', 'B =', 'C =', 'D =', 'VERIFY:',
            'FIRSTHIT:', 'sig =', '[firsthit ckpt]')
CASES = load('telemetry.json')
SCALARS = ('worker_ns', 'cell_ns', 'score_ns_est', 'sort_ns', 'complete_ns', 'score_calls', 'score_samples',
           'completions', 'hist_candidates', 'hist_nodes', 'total_nodes', 'unranked_nodes', 'unranked_count',
           'unranked_ns', 'cells_live_done', 'cells_partial', 'resume_replay_ns', 'resume_replay_score_ns_est',
           'resume_replay_sort_ns', 'resume_cells')


def semantic(out):
    return [re.sub(r' elapsed=.*', '', line) for line in out.splitlines() if line.startswith(SEMANTIC)]


def run_cli(n, sig, settings):
    env = {**ENV0, **{k: str(v) for k, v in settings.items()}}
    p = subprocess.run([str(wz.CLIPATH), str(n), *map(str, sig)], env=env, text=True, capture_output=True,
                       timeout=300)
    return p.returncode, p.stdout


def validate(r):
    """The aggregator's record rules (tools/aggregate_firsthit_telemetry.py:24-71)."""
    for key in ('version', 'n', 'stride') + SCALARS:
        assert type(r[key]) is int and r[key] >= 0, key
    assert r['version'] == 2 and r['n'] > 0 and r['stride'] in (1, 64)
    for key in ('depth_nodes', 'rank_nodes', 'rank_count', 'rank_ns', 'rank_aborts'):
        length = (r['n'] + 1) // 2 + 1 if key == 'depth_nodes' else 10
        assert len(r[key]) == length and all(type(v) is int and v >= 0 for v in r[key]), key
    assert sum(r['depth_nodes']) == r['hist_nodes']
    assert sum(r['rank_nodes']) + r['unranked_nodes'] == r['total_nodes']
    assert sum(r['rank_count']) + r['unranked_count'] == r['completions']
    assert sum(r['rank_ns']) + r['unranked_ns'] == r['complete_ns']
    assert all(a <= c for a, c in zip(r['rank_aborts'], r['rank_count']))
    assert r['hist_nodes'] <= r['total_nodes'] and r['hist_candidates'] <= r['completions']
    assert r['score_samples'] <= r['score_calls']
    if r['stride'] == 1:
        assert r['hist_nodes'] == r['total_nodes'] and r['hist_candidates'] == r['completions']
        assert r['score_samples'] == r['score_calls']
    assert r['worker_ns'] >= r['cell_ns'] >= r['sort_ns'] + r['complete_ns']
    replay, replay_sort, replay_score = r['resume_replay_ns'], r['resume_replay_sort_ns'], r['resume_replay_score_ns_est']
    gross = r['cell_ns'] - r['sort_ns'] - r['complete_ns']
    assert replay_sort <= r['sort_ns'] and replay_sort <= replay and replay_score <= r['score_ns_est']
    assert replay - replay_sort <= gross
    if r['stride'] == 1:
        assert r['score_ns_est'] <= gross
        assert replay_score <= replay - replay_sort
        assert gross - replay + replay_sort - r['score_ns_est'] + replay_score >= 0


def record(out):
    recs = [json.loads(l[len('FH_TELEM '):]) for l in out.splitlines() if l.startswith('FH_TELEM ')]
    assert len(recs) == 1
    return recs[0]


@pytest.mark.parametrize('case', CASES, ids=[c['name'].replace(' ', '_') for c in CASES])
def test_telemetry_record_matches_reference(case):
    n, sig, cfg = case['n'], case['sig'], case['settings']
    with tempfile.TemporaryDirectory() as td:
        ck = Path(td)
        if case['stop_after']:
            base = {k: v for k, v in cfg.items() if k != 'WZ_FH_TELEMETRY'}
            run_cli(n, sig, {**base, 'WZ_FH_CKPT_DIR': ck, 'WZ_FH_TEST_STOP_AFTER': case['stop_after']})
            p = ck / 'arm_0.ckpt'
            if case['forge']:
                t = p.read_text()
                for pat, rep in case['forge']:
                    t = re.sub(pat, rep, t)
                p.write_text(t)
            rc, out = run_cli(n, sig, {**cfg, 'WZ_FH_CKPT_DIR': ck})
        else:
            rc, out = run_cli(n, sig, cfg)
    assert rc == case['rc'], out[-800:]
    assert semantic(out) == case['lines']
    r = record(out)
    validate(r)
    assert {k: r[k] for k in case['telem']} == case['telem']
    if case['stop_after'] and not case['forge']:
        assert r['resume_cells'] == 1 and r['resume_replay_ns'] > 0
        if cfg.get('WZ_FH_CELL_ORDER', 1):
            assert r['resume_replay_sort_ns'] > 0
    if case['forge']:
        boundary = 'batch=1' in case['name']
        assert r['resume_cells'] == int(boundary) and r['resume_replay_sort_ns'] == 0
        assert (r['resume_replay_ns'] > 0) == boundary


def test_telemetry_changes_no_result_and_rejects_bad_strides():
    case = next(c for c in CASES if c['name'].startswith('case3 budget=30 stride=1'))
    off = {k: v for k, v in case['settings'].items() if k != 'WZ_FH_TELEMETRY'}
    rc, out = run_cli(case['n'], case['sig'], off)
    assert rc == case['rc'] and semantic(out) == case['lines'] and 'FH_TELEM' not in out
    rc, out = run_cli(case['n'], case['sig'], {**off, 'WZ_FH_TELEMETRY': 0})
    assert 'FH_TELEM' not in out
    rc, _ = run_cli(case['n'], case['sig'], {**off, 'WZ_FH_TELEMETRY': 7})
    assert rc == 2


def test_depth_histogram_sums_to_nodes_and_matches_split_walks():
    """wzfh_complete_depth walks each candidate whole: the same verdicts, nodes and A,B as the split
    completer, and per-candidate depth counts that sum to the nodes (budget aborts included)."""
    rec = next(c for c in load('complete_only.json') if c['n'] == 44)
    P = wz.Problem(44, rec['sig'], device=True)
    try:
        cases = [c for c in load('complete_only.json') if c['n'] == 44 and c['sig'] == rec['sig']][:24]
        cd = [(sum(1 << i for i, v in enumerate(c['C']) if v < 0), sum(1 << i for i, v in enumerate(c['D']) if v < 0))
              for c in cases]
        for budget in (2000000, 20000):
            split = P.complete(cd, budget=budget)
            whole, dep = P.complete_depth(cd, budget=budget)
            for a, b, d in zip(split, whole, dep):
                assert (a.verdict, a.nodes, a.A, a.B) == (b.verdict, b.nodes, b.A, b.B)
                assert sum(d) == b.nodes and len(d) == 23
    finally:
        P.close()
