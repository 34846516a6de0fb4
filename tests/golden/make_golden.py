#!/usr/bin/env python3
"""Writes the golden fixtures in this folder by running the reference solver.

The reference binary is oracle/_ref/wz_match, compiled by oracle/Makefile from the
reference's own single source file (src/solver/wz_match.cpp) exactly as the
reference's identity gates compile it (tools/test_cd_prune.py:41-45). It runs only
in the container that holds the reference; the fixtures it writes are plain JSON
(inputs and outputs), committed so the GPU host never needs the reference.

Fixtures:
  stream_cells.json   per-cell CELLSIZE/CDSTAT counts and a SHA-1 of each cell's
                      WZ_FH_DUMP stream lines (n = 6..13 every class, canon and
                      stream direction swept; n = 19 every class; ENDPOS cases)
  complete_only.json  WZ_FH_COMPLETE_C/_D verdicts, node counts and first A,B for
                      the six known-answer controls, sampled n = 44 candidates from
                      results/archive/n44_cands_20k.txt and sampled n = 13/19/29
                      stream candidates, including budget-abort cases
  firsthit_runs.json  full WZ_FIRSTHIT runs (stream -> sort -> complete) at n = 6..13
                      with WZ_FH_CAND_LOG: FIRSTHIT line, solution, summary counters
                      and the per-completion (idx, ci, pi, r, nodes) sequence
  firsthit_n29.json   the n = 29 known answer (WZ_FH_M6=1, default completer)
  solutions.json      the banked solution files (results/champions, results/reference)
                      and the quarantined invalid one (negative control)
  joins.json          WZ_JOIN22 runs (single thread): both sides' profile counts, the C,D
                      stream size, distinct keys, raw key hits, FOUND/EXHAUSTED and the
                      printed solution (n = 6..15 every class, n = 17 and 19 samples)
  resume.json         checkpoint / resume pairs: a run stopped by WZ_FH_TEST_STOP_AFTER, its
                      checkpoint bytes, an optional forgery of it, the resumed run's lines
                      and checkpoint bytes (the reference's own resume gates at n = 13, a
                      stale CFGSIG, n = 19 front-only interrupts)
  telemetry.json      FH_TELEM v2 records (deterministic fields: counts, node sums, depth and rank
                      histograms) of the reference's telemetry gate cases at strides 1 and 64
  a2_runs.json        full WZ_FIRSTHIT runs as firsthit_runs.json with WZ_FH_A2=2 (the exact
                      profile-row reachability prune), with the arm's A2SHADOW row counts
                      (n = 6..13 every class, n = 19 every class)
  qprune_cells.json   the Q-closure prune (WZ_FH_ORBIT_QPRUNE with the 64-element orbit
                      group): [qprune] counts, kept orbits, order-3 digests and the per-cell
                      streams of the cells left (n = 10..14 safe classes, n = 11, 13 ENDPOS)

Usage: python tests/golden/make_golden.py [--only NAME] [--jobs 8]
"""
import argparse
import concurrent.futures as cf
import hashlib
import itertools
import json
import os
import re
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
REF = REPO / 'oracle' / '_ref' / 'wz_match'
REFTREE = Path('/root/reference')
ENV0 = {k: v for k, v in os.environ.items() if not k.startswith(('WZ_', 'FH_'))}


def classes(n):
    """Signature classes a<=b, c<=d, nonnegative, parity-correct, norm 4n+2."""
    out = []
    for a, b, c, d in itertools.product(range(n + 2), range(n + 2), range(n + 1), range(n + 1)):
        if a > b or c > d:
            continue
        if (a - (n + 1)) % 2 or (b - (n + 1)) % 2 or (c - n) % 2 or (d - n) % 2:
            continue
        if a * a + b * b + c * c + d * d == 4 * n + 2:
            out.append((a, b, c, d))
    return out


def run(n, sig, settings, timeout=3600):
    env = {**ENV0, **{k: str(v) for k, v in settings.items()}}
    p = subprocess.run([str(REF), str(n), *map(str, sig)], env=env, text=True,
                       capture_output=True, timeout=timeout)
    return p.returncode, p.stdout


CELL = re.compile(r'^CELLSIZE pi=(\d+) cand=(\d+) capped=([01]) partial=([01]) sec=\S+ '
                  r'sec_at_buf=\S+ leaves=(\d+) hall_ok=(\d+)$', re.M)
CDSTAT = re.compile(r'^CDSTAT pi=(\d+) dfs_nodes=(\d+)', re.M)
SUMMARY_KEYS = ('candidates_streamed', 'pre_filter_rejected', 'backtracks_entered', 'clean_no_AB',
                'budget_aborted', 'total_AB_nodes', 'cells_done', 'cells_prof_dead', 'cells_orbit_dup',
                'cells_empty', 'cells_capped', 'pins_dropped', 'resume_pi', 'resume_batch', 'resume_k')


def summary(out):
    tail = out.split('=== FIRSTHIT SUMMARY')[-1]
    d = {}
    for k in SUMMARY_KEYS:
        m = re.search(r'\b' + k + r'=(\d+)', tail)
        if m:
            d[k] = int(m[1])
    m = re.search(r'^RESULT: (.*)$', tail, re.M)
    d['result'] = m[1] if m else None
    return d


def stream_case(n, sig, canon, rev, endpos=0, order=1):
    cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_PROF_ORDER=order,
               WZ_FH_ORBIT_CANON=canon, WZ_FH_STREAM_REV=rev, WZ_FH_PROG_SEC=99999,
               WZ_FH_CELLSIZE=10 ** 12)
    if endpos:
        cfg['WZ_FH_ENDPOS'] = 1
    with tempfile.TemporaryDirectory() as td:
        dump = Path(td) / 'dump.txt'
        rc, out = run(n, sig, {**cfg, 'WZ_FH_DUMP': dump})
        lines = dump.read_text().splitlines(keepends=True)
    assert rc in (0, 3), (n, sig, rc, out[-400:])
    cells = [tuple(int(g) for g in m.groups()) for m in CELL.finditer(out)]
    nodes = {int(m[1]): int(m[2]) for m in CDSTAT.finditer(out)}
    rows, pos = [], 0
    for pi, cand, capped, partial, leaves, hok in cells:
        h = hashlib.sha1(''.join(lines[pos:pos + cand]).encode()).hexdigest()
        pos += cand
        rows.append([pi, cand, leaves, hok, nodes[pi], h])
    assert pos == len(lines)
    m = re.search(r'\[firsthit\] profiles=(\d+)', out)
    kept = re.search(r'kept_orbits=(\d+)', out)
    rec = dict(n=n, sig=list(sig), canon=canon, rev=rev, endpos=endpos, order=order,
               profiles=int(m[1]), kept=int(kept[1]) if kept else None, cells=rows,
               summary=summary(out))
    if n <= 8:   # tiny: keep the literal stream for debugging
        rec['dump'] = ''.join(lines)
    return rec


def gen_stream(jobs):
    tasks = []
    for n in range(6, 14):
        for sig in classes(n):
            for canon, rev in itertools.product((0, 1), (0, 1)):
                tasks.append((n, sig, canon, rev, 0, 1))
    for sig in classes(19):
        tasks.append((19, sig, 1, 0, 0, 1))
    tasks.append((19, classes(19)[0], 1, 1, 0, 1))
    tasks.append((19, classes(19)[0], 0, 0, 0, 0))
    for n in (12, 13):
        for sig in classes(n)[:3]:
            tasks.append((n, sig, 0, 0, 1, 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: stream_case(*t), tasks))
    (HERE / 'stream_cells.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'stream_cells.json: {len(recs)} runs, {sum(len(r["cells"]) for r in recs)} cells, '
          f'{sum(r["summary"]["candidates_streamed"] for r in recs)} candidates')


def read_solution(path):
    rows = [l.split() for l in path.read_text().splitlines() if l.strip() and not l.startswith('#')]
    n = int(rows[0][0])
    return n, [int(x) for x in rows[1][:4]], [[int(x) for x in r] for r in rows[2:6]]


def complete_case(n, sig, C, D, budget, outer_k=3, m6=1):
    cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=m6, WZ_FH_AB_BUDGET=budget, WZ_FH_OUTER_K=outer_k,
               WZ_FH_COMPLETE_C=','.join(map(str, C)), WZ_FH_COMPLETE_D=','.join(map(str, D)))
    rc, out = run(n, sig, cfg)
    m = re.search(r'^COMPLETE_ONLY r=(\d) nodes=(\d+)', out, re.M)
    assert m, (n, sig, out[-300:])
    rec = dict(n=n, sig=list(sig), C=C, D=D, budget=budget, outer_k=outer_k, r=int(m[1]), nodes=int(m[2]))
    if rec['r'] == 0:
        rec['A'] = [int(x) for x in re.search(r'^A = \{([^}]*)\}', out, re.M)[1].split(',')]
        rec['B'] = [int(x) for x in re.search(r'^B = \{([^}]*)\}', out, re.M)[1].split(',')]
    return rec


def stream_candidates(n, sig, cap, end=400):
    cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_PROF_ORDER=1,
               WZ_FH_ORBIT_CANON=1, WZ_FH_PROG_SEC=99999, WZ_FH_CELLSIZE=cap, WZ_FH_PROF_END=end)
    with tempfile.TemporaryDirectory() as td:
        dump = Path(td) / 'dump.txt'
        run(n, sig, {**cfg, 'WZ_FH_DUMP': dump})
        out = []
        for line in dump.read_text().splitlines():
            c, d = line.split('|')
            out.append(([int(x) for x in c.split()], [int(x) for x in d.split()]))
    return out


def gen_complete(jobs):
    tasks = []
    six = {'ours-41': 'results/champions/champion_firsthit_bs42_41.txt',
           'ours-42': 'results/champions/champion_firsthit_bs43_42.txt',
           'ours-43': 'results/champions/champion_firsthit_bs44_43.txt',
           'WZ-41': 'results/reference/wz_table1_bs42_41.txt',
           'WZ-42': 'results/reference/wz_table1_bs43_42.txt',
           'WZ-43': 'results/reference/wz_table1_bs44_43.txt'}
    for name, f in six.items():
        n, sig, (A, B, C, D) = read_solution(REFTREE / f)
        asig = [abs(x) for x in sig]
        for k in (0, 3):
            tasks.append(('six:' + name, n, asig, C, D, 0, k))
    # n = 44 production-class candidates (class (1,7,8,8)), budget 2e6 and small abort budgets
    lines = (REFTREE / 'results/archive/n44_cands_20k.txt').read_text().splitlines()
    picks = list(range(0, 24)) + list(range(1000, 20000, 1000))
    for i in picks:
        c, d = lines[i].split('|')
        C, D = [int(x) for x in c.split()], [int(x) for x in d.split()]
        tasks.append((f'n44:{i}', 44, [1, 7, 8, 8], C, D, 2000000, 3))
    for i in range(0, 8):
        c, d = lines[i].split('|')
        C, D = [int(x) for x in c.split()], [int(x) for x in d.split()]
        tasks.append((f'n44abort:{i}', 44, [1, 7, 8, 8], C, D, 977, 3))
        tasks.append((f'n44k0:{i}', 44, [1, 7, 8, 8], C, D, 5000, 0))
    # smaller-n stream candidates, exact and capped budgets
    for n, sig, take in ((13, (2, 4, 3, 5), 40), (19, (0, 2, 5, 7), 40), (29, (0, 6, 1, 9), 24)):
        cands = stream_candidates(n, sig, 4)
        step = max(1, len(cands) // take)
        for j, (C, D) in enumerate(cands[::step][:take]):
            for budget in (0, 300):
                tasks.append((f'n{n}:{j}', n, list(sig), C, D, budget, 3))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(complete_case, *t[1:]) for t in tasks]
        recs = []
        for t, f in zip(tasks, futs):
            r = f.result()
            r['name'] = t[0]
            recs.append(r)
    (HERE / 'complete_only.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'complete_only.json: {len(recs)} completions, verdicts '
          f'{sorted(set((r["r"]) for r in recs))}')


CANDLINE = re.compile(r'^CAND idx=(\d+) ci=(-?\d+) pi=(\d+) r=(\d) nodes=(\d+)', re.M)
FHLINE = re.compile(r'^FIRSTHIT: idx=(\d+) profile_rank=(\d+) nodes_this_cand=(\d+) score=(\d+)', re.M)


def firsthit_case(n, sig, extra, outs=None):
    cfg = {**dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_PROF_ORDER=1,
                  WZ_FH_ORBIT_CANON=1, WZ_FH_PROG_SEC=99999, WZ_FH_CAND_LOG=1), **extra}
    rc, out = run(n, sig, cfg)
    if outs is not None:
        outs.append(out)
    rec = dict(n=n, sig=list(sig), settings={k: v for k, v in cfg.items() if k != 'WZ_FH_PROG_SEC'},
               rc=rc, summary=summary(out),
               cand=[[int(g) for g in m.groups()] for m in CANDLINE.finditer(out)])
    m = FHLINE.search(out)
    if m:
        rec['firsthit'] = [int(g) for g in m.groups()]
        for name in 'ABCD':
            rec[name] = [int(x) for x in re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)[1].split(',')]
    return rec


def gen_firsthit(jobs):
    tasks = []
    for n in range(6, 14):
        for sig in classes(n):
            for budget in (0, 30):
                tasks.append((n, sig, {'WZ_FH_AB_BUDGET': budget}))
            tasks.append((n, sig, {'WZ_FH_AB_BUDGET': 40, 'WZ_FH_BUF_CAP': 8, 'WZ_FH_DRAIN_TOP': 3,
                                   'WZ_FH_DRAIN_BATCHES': 2}))
            tasks.append((n, sig, {'WZ_FH_AB_BUDGET': 40, 'WZ_FH_STREAM_REV': 1, 'WZ_FH_ORBIT_CANON': 0}))
    for sig in classes(19):
        tasks.append((19, sig, {'WZ_FH_AB_BUDGET': 500, 'WZ_FH_BUF_CAP': 64, 'WZ_FH_DRAIN_TOP': 8,
                                'WZ_FH_MAX_CAND': 20000}))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: firsthit_case(*t), tasks))
    (HERE / 'firsthit_runs.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'firsthit_runs.json: {len(recs)} runs, {sum(len(r["cand"]) for r in recs)} completions, '
          f'{sum(1 for r in recs if "firsthit" in r)} hits')


A2LINE = re.compile(r'^A2SHADOW mode=(\d+) rows_tested=(\d+) rows_rejected=(\d+) cuts=(\d+) saved_nodes=(\d+) '
                    r'total_nodes=(\d+) hit_under_cut=(\d+)', re.M)


def a2_case(n, sig, extra):
    outs = []
    rec = firsthit_case(n, sig, {**extra, 'WZ_FH_A2': 2}, outs)
    m = A2LINE.findall(outs[0])
    assert len(m) == 1, out[-400:]
    rec['a2shadow'] = [int(g) for g in m[0]]
    return rec


def gen_a2(jobs):
    tasks = []
    for n in range(6, 14):
        for sig in classes(n):
            for budget in (0, 30):
                tasks.append((n, sig, {'WZ_FH_AB_BUDGET': budget}))
            tasks.append((n, sig, {'WZ_FH_AB_BUDGET': 40, 'WZ_FH_BUF_CAP': 8, 'WZ_FH_DRAIN_TOP': 3,
                                   'WZ_FH_DRAIN_BATCHES': 2}))
    for sig in classes(19):
        tasks.append((19, sig, {'WZ_FH_AB_BUDGET': 500, 'WZ_FH_BUF_CAP': 64, 'WZ_FH_DRAIN_TOP': 8,
                                'WZ_FH_MAX_CAND': 20000}))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: a2_case(*t), tasks))
    (HERE / 'a2_runs.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'a2_runs.json: {len(recs)} runs, {sum(len(r["cand"]) for r in recs)} completions, '
          f'{sum(1 for r in recs if "firsthit" in r)} hits, {sum(r["a2shadow"][2] for r in recs)} rows rejected')


def gen_n29(jobs):
    recs = []
    for k in (3, 0):
        cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_FH_OUTER_K=k, WZ_FH_PROG_SEC=99999)
        rc, out = run(29, (0, 6, 1, 9), cfg)
        m = FHLINE.search(out)
        rec = dict(n=29, sig=[0, 6, 1, 9], settings=cfg, rc=rc, summary=summary(out),
                   firsthit=[int(g) for g in m.groups()] if m else None)
        for name in 'ABCD':
            mm = re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)
            if mm:
                rec[name] = [int(x) for x in mm[1].split(',')]
        recs.append(rec)
    (HERE / 'firsthit_n29.json').write_text(json.dumps(recs, separators=(',', ':')))
    print('firsthit_n29.json:', [r['firsthit'] for r in recs])


def gen_solutions(jobs):
    """The reference's banked solution files (data: n, signature, A, B, C, D), plus the
    quarantined known-invalid one as a negative control."""
    recs = []
    for sub in ('results/champions', 'results/reference', 'results/quarantine'):
        for f in sorted((REFTREE / sub).glob('*.txt')):
            rows = [l.split() for l in f.read_text().splitlines() if l.strip() and not l.startswith('#')]
            try:
                n = int(rows[0][0])
                seqs = [[int(x) for x in r] for r in rows[2:6]]
            except (ValueError, IndexError):
                continue
            recs.append(dict(file=f'{sub}/{f.name}', n=n, sig=[int(x) for x in rows[1][:4]], A=seqs[0], B=seqs[1],
                             C=seqs[2], D=seqs[3], quarantined=sub.endswith('quarantine')))
    (HERE / 'solutions.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'solutions.json: {len(recs)} files')


JOIN_PROF = re.compile(r'^\[profiles\] A,B side surviving: (\d+)\s+C,D side surviving: (\d+)', re.M)
JOIN_BUILD = re.compile(r'^\[join22v2\] C,D stream=([0-9.e+]+)\s+distinct keys=(\d+)', re.M)
JOIN_HITS = re.compile(r'^\[join22v2\] phase 2 done: (\d+) raw key hits', re.M)


def join_case(n, sig, settings):
    cfg = {**dict(WZ_JOIN22=1, WZ_JOIN22_SLOTS_LOG2=22, OMP_NUM_THREADS=1), **settings}
    rc, out = run(n, sig, cfg)
    prof, build, hits = JOIN_PROF.search(out), JOIN_BUILD.search(out), JOIN_HITS.search(out)
    if not prof or not build:
        raise RuntimeError(f'join run n={n} sig={sig} {settings} rc={rc}: {out[-600:]}')
    rec = dict(n=n, sig=list(sig), settings=cfg, rc=rc, ab_profiles=int(prof[1]), cd_profiles=int(prof[2]),
               cd_stream=int(float(build[1])), cd_stream_exact='e' not in build[1],   # cout prints a double
               distinct_keys=int(build[2]), hits=int(hits[1]) if hits else 0,
               found='FOUND ***' in out, exhausted='JOIN22 EXHAUSTED' in out)
    for name in 'ABCD':
        mm = re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)
        if mm:
            rec[name] = [int(x) for x in mm[1].split(',')]
    return rec


def gen_join(jobs):
    tasks = []
    for n in range(6, 16):
        for sig in classes(n):
            tasks.append((n, sig, {}))
            tasks.append((n, sig, {'WZ_THM211B': 1, 'WZ_THM212': 1}))
    for n in (17, 19):
        for sig in classes(n)[:3]:
            tasks.append((n, sig, {'WZ_THM211B': 1, 'WZ_THM212': 1}))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: join_case(*t), tasks))
    (HERE / 'joins.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'joins.json: {len(recs)} runs, {sum(r["found"] for r in recs)} found, '
          f'{sum(r["cd_stream"] for r in recs)} C,D candidates, {sum(r["hits"] for r in recs)} hits')


def signed_variants(sig):
    """Every signing of a class representative (a, b, c, d) that keeps its norm: the join
    searches one SIGNED signature, and a negative is "no 2.2-normalized solution with this
    signed sig" (wz_match.cpp:2900-2904)."""
    out = []
    for mask in range(16):
        v = tuple(-x if (mask >> i) & 1 else x for i, x in enumerate(sig))
        if v not in out:
            out.append(v)
    return out


def gen_join_signed(jobs):
    tasks = [(n, v, {}) for n in range(6, 13) for sig in classes(n) for v in signed_variants(sig)[1:]]
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: join_case(*t), tasks))
    (HERE / 'joins_signed.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'joins_signed.json: {len(recs)} runs, {sum(r["found"] for r in recs)} found, '
          f'{sum(r["exhausted"] for r in recs)} exhausted')


SEMANTIC = ('RESULT:', 'candidates_streamed=', 'backtracks_entered=', 'A =', 'B =', 'C =', 'D =', 'VERIFY:',
            'FIRSTHIT:', 'sig =', '[firsthit ckpt]')


def semantic(out):
    """The lines a resume gate compares (tools/test_firsthit_telemetry.py:26-32), plus the
    checkpoint RESUME / STALE lines; timing fields dropped."""
    keep = []
    for line in out.splitlines():
        if line.startswith(SEMANTIC):
            keep.append(re.sub(r' elapsed=.*', '', line))
    return keep


def resume_case(name, n, sig, cfg, stop_after, forge=None):
    """One interrupt/resume pair in a fresh checkpoint directory: a run stopped by
    WZ_FH_TEST_STOP_AFTER (a SIGTERM after exactly that many completions, wz_match.cpp:2608-2612),
    its checkpoint bytes, an optional textual forgery of that checkpoint, then the resumed run
    and its checkpoint bytes. With stop_after = 0 the first run is uninterrupted."""
    with tempfile.TemporaryDirectory() as td:
        ck = Path(td)
        base = {**cfg, 'WZ_FH_CKPT_DIR': ck}
        first = {**base, 'WZ_FH_TEST_STOP_AFTER': stop_after} if stop_after else base
        rc1, out1 = run(n, sig, first)
        p = ck / 'arm_0.ckpt'
        ck1 = p.read_text() if p.exists() else None
        forged = None
        if forge and ck1 is not None:
            forged = ck1
            for pat, rep in forge:
                forged = re.sub(pat, rep, forged)
            p.write_text(forged)
        rc2, out2 = run(n, sig, base)
        ck2 = p.read_text() if p.exists() else None
    return dict(name=name, n=n, sig=list(sig), settings={k: v for k, v in cfg.items()}, stop_after=stop_after,
                forge=forge, first=dict(rc=rc1, lines=semantic(out1), ckpt=ck1), forged=forged,
                resumed=dict(rc=rc2, lines=semantic(out2), ckpt=ck2))


def gen_resume(jobs):
    """Checkpoint / resume fixtures: the reference's own resume gates at n = 13
    (tools/test_firsthit_telemetry.py:129-171 mid-drain interrupts and buffer-boundary forgeries,
    tools/test_resume_boundary.py:26-56 front-only forgeries at batch == DRAIN_BATCHES), a stale
    CFGSIG, and n = 19 front-only interrupts with the default completer."""
    sig13 = (2, 4, 3, 5)
    common = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_FH_AB_PROF=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_ORBIT_CANON=1,
                  WZ_FH_PROF_ORDER=1, WZ_FH_BUF_CAP=16, WZ_FH_DRAIN_TOP=5, WZ_FH_MAX_CAND=256,
                  WZ_FH_PROG_SEC=9999, WZ_FH_OUTER_K=0)
    tasks = []
    for rev, buffered, stop in ((0, 1, 3), (1, 1, 3), (0, 0, 3), (1, 0, 7), (0, 1, 7), (1, 1, 10)):
        cfg = {**common, 'WZ_FH_AB_BUDGET': 1, 'WZ_FH_STREAM_REV': rev, 'WZ_FH_CELL_ORDER': buffered}
        tasks.append((f'interrupt rev={rev} buffered={buffered} stop={stop}', 13, sig13, cfg, stop, None))
    for boundary in (0, 1):
        cfg = {**common, 'WZ_FH_AB_BUDGET': 1, 'WZ_FH_BUF_CAP': 2, 'WZ_FH_DRAIN_TOP': 0}
        tasks.append((f'buffer boundary batch={boundary}', 13, sig13, cfg, 2,
                      [(r'resume_batch=0\nresume_k=2\n', f'resume_batch={boundary}\nresume_k=0\n')]))
    front = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_FH_AB_PROF=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_ORBIT_CANON=1,
                 WZ_FH_PROF_ORDER=1, WZ_FH_AB_BUDGET=1, WZ_FH_PROG_SEC=9999)
    for buf, top, batches in ((4, 2, 1), (4, 2, 2), (6, 3, 1), (6, 3, 3), (8, 5, 2)):
        cfg = {**front, 'WZ_FH_BUF_CAP': buf, 'WZ_FH_DRAIN_TOP': top, 'WZ_FH_DRAIN_BATCHES': batches}
        tasks.append((f'front boundary buf={buf} top={top} batches={batches}', 13, sig13, cfg, 1,
                      [(r'resume_batch=\d+', f'resume_batch={batches}'), (r'resume_k=\d+', 'resume_k=0')]))
        tasks.append((f'front uninterrupted buf={buf} top={top} batches={batches}', 13, sig13, cfg, 0, None))
    cfg = {**common, 'WZ_FH_AB_BUDGET': 30}
    tasks.append(('stale cfgsig', 13, sig13, cfg, 4, [(r'CFGSIG=n13\.', 'CFGSIG=n13.x')]))
    tasks.append(('resume past a hit', 13, sig13, {**common, 'WZ_FH_AB_BUDGET': 0}, 0, None))
    for sig in classes(19)[:2]:
        cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_ORBIT_CANON=1, WZ_FH_PROF_ORDER=1,
                   WZ_FH_AB_BUDGET=12, WZ_FH_BUF_CAP=64, WZ_FH_DRAIN_TOP=8, WZ_FH_MAX_CAND=3000,
                   WZ_FH_PROG_SEC=9999)
        for stop in (5, 37):
            tasks.append((f'n19 front stop={stop}', 19, sig, cfg, stop, None))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: resume_case(*t), tasks))
    (HERE / 'resume.json').write_text(json.dumps(recs, indent=0))
    print(f'resume.json: {len(recs)} interrupt/resume pairs, '
          f'{sum(1 for r in recs if "RESUME" in " ".join(r["resumed"]["lines"]))} resumed from a checkpoint')


QPRUNE = re.compile(r'^\[qprune\] missing_Q_images=(\d+) not_realizable=(\d+) eq212=(\d+) nonintegral=(\d+) '
                    r'unknown=(\d+) dead_orbits=(\d+) cells_removed=(\d+)', re.M)
ORD3 = re.compile(r'^\[order\] ord3 cells=(\d+) kept=(\d+) .*digest=(\S+)', re.M)


def qprune_case(n, sig, endpos, order):
    """The Q-closure prune (WZ_FH_ORBIT_QPRUNE=1 with the 64-element group): the reference's
    [qprune] counts, the order-3 digest when order = 3, and the per-cell stream of every cell
    the arm streams (CELLSIZE/CDSTAT counts and DUMP SHA-1, as stream_cells.json)."""
    cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_PROF_ORDER=order,
               WZ_FH_ORBIT_CANON=1, WZ_FH_ORBIT_Q=1, WZ_FH_ORBIT_QPRUNE=1, WZ_FH_PROG_SEC=99999,
               WZ_FH_CELLSIZE=10 ** 12)
    if endpos:
        cfg['WZ_FH_ENDPOS'] = 1
    with tempfile.TemporaryDirectory() as td:
        dump = Path(td) / 'dump.txt'
        rc, out = run(n, sig, {**cfg, 'WZ_FH_DUMP': dump})
        lines = dump.read_text().splitlines(keepends=True)
    assert rc in (0, 3), (n, sig, rc, out[-400:])
    q = QPRUNE.search(out)
    assert q, out[-600:]
    cells = [tuple(int(g) for g in m.groups()) for m in CELL.finditer(out)]
    nodes = {int(m[1]): int(m[2]) for m in CDSTAT.finditer(out)}
    rows, pos = [], 0
    for pi, cand, capped, partial, leaves, hok in cells:
        h = hashlib.sha1(''.join(lines[pos:pos + cand]).encode()).hexdigest()
        pos += cand
        rows.append([pi, cand, leaves, hok, nodes[pi], h])
    m = re.search(r'\[firsthit\] profiles=(\d+)', out)
    kept = re.search(r'kept_orbits=(\d+)', out)
    o3 = ORD3.search(out)
    return dict(n=n, sig=list(sig), endpos=endpos, order=order, profiles=int(m[1]), kept=int(kept[1]),
                qprune=[int(g) for g in q.groups()], digest=o3[3] if o3 else None, cells=rows)


def gen_qprune(jobs):
    tasks = []
    for n in (10, 12, 14):
        for sig in classes(n):
            if (sig[2] + sig[3]) % 4 == 0:
                tasks.append((n, sig, 0, 1))
    for n in (11, 13):
        for sig in classes(n):
            tasks.append((n, sig, 1, 1))
    for n in (12, 16):
        for sig in classes(n)[:3]:
            if (sig[2] + sig[3]) % 4 == 0:
                tasks.append((n, sig, 0, 3))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: qprune_case(*t), tasks))
    (HERE / 'qprune_cells.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'qprune_cells.json: {len(recs)} runs, {sum(len(r["cells"]) for r in recs)} streamed cells, '
          f'{sum(r["qprune"][6] for r in recs)} cells removed by the prune')


TELEM_KEYS = ('version', 'n', 'stride', 'resume_cells', 'score_calls', 'score_samples', 'completions',
              'hist_candidates', 'hist_nodes', 'total_nodes', 'unranked_nodes', 'unranked_count', 'cells_live_done',
              'cells_partial', 'depth_nodes', 'rank_nodes', 'rank_count', 'rank_aborts')


def telem_record(out):
    recs = [json.loads(l[len('FH_TELEM '):]) for l in out.splitlines() if l.startswith('FH_TELEM ')]
    assert len(recs) == 1, out[-500:]
    return {k: recs[0][k] for k in TELEM_KEYS}


def telem_case(name, n, sig, cfg, stop_after=0, forge=None):
    """A WZ_FH_TELEMETRY run (or an interrupt/resume pair whose resumed run has telemetry on):
    the record's deterministic fields (counts, node sums, the depth and rank histograms; not the
    timers) and the result lines."""
    if stop_after:
        base = {k: v for k, v in cfg.items() if k != 'WZ_FH_TELEMETRY'}
        with tempfile.TemporaryDirectory() as td:
            ck = Path(td)
            run(n, sig, {**base, 'WZ_FH_CKPT_DIR': ck, 'WZ_FH_TEST_STOP_AFTER': stop_after})
            p = ck / 'arm_0.ckpt'
            if forge:
                t = p.read_text()
                for pat, rep in forge:
                    t = re.sub(pat, rep, t)
                p.write_text(t)
            rc, out = run(n, sig, {**cfg, 'WZ_FH_CKPT_DIR': ck})
    else:
        rc, out = run(n, sig, cfg)
    return dict(name=name, n=n, sig=list(sig), settings=cfg, stop_after=stop_after, forge=forge, rc=rc,
                lines=semantic(out), telem=telem_record(out))


def gen_telem(jobs):
    """FH_TELEM v2 fixtures: the reference's own telemetry gate cases
    (tools/test_firsthit_telemetry.py:94-171) at strides 1 and 64, with and without resume."""
    common = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_FH_AB_PROF=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_ORBIT_CANON=1,
                  WZ_FH_PROF_ORDER=1, WZ_FH_BUF_CAP=16, WZ_FH_DRAIN_TOP=5, WZ_FH_MAX_CAND=256,
                  WZ_FH_PROG_SEC=9999, WZ_FH_OUTER_K=0)
    cases = [((6, 5, 1, 0, 0), {}), ((11, 0, 6, 1, 3), {}), ((12, 1, 7, 0, 0), {}), ((13, 2, 4, 3, 5), {}),
             ((13, 2, 4, 3, 5), {'WZ_FH_CELL_ORDER': 0}), ((11, 0, 6, 1, 3), {'WZ_FH_PROF_END': 2}),
             ((13, 2, 4, 3, 5), {'WZ_FH_DRAIN_BATCHES': 2})]
    tasks = []
    for ci, (args, extra) in enumerate(cases):
        for budget in (0, 1, 30):
            for mode in (1, 64):
                cfg = {**common, **extra, 'WZ_FH_AB_BUDGET': budget, 'WZ_FH_STREAM_REV': ci % 2,
                       'WZ_FH_TELEMETRY': mode}
                tasks.append((f'case{ci} budget={budget} stride={mode}', args[0], args[1:], cfg, 0, None))
    for rev, buffered, stop in ((0, 1, 3), (1, 1, 3), (0, 0, 3), (1, 0, 7), (0, 1, 7), (1, 1, 10)):
        cfg = {**common, 'WZ_FH_AB_BUDGET': 1, 'WZ_FH_STREAM_REV': rev, 'WZ_FH_CELL_ORDER': buffered,
               'WZ_FH_TELEMETRY': 1}
        tasks.append((f'resume rev={rev} buffered={buffered} stop={stop}', 13, (2, 4, 3, 5), cfg, stop, None))
    for boundary in (0, 1):
        cfg = {**common, 'WZ_FH_AB_BUDGET': 1, 'WZ_FH_BUF_CAP': 2, 'WZ_FH_DRAIN_TOP': 0, 'WZ_FH_TELEMETRY': 1}
        tasks.append((f'boundary batch={boundary}', 13, (2, 4, 3, 5), cfg, 2,
                      [(r'resume_batch=0\nresume_k=2\n', f'resume_batch={boundary}\nresume_k=0\n')]))
    cfg = dict(WZ_FIRSTHIT=1, WZ_FH_M6=1, WZ_THM211B=1, WZ_THM212=1, WZ_FH_ORBIT_CANON=1, WZ_FH_PROF_ORDER=1,
               WZ_FH_AB_BUDGET=12, WZ_FH_BUF_CAP=64, WZ_FH_DRAIN_TOP=8, WZ_FH_MAX_CAND=3000,
               WZ_FH_PROG_SEC=9999)
    for sig in classes(19)[:2]:
        for mode in (1, 64):
            tasks.append((f'n19 stride={mode}', 19, sig, {**cfg, 'WZ_FH_TELEMETRY': mode}, 0, None))
    with cf.ThreadPoolExecutor(jobs) as ex:
        recs = list(ex.map(lambda t: telem_case(*t), tasks))
    (HERE / 'telemetry.json').write_text(json.dumps(recs, separators=(',', ':')))
    print(f'telemetry.json: {len(recs)} runs, {sum(r["telem"]["completions"] for r in recs)} completions')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--only', default='')
    ap.add_argument('--jobs', type=int, default=8)
    a = ap.parse_args()
    if not REF.exists():
        sys.exit(f'{REF} missing: run make -C oracle ref (needs the reference tree)')
    gens = dict(stream=gen_stream, complete=gen_complete, firsthit=gen_firsthit, n29=gen_n29,
                solutions=gen_solutions, join=gen_join, join_signed=gen_join_signed, resume=gen_resume,
                qprune=gen_qprune, telem=gen_telem, a2=gen_a2)
    for name, g in gens.items():
        if not a.only or name in a.only.split(','):
            g(a.jobs)


if __name__ == '__main__':
    main()
