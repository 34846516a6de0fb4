"""WZ_JOIN22 hash join on the GPU against the reference's own JOIN22 runs (tests/golden/joins.json,
single-threaded reference): both sides' profile counts, the C,D stream size, distinct keys, raw key
hits, FOUND / EXHAUSTED and the printed solution, through the library and through the CLI."""
import os
import re
import subprocess
import time

import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import load

pytestmark = pytest.mark.gpu

RUNS = load('joins.json')


def unpack(w, L):
    return [-1 if (int(w) >> i) & 1 else 1 for i in range(L)]


def counts(r):
    return dict(ab=r['ab_profiles'], cd=r['cd_profiles'], keys=r['distinct_keys'], hits=r['hits'], found=r['found'])


def stream_ok(got, want, exact):
    return got == want if exact else abs(got - want) <= 1e-5 * want   # the reference prints a double


@pytest.mark.parametrize('n', sorted({r['n'] for r in RUNS}))
def test_join_matches_reference(n):
    for r in (x for x in RUNS if x['n'] == n):
        st = r['settings']
        t0 = time.time()
        J = wz.Join(n, r['sig'], thm211b=int(st.get('WZ_THM211B', 0)), thm212=int(st.get('WZ_THM212', 0)))
        try:
            J.build()
            p = J.probe()
            got = dict(ab=p.ab_profiles, cd=p.cd_profiles, keys=p.distinct_keys, hits=p.hits, found=bool(p.found))
            assert got == counts(r), (n, r['sig'], st)
            assert stream_ok(p.cd_stream, r['cd_stream'], r['cd_stream_exact']), (n, r['sig'], p.cd_stream)
            if r['found']:
                sol = [unpack(p.A, n + 1), unpack(p.B, n + 1), unpack(p.C, n), unpack(p.D, n)]
                assert sol == [r['A'], r['B'], r['C'], r['D']], (n, r['sig'], st)
                assert wz.npaf_max(n, *sol) == 0
        finally:
            J.close()
        print(f'join n={n} sig={r["sig"]} {st.get("WZ_THM211B", 0)}: {time.time() - t0:.2f}s', flush=True)


def test_join_cli_lines_match_reference():
    cli = str(wz.CLIPATH)
    env0 = {k: v for k, v in os.environ.items() if not k.startswith(('WZ_', 'FH_'))}
    for r in [x for x in RUNS if x['n'] in (9, 13)][:6]:
        env = {**env0, **{k: str(v) for k, v in r['settings'].items()}}
        p = subprocess.run([cli, str(r['n']), *map(str, r['sig'])], env=env, text=True, capture_output=True, timeout=300)
        out = p.stdout
        assert p.returncode == (0 if r['found'] else 3), out[-400:]
        m = re.search(r'^\[profiles\] A,B side surviving: (\d+)\s+C,D side surviving: (\d+)', out, re.M)
        assert [int(m[1]), int(m[2])] == [r['ab_profiles'], r['cd_profiles']]
        m = re.search(r'^\[join22v2\] C,D stream=([0-9.e+]+)\s+distinct keys=(\d+)', out, re.M)
        assert stream_ok(int(float(m[1])), r['cd_stream'], r['cd_stream_exact']) and int(m[2]) == r['distinct_keys']
        assert int(re.search(r'phase 2 done: (\d+) raw key hits', out)[1]) == r['hits']
        if r['found']:
            assert out.count('FOUND ***') == 2 and 'NPAF==0 confirmed' in out
            for name in 'ABCD':
                line = re.search(r'^' + name + r' = \{([^}]*)\}', out, re.M)[1]
                assert [int(x) for x in line.split(',')] == r[name]
        else:
            assert 'JOIN22 EXHAUSTED' in out
