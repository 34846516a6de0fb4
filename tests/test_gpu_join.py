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


def test_join_table_at_2_28_slots_finds_exactly_the_planted_solutions():
    """The n = 44 join at the configuration's table size: 2^28 slots (44 int16 vector, tag, order and
    C,D per slot, about 32 GB of HBM) filled to load 1/2 with random C,D pairs, probed with as many
    random A,B pairs, with the images of the banked BS(45,44) planted on both sides. Distinct keys
    must be the random pairs plus the one planted vector, the hits exactly the planted A,B images,
    and the solution list exactly every planted A,B image with every planted C,D image, each with
    NPAF = 0."""
    from bs45_quantum_explorer_amd.workload import join_workload
    sol = next(r for r in load('solutions.json') if r['n'] == 44 and not r['quarantined'])
    nrand = (1 << 27) - 8
    cd, ab, pcd, pab = join_workload(44, (sol['A'], sol['B'], sol['C'], sol['D']), nrand, seed=4428)
    J = wz.Join(44, tuple(sol['sig']))
    try:
        t0 = time.time()
        r = J.build_from(cd, slots_log2=28)
        assert r.slots == 1 << 28 and r.cd_stream == len(cd)
        assert r.distinct_keys == nrand + 1
        r = J.probe_from(ab)
        assert r.hits == len(pab) and r.found
        sols = J.solutions()
        want = sorted((a, b, c, d) for a, b in pab for c, d in pcd)
        assert sols == want
        for A, B, C, D in sols:
            assert wz.npaf_max(44, unpack(A, 45), unpack(B, 45), unpack(C, 44), unpack(D, 44)) == 0
        bms, pms = J.kernel_ms()
        print(f'2^28-slot join: build {bms:.1f} ms, probe {pms:.1f} ms, {time.time() - t0:.1f} s wall', flush=True)
    finally:
        J.close()
