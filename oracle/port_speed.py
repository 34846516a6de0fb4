"""Speed of the CPU port (liboracle.so) against the compiled reference (oracle/_ref/wz_match),
single-threaded, on the same cells and the same unit. Test infrastructure: it runs only where
the reference tree is mounted (it needs oracle/_ref/wz_match, built by `make -C oracle`) and
writes the measurement that the benchmark's CPU baseline cites.

  config 2  n = 19, all four classes, every live mod-6 cell, stream only (Thm 2.11b/2.12 on,
            orbit canon on; the reference with WZ_FH_CELLSIZE so it counts without completing)
  config 3  n = 29, class (0,6,1,9), a range of live cells, the front-only unit with the whole
            cell buffered, top-64 and A,B budget 2^20 (WZ_FH_DRAIN_TOP=64)

Both sides also print per-cell counts, which are compared, so the two timed runs did the same
work. Usage: python oracle/port_speed.py [out.md]
"""
import os
import re
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / 'tests'))

from oracle_lib import OracleProblem  # noqa: E402

REF = HERE / '_ref' / 'wz_match'
BASE_ENV = {'WZ_FIRSTHIT': '1', 'WZ_FH_M6': '1', 'WZ_THM211B': '1', 'WZ_THM212': '1', 'WZ_FH_ORBIT_CANON': '1',
            'WZ_FH_PROF_ORDER': '1', 'OMP_NUM_THREADS': '1'}


def run_ref(n, sig, extra):
    env = dict(os.environ, **BASE_ENV, **extra)
    t = time.perf_counter()
    r = subprocess.run([str(REF), str(n), *map(str, sig)], env=env, capture_output=True, text=True, timeout=3600)
    return time.perf_counter() - t, r.stdout


def config2():
    rows, tref, tport, nodes = [], 0.0, 0.0, 0
    for sig in ((0, 2, 5, 7), (2, 4, 3, 7), (2, 8, 1, 3), (4, 6, 1, 5)):
        dt, out = run_ref(19, sig, {'WZ_FH_CELLSIZE': str(10 ** 12)})
        ref = {int(m.group(1)): int(m.group(2)) for m in re.finditer(r'CELLSIZE pi=(\d+) cand=(\d+)', out)}
        rn = {int(m.group(1)): int(m.group(2)) for m in re.finditer(r'CDSTAT pi=(\d+) dfs_nodes=(\d+)', out)}
        P = OracleProblem(19, sig, canon=1)
        t = time.perf_counter()
        port = {}
        for pi in range(P.num_cells()):
            _, _, kept, rws = P.cell(pi)
            if not kept or rws == -1:
                continue
            got, _, st = P.stream(pi, 0, 0)
            port[pi] = got
            assert st.dfs_nodes == rn[pi], (sig, pi)
            nodes += st.dfs_nodes
        tport += time.perf_counter() - t
        assert port == ref, sig
        tref += dt
        rows.append((sig, len(ref), sum(ref.values())))
    return tref, tport, nodes, rows


def config3(lo, hi):
    sig = (0, 6, 1, 9)
    budget = 1 << 20
    dt, out = run_ref(29, sig, {'WZ_FH_BUF_CAP': str(10 ** 9), 'WZ_FH_DRAIN_TOP': '64', 'WZ_FH_AB_BUDGET': str(budget),
                                'WZ_FH_PROF_SKIP': str(lo), 'WZ_FH_PROF_END': str(hi), 'WZ_FH_CAND_LOG': '1'})
    ref = {}
    for m in re.finditer(r'CAND idx=\d+ ci=(\d+) pi=(\d+) r=(\d) nodes=(\d+)', out):
        ref.setdefault(int(m.group(2)), []).append((int(m.group(1)), int(m.group(3)), int(m.group(4))))
    found = 'FOUND' in out
    P = OracleProblem(29, sig, canon=1)
    t = time.perf_counter()
    port = {}
    for pi in range(lo, hi):
        _, _, kept, rws = P.cell(pi)
        if not kept or rws == -1:
            continue
        r, per = P.search(pi, 0, 0, 64, budget)
        if r.completed:
            port[pi] = [(k, int(per['verdict'][k]), int(per['nodes'][k])) for k in range(r.completed)]
        if r.hit:
            break
    tport = time.perf_counter() - t
    assert port == ref, 'per-cell completions differ'
    return dt, tport, sum(len(v) for v in port.values()), found


def main():
    if not REF.exists():
        sys.exit(f'{REF} missing: the reference tree must be mounted and built (make -C oracle)')
    out = Path(sys.argv[1]) if len(sys.argv) > 1 else REPO / 'profiles' / 'cpu_port_vs_reference.md'
    r2 = config2()
    r3 = config3(0, 44)
    cpu = next((l.split(':', 1)[1].strip() for l in open('/proc/cpuinfo') if l.startswith('model name')), '?')
    text = f"""# CPU port vs the compiled reference (single thread)

Command: `python oracle/port_speed.py` (this file is its output). Host: {cpu}, one thread each,
the reference built by `oracle/Makefile` (`g++ -O3 -std=c++17`), the port `oracle/_build/liboracle.so`
(`g++ -O3 -std=c++17 -ffp-contract=off`). Both runs did the same work: per-cell counts (config 2) and
per-cell completion verdicts and node counts (config 3) are compared and equal. Reference times
include its per-(n, class) setup; the port's exclude its setup.

| input | reference s | port s | port / reference time | work |
|---|---:|---:|---:|---|
| config 2: n = 19, 4 classes, every live cell, stream only | {r2[0]:.2f} | {r2[1]:.2f} | {r2[1] / r2[0]:.3f} | {r2[2]:,} DFS nodes, cells/cands per class {[(s, c, k) for s, c, k in r2[3]]} |
| config 3: n = 29 (0,6,1,9), cells 0..43 of the list, top-64, A,B 2^20 | {r3[0]:.2f} | {r3[1]:.2f} | {r3[1] / r3[0]:.3f} | {r3[2]:,} completions{' (stopped at the first hit)' if r3[3] else ''} |

A ratio below 1 means the port is faster than the reference on that input; the benchmark's
`cpu_baseline` (the port on the GPU box's cores) is therefore an estimate of the reference's speed
on those cores scaled by these ratios.
"""
    out.write_text(text)
    print(text)


if __name__ == '__main__':
    main()
