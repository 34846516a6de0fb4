#!/usr/bin/env python3
"""Benchmark: cell searches per second at n = 44 on one MI355X, batched vs plain.

Workload (BS(45,44), the reference's n = 44 production setting): every signature class at
n = 44 (or --classes), mod-6 cells with Thm 2.11b/2.12, orbit canon, flat order; cells drawn
from each class's live list by a seeded splitmix64. One cell search = the front-only unit
of the reference arm: stream the cell's C,D candidates in DFS order (at most --node-budget
DFS nodes and --cand-cap candidates), stable-sort them by flatness, complete the top --top-k
with --ab-budget A,B nodes each under the cell's profile rows. Inputs are synthetic in the
sense that cells are sampled; every computation is the real one.

A step = one pass of the hot path over the batch of sampled cells:
  batched: all cells of a class in one persistent stream launch + one completer launch
  plain:   one cell per launch (the baseline single-cell GPU path)
`value` is the batched cells/s. The plain rate, candidate rates, per-cell latency p50/p99,
the stream kernel's roofline and a CPU baseline (the oracle port on the host cores, on a
bounded sample) are reported beside it on the same JSON line.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import bs45_quantum_explorer_amd as wz  # noqa: E402

# MI355X figures used for the roofline (published peaks; INT32 VALU is derived from the
# published CU count and clock: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz).
HBM_PEAK_GBS = 8000.0
INT_VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
OPS_PER_CD_NODE = 84          # per counted C,D DFS node (rec() entry), m = 6


def classes(n):
    out = []
    for a in range(n + 2):
        for b in range(a, n + 2):
            for c in range(n + 1):
                for d in range(c, n + 1):
                    if (a - n - 1) % 2 or (b - n - 1) % 2 or (c - n) % 2 or (d - n) % 2:
                        continue
                    if a * a + b * b + c * c + d * d == 4 * n + 2:
                        out.append((a, b, c, d))
    return out


def splitmix64(state):
    while True:
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        yield z ^ (z >> 31)


def sample_cells(live, k, seed):
    """k distinct cells from `live`, chosen by splitmix64 (partial Fisher-Yates)."""
    live = list(live)
    rng = splitmix64(seed)
    k = min(k, len(live))
    for i in range(k):
        j = i + next(rng) % (len(live) - i)
        live[i], live[j] = live[j], live[i]
    return sorted(live[:k])


_PROBS = []


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def percentile(v, q):
    return float(np.percentile(np.asarray(v, dtype=np.float64), q)) if len(v) else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=44)
    ap.add_argument('--classes', type=int, default=12, help='how many of the n=44 classes to use')
    ap.add_argument('--cells-per-class', type=int, default=256)
    ap.add_argument('--plain-cells', type=int, default=32, help='cells timed through the plain path')
    ap.add_argument('--node-budget', type=int, default=1 << 26)
    ap.add_argument('--cand-cap', type=int, default=0)
    ap.add_argument('--top-k', type=int, default=32)
    ap.add_argument('--ab-budget', type=int, default=1 << 20)
    ap.add_argument('--seed', type=int, default=44)
    ap.add_argument('--cpu-cells', type=int, default=6, help='cells per class in the CPU baseline sample')
    ap.add_argument('--cpu-threads', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({'error': 'this engine runs on one GPU; multi-GPU sharding is design-only '
                                   '(cells shard by pi mod G with a final min over (rank, idx))'}))
        sys.exit(2)

    t_setup = time.time()
    cls = classes(a.n)[:a.classes]
    probs, cells = _PROBS, []
    for i, sig in enumerate(cls):
        P = wz.Problem(a.n, sig, canon=1, order=1)
        live = P.live_cells()
        probs.append(P)
        cells.append(sample_cells(live, a.cells_per_class, a.seed * 1000 + i))
    setup_s = time.time() - t_setup
    total_cells = sum(len(c) for c in cells)
    log(f'setup: {len(cls)} classes, {total_cells} cells, {setup_s:.1f} s')

    def step_batched():
        res = []
        st_ms = cp_ms = 0.0
        st_units = 0
        for P, cl in zip(probs, cells):
            r = P.search(cl, a.top_k, a.ab_budget, cand_cap=a.cand_cap, node_budget=a.node_budget, batched=True)
            ms0, u0 = P.last_kernel(0)
            ms1, _ = P.last_kernel(1)
            st_ms += ms0
            cp_ms += ms1
            st_units += u0
            res.append(r)
        return res, st_ms, cp_ms, st_units

    for w in range(a.warmup):
        step_batched()
        log(f'warmup step {w + 1}/{a.warmup} done')
    times, kern = [], []
    results = None
    for _ in range(a.steps):
        t0 = time.perf_counter()
        results, st_ms, cp_ms, st_units = step_batched()
        times.append(time.perf_counter() - t0)
        kern.append((st_ms, cp_ms, st_units))
        log(f'step {len(times)}/{a.steps}: {times[-1] * 1e3:.1f} ms (stream kernels {st_ms:.1f} ms, '
            f'completer {cp_ms:.1f} ms)')
    ms_per_step = 1000.0 * float(np.mean(times))
    cells_per_s = total_cells / float(np.mean(times))
    streamed = sum(r.stream.cand for rr in results for r in rr)
    completed = sum(r.completed for rr in results for r in rr)
    hits = sum(r.hit for rr in results for r in rr)
    dfs_nodes = sum(r.stream.dfs_nodes for rr in results for r in rr)
    ab_nodes = sum(r.ab_nodes for rr in results for r in rr)

    # plain path: one cell per launch, same cells (first plain-cells of each class, round-robin)
    plain_lat, plain_res = [], {}
    per_class = max(1, a.plain_cells // max(1, len(probs)))
    t0 = time.perf_counter()
    nplain = 0
    for ci, (P, cl) in enumerate(zip(probs, cells)):
        for pi in cl[:per_class]:
            t1 = time.perf_counter()
            r = P.search([pi], a.top_k, a.ab_budget, cand_cap=a.cand_cap, node_budget=a.node_budget, batched=False)
            plain_lat.append(time.perf_counter() - t1)
            plain_res[(ci, pi)] = r[0]
            nplain += 1
        log(f'plain: class {ci + 1}/{len(probs)} done ({nplain} cells)')
    plain_s = time.perf_counter() - t0
    plain_cells_per_s = nplain / plain_s if plain_s > 0 else 0.0
    # the batched and plain paths must agree exactly on every cell they share
    mismatch = 0
    for ci, (P, cl) in enumerate(zip(probs, cells)):
        for j, pi in enumerate(cl):
            if (ci, pi) in plain_res:
                x, y = plain_res[(ci, pi)], results[ci][j]
                if (x.stream.cand, x.stream.dfs_nodes, x.completed, x.hit, x.ab_nodes, x.clean_no, x.aborted) != \
                   (y.stream.cand, y.stream.dfs_nodes, y.completed, y.hit, y.ab_nodes, y.clean_no, y.aborted):
                    mismatch += 1

    st_ms = float(np.mean([k[0] for k in kern]))
    cp_ms = float(np.mean([k[1] for k in kern]))
    st_units = float(np.mean([k[2] for k in kern]))
    nlaunch = len(probs)
    avg_ms = st_ms / nlaunch
    achieved_tops = OPS_PER_CD_NODE * (st_units / nlaunch) / (avg_ms / 1e3) / 1e12 if avg_ms > 0 else 0.0
    traffic = None
    tf = ROOT / 'profiles' / 'stream_kernel_traffic.json'
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get('bytes_per_launch')
        except ValueError:
            traffic = None
    roofline = {'kernel': 'wzfh::stream_kernel', 'bound': 'valu', 'achieved': round(achieved_tops, 3),
                'peak': round(INT_VALU_PEAK_TOPS, 2), 'unit': 'TOP/s (INT32 VALU)',
                'frac': round(achieved_tops / INT_VALU_PEAK_TOPS, 4), 'traffic': traffic,
                'per_unit': f'{OPS_PER_CD_NODE} INT ops per counted C,D DFS node',
                'units_per_launch': int(st_units / nlaunch), 'avg_launch_ms': round(avg_ms, 3),
                'completer_ms_per_step': round(cp_ms, 3)}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, cls, cells)

    line = {
        'metric': 'cells/sec & C,D cands/sec at n=44 (served vs plain GPU); p50/p99 cell latency',
        'value': round(cells_per_s, 3), 'unit': 'cells/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(ms_per_step, 3), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32+f64', 'data': 'synthetic (seeded cell sample of the real n=44 cell lists)',
        'config': {'workload': f'BS(45,44) front-only cell search, {len(cls)} classes x {a.cells_per_class} '
                               f'seeded cells, node budget {a.node_budget}, cand cap {a.cand_cap}, '
                               f'top-{a.top_k}, A,B budget {a.ab_budget}',
                   'n': a.n, 'classes': len(cls), 'cells_per_step': total_cells, 'node_budget': a.node_budget,
                   'cand_cap': a.cand_cap, 'top_k': a.top_k, 'ab_budget': a.ab_budget, 'seed': a.seed,
                   'parallelism': 'single GPU, persistent work-stealing launches'},
        'batched_cells_per_s': round(cells_per_s, 3),
        'plain_cells_per_s': round(plain_cells_per_s, 3),
        'batched_over_plain': round(cells_per_s / plain_cells_per_s, 2) if plain_cells_per_s else None,
        'cands_streamed_per_s': round(streamed / float(np.mean(times)), 1),
        'cands_completed_per_s': round(completed / float(np.mean(times)), 1),
        'plain_latency_ms': {'p50': round(1e3 * percentile(plain_lat, 50), 3),
                             'p99': round(1e3 * percentile(plain_lat, 99), 3)},
        'batched_latency_ms': {'p50': round(ms_per_step / max(1, len(probs)), 3),
                               'p99': round(1e3 * max(times), 3)},
        'hits': hits, 'dfs_nodes_per_step': dfs_nodes, 'ab_nodes_per_step': ab_nodes,
        'plain_vs_batched_mismatch': mismatch, 'setup_s': round(setup_s, 1),
        'roofline': roofline, 'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(a, cls, cells):
    """The oracle port (oracle/wz_oracle.cpp) on the host cores: the same cell unit on a
    bounded sample (the first --cpu-cells cells of each class), one cell per thread."""
    sys.path.insert(0, str(ROOT / 'tests'))
    try:
        from oracle_lib import OracleProblem
    except Exception as e:  # oracle not built: no baseline
        return {'error': f'oracle unavailable: {e}'}
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    work = []
    t_setup = time.time()
    oprobs = []
    for P_gpu, sig, cl in zip(_PROBS, cls, cells):
        pick = cl[:a.cpu_cells]
        P = OracleProblem(a.n, sig, canon=1, order=1, cells=[P_gpu.cell(pi)[:2] for pi in pick])
        oprobs.append(P)
        for j in range(len(pick)):
            work.append((P, j))
    setup = time.time() - t_setup
    cap = a.cand_cap if a.cand_cap > 0 else (1 << 22)

    def one(item):
        P, pi = item
        got, v, nd, idx, st = P.front_budget(pi, cap, a.node_budget, a.top_k, a.ab_budget)
        return got
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        cands = 0
        for k, c in enumerate(ex.map(one, work)):
            cands += c
            if (k + 1) % threads == 0:
                log(f'cpu baseline: {k + 1}/{len(work)} cells')
    dt = time.perf_counter() - t0
    return {'value': round(len(work) / dt, 4), 'unit': 'cells/s', 'cores': threads, 'kind': 'port',
            'sample': f'{len(work)} cells ({a.cpu_cells} per class) of the same workload, same unit',
            'cands_per_s': round(cands / dt, 1), 'seconds': round(dt, 2), 'setup_s': round(setup, 1),
            'cpu_model': _cpu_model()}


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == '__main__':
    main()
