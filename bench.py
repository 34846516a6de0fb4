#!/usr/bin/env python3
"""Benchmark: n = 44 cell searches per second on one MI355X, served vs plain GPU.

Workload (BS(45,44), the reference's n = 44 production setting): the 12 signature classes
at n = 44 (or --classes), mod-6 cells with Thm 2.11b/2.12 (forced at n >= 36), orbit canon,
flat order; cells drawn from each class's live list by a seeded splitmix64. One cell
search is the front-only unit of the reference arm: stream the cell's C,D candidates in
DFS order (at most --node-budget DFS nodes, a serving policy; the reference caps cells by
candidates or wall time only), stable-sort them by flatness, complete the top --top-k with
--ab-budget A,B nodes each under the cell's profile rows. Inputs are synthetic in the sense
that cells are sampled; every computation is the real one, results are exact.

A step = one pass of the hot path over the batch of sampled cells, served: every cell is
submitted to the continuous-batching server (one worker and HIP stream per class, requests
of a class admitted together) and the step ends when every result is back. `value` is
served cells/s. Beside it on the same JSON line:
  plain    one cell per search call (the single-cell GPU path), cells/s and latency
  batched  one persistent launch per class (the kernel-level view; roofline source)
  latency  open-loop Poisson arrivals at --load x the served rate: p50 / p99 cell latency
  roofline the stream kernel's INT32-VALU rate on counted DFS nodes, against peak
  cpu      the oracle port on the host cores on a bounded sample of the same cells
"""
import argparse
import json
import os
import random
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import bs45_quantum_explorer_amd as wz  # noqa: E402

# MI355X figures used for the roofline. INT32 VALU peak from the published CU count and
# clock: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (one 32-bit op per lane per cycle).
INT_VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_BPS = 8.0e12   # MI355X HBM3E
OPS_PER_CD_NODE = 84          # per counted C,D DFS node (rec() entry) at m = 6


def classes(n):
    """Signature classes (a <= b, c <= d) with a^2+b^2+c^2+d^2 = 4n+2 and the parities of
    lengths n+1 (A,B) and n (C,D)."""
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


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def pct(v, q):
    return round(1e3 * float(np.percentile(np.asarray(v, dtype=np.float64), q)), 3) if len(v) else None


def key(r):
    s = r.stream
    return (s.cand, s.dfs_nodes, r.completed, r.hit, r.hit_rank, r.ab_nodes, r.clean_no, r.aborted, r.A, r.B)


def served_pass(server, reqs, timeout_s=600):
    """Submit every (problem, cell) request, wait for all results; returns results by request
    index and the wall time."""
    t0 = time.perf_counter()
    tick = {}
    for i, (pi, c) in enumerate(reqs):
        tick[server.submit(pi, c)] = i
    got = [None] * len(reqs)
    n = 0
    deadline = time.monotonic() + timeout_s
    while n < len(reqs):
        if time.monotonic() > deadline:
            raise RuntimeError(f'served pass: {n} of {len(reqs)} results after {timeout_s} s')
        for r in server.poll(max_results=4096, min_results=1, timeout_ms=2000):
            if r.status != 0:
                raise RuntimeError(f'served request failed with status {r.status}')
            got[tick[r.ticket]] = r
            n += 1
    return got, time.perf_counter() - t0


def poisson_latency(server, reqs, rate, count, seed):
    """Open-loop Poisson arrivals at `rate` requests/s; latency = finish - submit."""
    rng = random.Random(seed)
    order = [reqs[rng.randrange(len(reqs))] for _ in range(count)]
    t0 = time.perf_counter()
    next_t = 0.0
    lat, pending, done = [], set(), 0
    i = 0
    while done < count:
        now = time.perf_counter() - t0
        while i < count and next_t <= now:
            pending.add(server.submit(*order[i]))
            i += 1
            next_t += rng.expovariate(rate)
        wait_ms = int(max(0.0, min(50.0, (next_t - now) * 1e3))) if i < count else 2000
        for r in server.poll(max_results=4096, min_results=1, timeout_ms=wait_ms):
            pending.discard(r.ticket)
            lat.append(r.finish_s - r.submit_s)
            done += 1
    return lat, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=44)
    ap.add_argument('--classes', type=int, default=12, help='how many of the n=44 classes to use')
    ap.add_argument('--cells-per-class', type=int, default=256)
    ap.add_argument('--node-budget', type=int, default=1 << 22)
    ap.add_argument('--cand-cap', type=int, default=0)
    ap.add_argument('--top-k', type=int, default=32)
    ap.add_argument('--ab-budget', type=int, default=1 << 20)
    ap.add_argument('--plain-cells', type=int, default=2, help='cells per class timed one per call')
    ap.add_argument('--latency-requests', type=int, default=1500)
    ap.add_argument('--load', type=float, default=0.6, help='Poisson rate as a fraction of the served rate')
    ap.add_argument('--max-batch', type=int, default=4096)
    ap.add_argument('--seed', type=int, default=44)
    ap.add_argument('--cpu-cells', type=int, default=48, help='cells per class in the CPU baseline sample')
    ap.add_argument('--cpu-threads', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch settings')
    ap.add_argument('--tune-cells', type=int, default=0,
                    help='cells of the autotuning sample (0: as many as a class batch, --cells-per-class)')
    ap.add_argument('--probe-traffic', action='store_true',
                    help='run the HBM-traffic probe for rocprofv3 --pmc (calibration kernel + one stream launch)')
    a = ap.parse_args()
    if a.probe_traffic:
        probe_traffic(a)
        return
    if a.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; cells shard by cell index mod G across GPUs '
                                   'with a final min over (rank, idx), which is not built here'}))
        sys.exit(2)

    t_setup = time.time()
    cls = classes(a.n)[:a.classes]
    probs, cells = [], []
    for i, sig in enumerate(cls):
        P = wz.Problem(a.n, sig, canon=1, order=1)
        probs.append(P)
        cells.append(sample_cells(P.live_cells(), a.cells_per_class, a.seed * 1000 + i))
        log(f'setup class {i + 1}/{len(cls)} {sig}: {P.num_cells()} cells')
    setup_s = time.time() - t_setup
    # launch settings (speed only): tuned once on a separate sample of the first class and
    # applied to every class (the tuner checks that each setting returns identical results)
    tune = None
    if not a.no_tune:
        t0 = time.time()
        sample = sample_cells(probs[0].live_cells(), a.tune_cells or a.cells_per_class, a.seed * 7919 + 1)
        t = probs[0].autotune(sample, a.top_k, a.ab_budget, cand_cap=a.cand_cap, node_budget=a.node_budget)
        for P in probs[1:]:
            P.set_tuning(t)
        tune = dict(t.as_dict(), seconds=round(time.time() - t0, 1), sample_cells=len(sample))
        log(f'autotune: {tune}')
    reqs = [(pi, c) for pi, cl in enumerate(cells) for c in cl]
    total_cells = len(reqs)
    log(f'setup: {len(cls)} classes, {total_cells} cells, {setup_s:.1f} s')
    search = dict(cand_cap=a.cand_cap, node_budget=a.node_budget)

    # ---- served (the measured value)
    server = wz.Server(probs, a.top_k, a.ab_budget, max_batch=a.max_batch, **search)
    try:
        for w in range(a.warmup):
            _, dt = served_pass(server, reqs)
            log(f'served warmup {w + 1}/{a.warmup}: {dt * 1e3:.1f} ms')
        times, served = [], None
        for s in range(a.steps):
            served, dt = served_pass(server, reqs)
            times.append(dt)
            log(f'served step {s + 1}/{a.steps}: {dt * 1e3:.1f} ms')
        step_s = float(np.mean(times))
        served_rate = total_cells / step_s
        # ---- open-loop latency at a fraction of the served rate
        rate = max(1.0, a.load * served_rate)
        lat, lat_wall = poisson_latency(server, reqs, rate, a.latency_requests, a.seed)
        log(f'latency: {len(lat)} Poisson requests at {rate:.1f}/s in {lat_wall:.1f} s')
    finally:
        server.close()

    # ---- batched: one persistent launch per class (kernel timing for the roofline)
    kern_ms, kern_units, batched_s = [], [], 0.0
    for P, cl in zip(probs, cells):
        t0 = time.perf_counter()
        P.search(cl, a.top_k, a.ab_budget, batched=True, **search)
        batched_s += time.perf_counter() - t0
        ms, units = P.last_kernel(0)
        kern_ms.append(ms)
        kern_units.append(units)
    log(f'batched: {len(probs)} launches, {batched_s * 1e3:.1f} ms')

    # ---- plain: one cell per call
    plain_lat, mismatch, nplain = [], 0, 0
    for ci, (P, cl) in enumerate(zip(probs, cells)):
        for j, c in enumerate(cl[:a.plain_cells]):
            t0 = time.perf_counter()
            r = P.search([c], a.top_k, a.ab_budget, batched=False, **search)[0]
            plain_lat.append(time.perf_counter() - t0)
            nplain += 1
            if key(r) != key(served[reqs.index((ci, c))].result):
                mismatch += 1
    plain_rate = nplain / sum(plain_lat) if plain_lat else 0.0
    log(f'plain: {nplain} cells, {plain_rate:.2f} cells/s, mismatches vs served {mismatch}')

    res = [r.result for r in served]
    streamed = sum(r.stream.cand for r in res)
    completed = sum(r.completed for r in res)
    dfs_nodes = sum(r.stream.dfs_nodes for r in res)
    ab_nodes = sum(r.ab_nodes for r in res)
    hits = sum(r.hit for r in res)

    # roofline of the dominant kernel (the stream kernel), from the batched launches
    avg_ms = float(np.mean(kern_ms))
    units_per_launch = float(np.mean(kern_units))
    achieved = OPS_PER_CD_NODE * units_per_launch / (avg_ms / 1e3) / 1e12 if avg_ms > 0 else 0.0
    traffic = None
    tf = ROOT / 'profiles' / 'stream_kernel_traffic.json'
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get('bytes_per_launch')
        except ValueError:
            traffic = None
    roofline = {'bound': 'valu', 'achieved': round(achieved, 3), 'peak': round(INT_VALU_PEAK_TOPS, 2),
                'unit': 'TOP/s (INT32 VALU)', 'frac': round(achieved / INT_VALU_PEAK_TOPS, 4), 'traffic': traffic,
                'kernel': 'wzfh::stream_kernel',
                'per_unit': f'{OPS_PER_CD_NODE} INT32 ops per counted C,D DFS node (rec() entry)',
                'traffic_source': 'profiles/stream_kernel_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, '
                                  'calibrated on a 256 MiB elementwise kernel)' if traffic else None,
                'hbm_gb_per_s': round(traffic / (avg_ms / 1e3) / 1e9, 2) if traffic and avg_ms > 0 else None,
                'hbm_frac': round(traffic / (avg_ms / 1e3) / HBM_PEAK_BPS, 5) if traffic and avg_ms > 0 else None,
                'units_per_launch': int(units_per_launch), 'avg_launch_ms': round(avg_ms, 3)}

    cpu = None if a.no_cpu else cpu_baseline(a, cls, probs, cells)

    line = {
        'metric': 'cells/sec & C,D cands/sec at n=44 (served vs plain GPU); p50/p99 cell latency',
        'value': round(served_rate, 3), 'unit': 'cells/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(1e3 * step_s, 3), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32 (DFS) + fp16x2/fp32 MFMA with fp64 recheck (leaf test)',
        'data': 'synthetic: seeded splitmix64 sample of the real n=44 cell lists',
        'config': {'workload': f'BS(45,44) front-only cell search: {len(cls)} classes x {a.cells_per_class} '
                               f'seeded cells, node budget {a.node_budget}, cand cap {a.cand_cap}, '
                               f'top-{a.top_k}, A,B budget {a.ab_budget}, served',
                   'n': a.n, 'classes': len(cls), 'cells_per_step': total_cells, 'node_budget': a.node_budget,
                   'cand_cap': a.cand_cap, 'top_k': a.top_k, 'ab_budget': a.ab_budget, 'seed': a.seed,
                   'parallelism': 'single GPU: continuous batching per class, persistent work-donating launches'},
        'served_cells_per_s': round(served_rate, 3),
        'plain_cells_per_s': round(plain_rate, 3),
        'served_over_plain': round(served_rate / plain_rate, 2) if plain_rate else None,
        'batched_cells_per_s': round(total_cells / batched_s, 3) if batched_s else None,
        'cands_streamed_per_s': round(streamed / step_s, 1),
        'cands_completed_per_s': round(completed / step_s, 1),
        'dfs_nodes_per_s': round(dfs_nodes / step_s, 1),
        'served_latency_ms': {'p50': pct(lat, 50), 'p99': pct(lat, 99), 'offered_per_s': round(rate, 2),
                              'requests': len(lat)},
        'plain_latency_ms': {'p50': pct(plain_lat, 50), 'p99': pct(plain_lat, 99), 'cells': nplain},
        'hits': hits, 'dfs_nodes_per_step': dfs_nodes, 'ab_nodes_per_step': ab_nodes,
        'plain_vs_served_mismatch': mismatch, 'setup_s': round(setup_s, 1),
        'roofline': roofline, 'cpu_baseline': cpu, 'autotune': tune,
    }
    print(json.dumps(line), flush=True)
    for P in probs:
        P.close()


def probe_traffic(a):
    """For `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`: first an elementwise kernel that
    reads and writes exactly 256 MiB each (the counters' calibration), then one stream launch of
    the bench workload (the first class, --cells-per-class seeded cells, the node budget) with the
    bench's tuned launch settings when profiles/stream_kernel_traffic.json names them."""
    import torch
    x = torch.zeros(64 << 20, dtype=torch.int32, device='cuda')
    y = torch.empty_like(x)
    torch.add(x, 1, out=y)
    torch.cuda.synchronize()
    P = wz.Problem(a.n, classes(a.n)[0], canon=1, order=1)
    cells = sample_cells(P.live_cells(), a.cells_per_class, a.seed * 1000)
    tf = ROOT / 'profiles' / 'stream_kernel_traffic.json'
    if tf.exists():
        t = json.loads(tf.read_text()).get('tuning')
        if t:
            P.set_tuning(wz.Tuning(*[t.get(k, d) for k, d in (('inner_moves', 16), ('donate_hops', 1),
                                                              ('static_window', 4), ('complete_split', 3),
                                                              ('idle_sleep', 64), ('blocks_per_cu', 4))]))
    st, _, _ = P.stream(cells, node_budget=a.node_budget, want_cands=False)
    ms, units = P.last_kernel(0)
    print(json.dumps({'calibration_bytes_each_way': 256 << 20, 'cells': len(cells), 'stream_ms': round(ms, 3),
                      'counted_nodes': units}), flush=True)
    P.close()


def cpu_baseline(a, cls, probs, cells):
    """The oracle port (oracle/wz_oracle.cpp) on the host cores: the same cell unit on a
    bounded sample (--cpu-cells cells of each class, evenly spaced in its seeded sample), one cell
    per thread."""
    sys.path.insert(0, str(ROOT / 'tests'))
    try:
        from oracle_lib import OracleProblem
    except Exception as e:  # oracle not built: no baseline
        return {'error': f'oracle unavailable: {e}'}
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    work, oprobs = [], []
    t_setup = time.time()
    for P_gpu, sig, cl in zip(probs, cls, cells):
        pick = cl[::max(1, len(cl) // max(1, a.cpu_cells))][:a.cpu_cells]   # spread over the seeded sample
        P = OracleProblem(a.n, sig, canon=1, order=1, cells=[P_gpu.cell(pi)[:2] for pi in pick])
        oprobs.append(P)
        work += [(P, j) for j in range(len(pick))]
    setup = time.time() - t_setup
    cap = a.cand_cap if a.cand_cap > 0 else (1 << 22)

    def one(item):
        P, j = item
        got, v, nd, idx, st = P.front_budget(j, cap, a.node_budget, a.top_k, a.ab_budget)
        return got, st.dfs_nodes

    t0 = time.perf_counter()
    cands = nodes = 0
    with ThreadPoolExecutor(threads) as ex:
        for k, (c, nd) in enumerate(ex.map(one, work)):
            cands += c
            nodes += nd
            if (k + 1) % threads == 0:
                log(f'cpu baseline: {k + 1}/{len(work)} cells')
    dt = time.perf_counter() - t0
    return {'value': round(len(work) / dt, 4), 'unit': 'cells/s', 'cores': threads, 'kind': 'port',
            'sample': f'{len(work)} cells ({a.cpu_cells} per class, evenly spaced in each class\'s seeded sample) '
                      f'of the same workload and unit',
            'cands_per_s': round(cands / dt, 1), 'dfs_nodes_per_s': round(nodes / dt, 1), 'seconds': round(dt, 2),
            'setup_s': round(setup, 1), 'cpu_model': _cpu_model()}


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
