#!/usr/bin/env python3
"""Benchmark: n = 44 cell searches per second on one MI355X, served vs plain GPU.

Workload (BS(45,44), the n = 44 configuration: 1,024 seeded cells from each of the 12
signature classes, A,B completion budget 2^24). Classes at n = 44 with mod-6 cells, Thm
2.11b/2.12 (forced at n >= 36), orbit canon, flat order; each class's 1,024 cells are drawn
from its live list by a seeded splitmix64. One cell search is the front-only unit of the
reference arm: stream the cell's C,D candidates in DFS order, stable-sort them by flatness,
complete the flattest --top-k (32; the configuration names no width, 32 is the serving
width) with the A,B budget each under the cell's profile rows, stopping at the first hit.
n = 44 cells hold ~10^11 DFS nodes each, so every cell is cut at --node-budget DFS nodes
(2^22): a policy of this build (the reference caps cells by candidates or wall time only),
labelled in config.workload. Inputs are synthetic in the sense that cells are sampled;
every computation is the real one, results are exact.

A step = one slice of that set: slice k holds cells j = k (mod 12) of every class's seeded
sample (~1,024 cells across the 12 classes), so 12 consecutive steps search the whole
12 x 1,024 set once; step s (warmup steps first) takes slice s mod 12. The steps are served:
each step's cells are submitted to the continuous-batching server as one burst, with eight steps
in flight (--in-flight): step s+8 is submitted when every result of step s is back, so the
stream launches of later steps run while earlier steps' completions finish, as under continuous
load; a stream launch admits at most two steps' cells. The timed region starts with nothing in flight
and ends when the last result of the last timed step is back. `value` is served cells/s over
the timed steps (cells / that wall time). Launch settings (speed only; every trial is checked to
return the same results) come from the library's autotuner on a step-sized sample of one class,
then the completer's item charge cap from a few served steps per candidate cap. On the same
JSON line:
  plain     one cell per search call (the single-cell GPU path), cells/s and latency
  batched   one persistent launch per class on slice 0 (the kernel-level view)
  latency   open-loop Poisson arrivals at --load x the served rate: p50 / p99 cell latency
  roofline  the dominant kernel's rate over the timed steps (HIP-event time of its launches,
            summed over the classes' handles) against the INT32-VALU peak; both kernels under
            rooflines{}
  verify    the matrix-core leaf test's false rejects / accepts against FP64 on a sample
  cpu       the oracle port on the host cores on a bounded sample of the same cells, with its
            per-cell results compared to the served ones (mismatches must be 0)
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

# One hardware queue per HIP stream for the server's streams (one per problem plus one per n):
# with HIP's default of 4 queues, the completer launches of the 12 classes share 4 queues and
# wait for each other (measured on the n = 44 bench: 653 served cells/s with 4 queues, 716 with 8,
# 845 with 16). Read by the HIP runtime when it starts, so it is set before the library loads.
os.environ.setdefault('GPU_MAX_HW_QUEUES', '16')

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import bs45_quantum_explorer_amd as wz  # noqa: E402
from bs45_quantum_explorer_amd.workload import classes, sample_cells  # noqa: E402

# MI355X figures used for the roofline. INT32 VALU peak from the published CU count and
# clock: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (one 32-bit op per lane per cycle).
INT_VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_BPS = 8.0e12   # MI355X HBM3E
OPS_PER_CD_NODE = 84          # per counted C,D DFS node (rec() entry) at m = 6


def ops_per_ab_node(n):
    """Per counted completer node: two O(L) placements updating Dab and Kab plus the n-shift
    bound scan, 6L + n INT ops (314 at n = 44)."""
    return 6 * (n + 1) + n


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def pct(v, q):
    return round(1e3 * float(np.percentile(np.asarray(v, dtype=np.float64), q)), 3) if len(v) else None


def key(r):
    s = r.stream
    return (s.cand, s.dfs_nodes, r.completed, r.hit, r.hit_rank, r.ab_nodes, r.clean_no, r.aborted, r.A, r.B)


def served_pass(server, reqs, timeout_s=600):
    """Submit every (problem, cell) request at once (one burst), wait for all results; returns
    results by request index and the wall time."""
    t0 = time.perf_counter()
    tick = {t: i for i, t in enumerate(server.submit_batch(reqs))}
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


def served_steps(server, steps, depth, timeout_s=600):
    """Serve a sequence of steps (lists of requests) back to back with at most `depth` steps in
    flight: steps 0..depth-1 are submitted at once, and step s+depth as soon as every result of
    step s is back, so the server always holds the next step while one completes (a server
    under continuous load: one step's stream launch runs while the previous step's completions
    finish). Returns per step its results by request index and its submit-to-last-result time,
    and the wall time of the whole sequence."""
    t0 = time.perf_counter()
    got = [[None] * len(st) for st in steps]
    left = [len(st) for st in steps]
    sub_t, dts = [0.0] * len(steps), [0.0] * len(steps)
    where = {}

    def submit(s):
        sub_t[s] = time.perf_counter()
        for i, t in enumerate(server.submit_batch(steps[s])):
            where[t] = (s, i)

    nxt = 0
    while nxt < min(depth, len(steps)):
        submit(nxt)
        nxt += 1
    done = 0
    deadline = time.monotonic() + timeout_s
    while done < len(steps):
        if time.monotonic() > deadline:
            raise RuntimeError(f'served steps: {done} of {len(steps)} steps after {timeout_s} s')
        for r in server.poll(max_results=4096, min_results=1, timeout_ms=2000):
            if r.status != 0:
                raise RuntimeError(f'served request failed with status {r.status}')
            s, i = where.pop(r.ticket)
            got[s][i] = r
            left[s] -= 1
            if left[s] == 0:
                dts[s] = time.perf_counter() - sub_t[s]
                done += 1
                if nxt < len(steps):
                    submit(nxt)
                    nxt += 1
    return got, dts, time.perf_counter() - t0


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
            if r.status != 0:
                raise RuntimeError(f'served request failed with status {r.status}')
            pending.discard(r.ticket)
            lat.append(r.finish_s - r.submit_s)
            done += 1
    return lat, time.perf_counter() - t0


def kernel_roofline(name, ops_per_unit, per_unit, ms, units, launches, wall_s, traffic=None):
    """Achieved INT32 rate of one kernel: algorithmic ops of its launches over their summed
    HIP-event durations (= per-launch ops / average launch duration)."""
    avg_ms = ms / launches if launches else 0.0
    achieved = ops_per_unit * units / (ms / 1e3) / 1e12 if ms > 0 else 0.0
    out = {'bound': 'valu', 'achieved': round(achieved, 4), 'peak': round(INT_VALU_PEAK_TOPS, 2),
           'unit': 'TOP/s (INT32 VALU)', 'frac': round(achieved / INT_VALU_PEAK_TOPS, 5),
           'traffic': traffic.get('bytes_per_launch') if traffic else None,
           'kernel': name, 'per_unit': per_unit, 'units_timed': int(units), 'launches_timed': int(launches),
           'avg_launch_ms': round(avg_ms, 3), 'units_per_launch': round(units / launches, 1) if launches else 0,
           # the same ops over the timed steps' wall clock (launches of several classes overlap)
           'device_rate_tops': round(ops_per_unit * units / wall_s / 1e12, 4) if wall_s > 0 else None}
    if traffic:
        out['traffic_source'] = traffic.get('source')
        if avg_ms > 0 and traffic.get('bytes_per_launch'):
            out['hbm_gb_per_s'] = round(traffic['bytes_per_launch'] / (avg_ms / 1e3) / 1e9, 3)
            out['hbm_frac'] = round(traffic['bytes_per_launch'] / (avg_ms / 1e3) / HBM_PEAK_BPS, 6)
    return out


def load_traffic():
    tf = ROOT / 'profiles' / 'kernel_traffic_n44.json'
    if not tf.exists():
        return {}
    try:
        return json.loads(tf.read_text()).get('kernels', {})
    except ValueError:
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=44)
    ap.add_argument('--classes', type=int, default=12, help='how many of the n=44 classes to use')
    ap.add_argument('--cells-per-class', type=int, default=1024)
    ap.add_argument('--slices', type=int, default=12, help='a step searches every slices-th cell of each class')
    ap.add_argument('--node-budget', type=int, default=1 << 22)
    ap.add_argument('--cand-cap', type=int, default=0)
    ap.add_argument('--top-k', type=int, default=32)
    ap.add_argument('--ab-budget', type=int, default=1 << 24)
    ap.add_argument('--plain-cells', type=int, default=2, help='cells per class timed one per call')
    ap.add_argument('--latency-requests', type=int, default=1500)
    ap.add_argument('--load', type=float, default=0.6, help='Poisson rate as a fraction of the served rate')
    ap.add_argument('--max-batch', type=int, default=0, help='cells per stream launch (0: two slices)')
    ap.add_argument('--in-flight', type=int, default=8, help='served steps in flight at once')
    ap.add_argument('--serve-caps', default='11,12,13,14',
                    help='completer charge caps (log2) tried on served steps after autotuning (empty: skip)')
    ap.add_argument('--serve-tune-steps', type=int, default=4, help='served steps per cap trial')
    ap.add_argument('--seed', type=int, default=44)
    ap.add_argument('--verify-cells', type=int, default=32, help='cells per class in the FP64 verify sample')
    ap.add_argument('--cpu-slices', type=int, default=2, help='CPU sample: the cells of slices 0 .. k-1')
    ap.add_argument('--cpu-stride', type=int, default=1, help='CPU sample: every k-th of those cells per class')
    ap.add_argument('--cpu-threads', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--join-slots-log2', type=int, default=28, help='hash-join table size (0: skip the join)')
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch settings')
    ap.add_argument('--tune-cells', type=int, default=0,
                    help='cells of the autotuning sample (0: as many as a step, the batch a served stream launch '
                         'sees: the server streams every class of one n in one launch)')
    ap.add_argument('--probe-traffic', action='store_true',
                    help='run the HBM-traffic probe for rocprofv3 --pmc (calibration kernel + two served steps)')
    ap.add_argument('--tuning', default='', help='launch settings (a bench line\'s autotune JSON) instead of autotuning')
    ap.add_argument('--traffic-summary', nargs=2, metavar=('FETCH_CSV', 'WRITE_CSV'),
                    help='summarize two rocprofv3 --pmc counter_collection.csv files into profiles/')
    ap.add_argument('--serve-sweep', action='store_true',
                    help='the serving configuration: Poisson requests with n ~ U{24..45}, offered rate swept to '
                         'saturation (its own JSON line)')
    ap.add_argument('--sweep-requests', type=int, default=600, help='requests per offered rate')
    ap.add_argument('--sweep-ab-budget', type=int, default=1 << 20)
    ap.add_argument('--sweep-cell-blocks', default='', help='comma list of WZFH_CELL_BLOCKS values to sweep too')
    a = ap.parse_args()
    if a.traffic_summary:
        traffic_summary(*a.traffic_summary)
        return
    if a.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; cells shard by cell index mod G across GPUs '
                                   'with a final min over (rank, idx), which is not built here'}))
        sys.exit(2)
    if a.probe_traffic:
        probe_traffic(a)
        return
    if a.serve_sweep:
        serve_sweep(a)
        return

    t_setup = time.time()
    cls = classes(a.n)[:a.classes]
    probs, samples = [], []
    for i, sig in enumerate(cls):
        P = wz.Problem(a.n, sig, canon=1, order=1)
        probs.append(P)
        samples.append(sample_cells(P.live_cells(), a.cells_per_class, a.seed * 1000 + i))
        log(f'setup class {i + 1}/{len(cls)} {sig}: {P.num_cells()} cells')
    setup_s = time.time() - t_setup
    slices = [[(ci, c) for ci, smp in enumerate(samples) for c in smp[k::a.slices]] for k in range(a.slices)]
    search = dict(cand_cap=a.cand_cap, node_budget=a.node_budget)
    # launch settings (speed only): tuned once on the first class's cells of one slice (the batch
    # a class launch sees) and applied to every class; the tuner checks that each setting
    # returns identical results
    tune = None
    if a.tuning:
        apply_tuning(probs, a.tuning)
        tune = dict(json.loads(a.tuning), source='--tuning')
        log(f'launch settings given: {tune}')
    elif not a.no_tune:
        t0 = time.time()
        sample = sample_cells(probs[0].live_cells(), a.tune_cells or len(slices[0]), a.seed * 7919 + 1)
        t = probs[0].autotune(sample, a.top_k, a.ab_budget, **search)
        for P in probs[1:]:
            P.set_tuning(t)
        tune = dict(t.as_dict(), seconds=round(time.time() - t0, 1), sample_cells=len(sample))
        log(f'autotune: {tune}')
    log(f'setup: {len(cls)} classes, {sum(map(len, samples))} cells in {a.slices} slices, {setup_s:.1f} s')

    # a stream launch admits at most two steps' cells (measured on the n = 44 set with 4-12 steps
    # in flight: 5-10 % more cells/s than one step per launch; three or four steps per launch lose)
    max_batch = a.max_batch or 2 * max(len(sl) for sl in slices)
    if a.serve_caps and not a.tuning:
        tune = served_cap_tune(a, probs, slices, search, max_batch, tune)

    # ---- served (the measured value)
    served = {}
    server = wz.Server(probs, a.top_k, a.ab_budget, max_batch=max_batch, **search)
    try:
        if a.warmup:
            wsteps = [slices[w % a.slices] for w in range(a.warmup)]
            got, dts, _ = served_steps(server, wsteps, a.in_flight)
            for w, (reqs, g) in enumerate(zip(wsteps, got)):
                served.update({rq: r.result for rq, r in zip(reqs, g)})
                log(f'served warmup {w + 1}/{a.warmup} (slice {w % a.slices}): {len(reqs)} cells, {dts[w] * 1e3:.1f} ms')
        server.kernel_totals(reset=True)
        tsteps = [slices[s % a.slices] for s in range(a.steps)]
        t_timed = time.perf_counter()
        got, dts, timed_wall = served_steps(server, tsteps, a.in_flight)
        tot = server.kernel_totals()
        ncells, res_timed = 0, []
        for s, (reqs, g) in enumerate(zip(tsteps, got)):
            ncells += len(reqs)
            res_timed += [r.result for r in g]
            served.update({rq: r.result for rq, r in zip(reqs, g)})
            log(f'served step {s + 1}/{a.steps} (slice {s % a.slices}): {len(reqs)} cells, {dts[s] * 1e3:.1f} ms '
                f'from submit to last result')
        log(f'served: {a.steps} steps, {ncells} cells in {timed_wall:.3f} s ({a.in_flight} in flight)')
        step_s = timed_wall / max(1, a.steps)
        served_rate = ncells / timed_wall
        # ---- open-loop latency at a fraction of the served rate, over the whole set
        everything = [rq for sl in slices for rq in sl]
        rate = max(1.0, a.load * served_rate)
        lat, lat_wall = poisson_latency(server, everything, rate, a.latency_requests, a.seed)
        log(f'latency: {len(lat)} Poisson requests at {rate:.1f}/s in {lat_wall:.1f} s')
    finally:
        server.close()

    # ---- batched: one persistent launch per class on slice 0
    batched_s, nbatched = 0.0, 0
    for ci, P in enumerate(probs):
        cl = [c for cj, c in slices[0] if cj == ci]
        t0 = time.perf_counter()
        P.search(cl, a.top_k, a.ab_budget, batched=True, **search)
        batched_s += time.perf_counter() - t0
        nbatched += len(cl)
    log(f'batched: {len(probs)} launches, {nbatched} cells, {batched_s * 1e3:.1f} ms')

    # ---- plain: one cell per call
    plain_lat, mismatch, nplain = [], 0, 0
    for ci, P in enumerate(probs):
        for c in [c for cj, c in slices[0] if cj == ci][:a.plain_cells]:
            t0 = time.perf_counter()
            r = P.search([c], a.top_k, a.ab_budget, batched=False, **search)[0]
            plain_lat.append(time.perf_counter() - t0)
            nplain += 1
            if key(r) != key(served[(ci, c)]):
                mismatch += 1
    plain_rate = nplain / sum(plain_lat) if plain_lat else 0.0
    log(f'plain: {nplain} cells, {plain_rate:.2f} cells/s, mismatches vs served {mismatch}')

    # ---- the leaf test against FP64 on a verify sample (slice 1): every certain decision
    verify = verify_sample(a, probs, slices[1 % a.slices])
    log(f'verify: {verify}')
    util = stream_utilization(probs[0], sample_cells(probs[0].live_cells(), len(slices[0]), a.seed * 7919 + 2), a)
    log(f'stream utilization: {util}')

    streamed = sum(r.stream.cand for r in res_timed)
    completed = sum(r.completed for r in res_timed)
    dfs_nodes = sum(r.stream.dfs_nodes for r in res_timed)
    ab_nodes = sum(r.ab_nodes for r in res_timed)
    hits = sum(r.hit for r in res_timed)

    traffic = load_traffic()
    sums = [list(tot[k]) for k in range(2)]
    roof = {
        'stream': dict(kernel_roofline('wzfh::stream_kernel', OPS_PER_CD_NODE,
                                       f'{OPS_PER_CD_NODE} INT32 ops per counted C,D DFS node (rec() entry)',
                                       *sums[0], timed_wall, traffic.get('wzfh::stream_kernel')), utilization=util),
        'completer': kernel_roofline('wzfh::complete_kernel', ops_per_ab_node(a.n),
                                     f'{ops_per_ab_node(a.n)} INT32 ops (6L+n) per counted completer node',
                                     *sums[1], timed_wall, traffic.get('wzfh::complete_kernel')),
    }
    dominant = max(roof, key=lambda k: roof[k]['avg_launch_ms'] * roof[k]['launches_timed'])

    join, jroof = (None, {}) if not a.join_slots_log2 else join_scale(a)
    roof.update(jroof)

    cpu = None if a.no_cpu else cpu_baseline(a, cls, probs, [rq for k in range(a.cpu_slices) for rq in slices[k]], served)

    line = {
        'metric': 'cells/sec & C,D cands/sec at n=44 (served vs plain GPU); p50/p99 cell latency',
        'value': round(served_rate, 3), 'unit': 'cells/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(1e3 * step_s, 3), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32 (DFS, completer) + fp16x2/fp32 MFMA with fp64 recheck (leaf test)',
        'data': 'synthetic: seeded splitmix64 sample of the real n=44 cell lists',
        'config': {'workload': f'BS(45,44): {len(cls)} classes x {a.cells_per_class} seeded cells, A,B budget '
                               f'{a.ab_budget}, top-{a.top_k}; per-cell DFS-node budget {a.node_budget} (a policy of '
                               f'this build: n=44 cells are ~1e11 nodes); a step = 1/{a.slices} of the set '
                               f'(cells j = step mod {a.slices} of every class), served',
                   'n': a.n, 'classes': len(cls), 'cells_per_class': a.cells_per_class, 'slices': a.slices,
                   'cells_per_step': round(ncells / max(1, a.steps), 1), 'node_budget': a.node_budget,
                   'cand_cap': a.cand_cap, 'top_k': a.top_k, 'ab_budget': a.ab_budget, 'seed': a.seed,
                   'parallelism': 'single GPU: continuous batching, persistent work-donating launches'},
        'served_cells_per_s': round(served_rate, 3),
        'plain_cells_per_s': round(plain_rate, 3),
        'served_over_plain': round(served_rate / plain_rate, 2) if plain_rate else None,
        'batched_cells_per_s': round(nbatched / batched_s, 3) if batched_s else None,
        'cands_streamed_per_s': round(streamed / timed_wall, 1),
        'cands_completed_per_s': round(completed / timed_wall, 1),
        'dfs_nodes_per_s': round(dfs_nodes / timed_wall, 1),
        'ab_nodes_per_s': round(ab_nodes / timed_wall, 1),
        'step_latency_ms': {'p50': pct(dts, 50), 'max': round(1e3 * max(dts), 1), 'in_flight': a.in_flight},
        'served_latency_ms': {'p50': pct(lat, 50), 'p99': pct(lat, 99), 'offered_per_s': round(rate, 2),
                              'load': a.load, 'requests': len(lat)},
        'plain_latency_ms': {'p50': pct(plain_lat, 50), 'p99': pct(plain_lat, 99), 'cells': nplain},
        'hits': hits, 'plain_vs_served_mismatch': mismatch, 'false_rejects': verify.get('false_decisions'),
        'verify': verify, 'setup_s': round(setup_s, 1), 'timed_wall_s': round(timed_wall, 3),
        'roofline': dict(roof[dominant], dominant=True), 'rooflines': roof, 'join': join,
        'cpu_baseline': cpu, 'autotune': tune,
    }
    print(json.dumps(line), flush=True)
    for P in probs:
        P.close()


def serve_sweep(a):
    """Serving configuration: every signature class of every n in 24..45 is a device-resident
    problem; a request draws n uniformly, then a class of that n uniformly, then a cell of the
    class's live list (seeded), and searches it with --node-budget DFS nodes, top --top-k and
    A,B budget --sweep-ab-budget. First the closed-loop capacity (all requests at once), then
    open-loop Poisson streams at fractions of it up to past saturation: achieved cells/s and
    p50 / p99 latency per offered rate; the plain path (one cell per launch, one at a time) on a
    sample beside it. Served results are compared with the same cells searched alone."""
    rng = random.Random(a.seed * 31 + 5)
    t0 = time.time()
    probs, index = [], {}
    for n in range(24, 46):
        for sig in classes(n):
            index[(n, sig)] = len(probs)
            probs.append(wz.Problem(n, sig, canon=1, order=1, m6=1, thm211b=1, thm212=1))
        log(f'sweep setup n={n}: {len(probs)} problems, {time.time() - t0:.0f} s')
    live = [P.live_cells() for P in probs]
    pool = []
    for _ in range(4096):
        n = rng.randrange(24, 46)
        sig = rng.choice(classes(n))
        pi = index[(n, sig)]
        pool.append((pi, live[pi][rng.randrange(len(live[pi]))]))
    setup_s = time.time() - t0
    search = dict(node_budget=a.node_budget, cand_cap=a.cand_cap)
    server = wz.Server(probs, a.top_k, a.sweep_ab_budget, max_batch=a.max_batch or 4096, **search)
    rows, served, variants = [], {}, {}
    try:
        for cb in [None] + [int(x) for x in a.sweep_cell_blocks.split(',') if x.strip()]:
            if cb is not None:
                os.environ['WZFH_CELL_BLOCKS'] = str(cb)
            got, dt = served_pass(server, pool[:a.sweep_requests])   # warm-up and closed-loop capacity
            got, dt = served_pass(server, pool[:a.sweep_requests])
            cap_cb = a.sweep_requests / dt
            served.update({rq: r.result for rq, r in zip(pool, got)})
            log(f'cell blocks {cb}: closed-loop capacity {cap_cb:.1f} cells/s ({a.sweep_requests} requests in {dt:.2f} s)')
            rows_cb = []
            for frac in (0.2, 0.4, 0.6, 0.8, 1.0, 1.25):
                rate = frac * cap_cb
                lat, wall = poisson_latency(server, pool, rate, a.sweep_requests, a.seed + int(100 * frac))
                rows_cb.append({'offered_frac': frac, 'offered_per_s': round(rate, 2),
                                'achieved_per_s': round(len(lat) / wall, 2), 'p50_ms': pct(lat, 50),
                                'p99_ms': pct(lat, 99)})
                log(f'cell blocks {cb}: offered {rate:.1f}/s: {rows_cb[-1]}')
            if cb is None:
                capacity, rows = cap_cb, rows_cb
            else:
                variants[cb] = {'closed_loop_cells_per_s': round(cap_cb, 3), 'sweep': rows_cb}
        os.environ.pop('WZFH_CELL_BLOCKS', None)
    finally:
        server.close()
    plain_lat, mism = [], 0
    for pi, c in pool[:48]:
        t1 = time.perf_counter()
        r = probs[pi].search([c], a.top_k, a.sweep_ab_budget, batched=False, **search)[0]
        plain_lat.append(time.perf_counter() - t1)
        if (pi, c) in served and key(r) != key(served[(pi, c)]):
            mism += 1
    plain_rate = len(plain_lat) / sum(plain_lat)
    line = {'metric': 'served cells/s and p50/p99 latency vs offered rate, n ~ U{24..45}',
            'value': round(capacity, 3), 'unit': 'cells/s', 'n_gpus': 1,
            'config': {'workload': f'serving: n ~ U{{24..45}}, class ~ U(classes(n)), seeded cells; node budget '
                                   f'{a.node_budget} per cell, top-{a.top_k}, A,B budget {a.sweep_ab_budget}',
                       'problems': len(probs), 'requests_per_rate': a.sweep_requests},
            'closed_loop_cells_per_s': round(capacity, 3), 'plain_cells_per_s': round(plain_rate, 3),
            'served_over_plain': round(capacity / plain_rate, 2),
            'plain_latency_ms': {'p50': pct(plain_lat, 50), 'p99': pct(plain_lat, 99), 'cells': len(plain_lat)},
            'sweep': rows, 'cell_blocks_variants': variants, 'plain_vs_served_mismatch': mism,
            'setup_s': round(setup_s, 1)}
    print(json.dumps(line), flush=True)
    for P in probs:
        P.close()


def join_scale(a):
    """The n = 44 hash join at the configuration's table size: 2^28 slots (44-int16 key vector, tag,
    order and C,D per slot: 120 B, 32 GB) filled to load 1/2 with random C,D pairs and probed with as
    many random A,B pairs, the banked BS(45,44)'s images planted on both sides (their hits and the
    solution list are checked). HBM-bound: algorithmic bytes per insert = the 120 B entry written +
    the 24 B of input (pair and order) read + one 8 B tag read; per probe = the 16 B pair read + the
    expected tag reads of an unsuccessful linear probe at load 1/2, (1 + 1/(1 - 1/2)^2) / 2 = 2.5
    slots x 8 B. Device time from HIP events around the table kernels."""
    from bs45_quantum_explorer_amd.workload import join_workload
    sol = next(r for r in json.loads((ROOT / 'tests' / 'golden' / 'solutions.json').read_text())
               if r['n'] == 44 and not r['quarantined'])
    lg = a.join_slots_log2
    nrand = (1 << (lg - 1)) - 8
    t0 = time.time()
    cd, ab, pcd, pab = join_workload(44, (sol['A'], sol['B'], sol['C'], sol['D']), nrand, seed=4428)
    J = wz.Join(44, tuple(sol['sig']))
    try:
        r = J.build_from(cd, slots_log2=lg)
        keys = r.distinct_keys
        r = J.probe_from(ab)
        hits = r.hits
        sols = J.solutions()
        bms, pms = J.kernel_ms()
    finally:
        J.close()
    ok = hits == len(pab) and keys == nrand + 1 and sorted(sols) == sorted((x, y, c, d) for x, y in pab for c, d in pcd)
    ins_b = 120 + 24 + 8
    prb_b = 16 + 2.5 * 8
    out = {'slots': 1 << lg, 'entry_bytes': 120, 'table_gb': round((1 << lg) * 120 / 1e9, 2), 'inserts': len(cd),
           'distinct_keys': keys, 'probes': len(ab), 'hits': hits, 'planted_solutions_found': ok,
           'build_ms': round(bms, 3), 'probe_ms': round(pms, 3),
           'inserts_per_s': round(len(cd) / (bms / 1e3), 1) if bms else None,
           'probes_per_s': round(len(ab) / (pms / 1e3), 1) if pms else None,
           'wall_s': round(time.time() - t0, 1)}
    roof = {}
    for name, nunits, per, ms in (('join_build', len(cd), ins_b, bms), ('join_probe', len(ab), prb_b, pms)):
        gbs = nunits * per / (ms / 1e3) / 1e9 if ms else 0.0
        roof[name] = {'bound': 'hbm', 'achieved': round(gbs, 2), 'peak': HBM_PEAK_BPS / 1e9, 'unit': 'GB/s',
                      'frac': round(gbs * 1e9 / HBM_PEAK_BPS, 5), 'traffic': None,
                      'per_unit': f'{per:g} algorithmic bytes per {"insert" if name == "join_build" else "probe"}',
                      'units_per_launch': nunits, 'avg_launch_ms': round(ms, 3)}
    log(f'join 2^{lg} slots: {out}')
    return out, roof


def verify_sample(a, probs, reqs):
    """Streams --verify-cells cells per class of one slice with WZFH_HALL_VERIFY=1: the FP64 test
    also runs on every leaf and every certain matrix-core decision is compared with it."""
    keep = {k: os.environ.get(k) for k in ('WZFH_HALL_VERIFY', 'WZFH_TELEMETRY')}
    os.environ['WZFH_HALL_VERIFY'] = '1'
    os.environ['WZFH_TELEMETRY'] = '1'
    leaves = uncertain = bad = cells = 0
    try:
        for ci, P in enumerate(probs):
            cl = [c for cj, c in reqs if cj == ci][:a.verify_cells]
            P.stream(cl, node_budget=a.node_budget, want_cands=False)
            leaves += P.last_kernel(5)[1]
            uncertain += P.last_kernel(6)[1]
            bad += P.last_kernel(7)[1]
            cells += len(cl)
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return {'cells': cells, 'leaves_tested': leaves, 'decided_in_fp64': uncertain, 'false_decisions': bad,
            'margin': round(probs[0].last_kernel(10)[0], 4)}


def stream_utilization(P, cells, a):
    """One stream launch of a step-sized batch of one class with the kernel's counters on
    (WZFH_TELEMETRY=1): DFS moves its lanes made, main-loop passes of its waves, the share of the
    lane slots of those passes that made a move, and counted nodes (inside the cells' budgets) per
    move. A move enters one live child; the dead children it skips are counted nodes too (the
    reference enters them and prunes), so a fully useful walk counts more than one node per move,
    and moves beyond the cuts (speculation, abandoned) lower the ratio."""
    keep = os.environ.get('WZFH_TELEMETRY')
    os.environ['WZFH_TELEMETRY'] = '1'
    try:
        st, _, _ = P.stream(cells, node_budget=a.node_budget, want_cands=False)
        moves, passes = P.last_kernel(11)[1], P.last_kernel(12)[1]
        ms = P.last_kernel(0)[0]
    finally:
        if keep is None:
            os.environ.pop('WZFH_TELEMETRY', None)
        else:
            os.environ['WZFH_TELEMETRY'] = keep
    inner = P.tuning().inner_moves
    counted = sum(x.dfs_nodes for x in st)
    return {'cells': len(cells), 'launch_ms': round(ms, 2), 'lane_moves': moves, 'wave_passes': passes,
            'moves_per_wave_pass': round(moves / passes, 2) if passes else None,
            'lane_slot_use': round(moves / (passes * 64 * inner), 5) if passes else None,
            'counted_nodes_per_move': round(counted / moves, 4) if moves else None}


def cpu_cores():
    """(threads used, what the host reports). The box gives one GPU a 16-core share."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return affinity, {'nproc_affinity': affinity, 'os_cpu_count': os.cpu_count()}


def cpu_baseline(a, cls, probs, reqs, served):
    """The oracle port (oracle/wz_oracle.cpp, wzo_cell_search) on the host cores: the same cell
    unit on a bounded sample of the same cells (every --cpu-stride-th cell of the first --cpu-slices
    slices, per class), one cell per thread, results compared with the served ones cell by cell."""
    sys.path.insert(0, str(ROOT / 'tests'))
    try:
        from oracle_lib import OracleProblem
    except Exception as e:  # oracle not built: no baseline
        return {'error': f'oracle unavailable: {e}'}
    avail, host = cpu_cores()
    threads = a.cpu_threads or min(16, avail)
    work, oprobs = [], []
    t_setup = time.time()
    for ci, (P_gpu, sig) in enumerate(zip(probs, cls)):
        pick = [c for cj, c in reqs if cj == ci][::a.cpu_stride]
        O = OracleProblem(a.n, sig, canon=1, order=1, cells=[P_gpu.cell(pi)[:2] for pi in pick])
        oprobs.append(O)
        work += [(ci, O, j, c) for j, c in enumerate(pick)]
    setup = time.time() - t_setup

    def one(item):
        ci, O, j, c = item
        r, _ = O.search(j, a.cand_cap, a.node_budget, a.top_k, a.ab_budget)
        return ci, c, r

    t0 = time.perf_counter()
    cands = nodes = mism = compared = 0
    with ThreadPoolExecutor(threads) as ex:
        for k, (ci, c, r) in enumerate(ex.map(one, work)):
            cands += r.cand
            nodes += r.dfs_nodes
            if (k + 1) % 64 == 0:
                log(f'cpu baseline: {k + 1}/{len(work)} cells')
            g = served.get((ci, c))
            if g is None:   # a cell the served steps did not reach: timed, not compared
                continue
            compared += 1
            want = (r.cand, r.dfs_nodes, r.completed, r.hit, r.hit_rank, r.ab_nodes, r.clean_no, r.aborted)
            got = (g.stream.cand, g.stream.dfs_nodes, g.completed, g.hit, g.hit_rank, g.ab_nodes, g.clean_no,
                   g.aborted)
            if want != got or (r.hit and (list(wz.unpack(g.A, a.n + 1)) != list(r.A[:a.n + 1]) or
                                          list(wz.unpack(g.B, a.n + 1)) != list(r.B[:a.n + 1]))):
                mism += 1
    dt = time.perf_counter() - t0
    model = _cpu_model()
    return {'value': round(len(work) / dt, 4), 'unit': 'cells/s', 'cores': threads, 'kind': 'port',
            'sample': f'{len(work)} cells ({"every" if a.cpu_stride == 1 else f"every {a.cpu_stride}th"} cell of slices '
                      f'0..{a.cpu_slices - 1} of each class) of the same '
                      f'workload and unit, one cell per thread',
            'mismatches_vs_served': mism, 'cells_compared': compared,
            'cands_per_s': round(cands / dt, 1), 'dfs_nodes_per_s': round(nodes / dt, 1), 'seconds': round(dt, 2),
            'setup_s': round(setup, 1), 'cpu_model': model, 'host': host,
            'port_vs_reference': 'profiles/cpu_port_vs_reference.md'}


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def served_cap_tune(a, probs, slices, search, max_batch, tune):
    """The completer's item charge cap for the served pipeline. The library's autotuner times one
    class's batched search, where a launch waits for its longest item; served steps of all classes
    overlap those tails, so a lower cap (shorter items, more launches) can serve faster. Each cap of
    --serve-caps (and the autotuner's) serves --serve-tune-steps steps (slices 0, 1, ...) with
    --in-flight steps in flight on a fresh server; the fastest is kept on every problem. Every cap
    must return the same per-cell results as the first, else this raises."""
    t0 = time.time()
    base = probs[0].tuning().as_dict()
    caps = [base['complete_cap_log2']] + [int(x) for x in a.serve_caps.split(',') if x.strip()]
    caps = list(dict.fromkeys(caps))
    steps = [slices[k % a.slices] for k in range(a.serve_tune_steps)]
    ref, rates = {}, {}
    for cap in caps:
        apply_tuning(probs, json.dumps(dict(base, complete_cap_log2=cap)))
        server = wz.Server(probs, a.top_k, a.ab_budget, max_batch=max_batch, **search)
        try:
            got, _, wall = served_steps(server, steps, a.in_flight)
        finally:
            server.close()
        for reqs, g in zip(steps, got):
            for rq, r in zip(reqs, g):
                if ref.setdefault(rq, key(r.result)) != key(r.result):
                    raise RuntimeError(f'served cap tuning: cap 2^{cap} changes the result of cell {rq}')
        rates[cap] = round(sum(map(len, steps)) / wall, 1)
        log(f'served cap 2^{cap}: {rates[cap]} cells/s')
    best = max(rates, key=rates.get)
    apply_tuning(probs, json.dumps(dict(base, complete_cap_log2=best)))
    out = dict(tune or {}, complete_cap_log2=best)
    out['served_cap_trials'] = {str(k): v for k, v in rates.items()}
    out['served_cap_seconds'] = round(time.time() - t0, 1)
    return out


def apply_tuning(probs, text):
    """Sets the launch settings named in `text` (a bench line's autotune JSON) on every problem."""
    want = json.loads(text)
    for P in probs:
        t = P.tuning()
        for k, v in want.items():
            if hasattr(t, k) and k not in ('trials', 'stream_ms', 'complete_ms'):
                setattr(t, k, v)
        P.set_tuning(t)


def probe_traffic(a):
    """For `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`: first an elementwise kernel that
    reads and writes exactly 256 MiB each (the counters' calibration), then two served steps of
    the benchmark (slices 0 and 1 of the n = 44 set, every class) with the launch settings given
    by --tuning (the autotune choice of a bench line), so every stream and completer dispatch in
    the profile is one of the timed region's kind."""
    import torch
    x = torch.zeros(64 << 20, dtype=torch.int32, device='cuda')
    y = torch.empty_like(x)
    torch.add(x, 1, out=y)
    torch.cuda.synchronize()
    cls = classes(a.n)[:a.classes]
    probs, samples = [], []
    for i, sig in enumerate(cls):
        P = wz.Problem(a.n, sig, canon=1, order=1)
        probs.append(P)
        samples.append(sample_cells(P.live_cells(), a.cells_per_class, a.seed * 1000 + i))
    if a.tuning:
        apply_tuning(probs, a.tuning)
    steps = [[(ci, c) for ci, smp in enumerate(samples) for c in smp[k::a.slices]] for k in range(2)]
    server = wz.Server(probs, a.top_k, a.ab_budget, max_batch=a.max_batch or 2 * max(map(len, steps)),
                       cand_cap=a.cand_cap, node_budget=a.node_budget)
    try:
        served_steps(server, steps, a.in_flight)   # as the timed region serves them
        tot = server.kernel_totals()
    finally:
        server.close()
    print(json.dumps({'calibration_bytes_each_way': 256 << 20, 'steps': 2,
                      'stream': {'ms': tot[0][0], 'units': tot[0][1], 'launches': tot[0][2]},
                      'completer': {'ms': tot[1][0], 'units': tot[1][1], 'launches': tot[1][2]},
                      'tuning': a.tuning}), flush=True)
    for P in probs:
        P.close()


def traffic_summary(fetch_csv, write_csv):
    """Per-kernel average FETCH_SIZE / WRITE_SIZE per dispatch from two rocprofv3 --pmc runs of
    --probe-traffic, scaled by what the counters read for the calibration kernel."""
    import csv

    def per_kernel(path, counter):
        acc = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get('Counter_Name') != counter:
                    continue
                name = row['Kernel_Name']
                acc.setdefault(name, []).append(float(row['Counter_Value']))
        return acc

    fetch, write = per_kernel(fetch_csv, 'FETCH_SIZE'), per_kernel(write_csv, 'WRITE_SIZE')
    cal_f = [v for k, vs in fetch.items() if 'elementwise' in k.lower() or 'add' in k.lower() for v in vs]
    cal_w = [v for k, vs in write.items() if 'elementwise' in k.lower() or 'add' in k.lower() for v in vs]
    ff = (256 << 10) / max(cal_f) if cal_f else 1.0
    fw = (256 << 10) / max(cal_w) if cal_w else 1.0
    out = {'source': 'rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --probe-traffic, calibrated on a '
                     '256 MiB elementwise kernel', 'fetch_factor': round(ff, 4), 'write_factor': round(fw, 4),
           'kernels': {}}
    for short in ('stream_kernel', 'complete_kernel', 'complete_reduce_kernel', 'stream_finalize'):
        fk = [v for k, vs in fetch.items() if short in k for v in vs]
        wk = [v for k, vs in write.items() if short in k for v in vs]
        if not fk or not wk:
            continue
        fb, wb = 1024 * ff * float(np.mean(fk)), 1024 * fw * float(np.mean(wk))
        out['kernels'][f'wzfh::{short}'] = {'dispatches': len(fk), 'fetch_kib_raw': round(float(np.mean(fk)), 1),
                                            'write_kib_raw': round(float(np.mean(wk)), 1),
                                            'bytes_per_launch': int(fb + wb), 'source': out['source']}
    dst = ROOT / 'profiles' / 'kernel_traffic_n44.json'
    dst.write_text(json.dumps(out, indent=1) + '\n')
    print(json.dumps(out, indent=1))


if __name__ == '__main__':
    main()
