"""The measured unit (wzfh_search_cells: stream -> flatness top-K -> completion under the
cell's profile rows) against the oracle's wzo_cell_search on the workloads the benchmark
configurations name, batched, plain (one cell per launch) and served.

  config 3  BS(30,29), class (0,6,1,9), 4,096 seeded cells in one persistent launch, whole
            cells, top-64, A,B budget 2^20; a spread subset is checked against the oracle.
  config 4  BS(45,44), all 12 classes, seeded cells of the bench's sample, 2^22 DFS-node
            budget per cell (a serving policy), top-32, A,B budget 2^24; the matrix-core leaf
            test's false-reject / false-accept count against FP64 over a sample of every class.
  config 5  served requests of mixed n in 24..45, class and cell seeded per request, 2^22
            DFS-node budget, top-32; every result equal to the oracle and to the cell alone.

Results must be identical field by field: streamed candidates, DFS nodes, leaves (when the
cell was not cut), capped flag, completions counted (the drain stops at the first hit), hit
rank and stream index, completer nodes, clean / aborted counts, and the hit's A, B, C, D.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bs45_quantum_explorer_amd as wz
from bs45_quantum_explorer_amd.workload import classes, sample_cells
from oracle_lib import OracleProblem

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def oracle_for(P, cells):
    """The oracle over exactly these cells of the GPU problem (same flags, rows built per cell)."""
    c = P.cfg
    return OracleProblem(P.n, P.sig, m6=c.m6, thm211b=c.thm211b, thm212=c.thm212, canon=c.canon, order=c.order,
                         rev=c.rev, endpos=c.endpos, outer_k=c.outer_k, cells=[P.cell(pi)[:2] for pi in cells])


def oracle_search(O, ncells, cand_cap, node_budget, top_k, budget):
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda j: O.search(j, cand_cap, node_budget, top_k, budget)[0], range(ncells)))


def same(n, r, o):
    """Field-by-field comparison of a wzfh_search_result with a wzo_search_result."""
    s = r.stream
    got = (s.cand, s.dfs_nodes, s.capped, s.hall_ok, r.completed, r.hit, r.hit_rank, r.hit_idx, r.ab_nodes,
           r.clean_no, r.aborted, r.internal)
    want = (o.cand, o.dfs_nodes, o.capped, o.hall_ok, o.completed, o.hit, o.hit_rank, o.hit_idx, o.ab_nodes,
            o.clean_no, o.aborted, o.internal)
    if got != want:
        return f'{got} != {want}'
    if not o.budget_cut and s.leaves != o.leaves:
        return f'leaves {s.leaves} != {o.leaves}'
    if o.hit:
        for k, ln in (('A', n + 1), ('B', n + 1), ('C', n), ('D', n)):
            if list(wz.unpack(getattr(r, k), ln)) != list(getattr(o, k)[:ln]):
                return f'hit {k} differs'
    return None


def test_config3_n29_one_launch_matches_oracle():
    n, sig = 29, (0, 6, 1, 9)
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1)
    live = P.live_cells()
    assert len(live) == 9887
    cells = sample_cells(live, 4096, 3029)
    print(f'config 3: {len(cells)} cells in one launch', flush=True)
    res = P.search(cells, 64, 1 << 20, batched=True)
    ms, units = P.last_kernel(0)
    assert units == sum(r.stream.dfs_nodes for r in res)
    assert all(not r.stream.capped for r in res)
    # a spread subset through the oracle: every 128th cell of the sample
    pick = list(range(0, len(cells), 128))
    O = oracle_for(P, [cells[j] for j in pick])
    want = oracle_search(O, len(pick), 0, 0, 64, 1 << 20)
    bad = [(cells[j], same(n, res[j], w)) for j, w in zip(pick, want) if same(n, res[j], w)]
    assert not bad, bad[:4]
    assert sum(w.completed for w in want) > 0 and sum(w.cand for w in want) > 0
    P.close()


@pytest.fixture(scope='module')
def n44():
    """The 12 n = 44 classes with the bench's seeded cell samples (1,024 per class, seed 44)."""
    probs, samples, busy = [], [], []
    for i, sig in enumerate(classes(44)):
        P = wz.Problem(44, sig, canon=1, order=1)
        probs.append(P)
        sample = sample_cells(P.live_cells(), 1024, 44 * 1000 + i)
        samples.append(sample)
        # cells of the first 256 that stream candidates under the 2^22 budget (most n = 44
        # cells stream none that early)
        st, _, _ = P.stream(sample[:256], node_budget=1 << 22, want_cands=False)
        busy.append([c for c, x in zip(sample[:256], st) if x.cand > 0])
    yield probs, samples, busy
    for P in probs:
        P.close()


def test_config4_n44_search_matches_oracle_every_class(n44):
    probs, samples, busy = n44
    nb, top, ab = 1 << 22, 32, 1 << 24
    checked = completed = 0
    for ci, (P, sample) in enumerate(zip(probs, samples)):
        # 16 cells spread over the class's sample and up to 6 that stream candidates
        spread = sample[::64]
        cells = sorted(set(spread) | set(busy[ci][:6]))
        print(f'config 4 class {ci} {P.sig}: {len(cells)} cells', flush=True)
        res = P.search(cells, top, ab, node_budget=nb, batched=True)
        O = oracle_for(P, cells)
        want = oracle_search(O, len(cells), 0, nb, top, ab)
        bad = [(c, same(44, r, w)) for c, r, w in zip(cells, res, want) if same(44, r, w)]
        assert not bad, (P.sig, bad[:4])
        # the plain path (one cell per launch) on the cells with candidates
        with_cands = [j for j, w in enumerate(want) if w.cand][:2]
        if with_cands:
            plain = P.search([cells[j] for j in with_cands], top, ab, node_budget=nb, batched=False)
            for j, r in zip(with_cands, plain):
                assert same(44, r, want[j]) is None, (P.sig, cells[j])
        checked += len(cells)
        completed += sum(w.completed for w in want)
        print(f"  {sum(w.cand for w in want)} candidates, {sum(w.completed for w in want)} completions", flush=True)
        O.close()
    assert checked >= 12 * 16 and completed > 0


def test_config4_n44_stream_order_matches_oracle(n44):
    """The candidate stream itself at n = 44 (L > 32: the leaf test's second K block of the
    matrix-core product), in DFS order, under the node budget."""
    probs, samples, busy = n44
    nb = 1 << 22
    total = 0
    for ci, (P, sample) in enumerate(zip(probs, samples)):
        cells = sorted(set(sample[7::256]) | set(busy[ci][-4:]))
        stats, offs, cd = P.stream(cells, node_budget=nb)
        O = oracle_for(P, cells)
        for j in range(len(cells)):
            got, cands, st = O.stream_budget(j, 0, nb, 1 << 16)
            assert (stats[j].cand, stats[j].dfs_nodes) == (got, st.dfs_nodes)
            mine = cd[offs[j]:offs[j + 1]]
            want = np.array([[wz.pack(c[0]), wz.pack(c[1])] for c in cands], dtype=np.uint64).reshape(-1, 2)
            assert np.array_equal(mine, want), (P.sig, cells[j])
            total += got
        O.close()
    assert total > 0


def test_config4_leaf_test_has_zero_false_rejects(n44):
    """WZFH_HALL_VERIFY=1 also runs the FP64 test on every leaf; every leaf the matrix-core test
    decided with certainty must agree with it (kind 7 counts disagreements)."""
    probs, samples, _ = n44
    old = {k: os.environ.get(k) for k in ('WZFH_HALL_VERIFY', 'WZFH_TELEMETRY')}
    os.environ['WZFH_HALL_VERIFY'] = '1'
    os.environ['WZFH_TELEMETRY'] = '1'
    try:
        leaves = 0
        for P, sample in zip(probs, samples):
            P.stream(sample[:128], node_budget=1 << 22, want_cands=False)
            _, tested = P.last_kernel(5)
            _, disagree = P.last_kernel(7)
            assert disagree == 0, P.sig
            leaves += tested
        print(f'config 4: {leaves} leaves tested, 0 disagreements with FP64', flush=True)
        assert leaves > 0
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_config5_served_mixed_n_matches_oracle_and_alone():
    import random
    rng = random.Random(2445)
    probs, reqs = [], []
    for n in range(24, 46):
        sig = rng.choice(classes(n))
        P = wz.Problem(n, sig, canon=1, order=1, m6=1, thm211b=1, thm212=1)
        probs.append(P)
        live = P.live_cells()
        for c in sample_cells(live, 2, rng.getrandbits(32)):
            reqs.append((len(probs) - 1, c))
    nb, top, ab = 1 << 22, 32, 1 << 20
    print(f'config 5: {len(reqs)} requests over n = 24..45', flush=True)
    srv = wz.Server(probs, top, ab, node_budget=nb, max_batch=64)
    try:
        tick = {srv.submit(p, c): i for i, (p, c) in enumerate(reqs)}
        got = [None] * len(reqs)
        while any(g is None for g in got):
            for r in srv.poll(max_results=256, min_results=1, timeout_ms=5000):
                assert r.status == 0
                got[tick[r.ticket]] = r.result
    finally:
        srv.close()
    for pidx, P in enumerate(probs):
        mine = [(i, c) for i, (p, c) in enumerate(reqs) if p == pidx]
        cells = [c for _, c in mine]
        O = oracle_for(P, cells)
        want = oracle_search(O, len(cells), 0, nb, top, ab)
        alone = P.search(cells, top, ab, node_budget=nb, batched=False)
        for (i, c), w, a in zip(mine, want, alone):
            assert same(P.n, got[i], w) is None, (P.n, P.sig, c, same(P.n, got[i], w))
            assert same(P.n, a, w) is None, (P.n, P.sig, c)
        O.close()
    for P in probs:
        P.close()
