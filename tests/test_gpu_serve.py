"""Serving: requests of mixed n admitted by continuous batching give exactly the result of the
same cell searched alone (one cell per launch), and every request is answered once, with one
or several completion workers per problem."""
import random
import time

import pytest

import bs45_quantum_explorer_amd as wz

pytestmark = pytest.mark.gpu

SPEC = [(13, (2, 4, 3, 5)), (19, (0, 2, 5, 7)), (29, (0, 6, 1, 9))]


def key(r):
    s = r.stream
    return (s.cand, s.dfs_nodes, s.hall_ok, r.completed, r.hit, r.hit_rank, r.hit_idx, r.ab_nodes, r.clean_no,
            r.aborted, r.A, r.B, r.C, r.D)


@pytest.mark.parametrize('node_budget,cand_cap,workers', [(200000, 0, 0), (0, 64, 0), (200000, 0, 1), (200000, 0, 3)])
def test_served_results_equal_alone(node_budget, cand_cap, workers):
    probs = [wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1) for n, sig in SPEC]
    rng = random.Random(7)
    reqs = []
    for pi, P in enumerate(probs):
        live = P.live_cells()
        for c in rng.sample(live, min(24, len(live))):
            reqs.append((pi, c))
    rng.shuffle(reqs)
    top_k, budget = 8, 20000
    alone = {}
    for pi, c in reqs:
        alone[(pi, c)] = key(probs[pi].search([c], top_k, budget, cand_cap=cand_cap, node_budget=node_budget,
                                             batched=False)[0])
    S = wz.Server(probs, top_k, budget, cand_cap=cand_cap, node_budget=node_budget, max_batch=16,
                  complete_workers=workers)
    try:
        tickets = {}
        for pi, c in reqs:
            tickets[S.submit(pi, c)] = (pi, c)
        got = {}
        deadline = time.monotonic() + 90
        while len(got) < len(reqs):
            assert time.monotonic() < deadline, f'{len(got)} of {len(reqs)} answered'
            for r in S.poll(max_results=256, min_results=1, timeout_ms=60000):
                assert r.status == 0
                assert r.ticket not in got
                got[r.ticket] = r
        for t, (pi, c) in tickets.items():
            r = got[t]
            assert (r.problem, r.cell) == (pi, c)
            assert key(r.result) == alone[(pi, c)], (pi, c)
            assert r.finish_s >= r.start_s >= r.submit_s - 1e-6
    finally:
        S.close()
        for P in probs:
            P.close()
