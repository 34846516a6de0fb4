"""Autotuner (wzfh_autotune): the launch settings it picks change speed only. Results of the
search unit after tuning equal the results before it, every trial inside the tuner is checked
against the first (the tuner raises otherwise), and the picked values are in range."""
import numpy as np
import pytest

import bs45_quantum_explorer_amd as wz

pytestmark = pytest.mark.gpu


def _key(r):
    s = r.stream
    return (s.cand, s.dfs_nodes, s.leaves, s.hall_ok, r.completed, r.hit, r.hit_rank, r.ab_nodes, r.clean_no,
            r.aborted, r.A, r.B)


def test_autotune_keeps_results_and_picks_valid_settings():
    P = wz.Problem(29, (0, 6, 1, 9), canon=1, order=1)
    try:
        live = P.live_cells()
        cells = live[:: max(1, len(live) // 16)][:16]
        unit = dict(top_k=8, budget=1 << 14, node_budget=1 << 16)
        before = [_key(r) for r in P.search(cells, unit['top_k'], unit['budget'], node_budget=unit['node_budget'])]
        t = P.autotune(cells, unit['top_k'], unit['budget'], node_budget=unit['node_budget'])
        assert t.trials >= 10 and t.stream_ms > 0
        assert t.inner_moves in (4, 8, 16, 32, 64) and t.donate_hops in (0, 1, 2)
        assert t.static_window in (2, 4, 8, 16) and t.complete_split in (3, 4, 5) and t.idle_sleep in (0, 16, 64, 128, 256)
        assert t.blocks_per_cu in (1, 2, 3, 4)
        got = P.tuning()
        assert (got.inner_moves, got.donate_hops, got.static_window, got.complete_split, got.idle_sleep,
                got.blocks_per_cu) == (t.inner_moves, t.donate_hops, t.static_window, t.complete_split, t.idle_sleep,
                                       t.blocks_per_cu)
        after = [_key(r) for r in P.search(cells, unit['top_k'], unit['budget'], node_budget=unit['node_budget'])]
        assert after == before
        bad = wz.Tuning(0, 1, 4, 3, 4)
        with pytest.raises(wz.WzfhError):
            P.set_tuning(bad)
    finally:
        P.close()
