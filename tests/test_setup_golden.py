"""Host setup (the product's per-(n, class) cell list, orbit canon, order, dead-cell rows)
against the reference's cell lists (tests/golden/stream_cells.json)."""
import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import load, stream_runs

RUNS = stream_runs()


@pytest.mark.parametrize('n', sorted(set(r['n'] for r in RUNS)))
def test_live_cell_list_matches_reference(n):
    for rec in (r for r in RUNS if r['n'] == n):
        P = wz.Problem(n, rec['sig'], device=False, m6=1, thm211b=1, thm212=1, canon=rec['canon'], rev=rec['rev'],
                       endpos=rec['endpos'], order=rec['order'])
        try:
            assert P.num_cells() == rec['profiles']
            assert P.live_cells() == [c[0] for c in rec['cells']]
        finally:
            P.close()


def test_n44_class_list_size_matches_reference_probe():
    # the reference's n=44 (1,7,8,8) setup reports 977,601 mod-6 cells
    P = wz.Problem(44, (1, 7, 8, 8), device=False)
    assert P.num_cells() == 977601
    px, py, kept, rows = P.cell(0)
    assert sum(px) == 8 and sum(py) == 8


def test_join_cd_profiles_match_reference():
    """survive_profiles on the C,D side (m = 3 with the mod-6 tighten; the JOIN22 and
    WZ_FH_M6=0 cell list) against the reference's '[profiles] ... C,D side surviving' counts,
    and every solution the reference's JOIN22 printed has NPAF = 0 at every shift."""
    for r in load('joins.json'):
        st = r['settings']
        P = wz.Problem(r['n'], r['sig'], device=False, m6=0, canon=0, order=0,
                       thm211b=int(st.get('WZ_THM211B', 0)), thm212=int(st.get('WZ_THM212', 0)))
        assert P.num_cells() == r['cd_profiles'], (r['n'], r['sig'], st)
        P.close()
        if r['found']:
            assert wz.npaf_max(r['n'], r['A'], r['B'], r['C'], r['D']) == 0


ORDER3 = load('order3_digests.json')


@pytest.mark.parametrize('n', sorted({r['n'] for r in ORDER3}))
def test_order3_list_digest_matches_reference(n):
    """WZ_FH_PROF_ORDER=3 (wz_match.cpp:1809-1872): the total order of the cell list under the
    orbit canon (32- or 64-element group) gives the reference's cell count, kept-orbit count and
    both FNV-1a digests ("<list>:<kept>") that WZ_FH_EXPECT_DIGEST checks."""
    for r in (x for x in ORDER3 if x['n'] == n):
        st = r['settings']
        P = wz.Problem(n, r['sig'], device=False, m6=1, canon=1, order=3, orbit_q=r['orbit_q'],
                       thm211b=int(st.get('WZ_THM211B', 0)), thm212=int(st.get('WZ_THM212', 0)))
        kept = sum(P.cell(pi)[2] for pi in range(P.num_cells()))
        assert (P.num_cells(), kept, P.digest()) == (r['cells'], r['kept'], r['digest']), (n, r['sig'], r['orbit_q'])
        P.close()


QPRUNE = load('qprune_cells.json')


@pytest.mark.parametrize('rec', QPRUNE, ids=[f"{r['n']}-{tuple(r['sig'])}-ep{r['endpos']}-o{r['order']}" for r in QPRUNE])
def test_qprune_cell_list_matches_reference(rec):
    """WZ_FH_ORBIT_QPRUNE (wz_match.cpp:1735-1800): the orbits the Q-closure prune proves empty, the
    kept orbits left, the order-3 digest over them and the cells the arm streams equal the
    reference's ([qprune] counts, kept_orbits, [order] digest, CELLSIZE cells)."""
    P = wz.Problem(rec['n'], rec['sig'], device=False, m6=1, thm211b=1, thm212=1, canon=1, orbit_q=1,
                   orbit_qprune=1, endpos=rec['endpos'], order=rec['order'])
    try:
        ran, st = P.qprune_stats()
        assert ran and st == rec['qprune']
        assert P.num_cells() == rec['profiles']
        assert sum(P.cell(pi)[2] for pi in range(P.num_cells())) == rec['kept']
        assert P.live_cells() == [c[0] for c in rec['cells']]
        if rec['digest']:
            assert P.digest() == rec['digest']
    finally:
        P.close()


def test_qprune_needs_the_quad_switch():
    """Without Q active (n odd and no ENDPOS) the prune does not run, and the kept set is the
    plain orbit canon's."""
    a = wz.Problem(13, (0, 2, 1, 7), device=False, m6=1, thm211b=1, thm212=1, canon=1, orbit_q=1, orbit_qprune=1)
    b = wz.Problem(13, (0, 2, 1, 7), device=False, m6=1, thm211b=1, thm212=1, canon=1, orbit_q=1)
    try:
        assert a.qprune_stats()[0] is False
        assert a.live_cells() == b.live_cells()
    finally:
        a.close()
        b.close()
