"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or shipped.
"""
import ctypes as C
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
LIB = REPO / 'oracle' / '_build' / 'liboracle.so'


class Config(C.Structure):
    _fields_ = [(k, C.c_int) for k in ('n', 'a', 'b', 'c', 'd', 'm6', 'thm211b', 'thm212', 'canon',
                                        'orbit_q', 'endpos', 'order', 'rev', 'ab_prof', 'outer_k',
                                        'no_canon_ab')] + [('abp_cap', C.c_longlong), ('no_cells', C.c_int), ('a2', C.c_int)]


class StreamStats(C.Structure):
    _fields_ = [('leaves', C.c_longlong), ('hall_ok', C.c_longlong), ('dfs_nodes', C.c_longlong),
                ('cand', C.c_longlong)]


class SearchResult(C.Structure):
    """wzo_search_result: one front-only cell search (the unit of wzfh_search_cells)."""
    _fields_ = [('cand', C.c_longlong), ('leaves', C.c_longlong), ('hall_ok', C.c_longlong),
                ('dfs_nodes', C.c_longlong), ('capped', C.c_int), ('budget_cut', C.c_int), ('completed', C.c_int),
                ('hit', C.c_int), ('hit_rank', C.c_longlong), ('hit_idx', C.c_longlong), ('ab_nodes', C.c_longlong),
                ('pre_rejected', C.c_int), ('clean_no', C.c_int), ('aborted', C.c_int), ('internal', C.c_int)] + \
               [(k, C.c_int8 * 64) for k in 'ABCD']


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise FileNotFoundError(f'{LIB} missing: run make -C oracle')
        L = C.CDLL(str(LIB))
        L.wzo_create.restype = C.c_void_p
        L.wzo_create.argtypes = [C.POINTER(Config)]
        L.wzo_destroy.argtypes = [C.c_void_p]
        L.wzo_a2_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.wzo_create_with_cells.restype = C.c_void_p
        L.wzo_create_with_cells.argtypes = [C.POINTER(Config), C.c_void_p, C.c_int]
        L.wzo_num_cells.argtypes = [C.c_void_p]
        L.wzo_cell_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.wzo_stream_cell.restype = C.c_longlong
        L.wzo_stream_cell.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_void_p, C.c_longlong,
                                      C.POINTER(StreamStats)]
        L.wzo_complete.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_longlong,
                                   C.POINTER(C.c_longlong), C.c_void_p, C.c_void_p]
        L.wzo_flat_score.restype = C.c_longlong
        L.wzo_flat_score.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.wzo_npaf_at.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.wzo_outer_tables.argtypes = [C.c_void_p, C.c_void_p]
        L.wzo_stream_cell_budget.restype = C.c_longlong
        L.wzo_stream_cell_budget.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_void_p,
                                             C.c_longlong, C.POINTER(StreamStats)]
        L.wzo_cell_search.restype = C.c_longlong
        L.wzo_cell_search.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_int, C.c_longlong,
                                      C.POINTER(SearchResult), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.wzo_solutions.restype = C.c_longlong
        L.wzo_solutions.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong]
        _lib = L
    return _lib


def _i8(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int8))


class OracleProblem:
    def __init__(self, n, sig, m6=1, thm211b=1, thm212=1, canon=0, orbit_q=0, endpos=0, order=1,
                 rev=0, ab_prof=-1, outer_k=3, no_canon_ab=0, abp_cap=20000, no_cells=0, cells=None, a2=0):
        self.n = n
        self.m = 6 if m6 else 3
        cfg = Config(n, *sig, m6, thm211b, thm212, canon, orbit_q, endpos, order, rev, ab_prof,
                     outer_k, no_canon_ab, abp_cap, no_cells, a2)
        if cells is None:
            self.h = lib().wzo_create(C.byref(cfg))
        else:   # explicit (px, py) list: cell i of this problem is cells[i]
            arr = np.zeros((len(cells), 16), dtype=np.int8)
            for i, (px, py) in enumerate(cells):
                arr[i, :len(px)] = px
                arr[i, 8:8 + len(py)] = py
            self.h = lib().wzo_create_with_cells(C.byref(cfg), arr.ctypes.data, len(cells))

    def close(self):
        if self.h:
            lib().wzo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def a2_counts(self):
        """(rows_tested, rows_rejected) of the A2 row test since the last call."""
        t, r = C.c_longlong(), C.c_longlong()
        lib().wzo_a2_counts(self.h, C.byref(t), C.byref(r))
        return t.value, r.value

    def num_cells(self):
        return lib().wzo_num_cells(self.h)

    def cell(self, pi):
        px = (C.c_int * 8)()
        py = (C.c_int * 8)()
        kept, rows = C.c_int(), C.c_int()
        lib().wzo_cell_info(self.h, pi, px, py, C.byref(kept), C.byref(rows))
        return list(px[:self.m]), list(py[:self.m]), kept.value, rows.value

    def stream(self, pi, cap=0, out_cap=0):
        st = StreamStats()
        out = np.zeros((max(out_cap, 1), 2, self.n), dtype=np.int8)
        got = lib().wzo_stream_cell(self.h, pi, cap, out.ctypes.data, out_cap, C.byref(st))
        return got, out[:min(got, out_cap)], st

    def stream_budget(self, pi, cap, node_budget, out_cap):
        st = StreamStats()
        out = np.zeros((max(out_cap, 1), 2, self.n), dtype=np.int8)
        got = lib().wzo_stream_cell_budget(self.h, pi, cap, node_budget, out.ctypes.data, out_cap, C.byref(st))
        return got, out[:min(got, out_cap)], st

    def search(self, pi, buf_cap, node_budget, top_k, budget):
        """Front-only cell search (wzo_cell_search): returns (result, per-rank arrays) with
        per-rank verdict, nodes, in-cell stream index and flat score of the completions run."""
        v = np.zeros(max(top_k, 1), np.int32)
        nd = np.zeros(max(top_k, 1), np.int64)
        idx = np.zeros(max(top_k, 1), np.int64)
        sc = np.zeros(max(top_k, 1), np.int64)
        res = SearchResult()
        lib().wzo_cell_search(self.h, pi, buf_cap, node_budget, top_k, budget, C.byref(res), v.ctypes.data,
                              nd.ctypes.data, idx.ctypes.data, sc.ctypes.data)
        k = res.completed
        return res, {'verdict': v[:k], 'nodes': nd[:k], 'idx': idx[:k], 'score': sc[:k]}

    def complete(self, C_, D_, budget, pi=-1):
        n = self.n
        A = np.zeros(n + 1, np.int8)
        B = np.zeros(n + 1, np.int8)
        nodes = C.c_longlong()
        Cc, Dc = _i8(C_), _i8(D_)
        r = lib().wzo_complete(self.h, pi, Cc.ctypes.data, Dc.ctypes.data, budget, C.byref(nodes),
                               A.ctypes.data, B.ctypes.data)
        return r, nodes.value, A, B


def solutions(n, sig):
    """Every solution of the signed signature in the space WZ_JOIN22 searches (wzo_solutions), as a
    sorted list of packed (A, B, C, D) word tuples."""
    a, b, c, d = sig
    cnt = lib().wzo_solutions(n, a, b, c, d, None, 0)
    out = np.zeros((max(cnt, 1), 4), np.uint64)
    lib().wzo_solutions(n, a, b, c, d, out.ctypes.data, cnt)
    return [tuple(int(v) for v in row) for row in out[:cnt]]


def flat_score(n, C_, D_):
    Cc, Dc = _i8(C_), _i8(D_)
    return lib().wzo_flat_score(n, Cc.ctypes.data, Dc.ctypes.data)


def dump_lines(cands):
    """WZ_FH_DUMP text of candidates (wz_match.cpp:2743-2748)."""
    out = []
    for cd in cands:
        out.append(''.join(f'{int(v)} ' for v in cd[0]) + '| ' + ''.join(f'{int(v)} ' for v in cd[1]) + '\n')
    return out
