"""Runs kernel-logic checks on the CPU emulation of libwzfh (tests/emu/build.py) against the
golden fixtures; prints one line per check and 'bad N' at the end. Invoked in a subprocess by
tests/test_emu_kernels.py so the emulated library never mixes with the real one."""
import hashlib
import os
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))
import numpy as np  # noqa: E402

import bs45_quantum_explorer_amd as wz  # noqa: E402
wz.LIBPATH = HERE / '_build' / 'libwzfh_emu.so'
from golden_data import complete_cases, stream_runs  # noqa: E402

bad = 0


T0 = time.time()


def report(ok, *msg):
    global bad
    bad += 0 if ok else 1
    print('ok ' if ok else 'BAD', f'[{time.time() - T0:6.1f} s]', *msg, flush=True)


# completer: budget aborts found through split items, capped items re-split deeper, the
# abort item walked whole; the device reduce checked against the host scan of the items
cases = [c for c in complete_cases() if c['n'] == 13 and c['name'] in ('n13:5', 'n13:23', 'n13:25', 'n13:37')]
P = wz.Problem(13, tuple(cases[0]['sig']), m6=0, thm211b=0, thm212=0, outer_k=cases[0]['outer_k'])
os.environ['WZFH_CHECK_REDUCE'] = '1'
os.environ['WZFH_COMPLETE_EXACT_ITEM'] = '1'
# capped items (charge cap 64 under budget 300: the abort falls inside a capped item's subtree)
# walk abort items of up to 2,048 charges whole; a smaller cap or exact-item threshold multiplies
# the emulated launches (cap 4: ~200 s on this emulation)
plan = [(2, 0, c) for c in cases] + [(1, 64, c) for c in cases if (c['name'], c['budget']) in (('n13:5', 0), ('n13:23', 300))]
# small caps: capped items inside the subtrees of capped items, each next round resuming from the
# sub-path where the capped item stopped
plan += [(2, 16, c) for c in cases if (c['name'], c['budget']) == ('n13:23', 0)]
plan += [(2, 8, c) for c in cases if (c['name'], c['budget']) == ('n13:25', 0)]
plan += [(2, 4, c) for c in cases if (c['name'], c['budget']) == ('n13:5', 0)]
for split, cap, c in plan:
    os.environ['WZFH_COMPLETE_SPLIT'] = str(split)
    os.environ['WZFH_COMPLETE_ITEM_CAP'] = str(cap)
    os.environ['WZFH_COMPLETE_EXACT_ITEM'] = '1' if cap == 0 else '2048'
    if True:
        cd = np.array([[wz.pack(c['C']), wz.pack(c['D'])]], dtype=np.uint64)
        o = P.complete(cd, budget=c['budget'])[0]
        ok = (o.verdict, o.nodes) == (c['r'], c['nodes'])
        if ok and c['r'] == 0:
            ok = list(wz.unpack(o.A, 14)) == c['A'] and list(wz.unpack(o.B, 14)) == c['B']
        report(ok, 'complete', c['name'], 'budget', c['budget'], 'split', split, 'cap', cap, (o.verdict, o.nodes),
               (c['r'], c['nodes']))
# the allocation-free entry point: the same completions with every buffer carved from the
# library's static scratch array (two calls in a row, the second smaller)
os.environ['WZFH_COMPLETE_SPLIT'] = '2'
os.environ['WZFH_COMPLETE_ITEM_CAP'] = '0'
os.environ['WZFH_COMPLETE_EXACT_ITEM'] = '1'
for batch in (cases[:2], cases[:1]):
    cd = np.array([[wz.pack(c['C']), wz.pack(c['D'])] for c in batch], dtype=np.uint64)
    for budget in sorted({c['budget'] for c in batch}):
        mine = [c for c in batch if c['budget'] == budget]
        sub = np.array([[wz.pack(c['C']), wz.pack(c['D'])] for c in mine], dtype=np.uint64)
        got = P.complete(sub, budget=budget, noalloc=True)
        ok = [(o.verdict, o.nodes) for o in got] == [(c['r'], c['nodes']) for c in mine]
        report(ok, 'complete_noalloc', len(batch), 'budget', budget, [(o.verdict, o.nodes) for o in got])
P.close()

# the profile-row filter with the A2 reachability prune (WZ_FH_A2=2) against the oracle: in-cell
# completions under their cells' rows, split 2, of stream candidates whose node counts the prune
# changes (n = 13: cell 3, candidates 24-31, exact and budget 112; n = 12: cell 43, hits)
from oracle_lib import OracleProblem  # noqa: E402
os.environ['WZFH_COMPLETE_SPLIT'] = '2'
for n, sig, pi, j0, budgets in ((13, (2, 4, 3, 5), 3, 24, (0, 112)), (12, (1, 7, 0, 0), 43, 0, (0,))):
    P = wz.Problem(n, sig, m6=1, thm211b=1, thm212=1, canon=1, order=1, a2=2)
    O = OracleProblem(n, sig, canon=1, order=1, a2=2)
    _, cands, _ = O.stream(pi, j0 + 8, j0 + 8)
    cds = [[wz.pack([int(v) for v in c[0]]), wz.pack([int(v) for v in c[1]])] for c in cands[j0:]]
    for budget in budgets:
        got = P.complete(np.array(cds, dtype=np.uint64), budget=budget, cell_of=[pi] * len(cds))
        want = [tuple(O.complete(wz.unpack(cd[0], n), wz.unpack(cd[1], n), budget, pi=pi)[:2]) for cd in cds]
        res = [(o.verdict, o.nodes) for o in got]
        report(res == want, 'complete rows a2 n', n, 'cell', pi, 'budget', budget, res)
    P.close()
    O.close()

# stream: per-cell counts and stream hash of small-n runs, with donation forced every 2 moves
os.environ['WZFH_DONATE_MOVES'] = '2'
os.environ['WZFH_DONATE_LOW'] = '1000000'
os.environ['WZFH_DONATE_ANY'] = '1'
for rec in [r for r in stream_runs() if r['n'] == 8][:1]:
    S = wz.Problem(rec['n'], rec['sig'], m6=1, thm211b=1, thm212=1, canon=rec['canon'], rev=rec['rev'],
                   endpos=rec['endpos'], order=rec['order'])
    cells = [c[0] for c in rec['cells']][:4]
    if cells:
        stats, offs, cd = S.stream(cells)
        for i, row in enumerate(rec['cells'][:4]):
            part = cd[offs[i]:offs[i + 1]]
            h = hashlib.sha1(''.join(wz.dump_lines(rec['n'], part)).encode()).hexdigest()
            st = stats[i]
            got = [row[0], st.cand, st.leaves, st.hall_ok, st.dfs_nodes, h]
            report(got == list(row), 'stream n8', rec['sig'], 'cell', row[0])
    S.close()
# the same, under a node budget with multi-level frontier donation (every level from the shallowest
# live one down to d-2 given away at once, deepest first), against the oracle's budgeted stream
from oracle_lib import OracleProblem  # noqa: E402
os.environ['WZFH_DONATE_KEEP'] = '1'
n12 = next(r for r in stream_runs() if r['n'] == 12 and r['canon'] == 1 and r['rev'] == 0)
S = wz.Problem(12, n12['sig'], m6=1, thm211b=1, thm212=1, canon=1, order=1)
O = OracleProblem(12, n12['sig'], canon=1, order=1)
cells = S.live_cells()[:2]
for budget in (37, 3000):
    st, offs, cd = S.stream(cells, node_budget=budget, split_depth=1, window=2)
    for i, pi in enumerate(cells):
        got, cands, ost = O.stream_budget(pi, 0, budget, 100000)
        lines = [''.join(f'{int(v)} ' for v in c[0]) + '| ' + ''.join(f'{int(v)} ' for v in c[1]) + '\n' for c in cands]
        ok = (st[i].cand, st[i].dfs_nodes) == (got, ost.dfs_nodes) and wz.dump_lines(12, cd[offs[i]:offs[i + 1]]) == lines
        report(ok, 'stream n12 multi-level donation, budget', budget, 'cell', pi, (st[i].cand, st[i].dfs_nodes),
               (got, ost.dfs_nodes))
S.close()
O.close()
os.environ.pop('WZFH_DONATE_KEEP', None)
print('bad', bad, flush=True)
