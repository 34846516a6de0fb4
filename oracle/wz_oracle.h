/*
 * wz_oracle.h -- CPU oracle for the BS(n+1,n) first-hit hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain, scalar C++ restatement of the
 * reference solver's WZ_FIRSTHIT path (src/solver/wz_match.cpp), used as the
 * checker for the HIP library and as the CPU baseline in bench.py. The shipped
 * library never links, loads or calls it; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg do.
 *
 * Every function cites the reference lines it restates. Pinning (how each part is
 * checked against something other than itself):
 *   - cell list, orbit canon, order      : golden per-cell fixtures from the
 *                                          compiled reference (stream_cells.json)
 *   - C,D stream, hall_ok, DFS node count: same fixtures (counts + stream SHA-1)
 *   - completer (verdict, nodes, A,B)    : complete_only.json (six known-answer
 *                                          controls, n=44 archive candidates,
 *                                          budget aborts) + firsthit_runs.json
 *   - npaf / flat score                  : closed form + 43 banked solutions
 *   - outer (A1) tables                  : brute-force enumeration in tests
 *   - A2 row reachability (a2)           : a2_runs.json (arm runs with WZ_FH_A2=2:
 *                                          CAND lines, nodes, A2SHADOW row counts)
 *   - solution set (wzo_solutions)       : brute force over every quadruple at
 *                                          n = 6..9 in tests, the banked solutions
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n, a, b, c, d;   /* signature as given on the reference command line  */
  int m6;              /* 1: mod-6 cells (survive_profiles6), 0: mod-3 cells */
  int thm211b, thm212; /* forced on at n >= 36 like wz_match.cpp:1388       */
  int canon;           /* WZ_FH_ORBIT_CANON                                  */
  int orbit_q;         /* WZ_FH_ORBIT_Q (1 = only where proven safe, 2 = force) */
  int endpos;          /* WZ_FH_ENDPOS                                       */
  int order;           /* WZ_FH_PROF_ORDER 0..3                              */
  int rev;             /* WZ_FH_STREAM_REV                                   */
  int ab_prof;         /* -1 = default (= m6)                                */
  int outer_k;         /* WZ_FH_OUTER_K (default 3)                          */
  int no_canon_ab;     /* WZ_FH_NO_CANON (completer root/reversal canon off) */
  long long abp_cap;   /* WZ_FH_AB_PROF_CAP (default 20000)                  */
  int no_cells;        /* 1: skip the cell list (completion-only use)        */
  int a2;              /* WZ_FH_A2: 0 off, 1 count rows only, 2 prune rows    */
} wzo_config;

typedef struct {
  long long leaves, hall_ok, dfs_nodes, cand;
} wzo_stream_stats;

void *wzo_create(const wzo_config *cfg);
/* A problem over an explicit cell list (px[8], py[8] as int8 per cell), with the A,B
 * profile rows built for exactly those cells (a sample of a larger class list). */
void *wzo_create_with_cells(const wzo_config *cfg, const int8_t *pxpy, int ncells);
void wzo_destroy(void *h);
int wzo_num_cells(void *h);
/* px, py: m ints each. kept: orbit-canon keep flag (1 when canon is off).
 * rows: number of A,B profile rows of the cell (-1 = proven dead, -2 = rows off). */
int wzo_cell_info(void *h, int pi, int *px, int *py, int *kept, int *rows);
/* Streams cell pi in DFS order; stops after `cap` candidates (cap <= 0: none).
 * Writes up to out_cap candidates as int8 C[n] then D[n] per record.
 * Returns the number of candidates streamed. */
long long wzo_stream_cell(void *h, int pi, long long cap, int8_t *out, long long out_cap,
                          wzo_stream_stats *st);
/* One A,B completion. pi >= 0 applies cell pi's profile rows (when rows are on and the
 * list is within the cap), pi < 0 runs unconstrained (WZ_FH_COMPLETE_C/_D mode).
 * Returns the verdict 0..4; nodes = charged+middle nodes (fh_nodes_total delta). */
int wzo_complete(void *h, int pi, const int8_t *C, const int8_t *D, long long budget,
                 long long *nodes, int8_t *A, int8_t *B);
long long wzo_flat_score(int n, const int8_t *C, const int8_t *D);
/* The A2 row counters summed over this problem's completions since the last call (the
 * rows_tested / rows_rejected of the reference's A2SHADOW line), then cleared. */
void wzo_a2_counts(void *h, long long *tested, long long *rejected);
/* Summed NPAF at shift s (wz_match.cpp:184-192). */
int wzo_npaf_at(const int8_t *A, const int8_t *B, int n1, const int8_t *C, const int8_t *D,
                int n2, int s);
/* Same as wzo_stream_cell with a DFS node budget: only nodes with DFS index < node_budget
 * are visited (the serving policy's per-cell budget; the reference has no such cap). */
long long wzo_stream_cell_budget(void *h, int pi, long long cap, long long node_budget, int8_t *out,
                                 long long out_cap, wzo_stream_stats *st);
/* One front-only cell search (the unit of wzfh_search_cells). */
typedef struct {
  long long cand, leaves, hall_ok, dfs_nodes;   /* stream of the cell (prefix under the caps)  */
  int capped, budget_cut;   /* a cap or budget cut the cell / the node budget did            */
  int completed;            /* completions run, in sorted order, up to and including a hit  */
  int hit;
  long long hit_rank, hit_idx;   /* sorted rank and in-cell stream index of the hit (-1)    */
  long long ab_nodes;       /* completer nodes of the completions run                       */
  int pre_rejected, clean_no, aborted, internal;
  int8_t A[64], B[64], C[64], D[64];   /* the hit (A,B: n+1 entries, C,D: n)              */
} wzo_search_result;
/* Front-only cell unit (drain of wz_match.cpp:2656-2714 with BUF_CAP = buf_cap, DRAIN_TOP =
 * top_k, DRAIN_BATCHES = 1): stream the first buf_cap candidates of cell pi (0 = all; at most
 * node_budget DFS nodes when > 0), stable-sort them by flat score, complete the sorted ranks
 * < top_k in order with `budget` under the cell's profile rows, stopping after the first hit.
 * Per completed rank: verdict, nodes, in-cell stream index, flat score (arrays of top_k, may
 * be NULL). Returns the candidates streamed. */
long long wzo_cell_search(void *h, int pi, long long buf_cap, long long node_budget, int top_k, long long budget,
                          wzo_search_result *res, int *verdicts, long long *nodes, long long *stream_idx,
                          long long *scores);
/* Every solution (A, B, C, D) of the signed signature in the space WZ_JOIN22 searches
 * (wz_match.cpp:2894-2904): the Thm 2.2 pair encoding, X[0] = +1 when X's sum is 0, exact
 * sums; enumerated per side and joined on exact autocorrelation vectors, with no profile list
 * and no hall test. Writes up to cap quadruples of packed words (A, B: n+1 elements, C, D: n;
 * bit i set = element i is -1) in ascending (A, B, C, D) order; returns the count. */
long long wzo_solutions(int n, int a, int b, int c, int d, uint64_t *out, long long cap);
/* A1 outer tables (wz_match.cpp:560-598) for test cross-checks. */
void wzo_outer_tables(uint8_t *outer2 /*16*5*9*/, uint8_t *outer3 /*128*5*9*13*/);

#ifdef __cplusplus
}
#endif
