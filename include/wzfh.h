/*
 * wzfh.h -- C ABI of the MI355X BS(n+1,n) first-hit engine (libwzfh.so).
 *
 * Drop-in boundary for the reference solver's WZ_FIRSTHIT path
 * (src/solver/wz_match.cpp:1529-2892). The reference is a process CLI; its
 * in-process seams are count_pairs22 (:678-784), fh_complete_ab (:1298-1341),
 * hall_ok* (:116-181), npaf_at (:184-192) and autocorr_key (:855-871). Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *   - Plain C types only. Sequences cross the boundary either as int8 arrays of
 *     +1/-1 (the reference's int arrays) or as packed 64-bit words with bit i set
 *     when element i is -1 (n <= 63 on this path).
 *   - Every call returns an int status (WZFH_OK or a negative WZFH_E* code); no
 *     call exits the process. wzfh_last_error() gives the text of the last error
 *     on the calling thread.
 *   - Caller-owned buffers are borrowed for the duration of the call only.
 *     Device memory is owned by the problem handle.
 *   - A problem handle is not reentrant: calls on one handle must not overlap
 *     (the reference's completer and stream are non-reentrant too, :951-975).
 *   - The library needs a HIP device; without one wzfh_init fails with
 *     WZFH_ENODEV. There is no CPU fallback.
 */
#ifndef WZFH_H
#define WZFH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WZFH_OK          0
#define WZFH_EINVAL     -1   /* bad argument (reference exit 1/2 cases)            */
#define WZFH_ENOMEM     -2   /* host or device allocation failed                   */
#define WZFH_EDEVICE    -3   /* a HIP call failed                                  */
#define WZFH_ENODEV     -4   /* no HIP device                                      */
#define WZFH_ECAPACITY  -5   /* caller buffer too small (required size reported)   */
#define WZFH_EINTERNAL  -6   /* completer FOUND failed the NPAF recheck (exit 5)   */
#define WZFH_ETRUNC     -7   /* cell list hit the 20M cap (reference exit 2)       */
#define WZFH_EKERNEL    -8   /* a kernel's own consistency check failed (stream
                                watchdog, invalid work record, unfinished item,
                                results that differ between launch settings); no
                                result of the call is valid                        */

/* Verdicts of one A,B completion (wz_match.cpp:1295-1297, 1330-1335). */
#define WZFH_HIT        0
#define WZFH_PREREJECT  1
#define WZFH_CLEAN_NO   2
#define WZFH_ABORTED    3
#define WZFH_INTERNAL   4

/* Problem configuration: argv <n> <a> <b> <c> <d> plus the stream-shaping WZ_* flags
 * of wz_match.cpp:1363-1388, 1554-1891, 2119-2169. */
typedef struct {
  int n, a, b, c, d;
  int m6;            /* WZ_FH_M6: 1 = mod-6 cells (default n>=36), 0 = mod-3       */
  int thm211b;       /* WZ_THM211B (forced on for n >= 36)                          */
  int thm212;        /* WZ_THM212  (forced on for n >= 36)                          */
  int canon;         /* WZ_FH_ORBIT_CANON                                           */
  int orbit_q;       /* WZ_FH_ORBIT_Q (1 = where proven safe, 2 = forced)          */
  int endpos;        /* WZ_FH_ENDPOS                                                */
  int order;         /* WZ_FH_PROF_ORDER 0..3                                       */
  int rev;           /* WZ_FH_STREAM_REV                                            */
  int ab_prof;       /* WZ_FH_AB_PROF (-1 = default: on for mod-6 cells)           */
  int outer_k;       /* WZ_FH_OUTER_K (0, 2, 3; default 3)                          */
  int no_canon_ab;   /* WZ_FH_NO_CANON                                              */
  long long abp_cap; /* WZ_FH_AB_PROF_CAP (default 20000)                           */
  int orbit_qprune;  /* WZ_FH_ORBIT_QPRUNE: Q-closure orbit elimination (needs canon,
                        the quad switch Q active and mod-6 cells; else ignored)        */
  int a2;            /* WZ_FH_A2: 0 = off, 2 = exact profile-row reachability prune of the
                        completer's row filter (wz_match.cpp:977-1078; needs rows); the
                        shadow mode 1 is rejected with WZFH_EINVAL                       */
} wzfh_config;

typedef struct wzfh_problem wzfh_problem;

/* Fills the reference defaults for a signature (wz_match.cpp:1355-1388, 1561). */
int wzfh_config_default(wzfh_config *cfg, int n, int a, int b, int c, int d);

/* Per-(n, class) setup: norm/autocorr sets, the profile-cell list with ENDPOS,
 * ordering and orbit canonicalization, the A,B profile rows, the hall tables, and
 * their device copies. Replaces wz_match.cpp:1391-1410, 1514-1522, 1588-1891,
 * 2318-2381. Returns WZFH_EINVAL for a bad signature norm (reference exit 1). */
int wzfh_init(const wzfh_config *cfg, wzfh_problem **out);
/* Host-only setup: the same per-(n, class) state without touching a device, for
 * inspecting the cell list (the reference's WZ_FH_LIST_ONLY, wz_match.cpp:1873).
 * Stream / complete / search calls on such a handle return WZFH_ENODEV. */
int wzfh_init_host(const wzfh_config *cfg, wzfh_problem **out);
void wzfh_destroy(wzfh_problem *p);
const char *wzfh_last_error(void);

int wzfh_num_cells(const wzfh_problem *p);
/* Cell pi of the ordered list: px, py (m ints each), kept (orbit-canon keep flag),
 * rows (A,B profile rows of the cell; -1 = proven dead, -2 = rows off). */
int wzfh_cell_info(const wzfh_problem *p, int pi, int *px, int *py, int *kept, int *rows);
/* Indices of the cells an arm streams (kept by orbit canon and not proven dead), in list
 * order, written to out[0..cap); returns the count (the cell loop's skips, wz_match.cpp:2499-2521). */
int wzfh_live_cells(const wzfh_problem *p, int *out, int cap);
/* 64-bit FNV digests "<list>:<kept>" of the ordered cell list under order 3
 * (wz_match.cpp:1834-1856); empty string otherwise. buf needs 34 bytes. */
int wzfh_order_digest(const wzfh_problem *p, char *buf, int buflen);
/* Q-closure prune of the cell list (WZ_FH_ORBIT_QPRUNE, wz_match.cpp:1735-1800): stats[0..5] =
 * missing Q images, not realizable, eq 2.12, nonintegral, unknown, dead orbits; stats[6] = cells
 * removed. Returns 1 when the prune ran (the arm's CFGSIG gains ".qp1"), 0 when it did not. */
int wzfh_qprune_stats(const wzfh_problem *p, long long *stats);

/* ---- C,D cell stream (count_pairs22 + hall_ok_pair_fast, wz_match.cpp:678-784,
 * 143-168, on the C,D side of a mod-m cell) ------------------------------------ */
typedef struct {
  long long cand_cap;     /* stop a cell after this many candidates (WZ_FH_CELLSIZE /
                             BUF_CAP front); 0 = whole cell                          */
  long long node_budget;  /* stop a cell after this many DFS nodes (a build-side
                             serving policy, no reference counterpart); 0 = none     */
  int split_depth;        /* frontier depth of the subtree split; 0 = autotuned      */
  int window;             /* subtrees of one cell in flight; 0 = autotuned           */
} wzfh_stream_opts;

typedef struct {
  long long cand;         /* candidates streamed (CELLSIZE cand=)                    */
  long long leaves;       /* exact-class-sum leaves (CELLSIZE leaves=); -1 when a
                             node budget cut the cell                                */
  long long hall_ok;      /* leaves passing hall_ok (CELLSIZE hall_ok=)             */
  long long dfs_nodes;    /* DFS node entries (CDSTAT dfs_nodes=)                    */
  int capped;             /* 1 when a cap or budget cut the cell                     */
  int reserved;
} wzfh_cell_stats;

/* Streams cells[0..ncells) (indices into the ordered list) on the GPU. The candidates
 * of cell i land in DFS order in cd[2*offsets[i] .. 2*offsets[i+1]) as pairs of
 * packed words (C, D). offsets has ncells+1 entries. cd may be NULL to count only.
 * Returns WZFH_ECAPACITY (and the needed count in offsets[ncells]) when
 * cd_capacity (in candidates) is too small. */
int wzfh_stream_cells(wzfh_problem *p, const int *cells, int ncells, const wzfh_stream_opts *opts,
                      wzfh_cell_stats *stats, uint64_t *cd, long long cd_capacity,
                      long long *offsets);

/* ---- A,B completion (fh_complete_ab, wz_match.cpp:1298-1341) ----------------- */
typedef struct {
  int verdict;            /* WZFH_HIT .. WZFH_INTERNAL                               */
  int reserved;
  long long nodes;        /* fh_nodes_total delta (charged + odd-L middle nodes)     */
  uint64_t A, B;          /* the first completion found (verdict 0)                  */
} wzfh_completion;

/* Completes ncand candidates (packed C,D word pairs) with a node budget each
 * (0 = exact). cell_of[i] >= 0 applies that cell's A,B profile rows (the in-stream
 * completer); -1 runs unconstrained (the WZ_FH_COMPLETE_C/_D mode). cell_of may be
 * NULL (all unconstrained). Returns WZFH_EINTERNAL if any verdict is 4. */
int wzfh_complete(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand,
                  long long budget, wzfh_completion *out);

/* wzfh_complete with the nodes of each candidate split by depth: depth_nodes[64 i + d] = nodes
 * counted at depth d (charges of the quad loop at depth d, the unwinding charges of an abort at
 * their ancestor depth, and odd-L middle trials at depth (n+1)/2), so they sum to out[i].nodes.
 * This is the depth histogram of the reference's FH_TELEM record (wz_match.cpp:1198, 1244).
 * Every candidate is walked whole by one wavefront (no split), so it is slower than
 * wzfh_complete; verdicts, nodes and A,B are the same. */
int wzfh_complete_depth(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                        wzfh_completion *out, long long *depth_nodes);

/* Allocation-free variant of wzfh_complete. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, kernels and results as wzfh_complete, but
 * every per-call device buffer (work items, per-entry reduce, counters, row scratch) lives in
 * one static device array of WZFH_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
 * ends; the call allocates and frees no device memory and clears what must start at zero with
 * a fill kernel. At most WZFH_NOALLOC_MAX_CAND candidates per call (else WZFH_EINVAL); a
 * launch that would not fit the array returns WZFH_ECAPACITY. The array keeps the last call's
 * contents and is shared by every handle on that GPU: calls that overlap on one GPU race. */
#define WZFH_NOALLOC_SCRATCH_BYTES (64u << 20)
#define WZFH_NOALLOC_MAX_CAND 1024
int wzfh_complete_noalloc(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                          wzfh_completion *out);

/* ---- Front-only cell search: stream -> flatness top-K -> complete ------------
 * The unit of the reference's front policy (WZ_FH_BUF_CAP=cand_cap,
 * WZ_FH_DRAIN_TOP=top_k, WZ_FH_DRAIN_BATCHES=1; wz_match.cpp:2656-2714), batched over
 * many cells in one persistent launch. Per cell: the first cand_cap candidates (and
 * at most node_budget DFS nodes) are stable-sorted by flat_score, the first top_k
 * are completed with `budget` nodes each under the cell's profile rows. The per-cell
 * completion counters follow the reference's drain, which stops at the first hit
 * (:2686): completed, ab_nodes, clean_no and aborted count the sorted ranks up to and
 * including the lowest-ranked hit (ranks past it are computed on the device in the same
 * launch and not counted). Returns WZFH_EINTERNAL when a completion's FOUND failed the
 * exact NPAF recheck (the reference exits 5 there, :2591-2595): the results are written
 * and the cells concerned carry internal = 1. Any other error leaves no result valid. */
typedef struct {
  wzfh_stream_opts stream;
  int top_k;
  int batched;            /* 1 = all cells in one persistent launch, 0 = one cell per
                             launch (the plain path)                                 */
  long long budget;       /* A,B node budget per candidate                           */
} wzfh_search_opts;

typedef struct {
  wzfh_cell_stats stream;
  int completed;          /* candidates completed (<= top_k)                         */
  int hit;                /* 1 if a completion hit                                   */
  long long hit_rank;     /* sorted rank of the first hit (-1 none)                  */
  long long hit_idx;      /* in-cell stream index of the first hit (-1 none)         */
  long long ab_nodes;     /* total completer nodes of the cell                       */
  int clean_no, aborted;
  uint64_t A, B, C, D;    /* the hit (if any)                                        */
  int internal;           /* 1 if a counted completion had verdict 4 (NPAF recheck
                             of a FOUND failed): the cell's result is not valid      */
  int reserved2;
} wzfh_search_result;

int wzfh_search_cells(wzfh_problem *p, const int *cells, int ncells, const wzfh_search_opts *opts,
                      wzfh_search_result *out);

/* ---- Autotuning ----------------------------------------------------------------
 * No reference counterpart. Launch settings that change only speed, never results: the
 * stream kernel's DFS moves per lane per pass, the capped launches' donation generations
 * and static-item window, and the completer's split depth and item charge cap. wzfh_autotune times them on a
 * sample of cells with the caller's search options (coordinate descent, one knob at a
 * time), checks that every trial returns the same per-cell results as the first (else
 * WZFH_EINTERNAL), keeps the fastest in the problem handle and returns it. Environment
 * variables (WZFH_INNER_MOVES, WZFH_DONATE_HOPS, WZFH_STATIC_WINDOW, WZFH_IDLE_SLEEP, WZFH_BLOCKS_PER_CU,
 * WZFH_COMPLETE_SPLIT)
 * still override the handle's settings. */
typedef struct {
  int inner_moves;        /* stream: DFS moves per lane per pass of the main loop (1..256) */
  int donate_hops;        /* capped stream: generations of frontier donations (0..15)     */
  int static_window;      /* capped stream: static items handed out ahead of the walk     */
  int complete_split;     /* completer: quad-path levels per split round (0..6, 0 = none)  */
  int idle_sleep;         /* stream: 64-cycle sleeps of a wave with nothing to walk (0..256)*/
  int blocks_per_cu;      /* stream: persistent blocks (4 waves each) per compute unit (1..4)*/
  int trials;             /* wzfh_autotune: configurations timed                           */
  int complete_cap_log2;  /* completer: split items stop after 2^this charges and their
                             subtrees are split again (0 = no cap)                        */
  double stream_ms;       /* wzfh_autotune: stream kernel time of the chosen settings      */
  double complete_ms;     /* wzfh_autotune: completer time of the chosen settings          */
} wzfh_tuning;

int wzfh_autotune(wzfh_problem *p, const int *cells, int ncells, const wzfh_search_opts *opts, wzfh_tuning *out);
int wzfh_get_tuning(const wzfh_problem *p, wzfh_tuning *out);
int wzfh_set_tuning(wzfh_problem *p, const wzfh_tuning *t);

/* ---- Serving: continuous batching of cell-search requests ------------------------
 * No reference counterpart (the reference runs fixed arm processes). Requests name a
 * problem (an initialized (n, class) handle, kept device-resident) and a cell. The problems
 * of one n share a stream stage that admits whatever has arrived, up to max_batch cells, into
 * one persistent stream launch and hands each cell's top-K candidates to its problem's
 * completion workers, which complete them on HIP streams of their own while the next launch
 * streams; different n run concurrently. A result equals the same cell searched alone.
 * Handles passed to a server must not be used by the caller until wzfh_serve_close returns. */
typedef struct wzfh_server wzfh_server;
typedef struct {
  wzfh_search_opts search;   /* per-cell unit (batched flag ignored: always batched) */
  int max_batch;             /* cells admitted into one stream launch                 */
  int complete_workers;      /* completion runs per problem at once (0: 1, at most 8)  */
} wzfh_serve_opts;
typedef struct {
  long long ticket;
  int problem, cell;
  int status;                /* WZFH_OK or the error of the batch                      */
  int batch_size;            /* cells in the launch that served this request          */
  double submit_s, start_s, finish_s;   /* seconds since the server opened             */
  wzfh_search_result result;
} wzfh_serve_result;
int wzfh_serve_open(wzfh_problem **problems, int nproblems, const wzfh_serve_opts *opts, wzfh_server **out);
int wzfh_serve_submit(wzfh_server *s, int problem, int cell, long long *ticket);
/* Submits count requests at once (problems[i], cells[i]); tickets (may be NULL) receives their
 * tickets in order. They arrive together, so one admission can take them all. */
int wzfh_serve_submit_batch(wzfh_server *s, int count, const int *problems, const int *cells, long long *tickets);
/* Waits up to timeout_ms for at least min_results finished requests; returns how many
 * were written to out (<= max), or a negative error. */
int wzfh_serve_poll(wzfh_server *s, wzfh_serve_result *out, int max, int min_results, int timeout_ms);
int wzfh_serve_close(wzfh_server *s);

/* ---- Hash join over the Thm 2.2 space (WZ_JOIN22) ---------------------------- */
/* Replaces the WZ_JOIN22 mode of main (wz_match.cpp:2905-3104): the C,D stream of every
 * C,D profile (survive_profiles at m = 3, :439-487, :1521-1522) is stored in an
 * open-addressed HBM table keyed by its exact autocorrelation vector (-N_C(s) - N_D(s),
 * s = 1..n; the reference stores a 64-bit FNV of it, :855-871), one entry per distinct
 * vector (the earliest C,D in stream order); then the A,B stream of every A,B profile
 * probes it with (N_A(s) + N_B(s)). Equal vectors mean NPAF = 0 at every shift, so every
 * hit is a solution (no resolve pass, no FNV shadowing). Config: n, a, b, c, d,
 * thm211b, thm212 and rev are used; m6, canon, order, endpos, ab_prof are not. */
typedef struct wzfh_join wzfh_join;
typedef struct {
  long long cd_profiles, ab_profiles;
  long long cd_stream;       /* C,D candidates over all C,D profiles                  */
  long long distinct_keys;   /* table entries (distinct exact vectors)               */
  long long slots;           /* table slots (a power of two)                         */
  long long ab_stream;       /* A,B candidates probed                                */
  long long hits;            /* A,B candidates whose vector is stored: each a solution */
  int found;                 /* hits > 0                                              */
  int reserved;
  uint64_t A, B, C, D;       /* reported solution, as the reference prints it: the
                                earliest C,D (profile order, DFS order) with a hit and
                                the last-streamed A,B hitting it                       */
} wzfh_join_result;
int wzfh_join_open(const wzfh_config *cfg, wzfh_join **out);
/* Phase 1 (BUILD, :2929-2968): stream the C,D side and fill the table. slots_log2 = 0
 * sizes the table to at least twice the stream. Fills cd_*, distinct_keys, slots. */
int wzfh_join_build(wzfh_join *j, int slots_log2, wzfh_join_result *out);
/* Phase 2 (STREAM + probe, :2970-3010): stream the A,B side against the table. Fills
 * ab_*, hits, found and the solution; *out must hold wzfh_join_build's fields. */
int wzfh_join_probe(wzfh_join *j, wzfh_join_result *out);
/* Every solution found by the probe: each A,B hit paired with every C,D of the stream whose
 * vector it matched (the reference resolves hits by re-enumerating C,D and rechecking NPAF,
 * :3039-3090, and prints the first; this lists them all -- the full solution set of the
 * signed signature in the Thm 2.2 space, each with NPAF = 0 at every shift). Writes *count
 * quadruples (A, B, C, D) of packed words to out in ascending order, or only counts when out
 * is NULL; WZFH_ECAPACITY (with *count set) when cap is smaller. Call after wzfh_join_probe. */
int wzfh_join_solutions(wzfh_join *j, uint64_t *out, long long cap, long long *count);
/* The same table from caller-given candidates instead of the profile streams: ncd C,D pairs
 * (packed words, n elements; the earlier of two equal vectors keeps the entry, as stream order
 * does) and nab A,B pairs (n+1 elements). These are the reference's insert and lookup loops
 * (:2939-2950, :2985-2996) over an explicit list, for joins whose sides come from elsewhere
 * (a cell search, a file of candidates) and for measuring the table at a given size. */
int wzfh_join_build_from(wzfh_join *j, const uint64_t *cd, long long ncd, int slots_log2, wzfh_join_result *out);
int wzfh_join_probe_from(wzfh_join *j, const uint64_t *ab, long long nab, wzfh_join_result *out);
/* Device time (HIP events on the join's stream) of the table kernels of the last build (claim,
 * write, verify rounds and the duplicate pass) and of the last probe kernel, in ms. No reference
 * counterpart. */
int wzfh_join_kernel_ms(const wzfh_join *j, double *build_ms, double *probe_ms);
void wzfh_join_close(wzfh_join *j);

/* ---- Exact checks ----------------------------------------------------------- */
/* Summed NPAF over s = 1..n of int8 +1/-1 sequences (npaf_at, wz_match.cpp:184-192,
 * and the VERIFY line :2640-2646). *max_abs gets max_s |NPAF[s]|. */
int wzfh_npaf_check(int n, const int8_t *A, const int8_t *B, const int8_t *C, const int8_t *D,
                    int *max_abs);
/* Sum_s |CD[s]| of packed C,D (the flat_score lambda, wz_match.cpp:2569-2583). */
long long wzfh_flat_score(int n, uint64_t C, uint64_t D);

/* ---- Device timing (for measurement) ----------------------------------------- */
/* Average device duration of the named kernel over the last wzfh_search_cells /
 * wzfh_stream_cells / wzfh_complete call, measured with HIP events on the stream the
 * kernel runs on. kind: 0 = stream kernel, 1 = completer kernel (units = DFS nodes).
 * kind 2..9 and 11..12 (with WZFH_TELEMETRY=1 in the environment): the last stream
 * launch's counters -- 2 wave cycles in the leaf test, 3 total wave cycles, 4 leaf-test
 * flushes, 5 leaves tested, 6 leaves the matrix-core test left uncertain (decided in FP64),
 * 7 -- with WZFH_HALL_VERIFY=1, which also runs the FP64 test on every leaf -- leaves whose
 * certain decision disagreed with FP64 (a false reject or false accept; must be 0),
 * 8 donations, 9 donated work items, 11 lane DFS moves, 12 wave passes of the main loop.
 * kind 10: the leaf test's margin. kind 13 (WZFH_TELEMETRY=1): completer nodes walked by all
 * work items of the last completion run, including the partial walks of items stopped at the
 * charge cap and walks past a budget abort (kind 1 counts each of the reference's nodes once). */
int wzfh_last_kernel_ms(const wzfh_problem *p, int kind, double *ms, long long *units);
/* Running totals of a handle since the last reset (or since wzfh_init), for measurement over
 * a timed region: summed HIP-event durations of the stream kernel's launches (index 0) and of
 * the completer kernel's launches without their per-entry reduce (index 1), the units they
 * processed (counted C,D DFS nodes; completer nodes) and the launch counts. reset != 0 zeroes
 * the totals after reading them. No reference counterpart (the reference times phases with
 * host clocks, wz_match.cpp:2532-2553). */
typedef struct {
  double ms[2];
  long long units[2];
  long long launches[2];
} wzfh_kernel_sums;
int wzfh_kernel_totals(wzfh_problem *p, wzfh_kernel_sums *out, int reset);
/* The same totals over a server: its stream launches (which run on per-group handles of the
 * server's own, several classes of one n side by side) and its problems' completer launches.
 * No reference counterpart. */
int wzfh_serve_kernel_totals(wzfh_server *s, wzfh_kernel_sums *out, int reset);

#ifdef __cplusplus
}
#endif
#endif
