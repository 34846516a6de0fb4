// wzfh_match -- drop-in for the reference arm CLI `wz_match <n> <a> <b> <c> <d>` under
// WZ_FIRSTHIT=1 (src/solver/wz_match.cpp:1343-1362, 1544-2892), running the hot path
// through libwzfh's C ABI.
//
// It reproduces the arm contract the lane driver and the reference's tests consume:
// the WZ_FH_* lane/cell flags, the cell loop with orbit-dup and proven-dead skips, the
// flatness-sorted drain buffer with the front-only policy, the CAND / CELLSIZE / CDSTAT /
// FIRSTHIT / summary / RESULT lines, WZ_FH_DUMP, the checkpoint file with CFGSIG and
// exact positional resume, and the exit codes 0 / 1 / 2 / 3 / 5.
//
// GPU work is batched without changing semantics: cells are streamed a window at a
// time (each cell's stream is exact and complete up to the cap the policy needs), and a
// drain batch's completions run in one launch; the host then replays the reference's
// sequential decisions (first hit in sorted order stops the arm, later completions of the
// batch are discarded and never counted).
#include <algorithm>
#include <chrono>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wzfh.h"

using Clock = std::chrono::steady_clock;

static long long env_ll(const char *k, long long dflt) {
  const char *e = getenv(k);
  return e ? atoll(e) : dflt;
}
static bool env_set(const char *k) { return getenv(k) != nullptr; }

static volatile sig_atomic_t g_sigterm = 0;
static void on_sigterm(int) { g_sigterm = 1; }

static std::vector<int> parse_list(const char *s) {
  std::vector<int> v;
  int val, cnt;
  while (s && sscanf(s, "%d%n", &val, &cnt) == 1) {
    v.push_back(val);
    s += cnt;
    while (*s == ',' || *s == ' ') s++;
  }
  return v;
}
static uint64_t pack(const std::vector<int> &x) {
  uint64_t w = 0;
  for (size_t i = 0; i < x.size(); i++) if (x[i] < 0) w |= 1ull << i;
  return w;
}
static std::vector<int> unpack(uint64_t w, int len) {
  std::vector<int> v(len);
  for (int i = 0; i < len; i++) v[i] = ((w >> i) & 1) ? -1 : 1;
  return v;
}
static void print_seq(const char *name, const std::vector<int> &v) {
  printf("%s = {", name);
  for (size_t i = 0; i < v.size(); i++) printf("%d%s", v[i], i + 1 < v.size() ? "," : "");
  printf("};\n");
}

static void print_solution(int n, const std::vector<int> &A, const std::vector<int> &B, const std::vector<int> &C,
                           const std::vector<int> &D) {
  int sa = 0, sb = 0, sc = 0, sd = 0;
  for (int v : A) sa += v;
  for (int v : B) sb += v;
  for (int v : C) sc += v;
  for (int v : D) sd += v;
  printf("\n*** BS(%d,%d) FOUND ***\nsig = (%d,%d,%d,%d)\n", n + 1, n, sa, sb, sc, sd);
  print_seq("A", A); print_seq("B", B); print_seq("C", C); print_seq("D", D);
  std::vector<int8_t> A8(A.begin(), A.end()), B8(B.begin(), B.end()), C8(C.begin(), C.end()), D8(D.begin(), D.end());
  int mx = 0;
  wzfh_npaf_check(n, A8.data(), B8.data(), C8.data(), D8.data(), &mx);
  printf("VERIFY: max |NPAF[s]| over s=1..%d = %d%s\n", n, mx, mx == 0 ? "  (NPAF==0 confirmed)" : "  (NONZERO!)");
}

// WZ_FH_TELEMETRY (wz_match.cpp:901-950): per-arm diagnostics printed as one FH_TELEM v2 record
// after the summary, never part of the CFGSIG or the checkpoint. stride 1: every completion's
// depth histogram and every score call timed; stride 64: a deterministic 1/64 sample (the
// splitmix64 finalizer of the call's ordinal, low 6 bits zero). Completion times are the wall
// time of each device slice shared out evenly over the candidates it decided.
struct Telem {
  int stride = 0;
  long long cell_ns = 0, score_ns_est = 0, sort_ns = 0, complete_ns = 0;
  long long resume_replay_ns = 0, resume_replay_score_ns_est = 0, resume_replay_sort_ns = 0, resume_cells = 0;
  long long score_calls = 0, score_samples = 0, completions = 0;
  long long hist_candidates = 0, hist_nodes = 0, total_nodes = 0;
  long long unranked_nodes = 0, unranked_count = 0, unranked_ns = 0;
  long long cells_live_done = 0, cells_partial = 0;
  std::vector<long long> depth_nodes;
  long long rank_nodes[10] = {0}, rank_count[10] = {0}, rank_ns[10] = {0}, rank_aborts[10] = {0};
  static long long ns_since(Clock::time_point t) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t).count();
  }
  bool selected(uint64_t ordinal) const {
    uint64_t x = ordinal + 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return ((x ^ (x >> 31)) & (uint64_t)(stride - 1)) == 0;
  }
  void emit(int n, long long worker_ns) const {
    if (!stride) return;
    printf("FH_TELEM {\"version\":2,\"n\":%d,\"stride\":%d,\"worker_ns\":%lld", n, stride, worker_ns);
    const std::pair<const char *, long long> f[] = {
        {"cell_ns", cell_ns}, {"score_ns_est", score_ns_est}, {"sort_ns", sort_ns},
        {"resume_replay_ns", resume_replay_ns}, {"resume_replay_score_ns_est", resume_replay_score_ns_est},
        {"resume_replay_sort_ns", resume_replay_sort_ns}, {"resume_cells", resume_cells},
        {"complete_ns", complete_ns}, {"score_calls", score_calls}, {"score_samples", score_samples},
        {"completions", completions}, {"hist_candidates", hist_candidates}, {"hist_nodes", hist_nodes},
        {"total_nodes", total_nodes}, {"unranked_nodes", unranked_nodes}, {"unranked_count", unranked_count},
        {"unranked_ns", unranked_ns}, {"cells_live_done", cells_live_done}, {"cells_partial", cells_partial}};
    for (const auto &kv : f) printf(",\"%s\":%lld", kv.first, kv.second);
    auto arr = [](const char *key, const long long *v, size_t k) {
      printf(",\"%s\":[", key);
      for (size_t i = 0; i < k; i++) printf("%s%lld", i ? "," : "", v[i]);
      printf("]");
    };
    arr("depth_nodes", depth_nodes.data(), depth_nodes.size());
    arr("rank_nodes", rank_nodes, 10);
    arr("rank_count", rank_count, 10);
    arr("rank_ns", rank_ns, 10);
    arr("rank_aborts", rank_aborts, 10);
    printf("}\n");
    fflush(stdout);
  }
};

// WZ_JOIN22 (wz_match.cpp:2905-3104): the complete hash join over the Thm 2.2 space, with
// the reference's output lines and exit codes (0 found, 3 exhausted).
static int run_join22(int n, int a, int b, int c, int d, Clock::time_point T0) {
  wzfh_config cfg;
  wzfh_config_default(&cfg, n, a, b, c, d);
  if (env_set("WZ_THM211B")) cfg.thm211b = 1;
  if (env_set("WZ_THM212")) cfg.thm212 = 1;
  cfg.rev = env_ll("WZ_FH_STREAM_REV", 0) != 0;
  printf("========================================================\n");
  printf("  wzfh_match (MI355X) -- HASH-JOIN BS(%d,%d), sig=(%d,%d,%d,%d)\n", n + 1, n, a, b, c, d);
  printf("========================================================\n");
  fflush(stdout);
  wzfh_join *J = nullptr;
  wzfh_join_result r{};
  int rc = wzfh_join_open(&cfg, &J);
  if (rc == WZFH_OK) rc = wzfh_join_build(J, (int)env_ll("WZ_JOIN22_SLOTS_LOG2", 0), &r);
  if (rc != WZFH_OK) {
    printf("RESULT: DEVICE ERROR (%s)\n", wzfh_last_error());
    wzfh_join_close(J);
    return 2;
  }
  int lg = 0;
  while ((1ll << lg) < r.slots) lg++;
  const double entry_bytes = 32.0 + 2.0 * ((n + 3) & ~3);
  const double load = r.slots ? (double)r.distinct_keys / (double)r.slots : 0.0;
  printf("[profiles] A,B side surviving: %lld   C,D side surviving: %lld\n", r.ab_profiles, r.cd_profiles);
  printf("[join22v2] key table: 2^%d slots = %g GB\n", lg, (double)r.slots * entry_bytes / 1e9);
  printf("[join22v2] C,D stream=%g  distinct keys=%lld  (dedup x%g)  table load=%g\n", (double)r.cd_stream,
         r.distinct_keys, r.distinct_keys ? (double)r.cd_stream / (double)r.distinct_keys : 0.0, load);
  fflush(stdout);
  rc = wzfh_join_probe(J, &r);
  if (rc != WZFH_OK) {
    printf("RESULT: DEVICE ERROR (%s)\n", wzfh_last_error());
    wzfh_join_close(J);
    return 2;
  }
  printf("[join22v2] phase 2 done: %lld raw key hits\n", r.hits);
  const double t = std::chrono::duration<double>(Clock::now() - T0).count();
  if (r.found) {
    const std::vector<int> A = unpack(r.A, n + 1), B = unpack(r.B, n + 1), C = unpack(r.C, n), D = unpack(r.D, n);
    print_solution(n, A, B, C, D);
    printf("[join22v2] banner flushed AT FIND TIME [%gs] — resolve loop draining, final banner + Time follow\n", t);
    print_solution(n, A, B, C, D);
  } else {
    printf("\n=== JOIN22 EXHAUSTED (n=%d) — no solution in the Thm-2.2 space for sig (%d,%d,%d,%d) ===\n", n, a, b, c, d);
  }
  printf("Time: %gs\n", std::chrono::duration<double>(Clock::now() - T0).count());
  fflush(stdout);
  wzfh_join_close(J);
  return r.found ? 0 : 3;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "Usage: %s <n> <a> <b> <c> <d>\n  e.g.  %s 10 5 1 4 0   (BS(11,10))\n", argv[0], argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]);
  if (n < 4) { fprintf(stderr, "ERROR: n must be >= 4\n"); return 1; }
  const int a = atoi(argv[2]), b = atoi(argv[3]), c = atoi(argv[4]), d = atoi(argv[5]);
  if (a * a + b * b + c * c + d * d != 4 * n + 2) {
    fprintf(stderr, "ERROR: sig norm = %d, expected %d\n", a * a + b * b + c * c + d * d, 4 * n + 2);
    return 1;
  }
  const auto T0 = Clock::now();
  if (env_set("WZ_JOIN22")) return run_join22(n, a, b, c, d, T0);
  if (!env_set("WZ_FIRSTHIT")) {
    fprintf(stderr, "wzfh_match implements the WZ_FIRSTHIT arm and WZ_JOIN22 only (set one of them)\n");
    return 2;
  }
  wzfh_config cfg;
  wzfh_config_default(&cfg, n, a, b, c, d);
  if (env_set("WZ_THM211B")) cfg.thm211b = 1;
  if (env_set("WZ_THM212")) cfg.thm212 = 1;
  if (n >= 36) { cfg.thm211b = 1; cfg.thm212 = 1; }
  cfg.m6 = n >= 36;
  if (const char *e = getenv("WZ_FH_M6")) cfg.m6 = atoi(e) != 0;
  cfg.endpos = env_ll("WZ_FH_ENDPOS", 0) != 0;
  cfg.canon = env_ll("WZ_FH_ORBIT_CANON", 0) != 0;
  cfg.orbit_q = (int)env_ll("WZ_FH_ORBIT_Q", 0);
  cfg.order = (int)env_ll("WZ_FH_PROF_ORDER", 0);
  cfg.rev = env_ll("WZ_FH_STREAM_REV", 0) != 0;
  cfg.outer_k = (int)env_ll("WZ_FH_OUTER_K", 3);
  cfg.no_canon_ab = env_set("WZ_FH_NO_CANON");
  cfg.abp_cap = env_ll("WZ_FH_AB_PROF_CAP", 20000);
  if (const char *e = getenv("WZ_FH_AB_PROF")) cfg.ab_prof = atoi(e) != 0;
  cfg.orbit_qprune = env_ll("WZ_FH_ORBIT_QPRUNE", 0) != 0;
  cfg.a2 = (int)env_ll("WZ_FH_A2", 0);   // wz_match.cpp:1370: 2 prunes profile rows (A2)
  if (env_ll("WZ_FH_CD_PRUNE", 0) || (cfg.a2 != 0 && cfg.a2 != 2) || env_ll("WZ_FH_WALL_SEC", 0) ||
      env_set("WZ_FH_TARGET_C")) {
    printf("RESULT: UNSUPPORTED FLAG (CD_PRUNE, A2 shadow mode, WALL_SEC, TARGET are not implemented)\n");
    return 2;
  }
  Telem TM;
  if (const char *e = getenv("WZ_FH_TELEMETRY")) {   // wz_match.cpp:1545-1552
    if (strcmp(e, "0") && strcmp(e, "1") && strcmp(e, "64")) {
      fprintf(stderr, "WZ_FH_TELEMETRY must be 0, 1 or 64\n");
      return 2;
    }
    TM.stride = atoi(e);
  }
  if (TM.stride) TM.depth_nodes.assign((n + 1) / 2 + 1, 0);
  const long long budget = env_ll("WZ_FH_AB_BUDGET", 200000);
  const long long max_cand = env_ll("WZ_FH_MAX_CAND", 0);
  const long long score_max = env_ll("WZ_FH_SCORE_MAX", 0);
  int shard = (int)env_ll("WZ_FH_SHARD", 0), nshard = (int)env_ll("WZ_FH_NSHARD", 1);
  if (nshard < 1) nshard = 1;
  printf("========================================================\n");
  printf("  wzfh_match (MI355X) -- BS(%d,%d), sig=(%d,%d,%d,%d)\n", n + 1, n, a, b, c, d);
  printf("========================================================\n");
  fflush(stdout);
  if (shard < 0 || shard >= nshard) { printf("[firsthit] BAD SHARD %d/%d\n", shard, nshard); return 2; }

  wzfh_problem *P = nullptr;
  int rc = wzfh_init(&cfg, &P);
  if (rc == WZFH_ETRUNC) {
    printf("RESULT: CELL LIST TRUNCATED at the 20M cap — refusing to search an incomplete tile\n");
    return 2;
  }
  if (rc != WZFH_OK) {
    printf("RESULT: SETUP FAILED (%s)\n", wzfh_last_error());
    return rc == WZFH_EINVAL ? 1 : 2;
  }
  const int m = cfg.m6 ? 6 : 3;
  const int ncells = wzfh_num_cells(P);
  long long qps[7] = {0, 0, 0, 0, 0, 0, 0};
  const bool qp_active = wzfh_qprune_stats(P, qps) == 1;
  if (cfg.canon && cfg.orbit_qprune && !qp_active && cfg.orbit_q &&
      (cfg.orbit_q == 2 || cfg.endpos || ((n % 2 == 0) && (((c + d) % 4 + 4) % 4 == 0))))
    printf("[qprune] DISABLED (needs mod-6 cells and an untruncated cell list)\n");
  if (qp_active)
    printf("[qprune] missing_Q_images=%lld not_realizable=%lld eq212=%lld nonintegral=%lld unknown=%lld dead_orbits=%lld"
           " cells_removed=%lld\n", qps[0], qps[1], qps[2], qps[3], qps[4], qps[5], qps[6]);
  const bool pins_dropped = cfg.canon && (c == 0 || d == 0);
  char digest[64] = {0};
  wzfh_order_digest(P, digest, sizeof digest);
  if (cfg.order == 3) {
    printf("[order] ord3 cells=%d digest=%s\n", ncells, digest);
    if (const char *e = getenv("WZ_FH_EXPECT_DIGEST"))
      if (strcmp(e, digest) != 0) {
        printf("RESULT: DIGEST MISMATCH (expected %s) — refusing to search or resume\n", e);
        return 2;
      }
    if (env_set("WZ_FH_LIST_ONLY")) return 0;
    if (!env_set("WZ_FH_EXPECT_DIGEST") && !env_set("WZ_FH_CELLSIZE") && !env_set("WZ_FH_ORD3_NODIGEST_OK")) {
      printf("RESULT: PROF_ORDER=3 search requires WZ_FH_EXPECT_DIGEST (from the manifest)\n");
      return 2;
    }
  }
  printf("[firsthit] profiles=%d m=%d order=%d ab_budget=%lld max_cand=%lld score_max=%lld shard=%d/%d\n", ncells, m,
         cfg.order, budget, max_cand, score_max, shard, nshard);
  fflush(stdout);

  // ---- complete-only mode (wz_match.cpp:2403-2455): unconstrained rows
  if (const char *cc = getenv("WZ_FH_COMPLETE_C")) {
    std::vector<int> C0 = parse_list(cc), D0 = parse_list(getenv("WZ_FH_COMPLETE_D"));
    if ((int)C0.size() != n || (int)D0.size() != n) { printf("COMPLETE_ONLY: bad C/D length\n"); return 2; }
    uint64_t cd[2] = {pack(C0), pack(D0)};
    wzfh_completion out;
    auto t0 = Clock::now();
    int r2 = wzfh_complete(P, cd, nullptr, 1, budget, &out);
    if (r2 != WZFH_OK && r2 != WZFH_EINTERNAL) { printf("RESULT: DEVICE ERROR (%s)\n", wzfh_last_error()); return 2; }
    printf("COMPLETE_ONLY r=%d nodes=%lld budget=%lld outer_k=%d a2=%d sec=%g\n", out.verdict, out.nodes, budget,
           cfg.outer_k, cfg.a2, std::chrono::duration<double>(Clock::now() - t0).count());
    // a2_print (wz_match.cpp:1039-1047): complete-only runs without rows, so A2 tests none
    if (cfg.a2)
      printf("A2SHADOW mode=%d rows_tested=0 rows_rejected=0 cuts=0 saved_nodes=0 total_nodes=%lld hit_under_cut=0 "
             "cuts_by_depth= saved_by_depth=\n", cfg.a2, out.nodes);
    if (out.verdict == 0) {
      printf("*** BS(%d,%d) FOUND ***  (complete-only)\n", n + 1, n);
      print_seq("A", unpack(out.A, n + 1)); print_seq("B", unpack(out.B, n + 1));
      print_seq("C", C0); print_seq("D", D0);
    }
    return 0;
  }

  // ---- arm options (wz_match.cpp:2098-2240)
  const long long fh_skip = env_ll("WZ_FH_PROF_SKIP", 0), fh_end = env_ll("WZ_FH_PROF_END", 0);
  const long long buf_cap = env_ll("WZ_FH_BUF_CAP", 500000);
  const long long drain_top = env_ll("WZ_FH_DRAIN_TOP", 0);
  const long long drain_batches = std::max(1LL, env_ll("WZ_FH_DRAIN_BATCHES", 1));
  const long long cellsize = std::max(0LL, env_ll("WZ_FH_CELLSIZE", 0));
  const bool cell_order = cfg.m6 ? env_ll("WZ_FH_CELL_ORDER", 1) != 0 : env_ll("WZ_FH_CELL_ORDER", 0) != 0;
  const bool ab_prof = cfg.ab_prof < 0 ? cfg.m6 != 0 : cfg.ab_prof != 0;
  const bool cand_log = env_ll("WZ_FH_CAND_LOG", 0) != 0;
  const long long test_stop = env_ll("WZ_FH_TEST_STOP_AFTER", 0);
  const bool test_ierr = env_set("WZ_FH_TEST_INTERNAL_ERROR");
  double prog_sec = 60;
  if (const char *e = getenv("WZ_FH_PROG_SEC")) prog_sec = atof(e);
  std::string ckpt_path;
  if (const char *e = getenv("WZ_FH_CKPT_DIR")) ckpt_path = std::string(e) + "/arm_" + std::to_string(shard) + ".ckpt";
  bool resume_on = env_ll("WZ_FH_RESUME", 1) != 0;
  if (cellsize > 0) {
    ckpt_path.clear();
    resume_on = false;
    printf("[cellsize] MEASUREMENT MODE cap=%lld (no completions, no checkpoints)\n", cellsize);
  }
  FILE *dump = nullptr;
  if (const char *e = getenv("WZ_FH_DUMP")) dump = fopen(e, "w");

  // CFGSIG (wz_match.cpp:2241-2287): same fields, same append-only-when-non-default rule
  char sig[600];
  snprintf(sig, sizeof sig, "n%d.a%d.b%d.c%d.d%d.ns%d.sh%d.ord%d.m%d.co%d.ap%d.sm%lld.cap%lld.sk%lld.t1%d.t2%d.oc%d.dt%lld", n, a,
           b, c, d, nshard, shard, cfg.order, m, cell_order ? 1 : 0, ab_prof ? 1 : 0, score_max, buf_cap, fh_skip,
           cfg.thm211b ? 1 : 0, cfg.thm212 ? 1 : 0, cfg.canon ? 1 : 0, drain_top);
  std::string cfgsig = sig;
  if (drain_batches > 1) cfgsig += ".db" + std::to_string(drain_batches);
  if (cfg.rev) cfgsig += ".sr1";
  if (pins_dropped) cfgsig += ".np1";
  if (cfg.endpos) cfgsig += ".ep1";
  if (cfg.canon && cfg.orbit_q) {
    const bool q_safe = cfg.endpos || ((n % 2 == 0) && (((c + d) % 4 + 4) % 4 == 0));
    if (cfg.orbit_q == 2 || q_safe) cfgsig += ".oq1";
  }
  if (qp_active) cfgsig += ".qp1";
  if (cfg.order == 3) cfgsig += ".bud" + std::to_string(budget) + ".dg" + (digest[0] ? std::string(digest) : "none");

  long long cand = 0, pre_rej = 0, score_rej = 0, clean_no = 0, aborted = 0, cells_done = 0, cells_empty = 0;
  long long cells_dup = 0, cells_dead = 0, cells_uncap = 0, cells_capped = 0, ab_nodes = 0;
  long long hit_idx = -1, tested_base = 0;
  long long cum_done_base = 0, cum_dup_base = 0, cum_dead_base = 0, cum_empty_base = 0;
  bool resuming = false;
  long long res_pi = 0, res_batch = 0, res_k = 0;
  if (!ckpt_path.empty() && resume_on) {
    if (FILE *cf = fopen(ckpt_path.c_str(), "r")) {
      char line[700], fsig[620] = {0};
      long long rpi = -1, rb = -1, rk = -1, tc = 0, cdc = 0, cdu = 0, cde = 0, cem = 0;
      while (fgets(line, sizeof line, cf)) {
        if (sscanf(line, "CFGSIG=%619s", fsig) == 1) continue;
        if (sscanf(line, "resume_pi=%lld", &rpi) == 1) continue;
        if (sscanf(line, "resume_batch=%lld", &rb) == 1) continue;
        if (sscanf(line, "resume_k=%lld", &rk) == 1) continue;
        if (sscanf(line, "tested_cum=%lld", &tc) == 1) continue;
        if (sscanf(line, "cells_done_cum=%lld", &cdc) == 1) continue;
        if (sscanf(line, "cells_dup_cum=%lld", &cdu) == 1) continue;
        if (sscanf(line, "cells_dead_cum=%lld", &cde) == 1) continue;
        sscanf(line, "cells_empty_cum=%lld", &cem);
      }
      fclose(cf);
      if (cfgsig == fsig && rpi >= 0 && rb >= 0 && rk >= 0) {
        resuming = true;
        res_pi = rpi; res_batch = rb; res_k = rk; tested_base = tc;
        cum_done_base = cdc; cum_dup_base = cdu; cum_dead_base = cde; cum_empty_base = cem;
        printf("[firsthit ckpt] RESUME pi=%lld batch=%lld k=%lld tested_cum=%lld\n", rpi, rb, rk, tc);
      } else {
        printf("[firsthit ckpt] STALE/MISMATCHED checkpoint IGNORED (%s != %s) — starting FRESH\n", fsig, cfgsig.c_str());
      }
    }
  }
  signal(SIGTERM, on_sigterm);

  long long ck_pi = shard + fh_skip * nshard, ck_batch = 0, ck_k = 0;
  if (resuming) { ck_pi = res_pi; ck_batch = res_batch; ck_k = res_k; }
  auto write_ckpt = [&]() {
    if (ckpt_path.empty()) return;
    std::string tmp = ckpt_path + ".tmp";
    FILE *cf = fopen(tmp.c_str(), "w");
    if (!cf) return;
    fprintf(cf, "CFGSIG=%s\nresume_pi=%lld\nresume_batch=%lld\nresume_k=%lld\ntested_cum=%lld\n", cfgsig.c_str(), ck_pi,
            ck_batch, ck_k, tested_base + clean_no + aborted + (hit_idx >= 0 ? 1LL : 0LL));
    fprintf(cf, "cells_done_cum=%lld\ncells_dup_cum=%lld\ncells_dead_cum=%lld\ncells_empty_cum=%lld\n",
            cum_done_base + cells_done, cum_dup_base + cells_dup, cum_dead_base + cells_dead, cum_empty_base + cells_empty);
    fclose(cf);
    rename(tmp.c_str(), ckpt_path.c_str());
  };
  long long ck_tick = 0;
  auto last_ckpt = T0, last_prog = T0;
  auto maybe_ckpt = [&]() {
    if (ckpt_path.empty()) return;
    if ((++ck_tick & 63) != 0) return;
    auto now = Clock::now();
    if (std::chrono::duration<double>(now - last_ckpt).count() >= prog_sec) { last_ckpt = now; write_ckpt(); }
  };

  bool fh_stop = false, found = false, internal = false;
  long long test_ct = 0;
  std::vector<int> solA, solB, solC, solD;
  long long hit_nodes = 0, hit_score = 0;
  int hit_pi = -1;

  const long long pi_end = fh_end > 0 ? std::min<long long>(ncells, shard + fh_end * nshard) : ncells;
  // the cells this arm streams, in order (orbit duplicates and proven-dead cells are skipped)
  struct Owned { int pi; int kind; };   // kind 0 = live, 1 = orbit dup, 2 = dead
  std::vector<Owned> owned;
  for (long long pi = shard + fh_skip * nshard; pi < pi_end; pi += nshard) {
    int px[8], py[8], kept, rows;
    wzfh_cell_info(P, (int)pi, px, py, &kept, &rows);
    int kind = 0;
    if (cfg.canon && !kept) kind = 1;
    else if (ab_prof && rows == -1) kind = 2;
    else if (ab_prof && rows > cfg.abp_cap) kind = 3;   // live, runs without its rows
    owned.push_back({(int)pi, kind});
  }
  // per-cell stream cap the policy needs (0 = whole cell)
  long long stream_cap = 0;
  if (cellsize > 0) stream_cap = cellsize;
  else if (cell_order && drain_top > 0) stream_cap = drain_top > 0 ? drain_batches * buf_cap : 0;

  const int kWindow = (int)std::max(1LL, env_ll("WZFH_PREFETCH_CELLS", 256));
  // completion slices of the drain (candidates per wzfh_complete call; adapted to slice_sec)
  const size_t slice_max = (size_t)std::max(1LL, env_ll("WZFH_DRAIN_SLICE_MAX", 4096));
  size_t slice_len = std::min<size_t>(16, slice_max);
  double slice_sec = 2.0;
  if (const char *e = getenv("WZFH_DRAIN_SLICE_SEC")) slice_sec = std::max(0.0, atof(e));
  std::vector<int> win_cells;
  std::vector<wzfh_cell_stats> win_stats;
  std::vector<long long> win_offs;
  std::vector<uint64_t> win_cd;
  size_t win_begin = 0, win_end = 0;   // owned-index range currently streamed
  auto stream_window = [&](size_t from, long long cap_now, int max_cells) -> int {
    win_cells.clear();
    size_t i = from;
    for (; i < owned.size() && (int)win_cells.size() < max_cells; i++)
      if ((owned[i].kind == 0 || owned[i].kind == 3) && !(resuming && owned[i].pi < res_pi)) win_cells.push_back(owned[i].pi);
    win_begin = from; win_end = i;
    if (win_cells.empty()) return WZFH_OK;
    wzfh_stream_opts so{};
    so.cand_cap = cap_now;
    win_stats.assign(win_cells.size(), wzfh_cell_stats{});
    win_offs.assign(win_cells.size() + 1, 0);
    int r = wzfh_stream_cells(P, win_cells.data(), (int)win_cells.size(), &so, win_stats.data(), nullptr, 0, win_offs.data());
    if (r != WZFH_OK) return r;
    win_cd.assign(2 * std::max<long long>(1, win_offs.back()), 0);
    return wzfh_stream_cells(P, win_cells.data(), (int)win_cells.size(), &so, win_stats.data(), win_cd.data(),
                             win_offs.back(), win_offs.data());
  };

  size_t oi = 0;
  while (oi < owned.size() && !fh_stop) {
    const int pi = owned[oi].pi;
    if (resuming && pi < res_pi) { oi++; continue; }
    if (resuming && pi > res_pi) resuming = false;
    if (owned[oi].kind == 1) {
      cells_dup++;
      if (resuming && pi == res_pi) resuming = false;
      ck_pi = pi + nshard; ck_batch = 0; ck_k = 0;
      oi++;
      continue;
    }
    if (owned[oi].kind == 2) {
      cells_dead++; cells_done++;
      if (resuming && pi == res_pi) resuming = false;
      ck_pi = pi + nshard; ck_batch = 0; ck_k = 0;
      oi++;
      continue;
    }
    if (owned[oi].kind == 3) cells_uncap++;
    // streamed window: cells [win_begin, win_end) of `owned`
    long long cap_now = stream_cap;
    // a cell resumed at drain batch res_batch replays the batches before it and streams one more
    // buffer even when res_batch is already past the front (the drain that finds the front done
    // runs when that buffer fills, :2664-2676)
    const bool resume_cell = resuming && pi == res_pi && cell_order && drain_top > 0 && stream_cap > 0;
    if (resume_cell) cap_now = std::max(drain_batches, res_batch + 1) * buf_cap;
    if (max_cand > 0) {   // enough to know whether a candidate follows the max_cand-th one
      long long need = max_cand - cand + 1;
      cap_now = cap_now > 0 ? std::min(cap_now, need) : need;
    }
    if (!(oi >= win_begin && oi < win_end) || max_cand > 0 || resume_cell) {
      int r = stream_window(oi, cap_now, resume_cell ? 1 : kWindow);
      if (r != WZFH_OK) { printf("RESULT: DEVICE ERROR (%s)\n", wzfh_last_error()); return 2; }
    }
    int wi = -1;
    for (size_t t = 0; t < win_cells.size(); t++) if (win_cells[t] == pi) { wi = (int)t; break; }
    const long long b0 = win_offs[wi], e0 = win_offs[wi + 1];
    const wzfh_cell_stats &cst = win_stats[wi];
    ck_pi = pi;
    if (!(resuming && pi == res_pi)) { ck_batch = 0; ck_k = 0; }
    const long long cand_at_start = cand;
    // telemetry of this cell (wz_match.cpp:2532-2546, 2848-2853): a resumed cell replays the
    // batches (or the ranks) before its checkpoint, and that replay is timed apart
    const auto tm_cell_start = Clock::now();
    bool tm_replay = TM.stride && resuming && pi == res_pi && (res_batch > 0 || res_k > 0);
    if (tm_replay) TM.resume_cells++;
    const long long tm_score0 = TM.score_ns_est, tm_sort0 = TM.sort_ns;
    auto tm_end_replay = [&]() {
      if (!tm_replay) return;
      TM.resume_replay_ns += Telem::ns_since(tm_cell_start);
      TM.resume_replay_score_ns_est += TM.score_ns_est - tm_score0;
      TM.resume_replay_sort_ns += TM.sort_ns - tm_sort0;
      tm_replay = false;
    };
    auto tm_cell_end = [&]() {
      if (!TM.stride) return;
      tm_end_replay();
      TM.cell_ns += Telem::ns_since(tm_cell_start);
      if (fh_stop) TM.cells_partial++;
      else if (cand > cand_at_start) TM.cells_live_done++;
    };

    if (cellsize > 0) {   // measurement: count only (wz_match.cpp:2749-2781, 2814-2836)
      for (long long j = b0; j < e0; j++) {
        cand++;
        if (dump) {
          for (int i = 0; i < n; i++) fprintf(dump, "%d ", ((win_cd[2 * j] >> i) & 1) ? -1 : 1);
          fprintf(dump, "| ");
          for (int i = 0; i < n; i++) fprintf(dump, "%d ", ((win_cd[2 * j + 1] >> i) & 1) ? -1 : 1);
          fprintf(dump, "\n");
        }
      }
      const long long cc2 = cand - cand_at_start;
      printf("CELLSIZE pi=%d cand=%lld capped=%d partial=0 sec=0 sec_at_buf=-1 leaves=%lld hall_ok=%lld\n", pi, cc2,
             cc2 >= cellsize ? 1 : 0, cst.leaves, cst.hall_ok);
      printf("CDSTAT pi=%d dfs_nodes=%lld cd_prune=0 hall_fast=1\n", pi, cst.dfs_nodes);
      cells_done++;
      if (cand == cand_at_start) cells_empty++;
      tm_cell_end();
      oi++;
      continue;
    }

    // ---- probe replay (wz_match.cpp:2715-2808) with the buffered drain (:2656-2714)
    struct Buf { long long sc; uint64_t C, D; long long pos; };
    std::vector<Buf> cellbuf;
    long long cur_batch = 0, cell_done_ct = 0;
    bool cell_stop = false;
    // The sorted ranks [start, stop_at) go to the device in slices, so SIGTERM, a hit and the
    // checkpoint clock are looked at between slices (the reference looks after every
    // completion, :2686-2697): a slice is sized to take about slice_sec of device time
    // (doubling from 16 candidates while slices finish faster, halving when one is slower).
    std::vector<long long> slice_depth;
    long long slice_ns = 0;             // share of the current slice's time for the candidate at hand
    const long long *slice_dn = nullptr;
    int tm_rank = -1;                   // decile of the sorted rank (-1: unbuffered, no rank)
    auto complete_batch = [&](const std::vector<Buf> &items, size_t start, size_t stop_at, auto &&on_each) {
      std::vector<uint64_t> cd;
      std::vector<int> cellof;
      std::vector<wzfh_completion> out;
      for (size_t s0 = start; s0 < stop_at;) {
        if (g_sigterm || found) return false;
        const size_t s1 = std::min(stop_at, s0 + slice_len);
        cd.clear();
        cellof.clear();
        for (size_t ci = s0; ci < s1; ci++) { cd.push_back(items[ci].C); cd.push_back(items[ci].D); cellof.push_back(ab_prof ? pi : -1); }
        out.assign(s1 - s0, wzfh_completion{});
        const auto t0 = Clock::now();
        int r;
        if (TM.stride) {   // whole walks with per-depth node counts (FH_TELEM depth_nodes)
          slice_depth.assign((s1 - s0) * 64, 0);
          r = wzfh_complete_depth(P, cd.data(), cellof.data(), (int)(s1 - s0), budget, out.data(), slice_depth.data());
        } else {
          r = wzfh_complete(P, cd.data(), cellof.data(), (int)(s1 - s0), budget, out.data());
        }
        if (r != WZFH_OK && r != WZFH_EINTERNAL) { printf("RESULT: DEVICE ERROR (%s)\n", wzfh_last_error()); exit(2); }
        const long long dns = Telem::ns_since(t0);
        const double dt = 1e-9 * (double)dns;
        if (dt < 0.5 * slice_sec && s1 - s0 == slice_len) slice_len = std::min<size_t>(slice_len * 2, slice_max);
        else if (dt > 2 * slice_sec) slice_len = std::max<size_t>(slice_len / 2, 1);
        const long long k = (long long)(s1 - s0);
        for (size_t ci = s0; ci < s1; ci++) {
          const long long i = (long long)(ci - s0);
          slice_ns = dns * (i + 1) / k - dns * i / k;   // an even share of the slice's wall time
          slice_dn = TM.stride ? &slice_depth[(size_t)i * 64] : nullptr;
          if (!on_each(ci, out[ci - s0])) return false;
        }
        s0 = s1;
      }
      return true;
    };
    auto complete_one = [&](size_t rank, const Buf &item, const wzfh_completion &o) -> bool {
      if (g_sigterm || found) return false;
      int r = o.verdict;
      if (r == 0 && test_ierr) r = 4;
      if (r == 4) {
        printf("FH_INTERNAL_ERROR: completer FOUND but NPAF != 0\n");
        printf("RESULT: INTERNAL ERROR at candidate idx=%lld pi=%d — arm stopped, candidate NOT marked attempted, checkpoint NOT advanced\n", cand, pi);
        fflush(stdout);
        internal = true;
        return false;
      }
      ab_nodes += o.nodes;
      if (TM.stride) {   // wz_match.cpp:2587-2614
        const bool hist = TM.selected((uint64_t)TM.completions);
        TM.complete_ns += slice_ns;
        TM.completions++;
        TM.total_nodes += o.nodes;
        if (hist) {
          TM.hist_candidates++;
          TM.hist_nodes += o.nodes;
          for (size_t d = 0; d < TM.depth_nodes.size(); d++) TM.depth_nodes[d] += slice_dn[d];
        }
        if (tm_rank >= 0) {
          TM.rank_count[tm_rank]++;
          TM.rank_nodes[tm_rank] += o.nodes;
          TM.rank_ns[tm_rank] += slice_ns;
          TM.rank_aborts[tm_rank] += (r == 3);
        } else {
          TM.unranked_count++; TM.unranked_nodes += o.nodes; TM.unranked_ns += slice_ns;
        }
      }
      if (cand_log) printf("CAND idx=%lld ci=%lld pi=%d r=%d nodes=%lld ns=0\n", cand, (long long)rank, pi, r, o.nodes);
      if (test_stop > 0 && ++test_ct >= test_stop) { g_sigterm = 1; fh_stop = true; }
      if (r == 1) pre_rej++;
      else if (r == 2) clean_no++;
      else if (r == 3) aborted++;
      else {
        hit_idx = cand; hit_pi = pi; found = true; fh_stop = true;
        hit_nodes = o.nodes;
        solA = unpack(o.A, n + 1); solB = unpack(o.B, n + 1); solC = unpack(item.C, n); solD = unpack(item.D, n);
        {   // the FIRSTHIT line's score is one more flat-score call (wz_match.cpp:2640)
          const bool sample = TM.stride && TM.selected((uint64_t)TM.score_calls++);
          const auto st0 = Clock::now();
          hit_score = wzfh_flat_score(n, item.C, item.D);
          if (sample) { TM.score_ns_est += Telem::ns_since(st0) * TM.stride; TM.score_samples++; }
        }
        return false;
      }
      return !g_sigterm;
    };
    auto drain = [&]() {
      if (resuming && pi == res_pi && cur_batch < res_batch) {
        cur_batch++;
        cellbuf.clear();
        if (cur_batch == res_batch && res_k == 0) tm_end_replay();   // the next buffer is new work
        return;
      }
      if (drain_top > 0 && cur_batch >= drain_batches) {
        resuming = false;
        tm_end_replay();
        cellbuf.clear();
        if (!cell_stop) { cell_stop = true; cells_capped++; }
        return;
      }
      size_t start = 0;
      if (resuming && pi == res_pi && cur_batch == res_batch) {
        start = (size_t)std::min<long long>((long long)cellbuf.size(), res_k);
        resuming = false;
      }
      const auto sort_t0 = Clock::now();
      std::stable_sort(cellbuf.begin(), cellbuf.end(), [](const Buf &x, const Buf &y) { return x.sc < y.sc; });
      if (TM.stride) TM.sort_ns += Telem::ns_since(sort_t0);
      if (!resuming) tm_end_replay();
      size_t stop_at = cellbuf.size();
      if (drain_top > 0 && cur_batch < drain_batches) stop_at = std::min(cellbuf.size(), (size_t)drain_top);
      complete_batch(cellbuf, start, stop_at, [&](size_t ci, const wzfh_completion &o) {
        tm_rank = TM.stride ? (int)(10 * ci / stop_at) : -1;
        bool go = complete_one(ci, cellbuf[ci], o);
        if (internal) return false;
        if (!found) { ck_batch = cur_batch; ck_k = (long long)ci + 1; maybe_ckpt(); }
        else { ck_batch = cur_batch; ck_k = (long long)ci + 1; }
        return go;
      });
      if (internal) return;
      if (!(found || g_sigterm)) {
        cur_batch++;
        ck_batch = cur_batch;
        ck_k = 0;
        if (drain_top > 0 && cur_batch >= drain_batches) { cell_stop = true; cells_capped++; }
      }
      cellbuf.clear();
    };
    std::vector<Buf> direct;   // non-buffered path (cell_order off)
    for (long long j = b0; j < e0 && !internal; j++) {
      if (fh_stop) { cell_stop = true; break; }
      if (cell_stop) break;
      cand++;
      if ((cand & 255) == 0) {
        auto nowp = Clock::now();
        if (std::chrono::duration<double>(nowp - last_prog).count() >= prog_sec) {
          last_prog = nowp;
          printf("[firsthit progress] candidates_streamed=%lld  pre_filter_rejected=%lld  score_rejected=%lld  clean_no_AB=%lld  budget_aborted=%lld  total_AB_nodes=%lld  [%gs]\n",
                 cand, pre_rej, score_rej, clean_no, aborted, ab_nodes, std::chrono::duration<double>(nowp - T0).count());
          fflush(stdout);
        }
      }
      const uint64_t Cw = win_cd[2 * j], Dw = win_cd[2 * j + 1];
      if (dump) {
        for (int i = 0; i < n; i++) fprintf(dump, "%d ", ((Cw >> i) & 1) ? -1 : 1);
        fprintf(dump, "| ");
        for (int i = 0; i < n; i++) fprintf(dump, "%d ", ((Dw >> i) & 1) ? -1 : 1);
        fprintf(dump, "\n");
      }
      long long sc = -1;
      if (score_max > 0 || cell_order) {
        const bool sample = TM.stride && TM.selected((uint64_t)TM.score_calls++);
        const auto st0 = Clock::now();
        sc = wzfh_flat_score(n, Cw, Dw);
        if (sample) { TM.score_ns_est += Telem::ns_since(st0) * TM.stride; TM.score_samples++; }
      }
      if (score_max > 0 && sc > score_max) { score_rej++; continue; }
      if (cell_order) {
        cellbuf.push_back({sc, Cw, Dw, j - b0});
        if ((long long)cellbuf.size() >= buf_cap) drain();
        if (max_cand > 0 && cand >= max_cand) fh_stop = true;
        continue;
      }
      // non-buffered: completion order == eligible-stream order
      if (resuming && pi == res_pi) {
        if (cell_done_ct < res_k) {
          cell_done_ct++;
          if (cell_done_ct == res_k) tm_end_replay();
          continue;
        }
        resuming = false;
      }
      tm_rank = -1;
      Buf item{sc, Cw, Dw, j - b0};
      std::vector<Buf> one{item};
      const long long rank = cell_done_ct;
      complete_batch(one, 0, 1, [&](size_t, const wzfh_completion &o) { return complete_one((size_t)rank, item, o); });
      cell_done_ct++;
      ck_batch = 0;
      ck_k = cell_done_ct;
      maybe_ckpt();
      if (max_cand > 0 && cand >= max_cand) fh_stop = true;
    }
    if (internal) return 5;
    if (cell_order && !cellbuf.empty() && !cell_stop) drain();
    if (internal) return 5;
    if (!fh_stop && !g_sigterm) {
      cells_done++;
      if (cand == cand_at_start) cells_empty++;
      ck_pi = pi + nshard; ck_batch = 0; ck_k = 0;
      write_ckpt();
    }
    tm_cell_end();
    oi++;
  }
  if (dump) fclose(dump);
  const long long completed = clean_no + aborted + (hit_idx >= 0 ? 1 : 0);
  write_ckpt();
  if (found) {
    const double t = std::chrono::duration<double>(Clock::now() - T0).count();
    int sa = 0, sb = 0, sc = 0, sd = 0;
    for (int v : solA) sa += v;
    for (int v : solB) sb += v;
    for (int v : solC) sc += v;
    for (int v : solD) sd += v;
    printf("\n*** BS(%d,%d) FOUND ***  (firsthit probe)\nsig = (%d,%d,%d,%d)\n", n + 1, n, sa, sb, sc, sd);
    print_seq("A", solA); print_seq("B", solB); print_seq("C", solC); print_seq("D", solD);
    std::vector<int8_t> A8(solA.begin(), solA.end()), B8(solB.begin(), solB.end()), C8(solC.begin(), solC.end()),
        D8(solD.begin(), solD.end());
    int mx = 0;
    wzfh_npaf_check(n, A8.data(), B8.data(), C8.data(), D8.data(), &mx);
    printf("VERIFY: max |NPAF[s]| over s=1..%d = %d%s\n", n, mx, mx == 0 ? "  (NPAF==0 confirmed)" : "  (NONZERO!)");
    printf("FIRSTHIT: idx=%lld profile_rank=%d nodes_this_cand=%lld score=%lld elapsed=%gs\n", hit_idx, hit_pi, hit_nodes,
           hit_score, t);
  }
  printf("\n=== FIRSTHIT SUMMARY (n=%d, sig %d,%d,%d,%d, shard %d/%d) ===\n", n, a, b, c, d, shard, nshard);
  printf("candidates_streamed=%lld  pre_filter_rejected=%lld  score_rejected=%lld\n", cand, pre_rej, score_rej);
  printf("backtracks_entered=%lld  clean_no_AB=%lld  budget_aborted=%lld  total_AB_nodes=%lld  cells_done=%lld (skip=%lld)"
         "  ab_prof=%d  cells_prof_dead=%lld  cells_prof_uncap=%lld  cells_orbit_dup=%lld  cells_empty=%lld  cum_done=%lld"
         "  cum_dup=%lld  cum_dead=%lld  cum_empty=%lld  cells_capped=%lld (drain_top=%lld batches=%lld)  cells_walled=0"
         " (wall_sec=0)  range_end=%lld  pins_dropped=%d  tested_cum=%lld  resume_pi=%lld  resume_batch=%lld  resume_k=%lld\n",
         completed, clean_no, aborted, ab_nodes, cells_done, fh_skip, ab_prof ? 1 : 0, cells_dead, cells_uncap, cells_dup,
         cells_empty, cum_done_base + cells_done, cum_dup_base + cells_dup, cum_dead_base + cells_dead,
         cum_empty_base + cells_empty, cells_capped, drain_top, drain_batches, fh_end, pins_dropped ? 1 : 0,
         tested_base + completed, ck_pi, ck_batch, ck_k);
  if (hit_idx >= 0) printf("RESULT: FOUND at idx=%lld\n", hit_idx);
  else if (g_sigterm) printf("RESULT: INTERRUPTED (SIGTERM) — counts are LOWER BOUNDS, stream NOT exhausted\n");
  else if (fh_end > 0 && !fh_stop) printf("RESULT: RANGE EXHAUSTED — this arm completed its owned cells [skip,end)\n");
  else printf("RESULT: NO HIT within probe budget — NOT a proof of absence\n");
  printf("Time: %gs\n", std::chrono::duration<double>(Clock::now() - T0).count());
  fflush(stdout);
  TM.emit(n, TM.stride ? Telem::ns_since(T0) : 0);
  wzfh_destroy(P);
  return hit_idx >= 0 ? 0 : 3;
}
