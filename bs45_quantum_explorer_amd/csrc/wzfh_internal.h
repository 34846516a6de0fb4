// wzfh_internal.h -- shared host-side structures of libwzfh.so.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/wzfh.h"
#include "wzfh_kernels.h"

namespace wzfh {

constexpr int kMaxL = 64;          // one 64-bit word per sequence
constexpr int kMaxDepth = 32;

struct HostCell {
  int8_t px[8];
  int8_t py[8];
};

// One allowed A,B mod-m class-sum row (AbpRow, wz_match.cpp:970).
struct AbRow {
  int8_t k[8];
  int8_t r[8];
};

struct DeviceState;                // wzfh_device.hip

// Thm 2.2 quad tables (wz_match.cpp:531-545): index i is the sign pattern
// (i&8, i&4, i&2, i&1) of (X[d], Y[d], X[L-1-d], Y[L-1-d]); POS/NEG keep ascending i.
struct Quads {
  int all[16][4], pos[8][4], neg[8][4];
  Quads();
};
extern const Quads kQuads;

// A1 outer-lag tables (wz_match.cpp:547-598), built from their definition.
struct OuterTables {
  uint8_t t2[16][5][9];
  uint8_t t3[128][5][9][13];
  int8_t pos_index[16];
  OuterTables();
};
extern const OuterTables kOuter;

std::string set_error(const std::string &msg);   // records the thread's last error

}  // namespace wzfh

struct wzfh_problem {
  wzfh_config cfg;
  int n = 0, L1 = 0, m = 6;
  bool pinC = false, pinD = false, pins_dropped = false, ab_prof = false;
  bool ab_side = false;                // the A,B side of the join: streams A,B (length n+1)
  int side_len() const { return ab_side ? L1 : n; }
  std::vector<wzfh::HostCell> cells;
  std::vector<uint8_t> kept;
  std::vector<int> cell_rows;          // per cell: index into row_off/row_len, -1 = dead
  std::vector<int> row_off, row_len;   // per row list
  std::vector<wzfh::AbRow> rows;
  std::string digest;                  // order-3 "<list>:<kept>"
  bool qp_active = false;              // the Q-closure prune ran
  long long qp_stats[7] = {0, 0, 0, 0, 0, 0, 0};
  double hcos[201][wzfh::kMaxL];
  double hsin[201][wzfh::kMaxL];
  wzfh::DeviceState *dev = nullptr;
};

namespace wzfh {
// Host setup (wzfh_setup.cpp).
int build_problem(wzfh_problem &P);
int side_profiles3(const wzfh_config &c, int n, bool ab_side, std::vector<HostCell> &out);
// The stream of a batch of cells (wzfh_api.hip): per-cell stats and the candidates in
// each cell's DFS order (CandRec::ell holds the global DFS index G of the leaf).
struct StreamOut {
  std::vector<CandRec> cands;                 // by cell slot, then G
  std::vector<long long> cell_begin;          // per slot, into cands (ncells+1)
};
// cell_prob (optional): per cell slot, the problem whose cell list cells[i] indexes (problems
// of one n with the same stream tables, see stream_key); null: every cell is P's.
// cell_blocks > 0: at most that many persistent blocks per cell (a launch of a few cells then
// leaves most of the GPU to concurrent launches of other problems); 0: the whole grid.
int run_stream(wzfh_problem &P, const int *cells, int ncells, const wzfh_stream_opts &o, wzfh_cell_stats *stats,
               StreamOut &out, const wzfh_problem *const *cell_prob = nullptr, int cell_blocks = 0);
// The flattest K candidates of every cell of a stream batch, in sorted order per cell.
struct TopK {
  std::vector<uint64_t> cd;          // packed (C, D) pairs
  std::vector<int> owner;            // cell slot
  std::vector<long long> rank, sidx; // sorted rank, in-cell stream index
};
void select_topk(int n, const StreamOut &so, int ncells, int K, TopK &t);
// Per-cell drain counters from the completions of t (results must hold the stream stats).
void aggregate_completions(const TopK &t, const wzfh_completion *comp, wzfh_search_result *res);
struct Carve;
// dev (optional): run on this device state (complete_state) instead of P's own.
int run_complete(wzfh_problem &P, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                 wzfh_completion *out, Carve *carve = nullptr, long long *depth_nodes = nullptr,
                 DeviceState *dev = nullptr);
// A completer run that candidates join while it runs (the serving loop's continuous batching):
// before every round more() appends the packed (C, D) pairs and cell indices of the candidates
// that arrived and returns how many; done(i, c) reports candidate i (in joining order) as soon as
// it is decided. The run ends when no candidate is left and more() adds none. A candidate's
// result does not depend on what else shares the run.
struct CompleteFeed {
  std::function<int(std::vector<uint64_t> &cd, std::vector<int> &cell_of)> more;
  std::function<void(int, const wzfh_completion &)> done;
};
int run_complete_feed(wzfh_problem &P, long long budget, DeviceState *dev, CompleteFeed &feed);
// A device state for completer runs of P only (own HIP stream and buffers, copies of P's
// completer tables), so several completion runs of one problem can run side by side.
int complete_state(const wzfh_problem &P, DeviceState **out);
void complete_state_free(DeviceState *S);
void complete_state_sync_tuning(const wzfh_problem &P, DeviceState *S);   // copies P's launch settings
int state_kernel_totals(DeviceState *S, wzfh_kernel_sums *out, int reset);
// Problems whose stream kernels use identical tables (same n, residue modulus, stream order,
// ENDPOS, side and pinned roots) can share one stream launch.
std::string stream_key(const wzfh_problem &P);
// A device state for stream launches only, set up from P's tables (no cells, no rows): the
// serving loop streams a group's cells on it while the problems' own states complete.
int stream_handle(const wzfh_problem &P, wzfh_problem **out);
// Batch size (cells) of one problem's search that fits its share of free device memory.
long long serve_cell_slots(wzfh_problem &P, int nshare, const wzfh_search_opts &o);
int problem_init(const wzfh_config &cfg, bool ab_side, bool with_device, wzfh_problem **out);
// Device side (wzfh_api.hip).
int device_init(wzfh_problem &P);
void device_free(wzfh_problem &P);
}  // namespace wzfh
