// wzfh_kernels.h -- kernel argument blocks shared by the launchers and kernels.
#pragma once
#include <cstdint>

namespace wzfh {

// One streamed candidate: packed C,D (bit i set = element i is -1), the DFS node
// index of its leaf inside its work item, the work item's record, and the leaf ordinal
// inside the item (for exact leaf counts under a candidate cap).
struct CandRec {
  uint64_t X, Y;
  uint64_t ell;
  uint32_t item;
  uint32_t leaf;
};

// A candidate placed in its cell's global DFS order (after the stream's walk).
struct KeptCand {
  uint64_t X, Y;
  uint64_t G;       // global DFS node index of the leaf inside its cell
  uint64_t leaf;    // global leaf index inside its cell
  uint64_t cell;    // cell slot of the launch
};

constexpr uint32_t kEndItem = 0xFFFFFFFFu;      // end of a cell's item list
constexpr uint64_t kUnsetOff = ~0ull;           // item offset not known (yet)

// C,D stream kernel arguments (count_pairs22 on the C,D side, wz_match.cpp:678-784).
struct StreamArgs {
  int L, half, m, d0, npaths, window, ncells, nch0;
  int root_pin_mask;           // DFS digit b at depth 0 not entered (pins)
  int odd, cm_mid;             // odd L: class of the middle position
  uint32_t depth_info[32];     // depth d: c1 | c2<<8 | R1<<16 | R2<<24 (R = rem at d+1)
  uint8_t quad8[8];            // digit b (depth >= 1) -> sign bits x1,y1,x2,y2 (bit3..bit0, 1 = -1)
  uint8_t quad16[16];          // digit b at depth 0 -> sign bits
  uint8_t tab_ne[256];         // (mx1,my1,mx2,my2) -> alive-children mask over digits
  uint8_t tab_eq[64];          // c1 == c2: (sx,sy) -> mask
  uint16_t tab0_ne[256];       // depth 0 (16 or 8 children), pins applied
  uint16_t tab0_eq[64];
  const uint64_t *cell_res;    // [ncells][2] residual words (tx+16, ty+16 packed 5 bits/class)
  const int *cell_rootok;      // [ncells]
  double lim;                  // 4n+2 + 0.5
  const double *hcos, *hsin;   // [201][64]
  const uint8_t *angle_order;  // [200] angle j-1 in evaluation order
  long long node_budget, cand_cap;
  unsigned long long *work_ctr;   // static items handed out
  // Work items. Records [0, ncells*npaths) are the static items (path p of cell slot cs
  // at cs*npaths + p); records from there up to rec_cap are items donated by running
  // lanes. Each cell's items form a list (it_next) in the reference's DFS order; record
  // states carry the launch epoch: (epoch << 2) | 1 published donation, | 2 done,
  // | 3 abandoned beyond the cell's cut.
  uint32_t rec_cap, epoch;
  uint64_t *it_nodes;             // nodes of the item's run (incl. its tail)
  uint64_t *it_noff, *it_loff, *it_coff;   // node / leaf / candidate offsets in the cell
  uint32_t *it_leaves, *it_cands, *it_state, *it_next;
  uint64_t *it_key;               // donated: digits of the levels above its base
  uint32_t *it_meta;              // donated: base level | lo << 8 | hi << 12 | hops << 16
  uint32_t *it_cell, *it_tail;    // donated: cell slot; dead entries after its subtrees
  unsigned long long *dq;         // [0] records allocated, [1] claimed, [2] items in hand,
                                  // [3] capped launches: cells still handing out static items,
                                  // [4] progress (items finished + donations), [5] error flag
                                  // (1 = no progress, 2 = invalid work record)
  unsigned long long stall_ticks; // watchdog: constant-clock ticks without progress
  uint32_t don_moves, don_low;    // donation period (DFS moves) and pending low-water
  uint32_t don_moves_front;       // donation period of a capped cell's frontier item
  uint32_t don_hops;              // capped: generations of frontier donations that donate again
  int don_any;                    // 1: every item may donate (also under a cap)
  int don_max_level;              // deepest level whose siblings are donated
  int don_keep;                   // capped: > 0 donates every level from the shallowest live one down
                                  // to depth d-1-don_keep (deep items first); 0 = the shallowest only
  int *cell_lock, *cell_crossed;  // walker try-lock; 1 once the cut is known
  // capped launches hand out a cell's static items through its own cursor, at most
  // static_win items ahead of the walk, so lanes work near each cell's frontier
  int *cell_open;                 // 1 while the cell still has static items to hand out
  uint32_t *cell_next, *cell_sp;  // next static item to hand out; static items walked past
  uint32_t static_win;
  int scan_max;                   // longest backoff (loop iterations) of idle polling
  int inner_moves;                // DFS moves a lane makes per pass of the kernel's main loop
  int idle_sleep;                 // s_sleep count of a wave with nothing to walk (0 = none)
  uint32_t *cell_wp;              // first item not yet walked
  unsigned long long *cell_pn, *cell_pl, *cell_pc;   // walked node / leaf / candidate sums
  CandRec *cands;
  unsigned long long *ncands;
  unsigned long long cand_capacity;
  unsigned int *angle_rej;     // [200] first-rejecting-angle histogram (angle order tuning)
  // Leaf test on the matrix cores: A-operand fragments of the split angle table (see
  // hall_mfma in wzfh_stream.hip), the certain-reject / uncertain thresholds on the pair
  // energy, the number of 32-position K blocks covering L, and a verify switch that also
  // runs the FP64 test on every leaf and counts disagreements with the certain decisions.
  const void *hall_frag;
  float hall_hi, hall_lo;
  int hall_nkb;
  int hall_verify;
  unsigned long long *telem;   // optional per-launch counters (null = off): [0] cycles in the
                               // leaf test, [1] total wave cycles, [2] flushes, [3] leaves tested,
                               // [4] leaves rechecked in FP64, [5] disagreements (verify mode),
                               // [6] donations, [7] items donated, [8] lane DFS moves,
                               // [9] wave passes of the main loop
};

// A,B completer kernel arguments (fh_complete_ab, wz_match.cpp:1092-1341).
struct CompleteArgs {
  int n, L, m, absA, absB, outer_k, canon;
  int ncand;                   // entries of the launch
  long long budget;
  // Work items. Entry e (a candidate, or a subtree of one below a quad path) has 8^split_ext
  // items: item u = e * 8^split_ext + x walks the subtree below path (ent_path[e] << 3 ext) | x
  // of depth ent_base[e] + split_ext (null arrays: path 0, depth 0), with the entry's budget
  // ent_budget[e] (null: budget). Per-item results go to verdict/nodes/outA/outB/charges at u.
  // Inputs cd/row_off/row_len are indexed by entry.
  int split_ext;
  int root_pair;                  // 1: the first digit only takes the root-canon quads 6 and 7 (x >> 3(ext-1) = 0, 1)
  int items_log2;                 // log2 of items per entry: 3 split_ext, or 3 split_ext - 2 with root_pair
  long long nwork;
  const long long *ent_path;
  const int *ent_base;
  const long long *ent_budget;
  long long item_cap;             // split: an item stops after this many charges (verdict 5; 0 = no cap)
  long long *charges;
  long long *item_charged;        // split: charges each item has reported so far (256 at a time)
  const uint64_t *cd;          // [ncand][2]
  const int *row_off;          // [ncand] offset into rows, -1 = unconstrained
  const int *row_len;          // [ncand]
  const uint64_t *rows;        // rows packed: k[8], r[8] as int8 (2 words per row)
  uint8_t rcap[32][8];         // [d][c]: class-c positions still unplaced after depth d's quad
  const uint8_t *outer2, *outer3;
  uint64_t pos_pack;           // pos_index[sb] + 1 in bits 4sb..4sb+3 (outer-lag pattern digit)
  uint64_t *scratch;           // per-wave row bitsets
  int scratch_words;           // words per level
  unsigned long long *work_ctr;
  int *verdict;
  long long *nodes;
  uint64_t *outA, *outB;
  long long *depth_out;        // unsplit launches: [item][64] nodes counted at each depth (null: off)
  // Resumed subtrees: items below ent_skip[e] were walked by the capped item whose subtree entry e
  // is (the host carries their charges and nodes), so they end at once with nothing counted (null:
  // none). snap_ext > 0: a capped item reports where it stopped for the next round's split of
  // snap_ext levels: outA = charges before the first node of that split's sub-path it was in,
  // plus the sub-path's index << 44; outB = the nodes counted before it.
  const int *ent_skip;
  int snap_ext;
  // Exact profile-row reachability (WZ_FH_A2=2, wz_match.cpp:977-1078): a2 != 0 also drops the
  // rows of the filter whose residuals the remaining positive quads of some mirror block cannot
  // reach. Block b covers residues a2_r[b], a2_s[b] (equal: self-reflected; a2_mid[b]: it holds
  // the free odd-L middle); a2_cnt[p][b] = quads of block b at depths p..half-1.
  int a2, a2_nblk;
  int8_t a2_r[8], a2_s[8], a2_mid[8];
  uint8_t a2_cnt[33][8];
};

}  // namespace wzfh
