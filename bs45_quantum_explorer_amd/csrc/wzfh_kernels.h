// wzfh_kernels.h -- kernel argument blocks shared by the launchers and kernels.
#pragma once
#include <cstdint>

namespace wzfh {

// One streamed candidate: packed C,D (bit i set = element i is -1), the DFS node
// index of its leaf inside its work item, the work item, and the leaf ordinal
// inside the item (for exact leaf counts under a candidate cap).
struct CandRec {
  uint64_t X, Y;
  uint64_t ell;
  uint32_t item;
  uint32_t leaf;
};

// A candidate placed in its cell's global DFS order (after stream_finalize).
struct KeptCand {
  uint64_t X, Y;
  uint64_t G;       // global DFS node index of the leaf inside its cell
  uint64_t leaf;    // global leaf index inside its cell
  uint64_t cell;    // cell slot of the launch
};

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
  unsigned long long *work_ctr;
  int *cut;                    // [ncells] first path index known to be beyond the cut
  uint64_t *item_nodes;        // [ncells * npaths]
  uint32_t *item_leaves;       // [ncells * npaths]
  uint32_t *item_cands;        // [ncells * npaths]
  uint32_t *item_done;         // [ncells * npaths] set once the item's counts are final
  int *cell_lock;              // [ncells] completed-prefix tracker (try-lock)
  int *cell_pp;                // [ncells] items in the completed prefix
  unsigned long long *cell_pn, *cell_pc;   // [ncells] its node and candidate sums
  CandRec *cands;
  unsigned long long *ncands;
  unsigned long long cand_capacity;
  unsigned int *angle_rej;     // [200] first-rejecting-angle histogram (angle order tuning)
  // Leaf test on the matrix cores: A-operand fragments of the split angle table (see
  // hall_mfma in wzfh_stream.hip), the certain-reject / uncertain thresholds on the pair
  // energy, the number of 16-position K blocks covering L, and a verify switch that also
  // runs the FP64 test on every leaf and counts disagreements with the certain decisions.
  const void *hall_frag;
  float hall_hi, hall_lo;
  int hall_nkb;
  int hall_verify;
  unsigned long long *telem;   // optional per-launch counters (null = off): [0] cycles in the
                               // leaf test, [1] total wave cycles, [2] flushes, [3] leaves tested,
                               // [4] leaves rechecked in FP64, [5] disagreements (verify mode)
};

// A,B completer kernel arguments (fh_complete_ab, wz_match.cpp:1092-1341).
struct CompleteArgs {
  int n, L, m, absA, absB, outer_k, canon;
  int ncand;
  long long budget;
  const uint64_t *cd;          // [ncand][2]
  const int *row_off;          // [ncand] offset into rows, -1 = unconstrained
  const int *row_len;          // [ncand]
  const uint64_t *rows;        // rows packed: k[8], r[8] as int8 (2 words per row)
  int tic[8];                  // class counts of length L at modulus m
  const uint8_t *outer2, *outer3;
  const int8_t *pos_index;     // [16]
  uint64_t *scratch;           // per-wave row bitsets
  int scratch_words;           // words per level
  unsigned long long *work_ctr;
  int *verdict;
  long long *nodes;
  uint64_t *outA, *outB;
};

}  // namespace wzfh
