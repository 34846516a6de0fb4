// wzfh_stream.hip -- C,D cell stream kernel for gfx950.
//
// Restates count_pairs22 on the C,D side (src/solver/wz_match.cpp:678-784) with
// the hall_ok_pair_fast leaf test (:143-168) as a persistent, work-stealing kernel:
//
//  * Work item = (cell, path p): the DFS of one cell is split at a fixed depth d0 into
//    nch0 * 8^(d0-1) ordered paths. One wavefront lane owns one item: it replays the
//    path's d0 quads (checking each node) and, when the depth-d0 node is alive,
//    walks its subtree with an iterative DFS. Items are grabbed window-major so that
//    a cell with a cap is explored front first.
//  * Node order and node counts equal the reference's: every node the reference
//    enters (rec() entry, G_CD_NODES) is counted in the item that owns it, and items
//    in path order concatenate to the reference's DFS order, so per-item counts plus
//    prefix sums give exact global node / candidate indices afterwards.
//  * The per-node capacity/parity prune only needs the two classes the quad touches:
//    parity is fixed per cell (checked at the root) and the other classes are
//    unchanged since the parent passed. The 8 children are tested at once by table
//    lookup on the two residuals of X and Y.
//  * Leaves go to a per-wave LDS queue; when 64 are queued the whole wave evaluates
//    them, one leaf per lane, with the reference's FP64 arithmetic (sequential sums,
//    no contraction), so divergent leaf tests never stall the DFS lanes.
#include <hip/hip_runtime.h>

#include "wzfh_kernels.h"

namespace wzfh {

namespace {

constexpr int kWaves = 4;       // waves per block
constexpr int kQueue = 128;     // leaf queue entries per wave

// lane states: no item, claimed a donated record not yet published, DFS moves, item
// finished (publish), item abandoned beyond its cell's cut
enum : int { ST_IDLE = 0, ST_WAIT = 1, ST_EXPAND = 2, ST_BACK = 3, ST_DONE = 4, ST_SKIP = 5 };

__device__ __forceinline__ uint32_t rec_tag(uint32_t epoch, uint32_t s) { return (epoch << 2) | s; }

// Broadcast of a 64-bit value held by `src` into scalar registers.
__device__ __forceinline__ unsigned long long bcast64(unsigned long long v, int src) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

// Item keys hold the DFS digits of the levels above an item's base: level 0 (radix 16)
// in bits 60..63, level j >= 1 (radix 8) in bits 60-3j..62-3j (j <= 20).
__device__ __forceinline__ int key_digit(uint64_t key, int j) {
  return j == 0 ? (int)(key >> 60) : (int)((key >> (60 - 3 * j)) & 7u);
}
__device__ __forceinline__ uint64_t key_put(int j, int v) {
  return j == 0 ? ((uint64_t)v << 60) : ((uint64_t)v << (60 - 3 * j));
}

struct Leaf {
  uint64_t X, Y, ell;
  uint32_t item, leaf;
};

__device__ __forceinline__ int field(uint64_t w, int c) { return (int)((w >> (5 * c)) & 31u) - 16; }

// bit0: element value +1 keeps |residual| <= R, bit1: value -1 does.
__device__ __forceinline__ unsigned ok_pm(int e, int R) {
  unsigned r = (e - 1 >= -R && e - 1 <= R) ? 1u : 0u;
  return r | ((e + 1 >= -R && e + 1 <= R) ? 2u : 0u);
}
// Same class at both ends of the quad: sums +2 (bit0), 0 (bit1), -2 (bit2).
__device__ __forceinline__ unsigned ok_sum(int e, int R) {
  unsigned r = (e - 2 >= -R && e - 2 <= R) ? 1u : 0u;
  r |= (e >= -R && e <= R) ? 2u : 0u;
  return r | ((e + 2 >= -R && e + 2 <= R) ? 4u : 0u);
}

template <typename T>
__device__ __forceinline__ unsigned child_mask(const T *tne, const T *teq, uint32_t di, uint64_t RX,
                                               uint64_t RY) {
  const int c1 = di & 255, c2 = (di >> 8) & 255, R1 = (di >> 16) & 255, R2 = di >> 24;
  const int e1x = field(RX, c1), e1y = field(RY, c1);
  if (c1 != c2) {
    const int e2x = field(RX, c2), e2y = field(RY, c2);
    unsigned idx = ok_pm(e1x, R1) | (ok_pm(e1y, R1) << 2) | (ok_pm(e2x, R2) << 4) | (ok_pm(e2y, R2) << 6);
    return tne[idx];
  }
  return teq[ok_sum(e1x, R1) | (ok_sum(e1y, R1) << 3)];
}

// Places quad sign bits sb (x1,y1,x2,y2 as bit3..bit0, set = -1) at depth d.
__device__ __forceinline__ void apply_quad(uint32_t di, int d, int L, unsigned sb, int dir, uint64_t &RX,
                                           uint64_t &RY, uint64_t &X, uint64_t &Y) {
  const int c1 = di & 255, c2 = (di >> 8) & 255;
  // residual -= value  <=>  field += (value < 0 ? +1 : -1); dir = +1 place, -1 undo
  auto delta = [&](unsigned neg, int c) -> uint64_t {
    long long v = neg ? 1 : -1;
    return (uint64_t)(v * dir) << (5 * c);
  };
  RX += delta((sb >> 3) & 1, c1) + delta((sb >> 1) & 1, c2);
  RY += delta((sb >> 2) & 1, c1) + delta(sb & 1, c2);
  const uint64_t b1 = 1ull << d, b2 = 1ull << (L - 1 - d);
  if (dir > 0) {
    X |= (((sb >> 3) & 1) ? b1 : 0) | (((sb >> 1) & 1) ? b2 : 0);
    Y |= (((sb >> 2) & 1) ? b1 : 0) | ((sb & 1) ? b2 : 0);
  } else {
    X &= ~(b1 | b2);
    Y &= ~(b1 | b2);
  }
}

// Residual part of apply_quad only (used when undoing a path to find untried siblings).
__device__ __forceinline__ void quad_res(uint32_t di, unsigned sb, int dir, uint64_t &RX, uint64_t &RY) {
  const int c1 = di & 255, c2 = (di >> 8) & 255;
  auto delta = [&](unsigned neg, int c) -> uint64_t {
    long long v = neg ? 1 : -1;
    return (uint64_t)(v * dir) << (5 * c);
  };
  RX += delta((sb >> 3) & 1, c1) + delta((sb >> 1) & 1, c2);
  RY += delta((sb >> 2) & 1, c1) + delta(sb & 1, c2);
}

// FP64 hall_ok decision of one leaf, evaluated exactly like the reference: for
// every angle the sums run i = 0..L-1 left to right, products and sums are rounded
// separately, and the pair energy is ((rx^2 + ix^2) + ry^2) + iy^2. A leaf is
// rejected at the first angle where X, Y or the pair exceeds 4n+2+0.5; the
// conjunction does not depend on the angle order. Loop bounds are wave-uniform so
// the table reads go through the scalar cache.
__device__ bool hall_pass(uint64_t X, uint64_t Y, bool active, int L, double lim, const double *hcos,
                          const double *hsin, const uint8_t *order, unsigned *s_rej) {
  bool undecided = active, pass = active;
  for (int t = 0; t < 200; t++) {
    if (!__any(undecided)) break;
    const int j = order[t] + 1;
    const double *cj = hcos + j * 64, *sj = hsin + j * 64;
    if (undecided) {
      double rx = 0.0, ix = 0.0;
      for (int i = 0; i < L; i++) {
        const bool ng = (X >> i) & 1;
        const double c = cj[i], s = sj[i];
        rx = __dadd_rn(rx, ng ? -c : c);
        ix = __dadd_rn(ix, ng ? -s : s);
      }
      const double px2 = __dadd_rn(__dmul_rn(rx, rx), __dmul_rn(ix, ix));
      bool fail = px2 > lim;
      if (!fail) {
        double ry = 0.0, iy = 0.0;
        for (int i = 0; i < L; i++) {
          const bool ng = (Y >> i) & 1;
          const double c = cj[i], s = sj[i];
          ry = __dadd_rn(ry, ng ? -c : c);
          iy = __dadd_rn(iy, ng ? -s : s);
        }
        const double ry2 = __dmul_rn(ry, ry), iy2 = __dmul_rn(iy, iy);
        const double py2 = __dadd_rn(ry2, iy2);
        const double pp = __dadd_rn(__dadd_rn(px2, ry2), iy2);
        fail = (py2 > lim) || (pp > lim);
      }
      if (fail) {
        undecided = false;
        pass = false;
        atomicAdd(&s_rej[t], 1u);
      }
    }
  }
  return pass;
}

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

// 8 consecutive positions base..base+7 of a packed sequence as fp16 +-1 (bit set = -1),
// 0 at positions >= L.
__device__ __forceinline__ half8_t seq_frag(uint64_t bits, int base, int L) {
  const unsigned byte = (unsigned)(bits >> base) & 0xFFu;
  const int nv = L - base;
  const unsigned vm = nv >= 8 ? 0xFFu : (nv <= 0 ? 0u : ((1u << nv) - 1u));
  u32x4_t w;
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const unsigned e0 = 2 * h, e1 = 2 * h + 1;
    const unsigned h0 = ((vm >> e0) & 1u) ? (0x3C00u | (((byte >> e0) & 1u) << 15)) : 0u;
    const unsigned h1 = ((vm >> e1) & 1u) ? (0x3C00u | (((byte >> e1) & 1u) << 15)) : 0u;
    w[h] = h0 | (h1 << 16);
  }
  return __builtin_bit_cast(half8_t, w);
}

constexpr int kHallGroups = 7;   // 16-angle row groups covering the 101 distinct angles

// Leaf test of the 64 queued leaves on the matrix cores. The reference rejects a leaf iff
// some angle j = 1..200 has P_X + P_Y > lim in FP64 (P_X > lim alone implies it), and
// angles j and 200-j carry the same energy, so 101 angles (j = 1..100 and 200) decide.
// D[angle][sequence] = T[angle][pos] * S[pos][sequence] runs as v_mfma_f32_16x16x32_f16
// with T = hi + lo, a two-term fp16 split of the FP64 table (|T - hi - lo| <= 2^-24), and
// the sequences as exact fp16 +-1. Rows are angles (7 groups of 16, cos and sin tiles),
// columns are the X and Y sequences of 16 leaves, so one lane holds the four sums of one
// (leaf, angle) and forms S = xc^2 + yc^2 + xs^2 + ys^2 in FP32. The host bounds
// |S - S_fp64| by a margin eps (table split, FP32 accumulation and evaluation, and the
// reference's own FP64 rounding) and passes hall_hi = lim + eps, hall_lo = lim - eps:
// S > hall_hi at any angle is a certain reject, S <= hall_lo at every angle a certain
// pass, anything else is uncertain and is decided by the FP64 test. Returns, for leaf
// `lane`: bit0 = certain reject, bit1 = some angle above hall_lo. Must be called with the
// whole wave active.
__device__ unsigned hall_mfma(const Leaf *q, int L, int nkb, const u32x4_t *frag, float thr_hi, float thr_lo,
                              int lane) {
  unsigned mine = 0;
#pragma unroll 1
  for (int lg = 0; lg < 4; lg++) {
    const Leaf &f = q[lg * 16 + (lane & 15)];
    const uint64_t xb = f.X, yb = f.Y;
    const int base = 8 * (lane >> 4);
    const half8_t bx0 = seq_frag(xb, base, L), by0 = seq_frag(yb, base, L);
    const half8_t bx1 = seq_frag(xb, 32 + base, L), by1 = seq_frag(yb, 32 + base, L);
    unsigned flags = 0;
#pragma unroll 1
    for (int g = 0; g < kHallGroups; g++) {
      floatx4_t xc = {}, yc = {}, xs = {}, ys = {};
      const u32x4_t *fg = frag + (size_t)g * 2 * 2 * 2 * 64 + lane;   // [t][kb][part][lane]
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const half8_t ac = __builtin_bit_cast(half8_t, fg[(0 * 4 + 0 * 2 + p) * 64]);
        const half8_t as = __builtin_bit_cast(half8_t, fg[(1 * 4 + 0 * 2 + p) * 64]);
        xc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac, bx0, xc, 0, 0, 0);
        yc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac, by0, yc, 0, 0, 0);
        xs = __builtin_amdgcn_mfma_f32_16x16x32_f16(as, bx0, xs, 0, 0, 0);
        ys = __builtin_amdgcn_mfma_f32_16x16x32_f16(as, by0, ys, 0, 0, 0);
      }
      if (nkb > 1) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
          const half8_t ac = __builtin_bit_cast(half8_t, fg[(0 * 4 + 1 * 2 + p) * 64]);
          const half8_t as = __builtin_bit_cast(half8_t, fg[(1 * 4 + 1 * 2 + p) * 64]);
          xc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac, bx1, xc, 0, 0, 0);
          yc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ac, by1, yc, 0, 0, 0);
          xs = __builtin_amdgcn_mfma_f32_16x16x32_f16(as, bx1, xs, 0, 0, 0);
          ys = __builtin_amdgcn_mfma_f32_16x16x32_f16(as, by1, ys, 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float S = ((xc[r] * xc[r] + yc[r] * yc[r]) + xs[r] * xs[r]) + ys[r] * ys[r];
        flags |= (S > thr_hi ? 1u : 0u) | (S > thr_lo ? 2u : 0u);
      }
    }
    // lanes l, l^16, l^32, l^48 hold the other angle rows of the same leaf
    flags |= (unsigned)__shfl_xor((int)flags, 16);
    flags |= (unsigned)__shfl_xor((int)flags, 32);
    if ((lane >> 4) == lg) mine = flags;
  }
  return mine;
}

// Walker of one cell's item list. Items finish in any order; the lane that takes the
// cell's lock follows the list from the first unwalked item through finished items,
// giving each its node / leaf / candidate offsets, and marks the cell cut as soon as the
// sums reach the node budget or the candidate cap: every item after that point is beyond
// the reference's stop and is skipped or abandoned. The first unfinished item is told
// where it starts, so it can stop itself exactly at the budget or cap. The lock is only
// tried (never waited on); a lane that sees the next item finished after releasing it
// tries again, so the walk never stalls behind a finished item.
__device__ void advance_prefix(const StreamArgs &a, int cs) {
  const uint32_t done = rec_tag(a.epoch, 2);
  const uint32_t nstatic = (uint32_t)a.ncells * (uint32_t)a.npaths;
  while (true) {   // every pass walks at least one item, so this ends
    if (atomicCAS(&a.cell_lock[cs], 0, 1) != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    uint32_t r = a.cell_wp[cs];
    unsigned long long pn = a.cell_pn[cs], pl = a.cell_pl[cs], pc = a.cell_pc[cs];
    bool crossed = false;
    while (r != kEndItem) {
      if (__hip_atomic_load(&a.it_state[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != done) break;
      a.it_loff[r] = pl;
      a.it_coff[r] = pc;
      __hip_atomic_store(&a.it_noff[r], (uint64_t)pn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pn += a.it_nodes[r];
      pl += a.it_leaves[r];
      pc += __hip_atomic_load(&a.it_cands[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (r < nstatic)
        __hip_atomic_store(&a.cell_sp[cs], r - (uint32_t)cs * (uint32_t)a.npaths + 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      r = a.it_next[r];
      // a node budget B is crossed once node index B exists (the cell then has more than B
      // nodes, so it is capped); a candidate cap once the cap-th candidate is found
      if ((a.node_budget > 0 && pn > (unsigned long long)a.node_budget) ||
          (a.cand_cap > 0 && pc >= (unsigned long long)a.cand_cap)) { crossed = true; break; }
    }
    a.cell_wp[cs] = crossed ? kEndItem : r;
    a.cell_pn[cs] = pn;
    a.cell_pl[cs] = pl;
    a.cell_pc[cs] = pc;
    if (crossed) {
      __hip_atomic_store(&a.cell_crossed[cs], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (atomicExch(&a.cell_open[cs], 0) != 0) atomicAdd(&a.dq[3], ~0ull);   // one fewer open cell
    } else if (r != kEndItem) {   // the first unfinished item learns its offsets
      a.it_coff[r] = pc;
      __hip_atomic_store(&a.it_noff[r], (uint64_t)pn, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicExch(&a.cell_lock[cs], 0);
    if (crossed || r == kEndItem) return;
    // Handoff with a lane that finished item r while the lock was held: it stores r's state,
    // fences, then tries the lock; this side releases the lock, fences, then reads r's state.
    // With a sequentially consistent fence on both sides at least one of them sees the
    // other's write, so a finished item is never left unwalked.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    if (__hip_atomic_load(&a.it_state[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != done) return;
  }
}

}  // namespace

// The stream kernel. Persistent waves; every lane owns at most one work item at a time.
//
// An item is a contiguous run of one cell's DFS (the reference's node order):
//  * static item p of a cell: the path of digits k_0..k_{d0-1}; its run is the subtree
//    below that path, preceded by the path's nodes that no earlier path enters (nodes
//    whose remaining digits are all zero belong to the zero path), so static items in
//    path order concatenate to the DFS;
//  * donated item (base level l, digits lo..hi): the subtrees of digits lo..hi at level l
//    below a fixed prefix, followed by `tail` dead entries (see below).
// A lane that has run an item for a while, when idle lanes have nothing to take, donates
// the untried siblings at the shallowest level l of its current path that still has live
// untried siblings: one item per live sibling (dead siblings join the item of the live
// one before them; the first item starts right after the lane's digit), inserted into the
// cell's list right after the donor. Shallower levels then have only dead untried siblings;
// their entries come after the donated subtrees in DFS order, so they move to the last
// donated item's tail, and the donor's run now ends with the subtree it is in. Item lists
// therefore stay in DFS order, each node is counted by exactly one item, and a candidate's
// global node index is its item's offset (from the walk) plus its index in the item.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void stream_kernel(StreamArgs a) {
  __shared__ uint8_t s_ne[256], s_eq[64];
  __shared__ uint16_t s0_ne[256], s0_eq[64];
  __shared__ uint32_t s_di[32];
  __shared__ uint8_t s_q8[8], s_q16[16], s_order[200];
  __shared__ unsigned s_rej[200];
  __shared__ Leaf s_q[kWaves][kQueue];
  __shared__ unsigned long long s_tel[10];   // block telemetry counters (see StreamArgs::telem)
  __shared__ unsigned long long s_wd[kWaves][2];   // watchdog: last progress seen, and when

  for (int i = threadIdx.x; i < 256; i += blockDim.x) { s_ne[i] = a.tab_ne[i]; s0_ne[i] = a.tab0_ne[i]; }
  if (threadIdx.x < 10) s_tel[threadIdx.x] = 0;
  if (threadIdx.x < kWaves) { s_wd[threadIdx.x][0] = ~0ull; s_wd[threadIdx.x][1] = 0; }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) { s_eq[i] = a.tab_eq[i]; s0_eq[i] = a.tab0_eq[i]; }
  for (int i = threadIdx.x; i < 200; i += blockDim.x) { s_order[i] = a.angle_order[i]; s_rej[i] = 0; }
  if (threadIdx.x < 32) s_di[threadIdx.x] = a.depth_info[threadIdx.x];
  if (threadIdx.x < 8) s_q8[threadIdx.x] = a.quad8[threadIdx.x];
  if (threadIdx.x < 16) s_q16[threadIdx.x] = a.quad16[threadIdx.x];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  Leaf *q = s_q[threadIdx.x >> 6];
  const int L = a.L, half = a.half, d0 = a.d0;
  const unsigned long long rounds = (unsigned long long)((a.npaths + a.window - 1) / a.window);
  const unsigned long long total = rounds * (unsigned long long)a.ncells * (unsigned long long)a.window;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t ep = a.epoch;
  const bool capped = a.node_budget > 0 || a.cand_cap > 0;

  int st = ST_IDLE;
  int cs = 0;
  uint32_t rec = 0, nleaf = 0, tail = 0, dmask = 0, since = 0, hops = 0;
  bool front = false;   // capped launches: the walk has reached this item (its offsets are known)
  // wave-uniform: demand for donations (static items gone, few donations pending), refreshed
  // every 64 iterations; claim polling with exponential backoff while nothing is pending
  bool hungry = false;
  int poll_cd = 0, backoff = 1, scan_cd = 0, scan_backoff = 1;
  unsigned pend = 0;
  unsigned iter = 0;
  uint32_t rr = (uint32_t)((blockIdx.x * kWaves + (threadIdx.x >> 6)) * 61u);   // capped: next cell to probe
  const unsigned long long lanes_total = (unsigned long long)gridDim.x * blockDim.x;
  int d = 0, base = 0, lo_b = 0, hi_b = 7;
  uint64_t dig = 0, RX = 0, RY = 0, X = 0, Y = 0, ell = 0;
  bool more = true;   // static items may remain
  int qn = 0;
  const bool telem = a.telem != nullptr;
  const unsigned long long t_start = telem ? clock64() : 0;

  while (true) {
    bool has_leaf = false;
    Leaf lf;
    if ((++iter & 127u) == 1u) {   // demand: idle lanes outnumber waiting donations
      int pd = 0, h = 0, op = 1, quit = 0;
      if (lane == 0) {
        // watchdog: no item finished and none donated anywhere for a_stall_ticks of the
        // constant clock means the launch cannot end; flag it and leave (the host reports
        // an internal error instead of waiting forever)
        const unsigned long long prog = __hip_atomic_load(&a.dq[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long now = __builtin_amdgcn_s_memrealtime();
        unsigned long long *wd = s_wd[threadIdx.x >> 6];
        if (prog != wd[0]) { wd[0] = prog; wd[1] = now; }
        else if (now - wd[1] > a.stall_ticks) atomicExch(&a.dq[5], 1ull);
        quit = __hip_atomic_load(&a.dq[5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 1 : 0;
        const unsigned long long al = __hip_atomic_load(&a.dq[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long hd = __hip_atomic_load(&a.dq[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long held = __hip_atomic_load(&a.dq[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long pq = al > hd ? al - hd : 0;
        pd = (int)min(pq, 0x7FFFFFFFull);
        h = (held + pq + (lanes_total >> 3) < lanes_total && pq < (unsigned long long)a.don_low) ? 1 : 0;
        if (capped) op = __hip_atomic_load(&a.dq[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 1 : 0;
      }
      pend = (unsigned)__builtin_amdgcn_readfirstlane(pd);
      hungry = __builtin_amdgcn_readfirstlane(h) != 0;
      if (capped && __builtin_amdgcn_readfirstlane(op) == 0) more = false;
      if (__builtin_amdgcn_readfirstlane(quit)) break;
    }
    // ---- 1a. once static items are gone, idle lanes claim published donations (rate-limited:
    // the queue counters are shared by every wave)
    bool tried_empty = false;
    {
      const unsigned long long need = __ballot(st == ST_IDLE);
      if (need && (pend > 0 || need == ~0ull || (!capped && !more))) {
        if (poll_cd > 0) {
          poll_cd--;
        } else {
          const int leader = __ffsll((long long)need) - 1;
          unsigned long long got = 0, h0 = 0;
          if (lane == leader) {
            const unsigned long long h = __hip_atomic_load(&a.dq[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long al = __hip_atomic_load(&a.dq[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (h < al) {
              const unsigned long long want = min((unsigned long long)__popcll(need), al - h);
              if (atomicCAS(&a.dq[1], h, h + want) == h) {
                got = want;
                h0 = h;
                atomicAdd(&a.dq[2], got);
              }
            }
          }
          got = bcast64(got, leader);
          h0 = bcast64(h0, leader);
          if (got) {
            backoff = 1;
          } else {
            backoff = backoff < a.scan_max ? 2 * backoff : a.scan_max;
            poll_cd = backoff;
            tried_empty = true;
          }
          if (st == ST_IDLE) {
            const unsigned long long rank = __popcll(need & lt_mask);
            if (rank < got) {
              rec = (uint32_t)(h0 + rank);
              st = ST_WAIT;
            }
          }
        }
      }
    }
    // ---- 1b. remaining idle lanes take static items: window-major over cells when nothing
    // caps the launch; under a cap each cell hands out its next item only while it is at
    // most static_win items ahead of the cell's walk (probing is rate-limited)
    if (more) {
      const unsigned long long need = __ballot(st == ST_IDLE);
      if (need) {
        bool started = false, take = false;
        int p = 0;
        if (!capped) {
          const int cnt = __popcll(need);
          const int leader = __ffsll((long long)need) - 1;
          unsigned long long gbase = 0;
          if (lane == leader) gbase = atomicAdd(a.work_ctr, (unsigned long long)cnt);
          gbase = bcast64(gbase, leader);
          if (gbase + cnt >= total) more = false;
          if (st == ST_IDLE) {
            const unsigned long long g = gbase + __popcll(need & lt_mask);
            if (g < total) {
              const unsigned long long w = g % a.window, t = g / a.window;
              cs = (int)(t % a.ncells);
              p = (int)((t / a.ncells) * a.window + w);
              take = p < a.npaths;
            }
          }
        } else if (scan_cd > 0) {
          scan_cd--;
        } else {
          // the wave's first idle lane claims up to (idle lanes) consecutive items of one
          // cell with a single CAS; the cell cursors are shared by every wave, so a wave
          // that finds nothing backs off up to scan_max iterations
          const int leader = __ffsll((long long)need) - 1;
          const int c = (int)(rr % (uint32_t)a.ncells);
          unsigned got = 0, base_p = 0;
          if (lane == leader && __hip_atomic_load(&a.cell_open[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            const uint32_t nx = __hip_atomic_load(&a.cell_next[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t sp = __hip_atomic_load(&a.cell_sp[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t lim = min((uint32_t)a.npaths, sp + a.static_win);
            if (nx < lim) {
              const uint32_t k = min(lim - nx, (uint32_t)__popcll(need));
              if (atomicCAS(&a.cell_next[c], nx, nx + k) == nx) {
                got = k;
                base_p = nx;
                if (nx + k == (uint32_t)a.npaths && atomicExch(&a.cell_open[c], 0) != 0)
                  atomicAdd(&a.dq[3], ~0ull);   // the cell's last static item is out
              }
            } else if (nx < (uint32_t)a.npaths && scan_backoff >= 16) {
              advance_prefix(a, c);   // window full for a while: make sure the walk is not behind
            }
          }
          got = (unsigned)__builtin_amdgcn_readlane((int)got, leader);
          base_p = (unsigned)__builtin_amdgcn_readlane((int)base_p, leader);
          rr += 1u;
          if (st == ST_IDLE) {
            const unsigned rank = (unsigned)__popcll(need & lt_mask);
            if (rank < got) {
              take = true;
              cs = c;
              p = (int)(base_p + rank);
            }
          }
          if (got) {
            scan_backoff = 1;
          } else {
            scan_backoff = scan_backoff < a.scan_max ? 2 * scan_backoff : a.scan_max;
            scan_cd = scan_backoff;
          }
        }
        if (st == ST_IDLE && take) {
          if (!__hip_atomic_load(&a.cell_crossed[cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            {
              // ---- replay path p: digits k_0 (radix nch0) then k_1..k_{d0-1} (radix 8)
              started = true;
              rec = (uint32_t)cs * (uint32_t)a.npaths + (uint32_t)p;
              RX = a.cell_res[2 * cs];
              RY = a.cell_res[2 * cs + 1];
              X = Y = 0;
              nleaf = 0;
              tail = 0;
              dmask = 0;
              since = 0;
              hops = 0;
              front = p == 0;   // a cell's first item starts at offset 0
              base = d0;
              lo_b = 0;
              hi_b = 7;
              int z = 0;   // trailing zero digits
              {
                int pp = p;
                bool nz = false;
                for (int j = d0 - 1; j >= 1; j--) { if (!nz && (pp & 7) == 0) z++; else nz = true; pp >>= 3; }
                if (!nz && pp == 0) z++;
              }
              uint64_t owned = (p == 0) ? 1 : 0;   // the root node belongs to path 0
              bool alive = a.cell_rootok[cs] != 0;
              for (int j = 0; j < d0 && alive; j++) {
                const int kj = (j == 0) ? (p >> (3 * (d0 - 1))) : ((p >> (3 * (d0 - 1 - j))) & 7);
                unsigned mask;
                if (j == 0) {
                  if ((a.root_pin_mask >> kj) & 1) { alive = false; break; }   // pinned: not entered
                  mask = child_mask(s0_ne, s0_eq, s_di[0], RX, RY);
                } else {
                  mask = child_mask(s_ne, s_eq, s_di[j], RX, RY);
                }
                if (j + 1 >= d0 - z) owned++;
                if (!((mask >> kj) & 1)) { alive = false; break; }
                apply_quad(s_di[j], j, L, (j == 0) ? s_q16[kj] : s_q8[kj], 1, RX, RY, X, Y);
              }
              ell = owned;
              d = d0;
              dig = 0;
              if (!alive) st = ST_DONE;
              else if (d0 == half) {
                has_leaf = true;
                st = ST_DONE;
              } else st = ST_EXPAND;
            }
          }
        }
        const unsigned long long sm = __ballot(started);
        if (sm && lane == __ffsll((long long)sm) - 1) atomicAdd(&a.dq[2], (unsigned long long)__popcll(sm));
      }
    }
    // ---- 1c. a claimed donation starts once its record is published
    if (st == ST_WAIT &&
        __hip_atomic_load(&a.it_state[rec], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == rec_tag(ep, 1)) {
      cs = (int)a.it_cell[rec];
      const uint32_t meta = a.it_meta[rec];
      base = (int)(meta & 255u);
      lo_b = (int)((meta >> 8) & 15u);
      hi_b = (int)((meta >> 12) & 15u);
      if ((unsigned)cs >= (unsigned)a.ncells || base < 1 || base >= half || lo_b > hi_b || hi_b > 7) {
        // a record that cannot have been written by a donor: report, touch nothing
        atomicExch(&a.dq[5], 2ull);
        cs = 0;
        st = ST_SKIP;
      } else if (__hip_atomic_load(&a.cell_crossed[cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        st = ST_SKIP;
      } else {
        const uint64_t key = a.it_key[rec];
        hops = (meta >> 16) & 15u;
        front = false;
        tail = a.it_tail[rec];
        RX = a.cell_res[2 * cs];
        RY = a.cell_res[2 * cs + 1];
        X = Y = 0;
        for (int j = 0; j < base; j++) {   // the prefix is live: the donor stood on it
          const int kj = key_digit(key, j);
          apply_quad(s_di[j], j, L, (j == 0) ? s_q16[kj] : s_q8[kj], 1, RX, RY, X, Y);
        }
        d = base;
        dig = 0;
        ell = 0;
        nleaf = 0;
        dmask = 0;
        since = 0;
        st = ST_EXPAND;
      }
    }

    // ---- 2. DFS moves: a working lane makes up to a.inner_moves moves per pass of this
    // loop, stopping at a leaf (queued below), at the end of its item or past its node
    // budget. At the item's base level only digits lo_b..hi_b are its own; a level
    // truncated by a donation ends at the lane's digit.
    uint32_t nmoves = 0;
    {
      bool go = (st == ST_EXPAND || st == ST_BACK) && !has_leaf;
      for (int mv = 0; mv < a.inner_moves; mv++) {
        if (go) {
          if (st == ST_EXPAND) {
            const int lo_e = (d == base) ? lo_b : 0, hi_e = (d == base) ? hi_b : 7;
            const unsigned rmask = (0xFFu >> (7 - hi_e)) & (0xFFu << lo_e);
            const unsigned mask = child_mask(s_ne, s_eq, s_di[d], RX, RY) & rmask;
            if (mask == 0) {
              ell += (uint64_t)(hi_e - lo_e + 1);
              st = ST_BACK;
            } else {
              const int b = __ffs(mask) - 1;
              ell += (uint64_t)(b - lo_e + 1);
              apply_quad(s_di[d], d, L, s_q8[b], 1, RX, RY, X, Y);
              dig |= (uint64_t)b << (3 * (d - base));
              d++;
              if (d == half) { has_leaf = true; st = ST_BACK; }
            }
          } else {
            while (true) {
              if (d == base) { st = ST_DONE; break; }
              d--;
              const int sh = 3 * (d - base);
              const int b = (int)((dig >> sh) & 7);
              dig &= ~(7ull << sh);
              apply_quad(s_di[d], d, L, s_q8[b], -1, RX, RY, X, Y);
              const int hi_e = ((dmask >> d) & 1u) ? b : ((d == base) ? hi_b : 7);
              const unsigned mask = child_mask(s_ne, s_eq, s_di[d], RX, RY) & (0xFEu << b) & (0xFFu >> (7 - hi_e));
              if (mask) {
                const int b2 = __ffs(mask) - 1;
                ell += (uint64_t)(b2 - b);
                apply_quad(s_di[d], d, L, s_q8[b2], 1, RX, RY, X, Y);
                dig |= (uint64_t)b2 << sh;
                d++;
                if (d == half) has_leaf = true;
                else st = ST_EXPAND;
                break;
              }
              ell += (uint64_t)(hi_e - b);
            }
          }
          nmoves++;
          go = !has_leaf && (st == ST_EXPAND || st == ST_BACK) &&
               !(a.node_budget > 0 && ell > (uint64_t)a.node_budget);
        }
        if (!__any(go)) break;
      }
      if (telem) {   // lane moves and passes of this wave
        unsigned long long wm = nmoves;
        for (int o = 32; o > 0; o >>= 1) wm += __shfl_xor(wm, o);
        if (lane == 0) { atomicAdd(&s_tel[8], wm); atomicAdd(&s_tel[9], 1ull); }
      }
    }
    if (has_leaf) {
      lf.X = X;
      lf.Y = Y;
      if (a.odd) {   // the free middle element takes the remaining class residual
        const uint64_t mb = 1ull << half;
        if (field(RX, a.cm_mid) < 0) lf.X |= mb;
        if (field(RY, a.cm_mid) < 0) lf.Y |= mb;
      }
      lf.ell = ell - 1;
      lf.item = rec;
      lf.leaf = nleaf++;
    }
    // ---- local caps, the cell's cut, and the item's own offsets once the walk reaches it
    if (st == ST_EXPAND || st == ST_BACK) {
      bool stop = a.node_budget > 0 && ell > (uint64_t)a.node_budget;
      if (a.cand_cap > 0 && has_leaf &&
          __hip_atomic_load(&a.it_cands[rec], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)a.cand_cap)
        stop = true;
      const uint32_t since0 = since;
      since += nmoves;
      if (capped && ((since0 >> 7) != (since >> 7) || (a.cand_cap > 0 && has_leaf))) {
        if (__hip_atomic_load(&a.cell_crossed[cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          st = ST_SKIP;
        } else {
          const uint64_t no = __hip_atomic_load(&a.it_noff[rec], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          front = no != kUnsetOff;
          if (front) {
            if (a.node_budget > 0 && no + ell > (uint64_t)a.node_budget) stop = true;
            if (a.cand_cap > 0 &&
                a.it_coff[rec] + __hip_atomic_load(&a.it_cands[rec], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                    (uint64_t)a.cand_cap)
              stop = true;
          }
        }
      }
      if (stop && st != ST_SKIP) st = ST_DONE;
    }
    // ---- donation: when static items are gone and few donations wait, a lane that has run
    // its item for a while gives away its untried siblings at the shallowest possible level.
    // Under a cap only the cell's frontier region splits: the item the walk has reached, and
    // a few generations of items donated from it (hops); elsewhere work would be speculative.
    if ((st == ST_EXPAND || st == ST_BACK) && hungry &&
        since >= ((capped && (front || hops > 0)) ? a.don_moves_front : a.don_moves) &&
        (!capped || front || hops > 0 || a.don_any)) {
      since = 0;
      {
        // remaining live siblings per level, from the residuals found by undoing the path
        // (8 bits per level, levels 0..31 in four words)
        uint64_t rx = RX, ry = RY;
        int best = -1;
        uint64_t lm0 = 0, lm1 = 0, lm2 = 0, lm3 = 0;
        for (int lev = d - 1; lev >= base; lev--) {
          const int b = (int)((dig >> (3 * (lev - base))) & 7);
          quad_res(s_di[lev], s_q8[b], -1, rx, ry);
          const int hi_e = ((dmask >> lev) & 1u) ? b : ((lev == base) ? hi_b : 7);
          const unsigned m = child_mask(s_ne, s_eq, s_di[lev], rx, ry) & (0xFEu << b) & (0xFFu >> (7 - hi_e));
          if (m) best = lev;
          const uint64_t put = (uint64_t)m << (8 * (lev & 7));
          if (lev < 8) lm0 |= put; else if (lev < 16) lm1 |= put; else if (lev < 24) lm2 |= put; else lm3 |= put;
        }
        auto lmask = [&](int lev) -> unsigned {
          const uint64_t w = lev < 8 ? lm0 : lev < 16 ? lm1 : lev < 24 ? lm2 : lm3;
          return (unsigned)((w >> (8 * (lev & 7))) & 255u);
        };
        // Under a cap the subtrees at the shallowest live level lie mostly beyond the cell's cut;
        // with don_keep > 0 every level from there down to the deepest live level above
        // d-1-don_keep is given away, deepest first, so the first new items are the work that
        // comes next in DFS order. The donor keeps its subtree and the levels below `top`.
        int top = best;
        if (best >= 0 && capped && a.don_keep > 0)
          for (int lev = min(d - 1 - a.don_keep, a.don_max_level); lev > best; lev--)
            if (lmask(lev)) { top = lev; break; }
        if (best >= 0 && best <= a.don_max_level) {
          int k = 0;
          for (int lev = top; lev >= best; lev--) k += __popc(lmask(lev));
          const unsigned long long al = __hip_atomic_load(&a.dq[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (al + (unsigned long long)k <= (unsigned long long)a.rec_cap &&
              atomicCAS(&a.dq[0], al, al + (unsigned long long)k) == al) {
            // prefix of the new items: the item's own prefix, then its digits down to best
            uint64_t key = 0;
            const uint32_t nstatic = (uint32_t)a.ncells * (uint32_t)a.npaths;
            if (rec < nstatic) {
              const int p = (int)(rec % (uint32_t)a.npaths);
              for (int j = 0; j < d0; j++)
                key |= key_put(j, (j == 0) ? (p >> (3 * (d0 - 1))) : ((p >> (3 * (d0 - 1 - j))) & 7));
            } else {
              key = a.it_key[rec];
            }
            uint32_t t_new = tail;   // dead entries after the last new item: levels above best
            for (int lev = base; lev < best; lev++) {
              const int b = (int)((dig >> (3 * (lev - base))) & 7);
              const int hi_e = ((dmask >> lev) & 1u) ? b : ((lev == base) ? hi_b : 7);
              t_new += (uint32_t)(hi_e - b);
            }
            const uint32_t after = a.it_next[rec];
            const uint32_t child_hops = front ? (uint32_t)a.don_hops : (hops > 0 ? hops - 1 : 0u);
            // items in DFS order: level top's live siblings, ..., level best's. Each covers one
            // live sibling and the dead ones up to the next live one at its level (the last one
            // of a level to the level's end); the dead entries of a level with no live sibling
            // between top and best become the tail of the item before them.
            uint32_t r = (uint32_t)al, last = (uint32_t)al, pend = 0;
            bool any = false;
            for (int lev = top; lev >= best; lev--) {
              const int bl = (int)((dig >> (3 * (lev - base))) & 7);
              const int hi_l = ((dmask >> lev) & 1u) ? bl : ((lev == base) ? hi_b : 7);
              unsigned m = lmask(lev);
              if (!m) { pend += (uint32_t)(hi_l - bl); continue; }
              if (any) a.it_tail[last] = pend;
              pend = 0;
              uint64_t key_l = key;   // the item's prefix, then the lane's digits down to lev
              for (int j = base; j < lev; j++) key_l |= key_put(j, (int)((dig >> (3 * (j - base))) & 7));
              int lo = bl + 1;
              while (m) {
                m &= m - 1;   // drop this item's live sibling; the next one bounds its range
                const int nxt = m ? (__ffs(m) - 1) : -1;
                const int hi = m ? nxt - 1 : hi_l;
                a.it_cell[r] = (uint32_t)cs;
                a.it_key[r] = key_l;
                a.it_meta[r] = (uint32_t)lev | ((uint32_t)lo << 8) | ((uint32_t)hi << 12) | (child_hops << 16);
                a.it_tail[r] = 0u;
                a.it_next[r] = r + 1;
                a.it_cands[r] = 0;
                a.it_noff[r] = kUnsetOff;
                last = r;
                any = true;
                r++;
                lo = nxt;
              }
            }
            a.it_tail[last] = t_new;
            a.it_next[last] = after;
            a.it_next[rec] = (uint32_t)al;   // the donor's run now ends below level top
            for (int i = 0; i < k; i++)
              __hip_atomic_store(&a.it_state[(uint32_t)al + (uint32_t)i], rec_tag(ep, 1), __ATOMIC_RELEASE,
                                 __HIP_MEMORY_SCOPE_AGENT);
            dmask |= ((2u << top) - 1u) & ~((1u << base) - 1u);
            tail = 0;
            atomicAdd(&a.dq[4], 1ull);   // progress
            if (telem) { atomicAdd(&s_tel[6], 1ull); atomicAdd(&s_tel[7], (unsigned long long)k); }
          }
        }
      }
    }
    // ---- 3. queue the new leaves
    const unsigned long long lm = __ballot(has_leaf);
    if (lm) {
      const int slot = qn + __popcll(lm & lt_mask);
      if (has_leaf) q[slot] = lf;
      qn += __popcll(lm);
    }
    // ---- 4. flush 64 leaves at a time through the leaf test. A finished item's candidate
    // count must be final before the item is published, so a wave holding such an item (or
    // with no DFS work left) drains its whole queue.
    const bool drain = (__ballot(st == ST_DONE && nleaf > 0) != 0) ||
                       (__ballot(st == ST_EXPAND || st == ST_BACK) == 0);
    while (qn >= 64 || (drain && qn > 0)) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int nf = qn < 64 ? qn : 64;
      const bool have = lane < nf;
      const unsigned long long t0 = telem ? clock64() : 0;
      Leaf e;
      if (have) e = q[lane];
      else { e.X = e.Y = e.ell = 0; e.item = e.leaf = 0; }
      const unsigned hv = hall_mfma(q, L, a.hall_nkb, (const u32x4_t *)a.hall_frag, a.hall_hi, a.hall_lo, lane);
      const bool sure_rej = (hv & 1u) != 0, unsure = !sure_rej && (hv & 2u);
      bool pass = have && !sure_rej && !unsure;
      // uncertain leaves (and, in verify mode, every leaf) take the FP64 test
      const bool re = have && (unsure || a.hall_verify);
      const unsigned long long rm = __ballot(re);
      if (rm) {
        const bool full = hall_pass(e.X, e.Y, re, L, a.lim, a.hcos, a.hsin, s_order, s_rej);
        if (telem) {
          const unsigned long long nr = __popcll(__ballot(have && unsure));
          const unsigned long long nb = a.hall_verify ? __popcll(__ballot(have && !unsure && full != pass)) : 0;
          if (lane == 0) { atomicAdd(&s_tel[4], nr); atomicAdd(&s_tel[5], nb); }
        }
        if (re) pass = full;
      }
      if (pass) atomicAdd(&a.it_cands[e.item], 1u);
      const unsigned long long pm = __ballot(pass);
      if (pm) {
        const int leader = __ffsll((long long)pm) - 1;
        unsigned long long cbase = 0;
        if (lane == leader) cbase = atomicAdd(a.ncands, (unsigned long long)__popcll(pm));
        cbase = __shfl(cbase, leader);
        if (pass) {
          const unsigned long long idx = cbase + __popcll(pm & lt_mask);
          if (idx < a.cand_capacity) {
            CandRec r;
            r.X = e.X; r.Y = e.Y; r.ell = e.ell; r.item = e.item; r.leaf = e.leaf;
            a.cands[idx] = r;
          }
        }
      }
      if (telem && lane == 0) {
        atomicAdd(&s_tel[0], (unsigned long long)(clock64() - t0));
        atomicAdd(&s_tel[2], 1ull);
        atomicAdd(&s_tel[3], (unsigned long long)nf);
      }
      const int rem = qn - nf;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < rem) q[lane] = q[nf + lane];
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      qn = rem;
    }
    // ---- 5. publish finished items and advance their cells' walk
    const bool fin = st == ST_DONE, skp = st == ST_SKIP;
    if (fin) {
      a.it_nodes[rec] = ell + tail;
      a.it_leaves[rec] = nleaf;
      __hip_atomic_store(&a.it_state[rec], rec_tag(ep, 2), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (capped) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");   // see advance_prefix
        advance_prefix(a, cs);
      }
    } else if (skp) {
      __hip_atomic_store(&a.it_state[rec], rec_tag(ep, 3), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned long long fm = __ballot(fin || skp);
    if (fm && lane == __ffsll((long long)fm) - 1) {
      atomicSub(&a.dq[2], (unsigned long long)__popcll(fm));
      atomicAdd(&a.dq[4], (unsigned long long)__popcll(fm));   // progress
    }
    if (fin || skp) st = ST_IDLE;
    // ---- 6. leave once nothing is in hand and no work can appear: static items are gone,
    // every donation is claimed and no item is held anywhere (only held items donate)
    if (!more && qn == 0 && __ballot(st != ST_IDLE) == 0) {
      if (tried_empty) {   // checked only right after a claim attempt found nothing
        int over = 0;
        if (lane == 0) {
          const unsigned long long h = __hip_atomic_load(&a.dq[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned long long al = __hip_atomic_load(&a.dq[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned long long held = __hip_atomic_load(&a.dq[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned long long open = capped ? __hip_atomic_load(&a.dq[3], __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT) : 0ull;
          over = (h >= al && held == 0 && open == 0) ? 1 : 0;
        }
        if (__shfl(over, 0)) break;
      }
      __builtin_amdgcn_s_sleep(32);
    } else if (qn == 0 && __ballot(st == ST_EXPAND || st == ST_BACK) == 0) {
      // nothing to walk in this wave (waiting for work): yield the SIMD's issue slots to
      // the waves that are walking
      for (int i = 0; i < a.idle_sleep; i++) __builtin_amdgcn_s_sleep(1);
    }
  }
  if (telem && lane == 0) atomicAdd(&s_tel[1], (unsigned long long)(clock64() - t_start));
  __syncthreads();
  for (int i = threadIdx.x; i < 200; i += blockDim.x)
    if (s_rej[i]) atomicAdd(&a.angle_rej[i], s_rej[i]);
  if (telem && threadIdx.x < 10) atomicAdd(&a.telem[threadIdx.x], s_tel[threadIdx.x]);
}

// Per-launch item state: static records start unwalked with no candidates, each linked to
// the next path of its cell; each cell's walk starts at its first path.
__global__ __launch_bounds__(256) void stream_init_kernel(StreamArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nstatic = (size_t)a.ncells * (size_t)a.npaths;
  if (i < (size_t)a.rec_cap) a.it_state[i] = 0;   // no record is published or done yet
  if (i < nstatic) {
    const uint32_t p = (uint32_t)(i % (uint32_t)a.npaths);
    a.it_next[i] = (p + 1 < (uint32_t)a.npaths) ? (uint32_t)i + 1 : kEndItem;
    a.it_cands[i] = 0;
    a.it_noff[i] = p == 0 ? 0ull : kUnsetOff;   // the first item of a cell starts at 0
    a.it_coff[i] = 0;
  }
  if (i < (size_t)a.ncells) {
    a.cell_wp[i] = (uint32_t)(i * (size_t)a.npaths);
    a.cell_lock[i] = 0;
    a.cell_crossed[i] = 0;
    a.cell_open[i] = 1;
    a.cell_next[i] = 0;
    a.cell_sp[i] = 0;
    a.cell_pn[i] = a.cell_pl[i] = a.cell_pc[i] = 0;
  }
  if (i < 6) a.dq[i] = i <= 1 ? (unsigned long long)nstatic : (i == 3 ? (unsigned long long)a.ncells : 0ull);
}

// The exact order after the stream: one thread per cell finishes its cell's walk (the
// whole list when nothing capped the launch). A cell that was not cut has every item of
// its list finished; an unfinished one is reported through cell_tot = kUnsetOff.
__global__ __launch_bounds__(64) void stream_walk_kernel(StreamArgs a, uint64_t *cell_tot) {
  const int cs = blockIdx.x * blockDim.x + threadIdx.x;
  if (cs >= a.ncells) return;
  unsigned long long pn = a.cell_pn[cs], pl = a.cell_pl[cs], pc = a.cell_pc[cs];
  bool bad = false;
  if (!a.cell_crossed[cs]) {
    const uint32_t done = rec_tag(a.epoch, 2);
    uint32_t r = a.cell_wp[cs];
    while (r != kEndItem) {
      if (a.it_state[r] != done) { bad = true; break; }
      a.it_noff[r] = pn; a.it_loff[r] = pl; a.it_coff[r] = pc;
      pn += a.it_nodes[r]; pl += a.it_leaves[r]; pc += a.it_cands[r];
      r = a.it_next[r];
    }
  }
  cell_tot[3 * cs] = bad ? kUnsetOff : pn;
  cell_tot[3 * cs + 1] = pl;
  cell_tot[3 * cs + 2] = pc;
}

// Every candidate of an item inside its cell's stream gets its global DFS node index G and
// leaf index; candidates of items beyond the cut (never walked) or at G >= the node
// budget are dropped.
__global__ __launch_bounds__(256) void stream_keep_kernel(StreamArgs a, unsigned long long ncands, KeptCand *kept,
                                                          unsigned long long *nkept) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncands) return;
  const CandRec r = a.cands[i];
  const uint64_t no = a.it_noff[r.item];
  if (no == kUnsetOff) return;
  const uint64_t G = no + r.ell;
  if (a.node_budget > 0 && G >= (uint64_t)a.node_budget) return;
  const uint32_t nstatic = (uint32_t)a.ncells * (uint32_t)a.npaths;
  const uint32_t cs = r.item < nstatic ? r.item / (uint32_t)a.npaths : a.it_cell[r.item];
  const unsigned long long k = atomicAdd(nkept, 1ull);
  KeptCand o;
  o.X = r.X; o.Y = r.Y; o.G = G; o.leaf = a.it_loff[r.item] + r.leaf; o.cell = (uint64_t)cs;
  kept[k] = o;
}

hipError_t launch_stream(const StreamArgs &a, int blocks, hipStream_t st) {
  const size_t nstatic = (size_t)a.ncells * (size_t)a.npaths;
  const size_t ninit = std::max((size_t)a.rec_cap, (size_t)std::max(a.ncells, 6));
  hipLaunchKernelGGL(stream_init_kernel, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(stream_kernel, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_stream_finalize(const StreamArgs &a, uint64_t *cell_tot, unsigned long long ncands, KeptCand *kept,
                                  unsigned long long *nkept, hipStream_t st) {
  hipLaunchKernelGGL(stream_walk_kernel, dim3((unsigned)((a.ncells + 63) / 64)), dim3(64), 0, st, a, cell_tot);
  if (ncands) {
    const unsigned blocks = (unsigned)((ncands + 255) / 256);
    hipLaunchKernelGGL(stream_keep_kernel, dim3(blocks), dim3(256), 0, st, a, ncands, kept, nkept);
  }
  return hipGetLastError();
}
}  // namespace wzfh
