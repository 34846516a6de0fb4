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

enum : int { ST_IDLE = 0, ST_EXPAND = 1, ST_BACK = 2, ST_DONE = 3 };

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

// Completed-prefix tracker of one cell. Items of a cell finish in any order; the lane
// that takes the cell's lock walks the run of finished items after the current prefix,
// summing their nodes and candidates, and publishes the cell's cut as soon as the sums
// reach the node budget or candidate cap: every later item is then certainly beyond the
// cut and is skipped or abandoned. The lock is only ever tried (never waited on), and a
// lane that sees new finished items after releasing it tries again a bounded number of
// times. The exact offsets are recomputed by stream_finalize_kernel, so this walk only
// decides how early a cell stops.
__device__ void advance_prefix(const StreamArgs &a, int cs) {
  const size_t base = (size_t)cs * a.npaths;
  for (int attempt = 0; attempt < 4; attempt++) {
    if (atomicCAS(&a.cell_lock[cs], 0, 1) != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    int pp = a.cell_pp[cs];
    unsigned long long pn = a.cell_pn[cs], pc = a.cell_pc[cs];
    bool crossed = false;
    while (pp < a.npaths) {
      if (!__hip_atomic_load(&a.item_done[base + pp], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) break;
      pn += a.item_nodes[base + pp];
      pc += __hip_atomic_load(&a.item_cands[base + pp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pp++;
      if ((a.node_budget > 0 && pn >= (unsigned long long)a.node_budget) ||
          (a.cand_cap > 0 && pc >= (unsigned long long)a.cand_cap)) { crossed = true; break; }
    }
    a.cell_pp[cs] = crossed ? a.npaths : pp;   // a crossed cell's walk is over
    a.cell_pn[cs] = pn;
    a.cell_pc[cs] = pc;
    if (crossed) atomicMin(&a.cut[cs], pp - 1);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicExch(&a.cell_lock[cs], 0);
    if (crossed || pp >= a.npaths) return;
    if (!__hip_atomic_load(&a.item_done[base + pp], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) return;
  }
}

}  // namespace

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void stream_kernel(StreamArgs a) {
  __shared__ uint8_t s_ne[256], s_eq[64];
  __shared__ uint16_t s0_ne[256], s0_eq[64];
  __shared__ uint32_t s_di[32];
  __shared__ uint8_t s_q8[8], s_q16[16], s_order[200];
  __shared__ unsigned s_rej[200];
  __shared__ Leaf s_q[kWaves][kQueue];

  for (int i = threadIdx.x; i < 256; i += blockDim.x) { s_ne[i] = a.tab_ne[i]; s0_ne[i] = a.tab0_ne[i]; }
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

  int st = ST_IDLE;
  int cs = 0, p = 0;
  uint32_t item = 0, nleaf = 0;
  int d = 0;
  uint64_t dig = 0, RX = 0, RY = 0, X = 0, Y = 0, ell = 0;
  bool more = true;
  int qn = 0;
  unsigned long long moves = 0;
  const bool telem = a.telem != nullptr;
  unsigned long long t_start = telem ? clock64() : 0, t_hall = 0, n_flush = 0, n_tested = 0, n_recheck = 0, n_bad = 0;

  while (true) {
    // ---- 1. idle lanes take the next items (one atomic per wave)
    bool has_leaf = false;
    Leaf lf;
    if (more) {
      const unsigned long long need = __ballot(st == ST_IDLE);
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.work_ctr, (unsigned long long)cnt);
        base = __shfl(base, leader);
        if (base + cnt >= total) more = false;
        if (st == ST_IDLE) {
          const unsigned long long g = base + __popcll(need & lt_mask);
          if (g < total) {
            const unsigned long long w = g % a.window, t = g / a.window;
            cs = (int)(t % a.ncells);
            p = (int)((t / a.ncells) * a.window + w);
            if (p < a.npaths && p <= a.cut[cs]) {
              // ---- replay path p: digits k_0 (radix nch0) then k_1..k_{d0-1} (radix 8)
              item = (uint32_t)cs * (uint32_t)a.npaths + (uint32_t)p;
              RX = a.cell_res[2 * cs];
              RY = a.cell_res[2 * cs + 1];
              X = Y = 0;
              nleaf = 0;
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
      }
    }

    // ---- 2. one DFS move per working lane
    if (st == ST_EXPAND) {
      const unsigned mask = child_mask(s_ne, s_eq, s_di[d], RX, RY);
      if (mask == 0) {
        ell += 8;
        st = ST_BACK;
      } else {
        const int b = __ffs(mask) - 1;
        ell += b + 1;
        apply_quad(s_di[d], d, L, s_q8[b], 1, RX, RY, X, Y);
        dig |= (uint64_t)b << (3 * (d - d0));
        d++;
        if (d == half) { has_leaf = true; st = ST_BACK; }
      }
    } else if (st == ST_BACK && !has_leaf) {
      while (true) {
        if (d == d0) { st = ST_DONE; break; }
        d--;
        const int sh = 3 * (d - d0);
        const int b = (int)((dig >> sh) & 7);
        dig &= ~(7ull << sh);
        apply_quad(s_di[d], d, L, s_q8[b], -1, RX, RY, X, Y);
        const unsigned mask = child_mask(s_ne, s_eq, s_di[d], RX, RY) & (0xFEu << b) & 0xFFu;
        if (mask) {
          const int b2 = __ffs(mask) - 1;
          ell += b2 - b;
          apply_quad(s_di[d], d, L, s_q8[b2], 1, RX, RY, X, Y);
          dig |= (uint64_t)b2 << sh;
          d++;
          if (d == half) has_leaf = true;
          else st = ST_EXPAND;
          break;
        }
        ell += 7 - b;
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
      lf.item = item;
      lf.leaf = nleaf++;
    }
    // ---- local caps and the per-cell cut (a cell's cut only moves down)
    if (st == ST_EXPAND || st == ST_BACK) {
      bool stop = false;
      if (a.node_budget > 0 && ell >= (uint64_t)a.node_budget) stop = true;
      if (a.cand_cap > 0 && has_leaf && __hip_atomic_load(&a.item_cands[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)a.cand_cap) stop = true;
      if (stop) { atomicMin(&a.cut[cs], p); st = ST_DONE; }
      else if ((++moves & 255) == 0 && __hip_atomic_load(&a.cut[cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < p) st = ST_DONE;
    }
    // ---- 3. queue the new leaves
    const unsigned long long lm = __ballot(has_leaf);
    if (lm) {
      const int slot = qn + __popcll(lm & lt_mask);
      if (has_leaf) q[slot] = lf;
      qn += __popcll(lm);
    }
    // ---- 4. flush 64 leaves at a time through the FP64 hall test. A finished item's
    // candidate count must be final before the item is published, so a wave holding such
    // an item (or out of work) drains its whole queue.
    const bool drain = (__ballot(st == ST_DONE && nleaf > 0) != 0) ||
                       (!more && __ballot(st == ST_EXPAND || st == ST_BACK) == 0);
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
          n_recheck += __popcll(__ballot(have && unsure));
          if (a.hall_verify) n_bad += __popcll(__ballot(have && !unsure && full != pass));
        }
        if (re) pass = full;
      }
      if (pass) atomicAdd(&a.item_cands[e.item], 1u);
      const unsigned long long pm = __ballot(pass);
      if (pm) {
        const int leader = __ffsll((long long)pm) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.ncands, (unsigned long long)__popcll(pm));
        base = __shfl(base, leader);
        if (pass) {
          const unsigned long long idx = base + __popcll(pm & lt_mask);
          if (idx < a.cand_capacity) {
            CandRec r;
            r.X = e.X; r.Y = e.Y; r.ell = e.ell; r.item = e.item; r.leaf = e.leaf;
            a.cands[idx] = r;
          }
        }
      }
      if (telem) { t_hall += clock64() - t0; n_flush++; n_tested += nf; }
      const int rem = qn - nf;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < rem) q[lane] = q[nf + lane];
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      qn = rem;
    }
    // ---- 5. publish finished items and advance their cells' completed prefix
    if (st == ST_DONE) {
      a.item_nodes[item] = ell;
      a.item_leaves[item] = nleaf;
      __hip_atomic_store(&a.item_done[item], (uint32_t)1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (a.node_budget > 0 || a.cand_cap > 0) advance_prefix(a, cs);
      st = ST_IDLE;
    }
    if (!more && qn == 0 && __ballot(st != ST_IDLE) == 0) break;
  }
  if (telem && lane == 0) {
    atomicAdd(&a.telem[0], t_hall);
    atomicAdd(&a.telem[1], (unsigned long long)(clock64() - t_start));
    atomicAdd(&a.telem[2], n_flush);
    atomicAdd(&a.telem[3], n_tested);
    atomicAdd(&a.telem[4], n_recheck);
    atomicAdd(&a.telem[5], n_bad);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 200; i += blockDim.x)
    if (s_rej[i]) atomicAdd(&a.angle_rej[i], s_rej[i]);
}

}  // namespace wzfh

namespace wzfh {

// Exact global order after the stream: one block per cell scans the cell's items in path
// order (the reference's DFS order) into node / leaf offsets and cell totals; then every
// candidate gets its global DFS node index G and leaf index, and the ones that can be in
// the cell's stream (path not beyond the cut, G below the node budget) are compacted.
__global__ __launch_bounds__(256) void stream_scan_kernel(StreamArgs a, uint64_t *noff, uint64_t *loff,
                                                          uint64_t *cell_tot) {
  __shared__ uint64_t s_n[256], s_l[256], s_c[256];
  const int cs = blockIdx.x, t = threadIdx.x;
  const size_t base = (size_t)cs * a.npaths;
  uint64_t carry_n = 0, carry_l = 0, carry_c = 0;
  for (int p0 = 0; p0 < a.npaths; p0 += 256) {
    const int p = p0 + t;
    const uint64_t vn = p < a.npaths ? a.item_nodes[base + p] : 0;
    const uint64_t vl = p < a.npaths ? a.item_leaves[base + p] : 0;
    const uint64_t vc = p < a.npaths ? a.item_cands[base + p] : 0;
    s_n[t] = vn; s_l[t] = vl; s_c[t] = vc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {   // inclusive Hillis-Steele scan
      uint64_t xn = 0, xl = 0, xc = 0;
      if (t >= off) { xn = s_n[t - off]; xl = s_l[t - off]; xc = s_c[t - off]; }
      __syncthreads();
      s_n[t] += xn; s_l[t] += xl; s_c[t] += xc;
      __syncthreads();
    }
    if (p < a.npaths) { noff[base + p] = carry_n + s_n[t] - vn; loff[base + p] = carry_l + s_l[t] - vl; }
    carry_n += s_n[255]; carry_l += s_l[255]; carry_c += s_c[255];
    __syncthreads();
  }
  if (t == 0) { cell_tot[3 * cs] = carry_n; cell_tot[3 * cs + 1] = carry_l; cell_tot[3 * cs + 2] = carry_c; }
}

__global__ __launch_bounds__(256) void stream_keep_kernel(StreamArgs a, const uint64_t *noff, const uint64_t *loff,
                                                          unsigned long long ncands, KeptCand *kept,
                                                          unsigned long long *nkept) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncands) return;
  const CandRec r = a.cands[i];
  const int cs = (int)(r.item / (uint32_t)a.npaths), p = (int)(r.item % (uint32_t)a.npaths);
  if (p > a.cut[cs]) return;
  const uint64_t G = noff[r.item] + r.ell;
  if (a.node_budget > 0 && G >= (uint64_t)a.node_budget) return;
  const unsigned long long k = atomicAdd(nkept, 1ull);
  KeptCand o;
  o.X = r.X; o.Y = r.Y; o.G = G; o.leaf = loff[r.item] + r.leaf; o.cell = (uint64_t)cs;
  kept[k] = o;
}

hipError_t launch_stream(const StreamArgs &a, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(stream_kernel, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_stream_finalize(const StreamArgs &a, uint64_t *noff, uint64_t *loff, uint64_t *cell_tot,
                                  unsigned long long ncands, KeptCand *kept, unsigned long long *nkept,
                                  hipStream_t st) {
  hipLaunchKernelGGL(stream_scan_kernel, dim3(a.ncells), dim3(256), 0, st, a, noff, loff, cell_tot);
  if (ncands) {
    const unsigned blocks = (unsigned)((ncands + 255) / 256);
    hipLaunchKernelGGL(stream_keep_kernel, dim3(blocks), dim3(256), 0, st, a, noff, loff, ncands, kept, nkept);
  }
  return hipGetLastError();
}
}  // namespace wzfh
