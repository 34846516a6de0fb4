// wzfh_complete.hip -- A,B completer kernel for gfx950.
//
// Restates fh_complete_ab / fh_ab_search (src/solver/wz_match.cpp:1092-1341) with one
// work item per wavefront: a whole candidate, or (split launches) the subtree of a candidate
// below one quad path, replayed from the root; items in path order concatenate to the
// reference's DFS, and a reduce kernel combines them. The mirror-quad DFS control is wave-uniform (scalar
// registers); lane s holds the shift-s state: the target T[s] = -CD[s], the partial
// A,B autocorrelation Dab[s], and the remaining-pair bound Kab[d][s] from an LDS table.
// A, B are packed 64-bit words (bit i set = element i is -1).
//
// A node is expanded in one vector pass: the new pairs that the quad (d, L-1-d) forms
// with the placed positions reduce to four per-lane partner sums U1, V1, U2, V2 and a
// cross term, so the Dab change of each of the 8 quads is a signed sum of those, and
// the bound scan of all 8 children is 8 compares and 8 ballots. The A1 mask, root and
// reversal canon, early outer check and sum capacity are scalar bit masks over the
// quads. The quad loop itself then runs on scalar masks: it charges the nodes the
// reference charges, in its order, and descends into the first child that also passes
// the profile-row filter (evaluated lazily, in quad order, like the reference).
//
// Profile rows: the surviving rows of each depth are bitsets kept in LDS (one word list
// plus a 64-row bitset per surviving word and depth); row data is read from HBM four
// words at a time so the loads overlap. When a launch's row lists do not fit in LDS the
// same code runs on a per-wave global scratch area.
//
// Semantics follow the reference exactly, including the order of the per-quad steps
// (A1 mask, root canon, reversal canon, charge + budget, early outer check, place,
// sum capacity, bound scan, row filter), the odd-L middle trials that count nodes
// without charging the budget, and the extra node the reference charges at each
// ancestor level while unwinding a budget abort.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wzfh_kernels.h"

namespace wzfh {

namespace {

constexpr int kLevels = 32;

__device__ __forceinline__ int sgn(uint64_t w, int q) { return 1 - 2 * (int)((w >> q) & 1u); }

// Makes a wave-uniform value visible as such to the compiler (an SGPR), so every branch
// of the DFS is a scalar branch and no loop exit is ever lowered to an exec mask.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ bool any_lane(bool p) { return uni(__any(p) ? 1 : 0) != 0; }
__device__ __forceinline__ long long uni64(long long v) {
  const unsigned lo = (unsigned)uni((int)(unsigned)(unsigned long long)v);
  const unsigned hi = (unsigned)uni((int)(unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Writes a wave-uniform value into lane `at` of a per-lane record (lane = DFS level).
__device__ __forceinline__ int put_lane(int val, int at, int old) {
  return ((int)(threadIdx.x & 63) == at) ? val : old;
}

// Quad k of depth d: (a1, b1, a2, b2) = signs of A[d], B[d], A[L-1-d], B[L-1-d]
// (P22_NEG / P22_POS of wz_match.cpp): a1 = +1 iff k&4, b1 = +1 iff k&2, a2 = +1 iff k&1,
// and b2 = a1 b1 a2 at depth >= 1 (quad product +1) or -a1 b1 a2 at depth 0 (product -1).
// Computed, not looked up: a table indexed by a run-time k is a chain of memory loads.
__device__ __forceinline__ void quad_signs(int d, int k, int &a1, int &b1, int &a2, int &b2) {
  a1 = (k & 4) ? 1 : -1;
  b1 = (k & 2) ? 1 : -1;
  a2 = (k & 1) ? 1 : -1;
  b2 = (d == 0 ? -1 : 1) * a1 * b1 * a2;
}

// Quad masks (bit k = quad k) of the +1 signs: a1, b1, a2, and b2 at depth >= 1.
constexpr unsigned kA1Pos = 0xF0u, kB1Pos = 0xCCu, kA2Pos = 0xAAu, kB2Pos = 0x96u;

// Per-lane partner sums of the quad at depth d (positions i1 = d, i2 = L-1-d) against
// the placed positions [0, d) and [L-d, L): lane s gets the sums of the A and B signs at
// the partners of i1 (U1, V1) and of i2 (U2, V2) at distance s, and x = 1 when the two
// new positions are themselves at distance s (their own pair).
struct Partners {
  int U1, V1, U2, V2, x;
};

__device__ __forceinline__ Partners partners(uint64_t A, uint64_t B, int d, int L, int s, bool valid) {
  Partners p{0, 0, 0, 0, 0};
  if (!valid) return p;
  const int i1 = d, i2 = L - 1 - d;
  int q = i1 - s;
  if (q >= 0) { p.U1 += sgn(A, q); p.V1 += sgn(B, q); }
  q = i1 + s;
  if (q >= L - d && q < L) { p.U1 += sgn(A, q); p.V1 += sgn(B, q); }
  q = i2 - s;
  if (q >= 0 && q < d) { p.U2 += sgn(A, q); p.V2 += sgn(B, q); }
  if (q == d) p.x = 1;
  q = i2 + s;
  if (q >= L - d && q < L) { p.U2 += sgn(A, q); p.V2 += sgn(B, q); }
  return p;
}

__device__ __forceinline__ int quad_delta(const Partners &p, int a1, int b1, int a2, int b2) {
  return a1 * p.U1 + b1 * p.V1 + a2 * p.U2 + b2 * p.V2 + p.x * (a1 * a2 + b1 * b2);
}

// a2_row_ok (wz_match.cpp:1017-1036): residual (u, v) of a self-reflected block with q quads
// left, each adding (+-2, +-2) or (0, 0)
__device__ __forceinline__ bool a2_self_ok(int u, int v, int q) {
  if ((u | v) & 1) return false;
  if (abs(u) > 2 * q || abs(v) > 2 * q) return false;
  return ((u - v) & 3) == 0;
}

// a2_row_ok: can the remaining positive quads of every mirror block reach the row (k, r words)
// from the partial sums (PA, PB words, 64-biased bytes) at depth p? A quad of a block of two
// residues adds a +-column of the 4x4 Hadamard matrix H, so with t = H (row - partial) the
// block's q quads reach it iff t = 0 mod 4, |t|_1 / 4 <= q and q - |t|_1 / 4 is even.
__device__ __forceinline__ bool a2_row_ok(const CompleteArgs &a, uint64_t kw, uint64_t rw, uint64_t PA,
                                          uint64_t PB, int p) {
  auto f = [](uint64_t w, int c) { return (int)(int8_t)((w >> (8 * c)) & 255u); };
  auto pf = [](uint64_t w, int c) { return (int)((w >> (8 * c)) & 255u) - 64; };
  for (int b = 0; b < a.a2_nblk; b++) {
    const int r = a.a2_r[b], sr = a.a2_s[b], q = a.a2_cnt[p][b];
    const int ur = f(kw, r) - pf(PA, r), vr = f(rw, r) - pf(PB, r);
    if (r != sr) {
      const int us = f(kw, sr) - pf(PA, sr), vs = f(rw, sr) - pf(PB, sr);
      const int t0 = ur + vr + us + vs, t1 = ur - vr + us - vs, t2 = ur + vr - us - vs, t3 = ur - vr - us + vs;
      if ((t0 | t1 | t2 | t3) & 3) return false;
      const int l1 = (abs(t0) + abs(t1) + abs(t2) + abs(t3)) >> 2;
      if (l1 > q || ((q - l1) & 1)) return false;
    } else if (a.a2_mid[b]) {
      if (!a2_self_ok(ur - 1, vr - 1, q) && !a2_self_ok(ur - 1, vr + 1, q) && !a2_self_ok(ur + 1, vr - 1, q) &&
          !a2_self_ok(ur + 1, vr + 1, q))
        return false;
    } else if (!a2_self_ok(ur, vr, q)) {
      return false;
    }
  }
  return true;
}

}  // namespace

// One wavefront per block. kLds: the row bitsets live in dynamic LDS; otherwise in the
// per-wave global scratch area a.scratch.
template <bool kLds>
__global__ __launch_bounds__(64) void complete_kernel(CompleteArgs a, const uint8_t *kab_table) {
  __shared__ int16_t s_dab[kLevels][64];
  __shared__ uint8_t s_kab[kLevels][64];
  HIP_DYNAMIC_SHARED(uint64_t, dyn_lds)
  for (int i = threadIdx.x; i < kLevels * 64; i += 64) s_kab[i / 64][i % 64] = kab_table[i];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int n = a.n, L = a.L, m = a.m, half = L / 2;
  const bool canon = a.canon != 0;
  const bool valid = lane >= 1 && lane <= n;
  const int s = lane;
  // row bitsets: [half + 1][W] words, then [half + 1][W] u16 word lists
  const int W = a.scratch_words;
  const size_t nlw = (size_t)(half + 1) * W;
  uint64_t *bits = kLds ? dyn_lds : a.scratch + (size_t)blockIdx.x * (nlw + (nlw + 3) / 4);
  uint16_t *wlist = (uint16_t *)(bits + nlw);
  const ulonglong2 *rows2 = (const ulonglong2 *)a.rows;
  auto class_field = [](uint64_t w, int c) { return (int)((w >> (8 * c)) & 255u) - 64; };

  // DFS states of one level (the iterative form of fh_ab_search's recursion)
  enum { S_ENTER = 0, S_NEXT = 1, S_RETURN = 2, S_FOUND = 3, S_ABORT = 4, S_EXHAUSTED = 5 };

  const int ext = a.split_ext;              // path digits this launch adds per entry
  while (true) {
    // next work item: the first active lane takes the ticket, readfirstlane broadcasts it
    const unsigned long long act = __ballot(1);
    unsigned long long u_l = 0;
    if (lane == __ffsll((long long)act) - 1) u_l = atomicAdd(a.work_ctr, 1ull);
    const long long u = (long long)((((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(u_l >> 32))) << 32) |
                                    (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u_l));
    if (u >= a.nwork) break;
    const int ci = (int)(u >> a.items_log2);                  // entry
    const int sub = (int)(u & ((1ll << a.items_log2) - 1));   // item of the entry
    // quad path below the entry; with root_pair the depth-0 digit is 6 or 7 (root canon: every
    // other depth-0 quad is cut before it is charged, so those paths carry no work)
    const int psub = a.root_pair ? (((6 | (sub >> (3 * (ext - 1)))) << (3 * (ext - 1))) | (sub & ((1 << (3 * (ext - 1))) - 1)))
                                 : sub;
    if (a.ent_skip && sub < uni(a.ent_skip[ci])) {   // walked already (see CompleteArgs::ent_skip)
      a.verdict[u] = 2;
      a.nodes[u] = 0;
      a.outA[u] = 0;
      a.outB[u] = 0;
      a.charges[u] = 0;
      continue;
    }
    const int base = uni((a.ent_base ? a.ent_base[ci] : 0) + ext);   // levels replayed along the path
    const long long path = uni64(((a.ent_path ? a.ent_path[ci] : 0ll) << (3 * ext)) | psub);   // depth l: bits 3(base-1-l)
    // a budgeted launch (a.budget > 0) gives each entry the charges left to it, possibly 0
    const bool limited = a.budget > 0;
    const long long bud = uni64(a.ent_budget ? a.ent_budget[ci] : a.budget);

    const uint64_t C = a.cd[2 * ci], D = a.cd[2 * ci + 1];
    // T[s] = -CD[s] (wz_match.cpp:1300-1304) from packed words: x_i x_{i+s} = 1 - 2 (b_i ^ b_{i+s}).
    int T = 0;
    if (s >= 1 && s < n) {
      const uint64_t mk = (1ull << (n - s)) - 1;
      const int dis = __popcll((C ^ (C >> s)) & mk) + __popcll((D ^ (D >> s)) & mk);
      T = -(2 * (n - s) - 2 * dis);
    }
    int verdict = 2;
    long long total = 0, cur = 0;   // nodes counted; nodes charged to the budget
    long long dn = 0;               // lane l: nodes counted at depth l (FH_TELEM depth_nodes)
    long long snap_c = 0, snap_n = 0;   // charges (| sub-path << 44) and nodes before the current sub-path
    uint64_t A = 0, B = 0;
    // pre-filter (wz_match.cpp:1305-1308): |T[s]| > 2 (L - s) for some s < L
    const bool prerej = any_lane(valid && (T > 2 * (L - s) || -T > 2 * (L - s)));
    if (!prerej) {
      const int roff = uni(a.row_off[ci]);
      const int rlen = uni(a.row_len[ci]);
      const bool rows = roff >= 0;
      const int nwords = rows ? (rlen + 63) / 64 : 0;
      int lvl_nw = 0;            // lane l: surviving-word count at depth l
      if (rows) {
        for (int w = lane; w < nwords; w += 64) {
          const int left = rlen - 64 * w;
          bits[w] = left >= 64 ? ~0ull : ((1ull << left) - 1);
          wlist[w] = (uint16_t)w;
        }
        lvl_nw = (lane == 0) ? nwords : 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      // uniform DFS state of the current node
      int d = 0, state = S_ENTER;
      int kcur = -1;             // last quad of this node's loop that was reached
      unsigned cmask = 0;        // quads passing the pre-charge filters (charged when reached)
      unsigned vmask = 0;        // ... that also pass the early check, sum capacity and bound
      int sumA = 0, sumB = 0;
      uint64_t PA = 0x4040404040404040ull, PB = PA;   // 8-bit class fields: class sum + 64
      int dab = 0;                // lane s: Dab[s] of the current node
      int a_tied = 1, a_cmp = 0, b_tied = 1, b_cmp = 0;
      int st_k = 0, st_mask = 0, st_ties = 0, st_sums = 0;   // lane l = level l records
      Partners pp{0, 0, 0, 0, 0};
      bool pp_ok = false;         // pp holds the current node's partner sums
      bool cap_stop = false;      // stopped at the launch's per-item charge cap
      // split items report their charges 256 at a time; an item whose entry's earlier items
      // have together charged more than the budget lies wholly past the abort point and stops
      // (its result is never used). An item before or at the abort point never stops early.
      auto charged_out = [&](long long c) -> bool {
        if (ext == 0 || !limited || (c & 255) != 0) return false;
        if (lane == 0) __hip_atomic_store(&a.item_charged[u], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((c & 4095) != 0 || sub == 0) return false;
        long long sum = 0;
        for (int q = lane; q < sub; q += 64)
          sum += __hip_atomic_load(&a.item_charged[u - sub + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        return uni((int)(sum > bud)) != 0;
      };
      while (state < S_FOUND) {
        if (state == S_ENTER) {
          if (d == half) {
            // ---- leaf (wz_match.cpp:1178-1206)
            bool ok_leaf = false;
            if (L & 1) {
              const int mid = half, cm = mid % m;
              for (int t = 0; t < 4 && !ok_leaf; t++) {
                const int av = (t & 2) ? -1 : 1, bv = (t & 1) ? -1 : 1;   // P22_4 order
                int delta = 0;
                if (valid) {
                  if (mid - s >= 0) delta += sgn(A, mid - s) * av + sgn(B, mid - s) * bv;
                  if (mid + s < L) delta += sgn(A, mid + s) * av + sgn(B, mid + s) * bv;
                }
                total++;
                dn += (lane == d) ? 1 : 0;
                bool ok = (abs(sumA + av) == a.absA) && (abs(sumB + bv) == a.absB);
                if (ok) ok = !any_lane(valid && (dab + delta != T));
                if (ok && rows) {
                  // fh_abp_leaf_ok (wz_match.cpp:1080-1090): a depth-half row equal to the
                  // class sums; only the middle's class can still differ
                  const int wantk = class_field(PA, cm) + av, wantr = class_field(PB, cm) + bv;
                  const int nw = uni(__builtin_amdgcn_readlane(lvl_nw, half));
                  const uint16_t *lst = wlist + (size_t)half * W;
                  const uint64_t *bts = bits + (size_t)half * W;
                  bool hitany = false;
                  for (int t0 = 0; t0 < nw && !hitany; t0 += 4) {
                    ulonglong2 r4[4];
                    bool on[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                      const int w = t0 + u < nw ? uni(lst[t0 + u]) : 0;
                      on[u] = t0 + u < nw && ((bts[w] >> lane) & 1);
                      if (on[u]) r4[u] = rows2[roff + 64 * w + lane];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                      bool hit = false;
                      if (on[u]) {
                        const int kk = (int)(int8_t)((r4[u].x >> (8 * cm)) & 255);
                        const int rr = (int)(int8_t)((r4[u].y >> (8 * cm)) & 255);
                        hit = (kk == wantk) && (rr == wantr);
                      }
                      hitany = hitany || any_lane(hit);
                    }
                  }
                  ok = hitany;
                }
                if (ok) {
                  ok_leaf = true;
                  if (av < 0) A |= 1ull << mid;
                  if (bv < 0) B |= 1ull << mid;
                  dab += delta;
                }
              }
            } else {
              bool ok = (abs(sumA) == a.absA) && (abs(sumB) == a.absB);
              if (ok) ok = !any_lane(valid && (dab != T));
              if (ok && rows) ok = uni(__builtin_amdgcn_readlane(lvl_nw, half)) > 0;
              ok_leaf = ok;
            }
            state = ok_leaf ? S_FOUND : S_RETURN;
            continue;
          }
          // ---- level entry. The A1 outer mask (wz_match.cpp:601-618, 1210-1211) is a table
          // entry: its load is issued first and used last, after the vector bound pass.
          const int Rl = T - dab;
          const uint8_t *optr = nullptr;
          bool ozero = false;
          if (a.outer_k && d > 0) {
            const int kk = a.outer_k;
            if (!(kk < 2 || d < kk || d + kk > L / 2 || (A & 1) || (B & 1))) {
              int pat = ((A >> (L - 1)) & 1) ? 0 : 1, mul = 2;
              bool skip = false;
              for (int t = 1; t < kk; t++) {
                const int sb = ((((A >> t) & 1) ? 0 : 1) << 3) | ((((B >> t) & 1) ? 0 : 1) << 2) |
                               ((((A >> (L - 1 - t)) & 1) ? 0 : 1) << 1) | (((B >> (L - 1 - t)) & 1) ? 0 : 1);
                const int q = (int)((a.pos_pack >> (4 * sb)) & 15u) - 1;
                if (q < 0) { skip = true; break; }
                pat += mul * q;
                mul *= 8;
              }
              if (!skip) {
                int R[3] = {0, 0, 0};
                for (int j = 0; j < kk; j++) {
                  R[j] = uni(__builtin_amdgcn_readlane(Rl, L - 1 - d - j));
                  if ((R[j] & 1) || R[j] < -4 * (j + 1) || R[j] > 4 * (j + 1)) ozero = true;
                }
                if (!ozero)
                  optr = (kk == 2) ? a.outer2 + (pat * 5 + (R[0] / 2 + 2)) * 9 + (R[1] / 2 + 4)
                                   : a.outer3 + ((pat * 5 + (R[0] / 2 + 2)) * 9 + (R[1] / 2 + 4)) * 13 + (R[2] / 2 + 6);
              }
            }
          }
          const unsigned oraw = optr ? (unsigned)*optr : 0xFFu;
          // root and reversal canon (wz_match.cpp:1214-1240), early outer check (:1250-1260)
          // and sum capacity (:1263-1267) as masks over the 8 quads (bit k = quad k)
          unsigned pmask = 0xFFu;
          if (d == 0) {
            if (canon) pmask = kA1Pos & kB1Pos;          // root canon: A[0] = B[0] = +1
          } else if (canon) {
            // while a prefix ties its reverse, a1 = -1 is dropped when a2 equals the tie sign
            if (a_tied) pmask &= ~(~kA1Pos & (a_cmp > 0 ? kA2Pos : ~kA2Pos));
            if (b_tied) pmask &= ~(~kB1Pos & (b_cmp > 0 ? kB2Pos : ~kB2Pos));
          }
          const unsigned b2pos = (d == 0) ? (~kB2Pos & 0xFFu) : kB2Pos;
          unsigned emask = 0xFFu;
          if (d > 0) {
            // A[0] a2 + B[0] b2 + A[L-1] a1 + B[L-1] b1 must equal the residual of shift L-1-d:
            // count, per quad, the terms that are +1 (bit-sliced over the 8 quads)
            const int Rse = uni(__builtin_amdgcn_readlane(Rl, L - 1 - d));
            const unsigned P1 = (A & 1) ? ~kA2Pos : kA2Pos, P2 = (B & 1) ? ~b2pos : b2pos;
            const unsigned P3 = ((A >> (L - 1)) & 1) ? ~kA1Pos : kA1Pos, P4 = ((B >> (L - 1)) & 1) ? ~kB1Pos : kB1Pos;
            const unsigned s1 = P1 ^ P2, c1 = P1 & P2, s2 = P3 ^ P4, c2 = P3 & P4;
            const unsigned b0 = s1 ^ s2, cr = s1 & s2;
            const unsigned h1 = c1 ^ c2 ^ cr, h2 = (c1 & c2) | (c1 & cr) | (c2 & cr);
            const int want = Rse + 4;   // 2 * (number of +1 terms)
            if ((want & 1) || want < 0 || want > 8) emask = 0;
            else {
              const int c = want / 2;
              emask = ((c & 1) ? b0 : ~b0) & ((c & 2) ? h1 : ~h1) & ((c & 4) ? h2 : ~h2);
            }
          }
          const int rem = L - 2 * (d + 1);
          auto cap_ok = [&](int absv, int v) { return abs(absv - v) <= rem || abs(-absv - v) <= rem; };
          const unsigned ap2 = kA1Pos & kA2Pos, am2 = ~kA1Pos & ~kA2Pos & 0xFFu;
          const unsigned bp2 = kB1Pos & b2pos, bm2 = ~kB1Pos & ~b2pos & 0xFFu;
          const unsigned amask = (cap_ok(a.absA, sumA + 2) ? ap2 : 0u) | (cap_ok(a.absA, sumA - 2) ? am2 : 0u) |
                                 (cap_ok(a.absA, sumA) ? (0xFFu & ~(ap2 | am2)) : 0u);
          const unsigned bmask_s = (cap_ok(a.absB, sumB + 2) ? bp2 : 0u) | (cap_ok(a.absB, sumB - 2) ? bm2 : 0u) |
                                   (cap_ok(a.absB, sumB) ? (0xFFu & ~(bp2 | bm2)) : 0u);
          const unsigned smask = pmask & emask & amask & bmask_s & 0xFFu;
          // bound scan (:1268-1270) of every remaining quad in one pass: lane s checks
          // |T[s] - Dab[s] - delta_k[s]| <= Kab[d+1][s]
          unsigned bmask = 0;
          if (smask) {
            pp = partners(A, B, d, L, s, valid);
            pp_ok = true;
            const int kb = valid ? (int)s_kab[d + 1][lane] : (1 << 20);
            const int sg = (d == 0) ? -1 : 1;
            const int V2s = sg * pp.V2, X2 = (1 + sg) * pp.x;
#pragma unroll
            for (int k = 0; k < 8; k++) {
              if ((smask >> k) & 1u) {
                const int a1 = (k & 4) ? 1 : -1, b1 = (k & 2) ? 1 : -1, a2 = (k & 1) ? 1 : -1;
                const int dk = a1 * pp.U1 + b1 * pp.V1 + a2 * pp.U2 + (a1 * b1 * a2) * V2s + (a1 * a2) * X2;
                const int e = Rl - dk;
                if (!any_lane((unsigned)(e + kb) > (unsigned)(2 * kb))) bmask |= 1u << k;
              }
            }
          }
          // the A1 mask gates the charges: with no quad left the level returns uncharged
          const unsigned omask = ozero ? 0u : (unsigned)uni((int)oraw);
          cmask = (omask == 0) ? 0u : (pmask & ((d > 0) ? omask : 0xFFu));
          vmask = cmask & smask & bmask;
          kcur = -1;
          state = S_NEXT;
          continue;
        }
        if (state == S_NEXT) {
          int k;
          if (d < base) {
            // replayed level of a split item: only the path's quad. Its charge belongs to
            // the first path below it (all later digits 0), so every charge of the levels
            // above the split is counted by exactly one item, in DFS order.
            const int sh = 3 * (base - 1 - d);
            k = (int)((path >> sh) & 7);
            if (!((cmask >> k) & 1u)) { state = S_EXHAUSTED; continue; }
            if ((path & ((1ll << sh) - 1)) == 0) {
              total++;
              dn += (lane == d) ? 1 : 0;
              if (limited && ++cur > bud) { state = S_ABORT; continue; }
            }
            if (charged_out(cur)) { state = S_ABORT; continue; }
            if (a.item_cap > 0 && cur >= a.item_cap) { cap_stop = true; state = S_ABORT; continue; }
            if (!((vmask >> k) & 1u)) { state = S_EXHAUSTED; continue; }
          } else {
            // the reference's quad loop on masks: every quad of cmask reached is charged;
            // one of vmask that passes the row filter is descended into
            const unsigned rest = cmask & (0xFFu << (kcur + 1)) & 0xFFu;
            if (rest == 0) { state = S_RETURN; continue; }
            k = __ffs((int)rest) - 1;
            kcur = k;
            if (d < base + a.snap_ext) {
              // the sub-path (digits of levels base..d, then 0) of the next round's split that this
              // charge belongs to; when it is a new one, this charge is its first node and what was
              // counted before belongs to earlier sub-paths (a charge at a deeper level with digit
              // 0 stays in the sub-path its shallower charge opened)
              int j = 0;
              for (int t = base; t < base + a.snap_ext; t++)
                j = (j << 3) | (t < d ? uni(__builtin_amdgcn_readlane(st_k, t)) : (t == d ? k : 0));
              if (j != (int)(snap_c >> 44)) {
                snap_c = cur | ((long long)j << 44);
                snap_n = total;
              }
            }
            total++;
            dn += (lane == d) ? 1 : 0;
            if (limited && ++cur > bud) { state = S_ABORT; continue; }
            if (charged_out(cur)) { state = S_ABORT; continue; }
            if (a.item_cap > 0 && cur >= a.item_cap) { cap_stop = true; state = S_ABORT; continue; }
            if (!((vmask >> k) & 1u)) continue;
          }
          kcur = k;
          int a1, b1, a2, b2;
          quad_signs(d, k, a1, b1, a2, b2);
          const int i1 = d, i2 = L - 1 - d;
          const int c1 = i1 % m, c2 = i2 % m;
          const uint64_t nPA = PA + ((uint64_t)(int64_t)a1 << (8 * c1)) + ((uint64_t)(int64_t)a2 << (8 * c2));
          const uint64_t nPB = PB + ((uint64_t)(int64_t)b1 << (8 * c1)) + ((uint64_t)(int64_t)b2 << (8 * c2));
          if (rows) {
            // fh_abp_filter (wz_match.cpp:1055-1078): rows of S_d inside the capacity box of
            // the two classes this quad changed (the other classes' boxes are unchanged)
            const int r1 = a.rcap[d][c1], r2 = a.rcap[d][c2];
            const int p1a = class_field(nPA, c1), p1b = class_field(nPB, c1);
            const int p2a = class_field(nPA, c2), p2b = class_field(nPB, c2);
            const int nw = uni(__builtin_amdgcn_readlane(lvl_nw, d));
            const uint16_t *lst = wlist + (size_t)d * W;
            const uint64_t *bts = bits + (size_t)d * W;
            uint16_t *nlst = wlist + (size_t)(d + 1) * W;
            uint64_t *nbts = bits + (size_t)(d + 1) * W;
            int cnt = 0;
            for (int t0 = 0; t0 < nw; t0 += 4) {
              ulonglong2 r4[4];
              bool on[4];
              int w4[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                w4[u] = t0 + u < nw ? uni(lst[t0 + u]) : 0;
                on[u] = t0 + u < nw && ((bts[w4[u]] >> lane) & 1);
                if (on[u]) r4[u] = rows2[roff + 64 * w4[u] + lane];
              }
#pragma unroll
              for (int u = 0; u < 4; u++) {
                bool ok = false;
                if (on[u]) {
                  const uint64_t kw = r4[u].x, rw = r4[u].y;
                  const int k1 = (int)(int8_t)((kw >> (8 * c1)) & 255), k2 = (int)(int8_t)((kw >> (8 * c2)) & 255);
                  const int q1 = (int)(int8_t)((rw >> (8 * c1)) & 255), q2 = (int)(int8_t)((rw >> (8 * c2)) & 255);
                  ok = abs(k1 - p1a) <= r1 && abs(q1 - p1b) <= r1 && abs(k2 - p2a) <= r2 && abs(q2 - p2b) <= r2;
                  if (ok && a.a2) ok = a2_row_ok(a, kw, rw, nPA, nPB, d + 1);
                }
                const uint64_t nb = __ballot(ok);
                if (t0 + u < nw) {
                  // every lane stores the same uniform values
                  nbts[w4[u]] = nb;
                  if (nb) { nlst[cnt] = (uint16_t)w4[u]; cnt++; }
                }
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            cnt = uni(cnt);
            lvl_nw = put_lane(cnt, d + 1, lvl_nw);
            if (cnt == 0) {
              if (d < base) state = S_EXHAUSTED;
              continue;
            }
          }
          // ---- descend (the recursion at wz_match.cpp:1281-1287)
          int na_tied = a_tied, na_cmp = a_cmp, nb_tied = b_tied, nb_cmp = b_cmp;
          if (canon) {
            if (d == 0) { na_cmp = a2; nb_cmp = b2; }
            else {
              if (a_tied && a1 != a_cmp * a2) na_tied = 0;
              if (b_tied && b1 != b_cmp * b2) nb_tied = 0;
            }
          }
          if (!pp_ok) pp = partners(A, B, d, L, s, valid);
          const int delta = quad_delta(pp, a1, b1, a2, b2);
          st_k = put_lane(k, d, st_k);
          st_mask = put_lane((int)(cmask | (vmask << 8)), d, st_mask);
          st_ties = put_lane(a_tied | ((a_cmp + 1) << 1) | (b_tied << 3) | ((b_cmp + 1) << 4), d, st_ties);
          st_sums = put_lane((sumA + 128) | ((sumB + 128) << 8), d, st_sums);
          s_dab[d][lane] = (int16_t)dab;
          A |= ((a1 < 0) ? (1ull << i1) : 0) | ((a2 < 0) ? (1ull << i2) : 0);
          B |= ((b1 < 0) ? (1ull << i1) : 0) | ((b2 < 0) ? (1ull << i2) : 0);
          PA = nPA; PB = nPB;
          sumA += a1 + a2; sumB += b1 + b2;
          dab += delta;
          a_tied = na_tied; a_cmp = na_cmp; b_tied = nb_tied; b_cmp = nb_cmp;
          d++;
          pp_ok = false;
          state = S_ENTER;
          continue;
        }
        // state == S_RETURN: this level is exhausted; resume the parent's quad loop (a split
        // item ends with its subtree)
        if (d == base) { state = S_EXHAUSTED; continue; }
        d--;
        kcur = uni(__builtin_amdgcn_readlane(st_k, d));
        const int mk = uni(__builtin_amdgcn_readlane(st_mask, d));
        cmask = (unsigned)mk & 0xFFu;
        vmask = ((unsigned)mk >> 8) & 0xFFu;
        const int ties = uni(__builtin_amdgcn_readlane(st_ties, d));
        a_tied = ties & 1; a_cmp = ((ties >> 1) & 3) - 1; b_tied = (ties >> 3) & 1; b_cmp = ((ties >> 4) & 3) - 1;
        const int sums = uni(__builtin_amdgcn_readlane(st_sums, d));
        sumA = (sums & 255) - 128; sumB = ((sums >> 8) & 255) - 128;
        dab = s_dab[d][lane];
        int a1, b1, a2, b2;
        quad_signs(d, kcur, a1, b1, a2, b2);
        const int i1 = d, i2 = L - 1 - d, c1 = i1 % m, c2 = i2 % m;
        A &= ~((1ull << i1) | (1ull << i2));
        B &= ~((1ull << i1) | (1ull << i2));
        PA -= ((uint64_t)(int64_t)a1 << (8 * c1)) + ((uint64_t)(int64_t)a2 << (8 * c2));
        PB -= ((uint64_t)(int64_t)b1 << (8 * c1)) + ((uint64_t)(int64_t)b2 << (8 * c2));
        pp_ok = false;
        state = S_NEXT;
      }

      if (state == S_FOUND) {
        // exact NPAF recheck (wz_match.cpp:1329-1335): AB[s] from the packed words
        int ab = 0;
        if (valid) {
          const uint64_t mk = (1ull << (L - s)) - 1;
          ab = 2 * (L - s) - 2 * (__popcll((A ^ (A >> s)) & mk) + __popcll((B ^ (B >> s)) & mk));
        }
        verdict = any_lane(valid && ab != T) ? 4 : 0;
      } else if (state == S_ABORT) {
        // unwinding charges (wz_match.cpp:1245-1247, 1282-1291): each ancestor level whose
        // loop still has a quad passing the pre-charge filters charges one more node (the
        // path's replayed levels included). In a split launch the item only reports that it
        // stopped: the abort point is then located and rerun as one item.
        for (int l = d - 1; l >= 0 && ext == 0; l--) {
          const int kl = uni(__builtin_amdgcn_readlane(st_k, l));
          const unsigned cm = (unsigned)uni(__builtin_amdgcn_readlane(st_mask, l)) & 0xFFu;
          if (cm & (0xFFu << (kl + 1)) & 0xFFu) { total++; dn += (lane == l) ? 1 : 0; }
        }
        verdict = cap_stop ? 5 : 3;
        A = cap_stop ? (uint64_t)snap_c : 0;
        B = cap_stop ? (uint64_t)snap_n : 0;
      } else {
        verdict = 2;
        A = B = 0;
      }
    } else {
      verdict = 1;
    }
    // every lane stores the same uniform values: no lane-0-only block inside the loop
    a.verdict[u] = verdict;
    a.nodes[u] = total;
    a.outA[u] = A;
    a.outB[u] = B;
    a.charges[u] = cur;
    if (a.depth_out) a.depth_out[u * 64 + lane] = dn;
  }
}

}  // namespace wzfh

namespace wzfh {
// Per-entry result of a split launch, one wavefront per entry: the items of an entry, in
// path order, are its DFS in order, so running sums of their counted and charged nodes give
// the reference's counts at every item boundary. The wave takes the items 64 at a time
// (lane = item), forms prefix sums of charges and nodes with shuffles, and the first lane
// with an event decides: pre-reject (1), a hit (0/4: nodes up to and including the item),
// the abort point (-1: the first item that stopped on its own budget or whose charges pass
// the entry's budget; reported with the charges and nodes before it and its own charges), or
// an item stopped at the launch's charge cap before any of those (-2: the host scans the
// entry's items and resolves the capped ones). A hit reports the nodes and charges through
// it. Every branch is on a ballot, so the loop exits uniformly.
__global__ __launch_bounds__(64) void complete_reduce_kernel(CompleteArgs a, int nent, int *fv, long long *fn,
                                                             uint64_t *fA, uint64_t *fB, int *fj, long long *fcc,
                                                             long long *fcur) {
  const int e = blockIdx.x;
  if (e >= nent) return;
  const int lane = threadIdx.x & 63;
  const long long np = 1ll << a.items_log2;
  const bool limited = a.budget > 0;
  const long long bud = a.ent_budget ? a.ent_budget[e] : a.budget;
  long long cum_n = 0, cum_c = 0;   // before the current chunk (uniform)
  int v_out = 2, j_out = -1;
  long long n_out = 0, c_out = 0, cj_out = 0;
  uint64_t A = 0, B = 0;
  for (long long p0 = 0; p0 < np; p0 += 64) {
    const long long p = p0 + lane;
    const bool in = p < np;
    const long long u = (long long)e * np + p;
    const int v = in ? a.verdict[u] : 2;
    const long long c = in ? a.charges[u] : 0, nd = in ? a.nodes[u] : 0;
    // inclusive prefix sums over the chunk
    long long pc = c, pn = nd;
    for (int o = 1; o < 64; o <<= 1) {
      const long long tc = __shfl_up(pc, o), tn = __shfl_up(pn, o);
      if (lane >= o) { pc += tc; pn += tn; }
    }
    const long long c_before = cum_c + pc - c, n_before = cum_n + pn - nd;
    const bool ev = in && (v == 1 || v == 3 || v == 5 || v == 0 || v == 4 || (limited && c_before + c > bud));
    const unsigned long long m = __ballot(ev);
    if (m) {
      const int f = __ffsll((long long)m) - 1;
      const int vf = __shfl(v, f);
      const long long cb = __shfl(c_before, f), nb = __shfl(n_before, f), cf = __shfl(c, f), nf = __shfl(nd, f);
      if (vf == 1) { v_out = 1; n_out = 0; c_out = cb; }
      else if (vf == 3 || (limited && cb + cf > bud)) { v_out = -1; j_out = (int)(p0 + f); n_out = nb; c_out = cb; cj_out = cf; }
      else if (vf == 5) { v_out = -2; n_out = nb; c_out = cb; }   // capped before a decision: split deeper
      else {   // a hit: nodes and charges through the hit
        v_out = vf; n_out = nb + nf; c_out = cb + cf;
        A = __shfl(in ? a.outA[u] : 0ull, f);
        B = __shfl(in ? a.outB[u] : 0ull, f);
      }
      break;
    }
    cum_c += __shfl(pc, 63);
    cum_n += __shfl(pn, 63);
    n_out = cum_n;
    c_out = cum_c;
  }
  if (lane == 0) {
    fv[e] = v_out;
    fn[e] = n_out;
    fA[e] = A;
    fB[e] = B;
    fj[e] = j_out;
    fcc[e] = c_out;
    fcur[e] = cj_out;
  }
}

hipError_t launch_complete_reduce(const CompleteArgs &a, int nent, int *fv, long long *fn, uint64_t *fA,
                                  uint64_t *fB, int *fj, long long *fcc, long long *fcur, hipStream_t st) {
  hipLaunchKernelGGL(complete_reduce_kernel, dim3(nent), dim3(64), 0, st, a, nent, fv, fn, fA, fB, fj, fcc, fcur);
  return hipGetLastError();
}

// Bytes of dynamic LDS one wave needs for row lists of `words` 64-row words.
size_t complete_lds_bytes(int L, int words) {
  const size_t nlw = (size_t)(L / 2 + 1) * (size_t)words;
  return nlw * sizeof(uint64_t) + ((nlw * sizeof(uint16_t) + 7) / 8) * 8;
}

hipError_t launch_complete(const CompleteArgs &a, const uint8_t *kab, int blocks, bool lds, size_t lds_bytes,
                           hipStream_t st) {
  if (lds) {
    hipLaunchKernelGGL(complete_kernel<true>, dim3(blocks), dim3(64), lds_bytes, st, a, kab);
  } else {
    hipLaunchKernelGGL(complete_kernel<false>, dim3(blocks), dim3(64), 0, st, a, kab);
  }
  return hipGetLastError();
}

// Zero-fills n bytes (n a multiple of 8) with vector stores.
__global__ void zero_kernel(uint64_t *p, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t st) {
  const size_t nw = (bytes + 7) / 8;
  const int blocks = (int)std::min<size_t>(1024, std::max<size_t>(1, (nw + 255) / 256));
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, st, (uint64_t *)p, nw);
  return hipGetLastError();
}

hipError_t complete_set_lds_limit(size_t bytes) {
  return hipFuncSetAttribute((const void *)complete_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}
}  // namespace wzfh
