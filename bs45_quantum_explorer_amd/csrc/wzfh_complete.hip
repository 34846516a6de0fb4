// wzfh_complete.hip -- A,B completer kernel for gfx950.
//
// Restates fh_complete_ab / fh_ab_search (src/solver/wz_match.cpp:1092-1341) with
// one candidate per wavefront. The mirror-quad DFS control is wave-uniform (scalar
// registers); lane s holds the shift-s state: the target T[s] = -CD[s], the partial
// A,B autocorrelation Dab[s], and reads the remaining-pair bound Kab from a table.
// A, B are packed 64-bit words (bit i set = element i is -1). A quad placement costs
// a handful of VALU ops per lane, the bound scan is one compare and a ballot, and the
// profile-row filter tests 64 rows per instruction against bitsets kept per depth.
//
// Semantics follow the reference exactly, including the order of the per-quad steps
// (A1 mask, root canon, reversal canon, charge + budget, early outer check, place,
// sum capacity, bound scan, row filter), the odd-L middle trials that count nodes
// without charging the budget, and the extra node the reference charges at each
// ancestor level while unwinding a budget abort.
#include <hip/hip_runtime.h>

#include "wzfh_kernels.h"

namespace wzfh {

namespace {

constexpr int kCW = 4;           // waves per block
constexpr int kLevels = 32;

__constant__ int8_t cNEG[8][4] = {{-1, -1, -1, 1}, {-1, -1, 1, -1}, {-1, 1, -1, -1}, {-1, 1, 1, 1},
                                  {1, -1, -1, -1}, {1, -1, 1, 1},   {1, 1, -1, 1},   {1, 1, 1, -1}};
__constant__ int8_t cPOS[8][4] = {{-1, -1, -1, -1}, {-1, -1, 1, 1}, {-1, 1, -1, 1}, {-1, 1, 1, -1},
                                  {1, -1, -1, 1},   {1, -1, 1, -1}, {1, 1, -1, -1}, {1, 1, 1, 1}};
__constant__ int8_t cMID[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};

__device__ __forceinline__ int sgn(uint64_t w, int q) { return 1 - 2 * (int)((w >> q) & 1u); }

// Makes a wave-uniform value visible as such to the compiler (an SGPR), so every branch
// of the DFS is a scalar branch and no loop exit is ever lowered to an exec mask.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ bool any_lane(bool p) { return uni(__any(p) ? 1 : 0) != 0; }

// Writes a wave-uniform value into lane `at` of a per-lane record (lane = DFS level).
__device__ __forceinline__ int put_lane(int val, int at, int old) {
  return ((int)(threadIdx.x & 63) == at) ? val : old;
}

// Pre-charge filters of quad k at a level: A1 mask, root canon, reversal canon.
// Returns false when the reference `continue`s before charging the node.
__device__ __forceinline__ bool precharge_ok(int d, int k, unsigned omask, bool canon, int a_tied, int a_cmp,
                                             int b_tied, int b_cmp) {
  if (d > 0 && !((omask >> k) & 1u)) return false;
  const int8_t *S = (d == 0) ? cNEG[k] : cPOS[k];
  if (d == 0) return !canon || (S[0] == 1 && S[1] == 1);
  if (canon) {
    if (a_tied && S[0] != a_cmp * S[2] && S[0] != 1) return false;
    if (b_tied && S[1] != b_cmp * S[3] && S[1] != 1) return false;
  }
  return true;
}

}  // namespace

__global__ __launch_bounds__(256) void complete_kernel(CompleteArgs a, const uint8_t *kab_table) {
  __shared__ int16_t s_dab[kCW][kLevels][64];
  __shared__ uint8_t s_kab[kLevels][64];
  for (int i = threadIdx.x; i < kLevels * 64; i += blockDim.x) s_kab[i / 64][i % 64] = kab_table[i];
  __syncthreads();

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = a.n, L = a.L, m = a.m, half = L / 2;
  const bool canon = a.canon != 0;
  const bool valid = lane >= 1 && lane <= n;
  const int s = lane;
  const uint64_t gwave = (uint64_t)blockIdx.x * kCW + wv;
  volatile uint64_t *sbits = a.scratch + gwave * (uint64_t)(kLevels * a.scratch_words * 2);
  volatile uint64_t *slist = sbits + (uint64_t)kLevels * a.scratch_words;   // u16 lists packed in u64 storage
  volatile uint16_t *list16 = (volatile uint16_t *)slist;
  auto class_field = [](uint64_t w, int c) { return (int)((w >> (8 * c)) & 255u) - 64; };

  // DFS states of one level (the iterative form of fh_ab_search's recursion)
  enum { S_ENTER = 0, S_NEXT = 1, S_RETURN = 2, S_FOUND = 3, S_ABORT = 4, S_EXHAUSTED = 5 };

  while (true) {
    // next candidate: the first active lane takes the ticket, readfirstlane broadcasts it
    const unsigned long long act = __ballot(1);
    int ci_l = 0;
    if (lane == __ffsll((long long)act) - 1) ci_l = (int)atomicAdd(a.work_ctr, 1ull);
    const int ci = uni(ci_l);   // the first active lane is the one that took the ticket
    if (ci >= a.ncand) break;

    const uint64_t C = a.cd[2 * ci], D = a.cd[2 * ci + 1];
    // T[s] = -CD[s] (wz_match.cpp:1300-1304) from packed words: x_i x_{i+s} = 1 - 2 (b_i ^ b_{i+s}).
    int T = 0;
    if (s >= 1 && s < n) {
      const uint64_t mk = (1ull << (n - s)) - 1;
      const int dis = __popcll((C ^ (C >> s)) & mk) + __popcll((D ^ (D >> s)) & mk);
      T = -(2 * (n - s) - 2 * dis);
    }
    int verdict = 2;
    long long total = 0;
    uint64_t A = 0, B = 0;
    // pre-filter (wz_match.cpp:1305-1308): |T[s]| > 2 (L - s) for some s < L
    const bool prerej = any_lane(valid && (T > 2 * (L - s) || -T > 2 * (L - s)));
    if (!prerej) {
      const int roff = uni(a.row_off[ci]);
      const int rlen = uni(a.row_len[ci]);
      const bool rows = roff >= 0;
      const int nwords = rows ? (rlen + 63) / 64 : 0;
      int lvl_nw = 0;            // lane l: nonzero-word count of S_l
      if (rows) {
        for (int w = lane; w < nwords; w += 64) {
          const int left = rlen - 64 * w;
          sbits[w] = left >= 64 ? ~0ull : ((1ull << left) - 1);
          list16[w] = (uint16_t)w;
        }
        lvl_nw = (lane == 0) ? nwords : 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      // uniform DFS state
      int d = 0, k = -1, state = S_ENTER;
      int sumA = 0, sumB = 0;
      uint64_t PA = 0x4040404040404040ull, PB = PA;   // 8-bit class fields: class sum + 64
      long long cur = 0;
      int dab = 0;                // lane s: Dab[s] of the current state
      int st_k = 0, st_omask = 0, st_ties = 0, st_sums = 0;   // lane l = level l records
      unsigned omask = 0xFFu;
      int a_tied = 1, a_cmp = 0, b_tied = 1, b_cmp = 0;

      while (state < S_FOUND) {
        if (state == S_ENTER) {
          if (d == half) {
            // ---- leaf (wz_match.cpp:1178-1206)
            bool ok_leaf = false;
            if (L & 1) {
              const int mid = half, cm = mid % m;
              for (int t = 0; t < 4 && !ok_leaf; t++) {
                const int av = cMID[t][0], bv = cMID[t][1];
                int delta = 0;
                if (valid) {
                  if (mid - s >= 0) delta += sgn(A, mid - s) * av + sgn(B, mid - s) * bv;
                  if (mid + s < L) delta += sgn(A, mid + s) * av + sgn(B, mid + s) * bv;
                }
                total++;
                bool ok = (abs(sumA + av) == a.absA) && (abs(sumB + bv) == a.absB);
                if (ok) ok = !any_lane(valid && (dab + delta != T));
                if (ok && rows) {
                  const int wantk = class_field(PA, cm) + av, wantr = class_field(PB, cm) + bv;
                  const int nw = uni(__builtin_amdgcn_readlane(lvl_nw, half));
                  bool hitany = false;
                  for (int t2 = 0; t2 < nw && !hitany; t2++) {
                    const int w = uni(list16[half * a.scratch_words * 4 + t2]);
                    const uint64_t bits = sbits[half * a.scratch_words + w];
                    bool hit = false;
                    if ((bits >> lane) & 1) {
                      const uint64_t kw = a.rows[2 * (roff + 64 * w + lane)], rw = a.rows[2 * (roff + 64 * w + lane) + 1];
                      const int kk = (int)(int8_t)((kw >> (8 * cm)) & 255), rr = (int)(int8_t)((rw >> (8 * cm)) & 255);
                      hit = (kk == wantk) && (rr == wantr);
                    }
                    hitany = any_lane(hit);
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
          // ---- level entry: A1 outer mask (wz_match.cpp:601-618, 1210-1211)
          omask = 0xFFu;
          if (a.outer_k && d > 0) {
            const int kk = a.outer_k;
            if (!(kk < 2 || d < kk || d + kk > L / 2 || (A & 1) || (B & 1))) {
              int pat = ((A >> (L - 1)) & 1) ? 0 : 1, mul = 2;
              bool skip = false;
              for (int t = 1; t < kk; t++) {
                const int sb = ((((A >> t) & 1) ? 0 : 1) << 3) | ((((B >> t) & 1) ? 0 : 1) << 2) |
                               ((((A >> (L - 1 - t)) & 1) ? 0 : 1) << 1) | (((B >> (L - 1 - t)) & 1) ? 0 : 1);
                const int q = a.pos_index[sb];
                if (q < 0) { skip = true; break; }
                pat += mul * q;
                mul *= 8;
              }
              if (!skip) {
                int R[3] = {0, 0, 0};
                bool zero = false;
                for (int j = 0; j < kk; j++) {
                  const int sj = L - 1 - d - j;
                  R[j] = uni(__builtin_amdgcn_readlane(T - dab, sj));
                  if ((R[j] & 1) || R[j] < -4 * (j + 1) || R[j] > 4 * (j + 1)) zero = true;
                }
                if (zero) omask = 0;
                else if (kk == 2) omask = a.outer2[(pat * 5 + (R[0] / 2 + 2)) * 9 + (R[1] / 2 + 4)];
                else omask = a.outer3[((pat * 5 + (R[0] / 2 + 2)) * 9 + (R[1] / 2 + 4)) * 13 + (R[2] / 2 + 6)];
              }
            }
          }
          omask = (unsigned)uni((int)omask);
          if (omask == 0) { state = S_RETURN; continue; }
          k = -1;
          st_omask = put_lane((int)omask, d, st_omask);
          st_ties = put_lane(a_tied | ((a_cmp + 1) << 1) | (b_tied << 3) | ((b_cmp + 1) << 4), d, st_ties);
          st_sums = put_lane((sumA + 128) | ((sumB + 128) << 8), d, st_sums);
          state = S_NEXT;
          continue;
        }
        if (state == S_NEXT) {
          k++;
          if (k >= 8) { state = S_RETURN; continue; }
          const int8_t *S = (d == 0) ? cNEG[k] : cPOS[k];
          const int a1 = S[0], b1 = S[1], a2 = S[2], b2 = S[3];
          if (!precharge_ok(d, k, omask, canon, a_tied, a_cmp, b_tied, b_cmp)) continue;
          int na_tied = a_tied, na_cmp = a_cmp, nb_tied = b_tied, nb_cmp = b_cmp;
          if (canon) {
            if (d == 0) { na_cmp = a2; nb_cmp = b2; }
            else {
              if (a_tied && a1 != a_cmp * a2) na_tied = 0;
              if (b_tied && b1 != b_cmp * b2) nb_tied = 0;
            }
          }
          total++;
          if (a.budget > 0 && ++cur > a.budget) { state = S_ABORT; continue; }
          const int i1 = d, i2 = L - 1 - d;
          if (d > 0) {   // early outer-correlation check (wz_match.cpp:1256-1259)
            const int se = L - 1 - d;
            const int lhs = sgn(A, 0) * a2 + sgn(B, 0) * b2 + sgn(A, L - 1) * a1 + sgn(B, L - 1) * b1;
            if (lhs != uni(__builtin_amdgcn_readlane(T - dab, se))) continue;
          }
          // place i1 then i2 (fh_place): new pairs of each position with placed ones
          const uint64_t nA = A | ((a1 < 0) ? (1ull << i1) : 0) | ((a2 < 0) ? (1ull << i2) : 0);
          const uint64_t nB = B | ((b1 < 0) ? (1ull << i1) : 0) | ((b2 < 0) ? (1ull << i2) : 0);
          int delta = 0;
          if (valid) {
            int q = i1 - s;
            if (q >= 0) delta += sgn(nA, q) * a1 + sgn(nB, q) * b1;
            q = i1 + s;
            if (q >= L - d && q < L) delta += sgn(nA, q) * a1 + sgn(nB, q) * b1;
            q = i2 - s;
            if (q >= 0 && q <= d) delta += sgn(nA, q) * a2 + sgn(nB, q) * b2;
            q = i2 + s;
            if (q >= L - d && q < L) delta += sgn(nA, q) * a2 + sgn(nB, q) * b2;
          }
          const int ndab = dab + delta;
          const int nsA = sumA + a1 + a2, nsB = sumB + b1 + b2, rem = L - 2 * (d + 1);
          bool prune = !(((abs(a.absA - nsA) <= rem) || (abs(-a.absA - nsA) <= rem)) &&
                         ((abs(a.absB - nsB) <= rem) || (abs(-a.absB - nsB) <= rem)));
          if (!prune) {
            const int kab = s_kab[d + 1][lane];
            prune = any_lane(valid && abs(T - ndab) > kab);
          }
          const int c1 = i1 % m, c2 = i2 % m;
          const uint64_t nPA = PA + ((uint64_t)(int64_t)a1 << (8 * c1)) + ((uint64_t)(int64_t)a2 << (8 * c2));
          const uint64_t nPB = PB + ((uint64_t)(int64_t)b1 << (8 * c1)) + ((uint64_t)(int64_t)b2 << (8 * c2));
          if (!prune && rows) {
            // fh_abp_filter (wz_match.cpp:1055-1078): rows of S_d inside the capacity box of
            // the two classes this quad changed (the other classes' boxes are unchanged).
            int r1 = a.tic[c1], r2 = a.tic[c2];
            for (int i = c1; i < L; i += m) if (i <= d || i >= L - 1 - d) r1--;
            for (int i = c2; i < L; i += m) if (i <= d || i >= L - 1 - d) r2--;
            const int p1a = class_field(nPA, c1), p1b = class_field(nPB, c1);
            const int p2a = class_field(nPA, c2), p2b = class_field(nPB, c2);
            const int nw = uni(__builtin_amdgcn_readlane(lvl_nw, d));
            int cnt = 0;
            for (int t = 0; t < nw; t++) {
              const int w = uni(list16[d * a.scratch_words * 4 + t]);
              const uint64_t bits = sbits[d * a.scratch_words + w];
              bool ok = false;
              if ((bits >> lane) & 1) {
                const uint64_t kw = a.rows[2 * (roff + 64 * w + lane)], rw = a.rows[2 * (roff + 64 * w + lane) + 1];
                const int k1 = (int)(int8_t)((kw >> (8 * c1)) & 255), k2 = (int)(int8_t)((kw >> (8 * c2)) & 255);
                const int q1 = (int)(int8_t)((rw >> (8 * c1)) & 255), q2 = (int)(int8_t)((rw >> (8 * c2)) & 255);
                ok = abs(k1 - p1a) <= r1 && abs(q1 - p1b) <= r1 && abs(k2 - p2a) <= r2 && abs(q2 - p2b) <= r2;
              }
              const uint64_t nb = __ballot(ok);
              if (lane == 0) {
                sbits[(d + 1) * a.scratch_words + w] = nb;
                if (nb) list16[(d + 1) * a.scratch_words * 4 + cnt] = (uint16_t)w;
              }
              cnt += (nb != 0) ? 1 : 0;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            cnt = uni(cnt);
            lvl_nw = put_lane(cnt, d + 1, lvl_nw);
            if (cnt == 0) prune = true;
          }
          if (prune) continue;
          // ---- descend (the recursion at wz_match.cpp:1281-1287)
          st_k = put_lane(k, d, st_k);
          s_dab[wv][d][lane] = (int16_t)dab;
          A = nA; B = nB; PA = nPA; PB = nPB;
          sumA = nsA; sumB = nsB;
          dab = ndab;
          a_tied = na_tied; a_cmp = na_cmp; b_tied = nb_tied; b_cmp = nb_cmp;
          d++;
          state = S_ENTER;
          continue;
        }
        // state == S_RETURN: this level is exhausted; resume the parent's quad loop
        if (d == 0) { state = S_EXHAUSTED; continue; }
        d--;
        k = uni(__builtin_amdgcn_readlane(st_k, d));
        omask = (unsigned)uni(__builtin_amdgcn_readlane(st_omask, d));
        const int ties = uni(__builtin_amdgcn_readlane(st_ties, d));
        a_tied = ties & 1; a_cmp = ((ties >> 1) & 3) - 1; b_tied = (ties >> 3) & 1; b_cmp = ((ties >> 4) & 3) - 1;
        const int sums = uni(__builtin_amdgcn_readlane(st_sums, d));
        sumA = (sums & 255) - 128; sumB = ((sums >> 8) & 255) - 128;
        dab = s_dab[wv][d][lane];
        const int8_t *S = (d == 0) ? cNEG[k] : cPOS[k];
        const int i1 = d, i2 = L - 1 - d, c1 = i1 % m, c2 = i2 % m;
        A &= ~((1ull << i1) | (1ull << i2));
        B &= ~((1ull << i1) | (1ull << i2));
        PA -= ((uint64_t)(int64_t)S[0] << (8 * c1)) + ((uint64_t)(int64_t)S[2] << (8 * c2));
        PB -= ((uint64_t)(int64_t)S[1] << (8 * c1)) + ((uint64_t)(int64_t)S[3] << (8 * c2));
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
        // loop still has a quad passing the pre-charge filters charges one more node
        for (int l = d - 1; l >= 0; l--) {
          const int kl = uni(__builtin_amdgcn_readlane(st_k, l));
          const unsigned om = (unsigned)uni(__builtin_amdgcn_readlane(st_omask, l));
          const int ties = uni(__builtin_amdgcn_readlane(st_ties, l));
          const int at = ties & 1, ac = ((ties >> 1) & 3) - 1, bt = (ties >> 3) & 1, bc = ((ties >> 4) & 3) - 1;
          for (int k2 = kl + 1; k2 < 8; k2++)
            if (precharge_ok(l, k2, om, canon, at, ac, bt, bc)) { total++; break; }
        }
        verdict = 3;
        A = B = 0;
      } else {
        verdict = 2;
        A = B = 0;
      }
    } else {
      verdict = 1;
    }
    // every lane stores the same uniform values: no lane-0-only block inside the loop
    a.verdict[ci] = verdict;
    a.nodes[ci] = total;
    a.outA[ci] = A;
    a.outB[ci] = B;
  }
}

}  // namespace wzfh

namespace wzfh {
hipError_t launch_complete(const CompleteArgs &a, const uint8_t *kab, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(complete_kernel, dim3(blocks), dim3(256), 0, st, a, kab);
  return hipGetLastError();
}
}  // namespace wzfh
