/*
 * wz_oracle.cpp -- CPU oracle (test infrastructure only; see wz_oracle.h).
 *
 * A scalar restatement of the reference's WZ_FIRSTHIT hot path. It keeps the
 * reference's arithmetic exactly (int class sums, FP64 hall_ok in the same
 * summation order without contraction, the same DFS orders and node charging),
 * because the HIP library is checked against it bit for bit.
 */
#include "wz_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

using Vec = std::vector<int>;

// ---- Thm 2.2 quad tables: wz_match.cpp:531-545. Quad index i in 0..15 is the
// sign pattern (i&8, i&4, i&2, i&1) of (X[d], Y[d], X[L-1-d], Y[L-1-d]), MSB first;
// POS / NEG keep ascending i with product +1 / -1. The order fixes the DFS order.
struct QuadTables {
  int all[16][4], pos[8][4], neg[8][4], mid[4][2];
  QuadTables() {
    int np = 0, nn = 0;
    for (int i = 0; i < 16; i++) {
      int v[4] = {(i & 8) ? 1 : -1, (i & 4) ? 1 : -1, (i & 2) ? 1 : -1, (i & 1) ? 1 : -1};
      std::memcpy(all[i], v, sizeof v);
      if (v[0] * v[1] * v[2] * v[3] == 1) std::memcpy(pos[np++], v, sizeof v);
      else std::memcpy(neg[nn++], v, sizeof v);
    }
    const int m4[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    std::memcpy(mid, m4, sizeof m4);
  }
};
const QuadTables QT;

// ---- A1 outer-lag tables: wz_match.cpp:547-598 (built here from the definition:
// F_j = sum_{r<=j} A[j-r]*aR_r + B[j-r]*bR_r + A[L-1-j+r]*aL_r + B[L-1-j+r]*bL_r).
struct OuterTables {
  uint8_t t2[16][5][9];
  uint8_t t3[128][5][9][13];
  int8_t pos_index[16];
  int root_index[2];
  static int bits(int a, int b, int c, int d) { return ((a > 0) << 3) | ((b > 0) << 2) | ((c > 0) << 1) | (d > 0); }
  OuterTables() {
    for (int i = 0; i < 16; i++) pos_index[i] = -1;
    for (int q = 0; q < 8; q++) pos_index[bits(QT.pos[q][0], QT.pos[q][1], QT.pos[q][2], QT.pos[q][3])] = (int8_t)q;
    for (int q = 0; q < 8; q++)
      if (QT.neg[q][0] == 1 && QT.neg[q][1] == 1) root_index[QT.neg[q][2] > 0] = q;
    std::memset(t2, 0, sizeof t2);
    std::memset(t3, 0, sizeof t3);
    for (int k = 2; k <= 3; k++) {
      int npat = 2 * (k == 2 ? 8 : 64);
      for (int pat = 0; pat < npat; pat++) {
        int lA[3], lB[3], rA[3], rB[3];   // boundary signs at depths 0..k-1 (left / right)
        const int *root = QT.neg[root_index[pat & 1]];
        lA[0] = root[0]; lB[0] = root[1]; rA[0] = root[2]; rB[0] = root[3];
        int rest = pat >> 1;
        for (int t = 1; t < k; t++) {
          const int *q = QT.pos[rest % 8];
          rest /= 8;
          lA[t] = q[0]; lB[t] = q[1]; rA[t] = q[2]; rB[t] = q[3];
        }
        int total = 1;
        for (int r = 0; r < k; r++) total *= 8;
        for (int asg = 0; asg < total; asg++) {
          int pick[3], a = asg;
          for (int r = 0; r < k; r++) { pick[r] = a % 8; a /= 8; }
          int F[3] = {0, 0, 0};
          for (int j = 0; j < k; j++)
            for (int r = 0; r <= j; r++) {
              const int *v = QT.pos[pick[r]];
              F[j] += lA[j - r] * v[2] + lB[j - r] * v[3] + rA[j - r] * v[0] + rB[j - r] * v[1];
            }
          if (k == 2) t2[pat][F[0] / 2 + 2][F[1] / 2 + 4] |= (uint8_t)(1u << pick[0]);
          else t3[pat][F[0] / 2 + 2][F[1] / 2 + 4][F[2] / 2 + 6] |= (uint8_t)(1u << pick[0]);
        }
      }
    }
  }
};
const OuterTables OT;

int class_count(int L, int c, int m) { return c < L ? (L - c + m - 1) / m : 0; }

// wz_match.cpp:203-224 -- class-sum tuples in lexicographic (ascending) order.
std::vector<Vec> class_sum_tuples(int L, int target, int m) {
  int cnt[8];
  for (int c = 0; c < m; c++) cnt[c] = class_count(L, c, m);
  std::vector<Vec> out;
  Vec cur(m, 0);
  struct Rec {
    static void go(int c, int acc, int m, int target, const int *cnt, Vec &cur, std::vector<Vec> &out) {
      if (c == m - 1) {
        int last = target - acc;
        if (last < -cnt[c] || last > cnt[c] || ((last + cnt[c]) & 1)) return;
        cur[c] = last;
        out.push_back(cur);
        return;
      }
      for (int v = -cnt[c]; v <= cnt[c]; v += 2) {
        cur[c] = v;
        go(c + 1, acc + v, m, target, cnt, cur, out);
      }
    }
  };
  Rec::go(0, 0, m, target, cnt, cur, out);
  return out;
}

int sqnorm(const Vec &v) { int s = 0; for (int x : v) s += x * x; return s; }
int acorr(const Vec &v, int s) {               // wz_match.cpp:268-273
  int m = (int)v.size(), t = 0;
  for (int i = 0; i + s < m; i++) t += v[i] * v[i + s];
  return t;
}
int pair_auto(const Vec &x, const Vec &y, int s) {   // wz_match.cpp:277-281
  int m = (int)x.size();
  return acorr(x, s) + acorr(y, s) + acorr(x, m - s) + acorr(y, m - s);
}
uint64_t auto_key(const Vec &T, int norm) {          // wz_match.cpp:284-289
  uint64_t k = (uint64_t)(uint16_t)norm;
  for (size_t i = 0; i < T.size() && i < 3; i++) k |= ((uint64_t)(uint16_t)(T[i] + 8192)) << (16 * (i + 1));
  return k;
}
// wz_match.cpp:302-322 -- Thm 2.3 eq (18) mod-4 conditions.
bool thm212_ok(const Vec &x, const Vec &y, int m, bool abSide, int n) {
  auto m4 = [](int v) { return ((v % 4) + 4) % 4; };
  if (abSide) {
    int c1 = n % m, want = (n % m) ? 2 : 0;
    if (m4(x[0] + y[0] + x[c1] + y[c1]) != want) return false;
    for (int j = 2; j <= m; j++) {
      if (j == c1 + 1) continue;
      int c = ((n + 1 - j) % m + m) % m;
      if (m4(x[j - 1] + y[j - 1] + x[c] + y[c]) != 0) return false;
    }
  } else {
    for (int j = 1; j <= m; j++) {
      int c = ((n - j) % m + m) % m;
      if (m4(x[j - 1] + y[j - 1] + x[c] + y[c]) != 0) return false;
    }
  }
  return true;
}

// wz_match.cpp:233-249 and 324-350.
struct NormSet {
  std::unordered_set<int> s;
  void build(int L, int a, int b, int m, int n) {
    auto K = class_sum_tuples(L, a, m), R = class_sum_tuples(L, b, m);
    int tgt = 4 * n + 2;
    for (auto &k : K) {
      int kn = sqnorm(k);
      if (kn > tgt) continue;
      for (auto &r : R) { int rn = sqnorm(r); if (kn + rn <= tgt) s.insert(kn + rn); }
    }
  }
  bool has(int v) const { return s.count(v) != 0; }
};
struct AutoSet {
  std::unordered_set<uint64_t> s;
  void build(int L, int a, int b, int m, int n, bool t212, bool abSide) {
    auto K = class_sum_tuples(L, a, m), R = class_sum_tuples(L, b, m);
    int tgt = 4 * n + 2, half = m / 2;
    Vec T(half);
    for (auto &k : K) {
      int kn = sqnorm(k);
      if (kn > tgt) continue;
      for (auto &r : R) {
        int rn = sqnorm(r);
        if (kn + rn > tgt) continue;
        if (t212 && !thm212_ok(k, r, m, abSide, n)) continue;
        for (int s = 1; s <= half; s++) T[s - 1] = pair_auto(k, r, s);
        this->s.insert(auto_key(T, kn + rn));
      }
    }
  }
  bool has(const Vec &T, int need) const { return s.count(auto_key(T, need)) != 0; }
};

Vec reduce63(const Vec &v) { return {v[0] + v[3], v[1] + v[4], v[2] + v[5]}; }

struct Cell { Vec px, py; };

struct Row { int8_t k[8], r[8]; };

struct Problem {
  wzo_config cfg;
  int n, L1, m;
  bool pinC, pinD;
  std::vector<Cell> cells;
  std::vector<char> kept;
  std::vector<int> rowkey;                      // cell -> index into rowlists (-1 = dead)
  std::vector<std::vector<Row>> rowlists;
  bool ab_prof;
  double hcos[201][64], hsin[201][64];
  // A2 (wz_match.cpp:991-1016): mirror blocks of residues and quads left per block and depth
  int a2_nblk = 0, a2_r[8] = {0}, a2_s[8] = {0}, a2_mid[8] = {0}, a2_cnt[40][8] = {{0}};
  mutable long long a2_tested = 0, a2_rejected = 0;
};

// init_a2_tables (wz_match.cpp:1001-1016): residue c and its reflection (L-1-c) mod m form a
// block; an odd L's middle residue is self-reflected and free until the leaf.
void init_a2(Problem &P) {
  int L = P.L1, m = P.m, half = L / 2, blk_of[8];
  for (int c = 0; c < 8; c++) blk_of[c] = -1;
  P.a2_nblk = 0;
  for (int c = 0; c < m; c++) {
    if (blk_of[c] >= 0) continue;
    int s = ((L - 1 - c) % m + m) % m;
    P.a2_r[P.a2_nblk] = c; P.a2_s[P.a2_nblk] = s; P.a2_mid[P.a2_nblk] = 0;
    blk_of[c] = blk_of[s] = P.a2_nblk++;
  }
  if (L % 2 == 1) P.a2_mid[blk_of[half % m]] = 1;
  for (int p = 0; p <= half + 1 && p < 40; p++) {
    for (int b = 0; b < 8; b++) P.a2_cnt[p][b] = 0;
    for (int i = p; i < half; i++) P.a2_cnt[p][blk_of[i % m]]++;
  }
}

std::string cell_key(const Vec &px, const Vec &py) {   // wz_match.cpp:1665-1671
  std::string k;
  for (int v : px) { k += std::to_string(v); k += ','; }
  k += '|';
  for (int v : py) { k += std::to_string(v); k += ','; }
  return k;
}

long long prof_score(const Cell &p) {
  long long s = 0;
  for (int v : p.px) s += std::abs(v);
  for (int v : p.py) s += std::abs(v);
  return s;
}

void build_cells(Problem &P) {
  const wzo_config &c = P.cfg;
  int n = P.n, L1 = P.L1, tgt = 4 * n + 2;
  NormSet ab3, ab6;
  ab3.build(L1, c.a, c.b, 3, n);
  ab6.build(L1, c.a, c.b, 6, n);
  AutoSet ab6a;
  if (c.thm211b) ab6a.build(L1, c.a, c.b, 6, n, c.thm212, true);
  std::vector<Cell> cells;
  if (c.m6) {
    // survive_profiles6: wz_match.cpp:792-837 (C,D side: L = n, sigs c,d)
    auto PX = class_sum_tuples(n, c.c, 6), PY = class_sum_tuples(n, c.d, 6);
    for (auto &px : PX) {
      int npx = sqnorm(px);
      if (npx > tgt) continue;
      Vec px3 = reduce63(px);
      int npx3 = sqnorm(px3);
      if (npx3 > tgt) continue;
      for (auto &py : PY) {
        int need = tgt - npx - sqnorm(py);
        if (need < 0 || !ab6.has(need)) continue;
        if (c.thm212 && !thm212_ok(px, py, 6, false, n)) continue;
        int need3 = tgt - npx3 - sqnorm(reduce63(py));
        if (need3 < 0 || !ab3.has(need3)) continue;
        if (c.thm211b) {
          Vec T(3);
          for (int s = 1; s <= 3; s++) T[s - 1] = -pair_auto(px, py, s);
          if (!ab6a.has(T, need)) continue;
        }
        cells.push_back({px, py});
      }
    }
    if (c.endpos) {   // wz_match.cpp:1598-1622
      std::vector<Cell> keep;
      for (auto &p : cells) {
        int s = 0;
        for (int q = 0; q < 6; q++) s += p.px[q] + p.py[q];
        int mid = (n % 2 == 1) ? p.px[(n / 2) % 6] + p.py[(n / 2) % 6] : 0;
        if ((((s - mid) % 4) + 4) % 4 == 0) keep.push_back(p);
      }
      cells.swap(keep);
    }
  } else {
    // survive_profiles (mod-3 with mod-6 tighten): wz_match.cpp:439-487
    auto PX3 = class_sum_tuples(n, c.c, 3), PY3 = class_sum_tuples(n, c.d, 3);
    auto PX6 = class_sum_tuples(n, c.c, 6), PY6 = class_sum_tuples(n, c.d, 6);
    for (auto &px : PX3) {
      int npx = sqnorm(px);
      if (npx > tgt) continue;
      for (auto &py : PY3) {
        int need = tgt - npx - sqnorm(py);
        if (need < 0 || !ab3.has(need)) continue;
        if (c.thm212 && !thm212_ok(px, py, 3, false, n)) continue;
        bool ok6 = false;
        for (auto &px6 : PX6) {
          if (reduce63(px6) != px) continue;
          int npx6 = sqnorm(px6);
          if (npx6 > tgt) continue;
          for (auto &py6 : PY6) {
            if (reduce63(py6) != py) continue;
            int need6 = tgt - npx6 - sqnorm(py6);
            if (need6 < 0) continue;
            if (c.thm212 && !thm212_ok(px6, py6, 6, false, n)) continue;
            if (c.thm211b) {
              Vec T(3);
              for (int s = 1; s <= 3; s++) T[s - 1] = -pair_auto(px6, py6, s);
              if (ab6a.has(T, need6)) { ok6 = true; break; }
            } else if (ab6.has(need6)) { ok6 = true; break; }
          }
          if (ok6) break;
        }
        if (ok6) cells.push_back({px, py});
      }
    }
  }
  int m = P.m;
  if (c.order == 1)
    std::sort(cells.begin(), cells.end(), [](const Cell &x, const Cell &y) { return prof_score(x) < prof_score(y); });
  else if (c.order == 2)
    std::sort(cells.begin(), cells.end(), [](const Cell &x, const Cell &y) { return prof_score(x) > prof_score(y); });

  P.kept.assign(cells.size(), 1);
  if (c.canon) {   // wz_match.cpp:1712-1808 (QPRUNE not restated)
    bool q_safe = c.endpos || ((n % 2 == 0) && (((c.c + c.d) % 4 + 4) % 4 == 0));
    bool orbit_q = (c.orbit_q == 2) || (c.orbit_q == 1 && q_safe);
    int nvar = orbit_q ? 64 : 32;
    auto revp = [&](const Vec &p) { Vec q(m); for (int k = 0; k < m; k++) q[((n - 1 - k) % m + m) % m] = p[k]; return q; };
    auto negp = [](Vec p) { for (auto &x : p) x = -x; return p; };
    auto quadq = [&](Vec &px, Vec &py) {
      Vec rx = revp(px), ry = revp(py), nx(m), ny(m);
      for (int k = 0; k < m; k++) {
        nx[k] = (px[k] + py[k] + rx[k] - ry[k]) / 2;
        ny[k] = (px[k] + py[k] - rx[k] + ry[k]) / 2;
      }
      px.swap(nx); py.swap(ny);
    };
    std::unordered_map<std::string, std::string> orbit_min;
    std::vector<std::string> own(cells.size()), oid(cells.size());
    for (size_t i = 0; i < cells.size(); i++) {
      own[i] = cell_key(cells[i].px, cells[i].py);
      std::string best;
      for (int var = 0; var < nvar; var++) {
        Vec px = cells[i].px, py = cells[i].py;
        if (var & 32) quadq(px, py);
        if (var & 1) px = negp(px);
        if (var & 2) py = negp(py);
        if (var & 4) px = revp(px);
        if (var & 8) py = revp(py);
        if (var & 16) std::swap(px, py);
        std::string k = cell_key(px, py);
        if (best.empty() || k < best) best = k;
      }
      oid[i] = best;
      auto it = orbit_min.find(best);
      if (it == orbit_min.end() || own[i] < it->second) orbit_min[best] = own[i];
    }
    if (c.order == 3) {   // wz_match.cpp:1816-1833 (total order)
      std::unordered_map<std::string, long long> omin;
      for (size_t i = 0; i < cells.size(); i++) {
        long long sc = prof_score(cells[i]);
        auto it = omin.find(oid[i]);
        if (it == omin.end() || sc < it->second) omin[oid[i]] = sc;
      }
      std::vector<size_t> perm(cells.size());
      for (size_t i = 0; i < perm.size(); i++) perm[i] = i;
      std::sort(perm.begin(), perm.end(), [&](size_t x, size_t y) {
        long long sx = omin[oid[x]], sy = omin[oid[y]];
        if (sx != sy) return sx < sy;
        if (oid[x] != oid[y]) return oid[x] < oid[y];
        return own[x] < own[y];
      });
      std::vector<Cell> sorted(cells.size());
      std::vector<std::string> own2(cells.size()), oid2(cells.size());
      for (size_t i = 0; i < perm.size(); i++) { sorted[i] = cells[perm[i]]; own2[i] = own[perm[i]]; oid2[i] = oid[perm[i]]; }
      cells.swap(sorted); own.swap(own2); oid.swap(oid2);
    }
    for (size_t i = 0; i < cells.size(); i++) P.kept[i] = orbit_min[oid[i]] == own[i];
  }
  P.cells.swap(cells);
}

// wz_match.cpp:2318-2378 -- A,B profile rows per cell key (need, needT).
void build_rows(Problem &P) {
  int n = P.n, m = P.m, half_m = m / 2, tgt = 4 * n + 2;
  std::unordered_map<uint64_t, int> keyidx;
  P.rowkey.assign(P.cells.size(), -1);
  std::vector<uint64_t> ckey(P.cells.size());
  for (size_t i = 0; i < P.cells.size(); i++) {
    const Cell &c = P.cells[i];
    int need = tgt - sqnorm(c.px) - sqnorm(c.py);
    Vec T(half_m);
    for (int s = 1; s <= half_m; s++) T[s - 1] = -pair_auto(c.px, c.py, s);
    ckey[i] = auto_key(T, need);
    if (!P.kept[i]) continue;
    auto it = keyidx.find(ckey[i]);
    if (it == keyidx.end()) { keyidx.emplace(ckey[i], (int)P.rowlists.size()); P.rowlists.emplace_back(); }
  }
  struct Tup { Vec v; int norm; int ca[3]; };
  auto tups = [&](int sig) {
    std::vector<Tup> out;
    for (int sgn = 1; sgn >= (sig ? -1 : 1); sgn -= 2)
      for (auto &v : class_sum_tuples(P.L1, sgn * sig, m)) {
        Tup t; t.v = v; t.norm = sqnorm(v);
        if (t.norm > tgt) continue;
        for (int s = 1; s <= half_m; s++) t.ca[s - 1] = acorr(v, s) + acorr(v, m - s);
        out.push_back(std::move(t));
      }
    return out;
  };
  auto Kt = tups(P.cfg.a), Rt = tups(P.cfg.b);
  Vec Ts(half_m);
  for (auto &kt : Kt)
    for (auto &rt : Rt) {
      int nn = kt.norm + rt.norm;
      if (nn > tgt) continue;
      for (int s = 0; s < half_m; s++) Ts[s] = kt.ca[s] + rt.ca[s];
      auto it = keyidx.find(auto_key(Ts, nn));
      if (it == keyidx.end()) continue;
      if (!thm212_ok(kt.v, rt.v, m, true, n)) continue;
      Row row{};
      for (int c = 0; c < m; c++) { row.k[c] = (int8_t)kt.v[c]; row.r[c] = (int8_t)rt.v[c]; }
      P.rowlists[it->second].push_back(row);
    }
  for (size_t i = 0; i < P.cells.size(); i++) {
    auto it = keyidx.find(ckey[i]);
    P.rowkey[i] = (it == keyidx.end() || P.rowlists[it->second].empty()) ? -1 : it->second;
  }
}

// ---------------------------------------------------------------------------
// C,D stream: count_pairs22 (wz_match.cpp:678-784) on the C,D side, m classes.
struct Stream {
  const Problem &P;
  int L, m, half;
  int tx[8], ty[8], tot[8], px[8] = {0}, py[8] = {0}, placed[8] = {0};
  int X[64], Y[64];
  long long cap, nout = 0, leaves = 0, ok = 0, nodes = 0, node_budget = 0;
  bool stop = false, budget_cut = false;
  int8_t *out; long long out_cap;
  std::vector<int8_t> *grow = nullptr;   // when set, every candidate is appended here instead
  Stream(const Problem &p, const Cell &cell, long long cap_, int8_t *o, long long oc)
      : P(p), L(p.n), m(p.m), half(p.n / 2), cap(cap_), out(o), out_cap(oc) {
    for (int c = 0; c < m; c++) { tx[c] = cell.px[c]; ty[c] = cell.py[c]; tot[c] = class_count(L, c, m); }
    std::memset(X, 0, sizeof X); std::memset(Y, 0, sizeof Y);
  }
  // hall_ok_pair_fast decision (wz_match.cpp:143-168): reject iff some angle j has
  // P_X > lim+0.5, P_Y > lim+0.5 or P_X+P_Y > lim+0.5, with P = rx*rx + ix*ix and the
  // sums accumulated left to right in FP64. The decision is a conjunction over angles, so
  // the visiting order does not change it; like the reference, angles are visited most-
  // rejecting first (per thread, re-sorted every 4096 rejections, :158-162).
  struct HallOrder {
    int order[200];
    long long fail[201] = {0}, rej = 0;
    HallOrder() { for (int t = 0; t < 200; t++) order[t] = t + 1; }
  };
  static HallOrder &hall_order() { static thread_local HallOrder h; return h; }
  bool hall() const {
    const double lim = 4.0 * P.n + 2.0 + 0.5;
    HallOrder &H = hall_order();
    for (int t = 0; t < 200; t++) {
      const int j = H.order[t];
      const double *cj = P.hcos[j], *sj = P.hsin[j];
      double rx = 0, ix = 0, ry = 0, iy = 0;
      for (int i = 0; i < L; i++) { rx += X[i] * cj[i]; ix += X[i] * sj[i]; }
      bool fail = rx * rx + ix * ix > lim;
      if (!fail) {
        for (int i = 0; i < L; i++) { ry += Y[i] * cj[i]; iy += Y[i] * sj[i]; }
        fail = ry * ry + iy * iy > lim || rx * rx + ix * ix + ry * ry + iy * iy > lim;
      }
      if (fail) {
        H.fail[j]++;
        if ((++H.rej & 4095) == 0)
          std::stable_sort(H.order, H.order + 200, [&](int a, int b) { return H.fail[a] > H.fail[b]; });
        return false;
      }
    }
    return true;
  }
  void leaf() {
    leaves++;
    if (!hall()) return;
    ok++;
    if (grow) {
      for (int i = 0; i < L; i++) grow->push_back((int8_t)X[i]);
      for (int i = 0; i < L; i++) grow->push_back((int8_t)Y[i]);
    } else if (nout < out_cap && out) {
      int8_t *o = out + nout * 2 * L;
      for (int i = 0; i < L; i++) { o[i] = (int8_t)X[i]; o[L + i] = (int8_t)Y[i]; }
    }
    nout++;
    if (cap > 0 && nout >= cap) stop = true;
  }
  void rec(int d) {
    if (stop) return;
    if (node_budget > 0 && nodes >= node_budget) { stop = budget_cut = true; return; }   // node index >= budget
    nodes++;
    for (int c = 0; c < m; c++) {
      int rem = tot[c] - placed[c], dx = tx[c] - px[c], dy = ty[c] - py[c];
      if (dx < -rem || dx > rem || dy < -rem || dy > rem) return;
      if (((dx + rem) & 1) || ((dy + rem) & 1)) return;
    }
    if (d == half) {
      if (L % 2 == 1) {
        int mid = half, c = mid % m;
        for (int k = 0; k < 4; k++) {
          X[mid] = QT.mid[k][0]; Y[mid] = QT.mid[k][1];
          px[c] += X[mid]; py[c] += Y[mid]; placed[c]++;
          bool exact = true;
          for (int q = 0; q < m; q++) if (px[q] != tx[q] || py[q] != ty[q]) exact = false;
          if (exact) leaf();
          px[c] -= X[mid]; py[c] -= Y[mid]; placed[c]--;
        }
        X[mid] = Y[mid] = 0;
      } else {
        bool exact = true;
        for (int q = 0; q < m; q++) if (px[q] != tx[q] || py[q] != ty[q]) exact = false;
        if (exact) leaf();
      }
      return;
    }
    int i1 = d, i2 = L - 1 - d, c1 = i1 % m, c2 = i2 % m;
    const int (*S)[4] = (d == 0) ? (P.cfg.endpos ? QT.pos : QT.all) : QT.pos;
    int ns = (d == 0 && !P.cfg.endpos) ? 16 : 8;
    for (int kk = 0; kk < ns; kk++) {
      int k = P.cfg.rev ? ns - 1 - kk : kk;
      if (d == 0 && P.pinC && S[k][0] != 1) continue;
      if (d == 0 && P.pinD && S[k][1] != 1) continue;
      X[i1] = S[k][0]; Y[i1] = S[k][1]; X[i2] = S[k][2]; Y[i2] = S[k][3];
      px[c1] += S[k][0]; py[c1] += S[k][1]; px[c2] += S[k][2]; py[c2] += S[k][3];
      placed[c1]++; placed[c2]++;
      rec(d + 1);
      px[c1] -= S[k][0]; py[c1] -= S[k][1]; px[c2] -= S[k][2]; py[c2] -= S[k][3];
      placed[c1]--; placed[c2]--;
    }
    X[i1] = Y[i1] = X[i2] = Y[i2] = 0;
  }
};

// ---------------------------------------------------------------------------
// A,B completer: fh_complete_ab / fh_ab_search (wz_match.cpp:1092-1341) with A1 outer
// tables, the early outer-correlation check, root + reversal canon, profile rows.
struct Completer {
  const Problem &P;
  int n, L, m, absA, absB;
  long long budget, cur = 0, total = 0;
  bool aborted = false;
  int T[64], A[64], B[64], Dab[64], Kab[64];
  int PA[8], PB[8], PL[8], TIC[8];
  const std::vector<Row> *rows = nullptr;
  std::vector<std::vector<int>> stack;
  Completer(const Problem &p, long long bud) : P(p), n(p.n), L(p.L1), m(p.m), budget(bud) {
    absA = std::abs(p.cfg.a); absB = std::abs(p.cfg.b);
    for (int c = 0; c < m; c++) TIC[c] = class_count(L, c, m);
  }
  void place(int p, int av, int bv) {   // fh_place V1 (wz_match.cpp:1128-1137)
    for (int q = 0; q < L; q++) {
      if (A[q] == 0 || q == p) continue;
      int s = q > p ? q - p : p - q;
      Dab[s] += A[q] * av + B[q] * bv;
      Kab[s] -= 2;
    }
    A[p] = av; B[p] = bv;
    int pc = p % m;
    PA[pc] += av; PB[pc] += bv; PL[pc]++;
  }
  void unplace(int p) {
    int av = A[p], bv = B[p];
    A[p] = 0; B[p] = 0;
    int pc = p % m;
    PA[pc] -= av; PB[pc] -= bv; PL[pc]--;
    for (int q = 0; q < L; q++) {
      if (A[q] == 0 || q == p) continue;
      int s = q > p ? q - p : p - q;
      Dab[s] -= A[q] * av + B[q] * bv;
      Kab[s] += 2;
    }
  }
  unsigned outer_mask(int d) const {   // wz_match.cpp:601-618
    int k = P.cfg.outer_k;
    if (k < 2 || d < k || d + k > L / 2 || A[0] != 1 || B[0] != 1) return 0xFFu;
    int pat = (A[L - 1] > 0), mul = 2;
    for (int t = 1; t < k; t++) {
      int q = OT.pos_index[OuterTables::bits(A[t], B[t], A[L - 1 - t], B[L - 1 - t])];
      if (q < 0) return 0xFFu;
      pat += mul * q; mul *= 8;
    }
    int R[3];
    for (int j = 0; j < k; j++) {
      int sj = L - 1 - d - j;
      R[j] = T[sj] - Dab[sj];
      if ((R[j] & 1) || R[j] < -4 * (j + 1) || R[j] > 4 * (j + 1)) return 0;
    }
    if (k == 2) return OT.t2[pat][R[0] / 2 + 2][R[1] / 2 + 4];
    return OT.t3[pat][R[0] / 2 + 2][R[1] / 2 + 4][R[2] / 2 + 6];
  }
  static bool a2_self_ok(int u, int v, int q) {   // wz_match.cpp:1017-1021
    if ((u | v) & 1) return false;
    if (std::abs(u) > 2 * q || std::abs(v) > 2 * q) return false;
    return ((u - v) & 3) == 0;
  }
  bool a2_row_ok(const Row &row, int p) const {   // wz_match.cpp:1022-1038
    for (int b = 0; b < P.a2_nblk; b++) {
      int r = P.a2_r[b], s = P.a2_s[b], q = P.a2_cnt[p][b];
      int ur = row.k[r] - PA[r], vr = row.r[r] - PB[r];
      if (r != s) {
        int us = row.k[s] - PA[s], vs = row.r[s] - PB[s];
        int t0 = ur + vr + us + vs, t1 = ur - vr + us - vs, t2 = ur + vr - us - vs, t3 = ur - vr - us + vs;
        if ((t0 | t1 | t2 | t3) & 3) return false;
        int l1 = (std::abs(t0) + std::abs(t1) + std::abs(t2) + std::abs(t3)) >> 2;
        if (l1 > q || ((q - l1) & 1)) return false;
      } else if (P.a2_mid[b]) {
        if (!a2_self_ok(ur - 1, vr - 1, q) && !a2_self_ok(ur - 1, vr + 1, q) &&
            !a2_self_ok(ur + 1, vr - 1, q) && !a2_self_ok(ur + 1, vr + 1, q)) return false;
      } else if (!a2_self_ok(ur, vr, q)) {
        return false;
      }
    }
    return true;
  }
  bool row_filter(int d) {   // fh_abp_filter (wz_match.cpp:1055-1078), A2 after the capacity box
    auto &src = stack[d];
    auto &dst = stack[d + 1];
    dst.clear();
    for (int idx : src) {
      const Row &row = (*rows)[idx];
      bool ok = true;
      for (int c = 0; c < m; c++) {
        int rem = TIC[c] - PL[c], dk = row.k[c] - PA[c], dr = row.r[c] - PB[c];
        if (dk < -rem || dk > rem || dr < -rem || dr > rem) { ok = false; break; }
      }
      if (ok && P.cfg.a2) {
        P.a2_tested++;
        if (!a2_row_ok(row, d + 1)) { P.a2_rejected++; if (P.cfg.a2 == 2) ok = false; }
      }
      if (ok) dst.push_back(idx);
    }
    return !dst.empty();
  }
  bool leaf_rows_ok() const {
    for (int idx : stack[L / 2]) {
      const Row &row = (*rows)[idx];
      bool eq = true;
      for (int c = 0; c < m; c++) if (row.k[c] != PA[c] || row.r[c] != PB[c]) { eq = false; break; }
      if (eq) return true;
    }
    return false;
  }
  bool final_ok(int sA, int sB) const {
    if (std::abs(sA) != absA || std::abs(sB) != absB) return false;
    for (int s = 1; s <= n; s++) if (Dab[s] != T[s]) return false;
    if (rows && !leaf_rows_ok()) return false;
    return true;
  }
  bool search(int d, int sumA, int sumB, int a_tied, int a_cmp, int b_tied, int b_cmp) {
    if (aborted) return false;
    int half = L / 2;
    bool canon = !P.cfg.no_canon_ab;
    if (d == half) {
      if (L % 2 == 1) {
        int mid = half;
        for (int k = 0; k < 4; k++) {
          int av = QT.mid[k][0], bv = QT.mid[k][1];
          place(mid, av, bv);
          total++;
          if (final_ok(sumA + av, sumB + bv)) return true;
          unplace(mid);
        }
        return false;
      }
      return final_ok(sumA, sumB);
    }
    int i1 = d, i2 = L - 1 - d;
    const int (*S)[4] = (d == 0) ? QT.neg : QT.pos;
    unsigned omask = (P.cfg.outer_k && d > 0) ? outer_mask(d) : 0xFFu;
    if (omask == 0) return false;
    for (int k = 0; k < 8; k++) {
      if (d > 0 && !((omask >> k) & 1u)) continue;
      int a1 = S[k][0], b1 = S[k][1], a2 = S[k][2], b2 = S[k][3];
      if (d == 0 && canon && (a1 != 1 || b1 != 1)) continue;
      int na_tied = a_tied, na_cmp = a_cmp, nb_tied = b_tied, nb_cmp = b_cmp;
      if (canon) {
        if (d == 0) { na_cmp = a2; nb_cmp = b2; }
        else {
          if (a_tied && a1 != a_cmp * a2) { if (a1 != 1) continue; na_tied = 0; }
          if (b_tied && b1 != b_cmp * b2) { if (b1 != 1) continue; nb_tied = 0; }
        }
      }
      total++;
      if (budget > 0 && ++cur > budget) { aborted = true; return false; }
      if (d > 0) {   // early outer-correlation check (default on)
        int s = L - 1 - d;
        if (A[0] * a2 + B[0] * b2 + A[L - 1] * a1 + B[L - 1] * b1 != T[s] - Dab[s]) continue;
      }
      place(i1, a1, b1);
      place(i2, a2, b2);
      int nsA = sumA + a1 + a2, nsB = sumB + b1 + b2, rem = L - 2 * (d + 1);
      bool prune = !(((std::abs(absA - nsA) <= rem) || (std::abs(-absA - nsA) <= rem)) &&
                     ((std::abs(absB - nsB) <= rem) || (std::abs(-absB - nsB) <= rem)));
      if (!prune)
        for (int s = 1; s <= n; s++)
          if (std::abs(T[s] - Dab[s]) > Kab[s]) { prune = true; break; }
      if (!prune && rows && !row_filter(d)) prune = true;
      if (!prune && search(d + 1, nsA, nsB, na_tied, na_cmp, nb_tied, nb_cmp)) return true;
      unplace(i2);
      unplace(i1);
    }
    return false;
  }
  int run(const int8_t *C, const int8_t *D, int8_t *Aout, int8_t *Bout) {
    for (int s = 1; s <= n; s++) {
      int cd = 0;
      for (int i = 0; i + s < n; i++) cd += C[i] * C[i + s] + D[i] * D[i + s];
      T[s] = -cd;
    }
    for (int s = 1; s < L; s++) if (std::abs(T[s]) > 2 * (L - s)) return 1;
    std::memset(A, 0, sizeof A); std::memset(B, 0, sizeof B); std::memset(Dab, 0, sizeof Dab);
    for (int s = 0; s <= n; s++) Kab[s] = (s >= 1 && s < L) ? 2 * (L - s) : 0;
    std::memset(PA, 0, sizeof PA); std::memset(PB, 0, sizeof PB); std::memset(PL, 0, sizeof PL);
    cur = 0; aborted = false;
    if (rows) {
      stack.assign(L / 2 + 2, {});
      stack[0].resize(rows->size());
      for (size_t i = 0; i < rows->size(); i++) stack[0][i] = (int)i;
    }
    bool found = search(0, 0, 0, 1, 0, 1, 0);
    if (!found) return aborted ? 3 : 2;
    for (int s = 1; s <= n; s++) {
      int c = 0;
      for (int i = 0; i < L - s; i++) c += A[i] * A[i + s] + B[i] * B[i + s];
      for (int i = 0; i < n - s; i++) c += C[i] * C[i + s] + D[i] * D[i + s];
      if (c != 0) return 4;
    }
    if (Aout) for (int i = 0; i < L; i++) Aout[i] = (int8_t)A[i];
    if (Bout) for (int i = 0; i < L; i++) Bout[i] = (int8_t)B[i];
    return 0;
  }
};

// ---- Every solution in the space WZ_JOIN22 searches (wz_match.cpp:2894-2904): both sides
// enumerated straight from the Thm 2.2 pair encoding (:525-530: quads (X[d], Y[d], X[L-1-d],
// Y[L-1-d]) with product +1 at depth >= 1; product -1 at depth 0 on the A,B side, any on the
// C,D side; a free odd middle), with X[0] = +1 pinned when X's sum is 0 (:1526-1527) and the
// exact signed sums of the signature, then joined on exact autocorrelation vectors. Unlike
// the reference it uses no profile list and no hall test: both are necessary conditions for a
// solution, so the set is the same, and this enumeration checks that independently.
struct SideEnum {
  int L, half, tx, ty;
  bool ab, pinx, piny;
  std::function<void(uint64_t, uint64_t)> sink;   // (X, Y) packed: bit i set = element i is -1
  void rec(int d, uint64_t X, uint64_t Y, int sx, int sy, int root = -1) {
    const int rem = L - 2 * d;
    if (std::abs(tx - sx) > rem || std::abs(ty - sy) > rem) return;
    if (d == half) {
      if (L & 1) {
        for (int t = 0; t < 4; t++) {
          const int x = (t & 2) ? -1 : 1, y = (t & 1) ? -1 : 1;
          if (sx + x == tx && sy + y == ty) sink(X | (x < 0 ? 1ull << d : 0), Y | (y < 0 ? 1ull << d : 0));
        }
      } else if (sx == tx && sy == ty) {
        sink(X, Y);
      }
      return;
    }
    const int i1 = d, i2 = L - 1 - d;
    for (int q = 0; q < 16; q++) {
      if (d == 0 && root >= 0 && q != root) continue;
      const int x1 = (q & 8) ? 1 : -1, y1 = (q & 4) ? 1 : -1, x2 = (q & 2) ? 1 : -1, y2 = (q & 1) ? 1 : -1;
      const int prod = x1 * y1 * x2 * y2;
      if (d == 0) {
        if (ab && prod != -1) continue;
        if ((pinx && x1 != 1) || (piny && y1 != 1)) continue;
      } else if (prod != 1) {
        continue;
      }
      const uint64_t nX = X | (x1 < 0 ? 1ull << i1 : 0) | (x2 < 0 ? 1ull << i2 : 0);
      const uint64_t nY = Y | (y1 < 0 ? 1ull << i1 : 0) | (y2 < 0 ? 1ull << i2 : 0);
      rec(d + 1, nX, nY, sx + x1 + x2, sy + y1 + y2);
    }
  }
};

// N_X(s) of a packed sequence of length len: (len - s) - 2 * (number of sign changes at
// distance s), i.e. the sum of x_i x_{i+s} with x_i x_{i+s} = 1 - 2 (b_i xor b_{i+s}).
inline int npaf_packed(uint64_t x, int len, int s) {
  if (s >= len) return 0;
  return (len - s) - 2 * __builtin_popcountll((x ^ (x >> s)) & ((1ull << (len - s)) - 1));
}

}  // namespace

extern "C" {

void *wzo_create(const wzo_config *cfg) {
  Problem *P = new Problem();
  P->cfg = *cfg;
  if (P->cfg.outer_k < 0) P->cfg.outer_k = 3;
  if (P->cfg.abp_cap <= 0) P->cfg.abp_cap = 20000;
  P->n = cfg->n; P->L1 = cfg->n + 1; P->m = cfg->m6 ? 6 : 3;
  if (P->n >= 36) { P->cfg.thm211b = 1; P->cfg.thm212 = 1; }
  P->pinC = (cfg->c == 0) && !cfg->canon;
  P->pinD = (cfg->d == 0) && !cfg->canon;
  P->ab_prof = cfg->ab_prof < 0 ? (cfg->m6 != 0) : (cfg->ab_prof != 0);
  init_a2(*P);
  for (int j = 1; j <= 200; j++) {   // init_hall_tables: wz_match.cpp:106-114
    double th = j * M_PI / 100.0;
    for (int i = 0; i < P->L1 && i < 64; i++) { P->hcos[j][i] = cos(i * th); P->hsin[j][i] = sin(i * th); }
  }
  if (!P->cfg.no_cells) {
    build_cells(*P);
    if (P->ab_prof) build_rows(*P);
  } else {
    P->ab_prof = false;
  }
  return P;
}

void *wzo_create_with_cells(const wzo_config *cfg, const int8_t *pxpy, int ncells) {
  wzo_config c2 = *cfg;
  c2.no_cells = 1;
  Problem *P = (Problem *)wzo_create(&c2);
  P->cfg.no_cells = 0;
  P->ab_prof = cfg->ab_prof < 0 ? (cfg->m6 != 0) : (cfg->ab_prof != 0);
  for (int i = 0; i < ncells; i++) {
    Cell c;
    c.px.assign(pxpy + 16 * i, pxpy + 16 * i + P->m);
    c.py.assign(pxpy + 16 * i + 8, pxpy + 16 * i + 8 + P->m);
    P->cells.push_back(c);
  }
  P->kept.assign(P->cells.size(), 1);
  if (P->ab_prof) build_rows(*P);
  return P;
}

void wzo_destroy(void *h) { delete (Problem *)h; }

int wzo_num_cells(void *h) { return (int)((Problem *)h)->cells.size(); }

int wzo_cell_info(void *h, int pi, int *px, int *py, int *kept, int *rows) {
  Problem &P = *(Problem *)h;
  if (pi < 0 || pi >= (int)P.cells.size()) return -1;
  for (int c = 0; c < P.m; c++) { px[c] = P.cells[pi].px[c]; py[c] = P.cells[pi].py[c]; }
  *kept = P.kept[pi];
  if (!P.ab_prof) *rows = -2;
  else *rows = P.rowkey[pi] < 0 ? -1 : (int)P.rowlists[P.rowkey[pi]].size();
  return 0;
}

long long wzo_stream_cell_budget(void *h, int pi, long long cap, long long node_budget, int8_t *out,
                                 long long out_cap, wzo_stream_stats *st) {
  Problem &P = *(Problem *)h;
  Stream S(P, P.cells[pi], cap, out, out_cap);
  S.node_budget = node_budget;
  S.rec(0);
  if (st) { st->leaves = S.leaves; st->hall_ok = S.ok; st->dfs_nodes = S.nodes; st->cand = S.nout; }
  return S.nout;
}

long long wzo_stream_cell(void *h, int pi, long long cap, int8_t *out, long long out_cap,
                          wzo_stream_stats *st) {
  Problem &P = *(Problem *)h;
  Stream S(P, P.cells[pi], cap, out, out_cap);
  S.rec(0);
  if (st) { st->leaves = S.leaves; st->hall_ok = S.ok; st->dfs_nodes = S.nodes; st->cand = S.nout; }
  return S.nout;
}

int wzo_complete(void *h, int pi, const int8_t *C, const int8_t *D, long long budget,
                 long long *nodes, int8_t *A, int8_t *B) {
  Problem &P = *(Problem *)h;
  Completer K(P, budget);
  if (pi >= 0 && P.ab_prof && P.rowkey[pi] >= 0 &&
      (long long)P.rowlists[P.rowkey[pi]].size() <= P.cfg.abp_cap)
    K.rows = &P.rowlists[P.rowkey[pi]];
  int r = K.run(C, D, A, B);
  if (nodes) *nodes = K.total;
  return r;
}

void wzo_a2_counts(void *h, long long *tested, long long *rejected) {
  Problem &P = *(Problem *)h;
  *tested = P.a2_tested; *rejected = P.a2_rejected;
  P.a2_tested = P.a2_rejected = 0;
}

long long wzo_flat_score(int n, const int8_t *C, const int8_t *D) {   // wz_match.cpp:2569-2583
  long long sc = 0;
  for (int s = 1; s <= n; s++) {
    int cd = 0;
    for (int i = 0; i + s < n; i++) cd += C[i] * C[i + s] + D[i] * D[i + s];
    sc += std::abs(cd);
  }
  return sc;
}

int wzo_npaf_at(const int8_t *A, const int8_t *B, int n1, const int8_t *C, const int8_t *D,
                int n2, int s) {
  int c = 0;
  if (s < n1) for (int i = 0; i < n1 - s; i++) c += A[i] * A[i + s] + B[i] * B[i + s];
  if (s < n2) for (int i = 0; i < n2 - s; i++) c += C[i] * C[i + s] + D[i] * D[i + s];
  return c;
}

// The front-only cell unit (the reference's drain, wz_match.cpp:2656-2714, with
// WZ_FH_BUF_CAP = buf_cap, WZ_FH_DRAIN_TOP = top_k, WZ_FH_DRAIN_BATCHES = 1): stream the
// first buf_cap candidates of cell pi (and, when node_budget > 0, only nodes with DFS index
// below it), stable-sort them by flat score (:2654-2655), complete the sorted ranks in order
// under the cell's profile rows and stop after the first hit (:2686, the drain's break).
// Per rank r < res->completed: verdict, nodes, in-cell stream index and flat score.
long long wzo_cell_search(void *h, int pi, long long buf_cap, long long node_budget, int top_k, long long budget,
                          wzo_search_result *res, int *verdicts, long long *nodes, long long *stream_idx,
                          long long *scores) {
  Problem &P = *(Problem *)h;
  const int n = P.n;
  std::vector<int8_t> buf;
  Stream S(P, P.cells[pi], buf_cap, nullptr, 0);
  S.grow = &buf;
  S.node_budget = node_budget;
  S.rec(0);
  const long long got = S.nout;
  std::memset(res, 0, sizeof *res);
  res->cand = got; res->leaves = S.leaves; res->hall_ok = S.ok; res->dfs_nodes = S.nodes;
  res->capped = (buf_cap > 0 && got >= buf_cap) || S.budget_cut;
  res->budget_cut = S.budget_cut;
  res->hit_rank = res->hit_idx = -1;
  std::vector<std::pair<long long, long long>> key(got);
  for (long long i = 0; i < got; i++) key[i] = {wzo_flat_score(n, &buf[i * 2 * n], &buf[i * 2 * n + n]), i};
  std::stable_sort(key.begin(), key.end(), [](auto &x, auto &y) { return x.first < y.first; });
  const long long stop = std::min<long long>(got, top_k);
  std::vector<int8_t> A(n + 1), B(n + 1);
  for (long long r = 0; r < stop; r++) {
    const long long i = key[r].second;
    long long nd = 0;
    const int v = wzo_complete(h, pi, &buf[i * 2 * n], &buf[i * 2 * n + n], budget, &nd, A.data(), B.data());
    if (verdicts) verdicts[r] = v;
    if (nodes) nodes[r] = nd;
    if (stream_idx) stream_idx[r] = i;
    if (scores) scores[r] = key[r].first;
    res->completed++;
    res->ab_nodes += nd;
    if (v == 1) res->pre_rejected++;
    else if (v == 2) res->clean_no++;
    else if (v == 3) res->aborted++;
    else if (v == 4) res->internal = 1;
    if (v == 4) break;   // the reference exits here (wz_match.cpp:2591-2595)
    if (v == 0) {
      res->hit = 1; res->hit_rank = r; res->hit_idx = i;
      for (int t = 0; t <= n; t++) { res->A[t] = A[t]; res->B[t] = B[t]; }
      for (int t = 0; t < n; t++) { res->C[t] = buf[i * 2 * n + t]; res->D[t] = buf[i * 2 * n + n + t]; }
      break;
    }
  }
  return got;
}

long long wzo_solutions(int n, int a, int b, int c, int d, uint64_t *out, long long cap) {
  // C,D side: every entry keyed by -(N_C(s) + N_D(s)), s = 1..n (a 64-bit hash indexes exact
  // vectors); an A,B pair with N_A(s) + N_B(s) equal to it has NPAF = 0 at every shift
  std::vector<std::pair<uint64_t, uint64_t>> cds;
  std::vector<int16_t> cdvec;
  std::unordered_map<uint64_t, std::vector<uint32_t>> byhash;
  auto vec_hash = [n](const int16_t *v) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int s = 0; s < n; s++) { h ^= (uint16_t)v[s]; h *= 0x100000001b3ull; }
    return h;
  };
  SideEnum cdside{n, n / 2, c, d, false, c == 0, d == 0, {}};
  cdside.sink = [&](uint64_t X, uint64_t Y) {
    const size_t at = cdvec.size();
    for (int s = 1; s <= n; s++) cdvec.push_back((int16_t)-(npaf_packed(X, n, s) + npaf_packed(Y, n, s)));
    byhash[vec_hash(&cdvec[at])].push_back((uint32_t)cds.size());
    cds.push_back({X, Y});
  };
  cdside.rec(0, 0, 0, 0, 0);
  // A,B side: one thread per admissible root quad, each probing at its leaves
  std::vector<std::vector<std::array<uint64_t, 4>>> part(16);
  std::vector<std::thread> th;
  for (int root = 0; root < 16; root++) {
    th.emplace_back([&, root]() {
      SideEnum abside{n + 1, (n + 1) / 2, a, b, true, a == 0, b == 0, {}};
      int16_t v[64];
      abside.sink = [&](uint64_t X, uint64_t Y) {
        for (int s = 1; s <= n; s++) v[s - 1] = (int16_t)(npaf_packed(X, n + 1, s) + npaf_packed(Y, n + 1, s));
        auto it = byhash.find(vec_hash(v));
        if (it == byhash.end()) return;
        for (uint32_t j : it->second)
          if (std::memcmp(&cdvec[(size_t)j * n], v, 2 * (size_t)n) == 0)
            part[root].push_back({X, Y, cds[j].first, cds[j].second});
      };
      abside.rec(0, 0, 0, 0, 0, root);
    });
  }
  for (auto &t : th) t.join();
  std::vector<std::array<uint64_t, 4>> sol;
  for (auto &p : part) sol.insert(sol.end(), p.begin(), p.end());
  std::sort(sol.begin(), sol.end());
  for (long long i = 0; i < (long long)sol.size() && i < cap; i++)
    for (int t = 0; t < 4; t++) out[4 * i + t] = sol[i][t];
  return (long long)sol.size();
}

void wzo_outer_tables(uint8_t *outer2, uint8_t *outer3) {
  std::memcpy(outer2, OT.t2, sizeof OT.t2);
  std::memcpy(outer3, OT.t3, sizeof OT.t3);
}

}  // extern "C"
