// wzfh_setup.cpp -- per-(n, class) host setup of the first-hit engine.
//
// Builds what the reference's main() builds before its cell loop
// (src/solver/wz_match.cpp:1391-1410, 1514-1522, 1588-1891, 2318-2381): the
// complement norm / residue-autocorrelation sets, the ordered and canonicalized
// mod-m profile-cell list, and the per-cell A,B profile rows. Results are identical
// to the reference's; the work is reorganized so n = 44 classes set up in about a
// second: the Thm 2.3 eq-(18) mod-4 test is separable per side, so the pair loop
// visits only the compatible residue bucket of each C profile, and cell keys are
// packed integers ordered exactly like the reference's string keys.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "wzfh_internal.h"

namespace wzfh {

Quads::Quads() {
  int np = 0, nn = 0;
  for (int i = 0; i < 16; i++) {
    int v[4] = {(i & 8) ? 1 : -1, (i & 4) ? 1 : -1, (i & 2) ? 1 : -1, (i & 1) ? 1 : -1};
    for (int t = 0; t < 4; t++) all[i][t] = v[t];
    int *dst = (v[0] * v[1] * v[2] * v[3] == 1) ? pos[np++] : neg[nn++];
    for (int t = 0; t < 4; t++) dst[t] = v[t];
  }
}
const Quads kQuads;

static int sign_bits(int a, int b, int c, int d) {
  return ((a > 0) << 3) | ((b > 0) << 2) | ((c > 0) << 1) | (d > 0);
}

// Reachable outer-lag residual tuples (wz_match.cpp:547-598). With the boundary
// quads of depths 0..k-1 fixed (root from NEG with A0=B0=+1, then POS quads), the k
// outermost open lags L-1-d-j receive F_j = sum_{r<=j} left[j-r].(right half of
// quad r) + right[j-r].(left half of quad r); the table marks which first quads
// reach each tuple.
OuterTables::OuterTables() {
  std::memset(t2, 0, sizeof t2);
  std::memset(t3, 0, sizeof t3);
  for (int i = 0; i < 16; i++) pos_index[i] = -1;
  for (int q = 0; q < 8; q++)
    pos_index[sign_bits(kQuads.pos[q][0], kQuads.pos[q][1], kQuads.pos[q][2], kQuads.pos[q][3])] = (int8_t)q;
  int root_of[2] = {0, 0};
  for (int q = 0; q < 8; q++)
    if (kQuads.neg[q][0] == 1 && kQuads.neg[q][1] == 1) root_of[kQuads.neg[q][2] > 0] = q;
  for (int k = 2; k <= 3; k++) {
    const int npat = (k == 2) ? 16 : 128;
    for (int pat = 0; pat < npat; pat++) {
      int bl[3][2], br[3][2];   // boundary (A,B) on the left / right at depth t
      const int *root = kQuads.neg[root_of[pat & 1]];
      bl[0][0] = root[0]; bl[0][1] = root[1]; br[0][0] = root[2]; br[0][1] = root[3];
      int rest = pat >> 1;
      for (int t = 1; t < k; t++) {
        const int *q = kQuads.pos[rest & 7];
        rest >>= 3;
        bl[t][0] = q[0]; bl[t][1] = q[1]; br[t][0] = q[2]; br[t][1] = q[3];
      }
      const int total = (k == 2) ? 64 : 512;
      for (int asg = 0; asg < total; asg++) {
        int pick[3] = {asg & 7, (asg >> 3) & 7, (asg >> 6) & 7};
        int F[3] = {0, 0, 0};
        for (int j = 0; j < k; j++)
          for (int r = 0; r <= j; r++) {
            const int *v = kQuads.pos[pick[r]];
            F[j] += bl[j - r][0] * v[2] + bl[j - r][1] * v[3] + br[j - r][0] * v[0] + br[j - r][1] * v[1];
          }
        if (k == 2) t2[pat][F[0] / 2 + 2][F[1] / 2 + 4] |= (uint8_t)(1u << pick[0]);
        else t3[pat][F[0] / 2 + 2][F[1] / 2 + 4][F[2] / 2 + 6] |= (uint8_t)(1u << pick[0]);
      }
    }
  }
}
const OuterTables kOuter;

namespace {

using Tup = std::vector<int8_t>;

int class_count(int L, int c, int m) { return c < L ? (L - c + m - 1) / m : 0; }

// All length-m class-sum tuples of a +/-1 sequence of length L with the given total,
// lexicographically ascending (enum_class_sums, wz_match.cpp:203-224).
std::vector<Tup> class_tuples(int L, int total, int m) {
  int cnt[8];
  for (int c = 0; c < m; c++) cnt[c] = class_count(L, c, m);
  std::vector<Tup> out;
  Tup cur(m, 0);
  struct Walk {
    int m, total;
    const int *cnt;
    Tup &cur;
    std::vector<Tup> &out;
    void go(int c, int acc) {
      if (c == m - 1) {
        int last = total - acc;
        if (last < -cnt[c] || last > cnt[c] || ((last + cnt[c]) & 1)) return;
        cur[c] = (int8_t)last;
        out.push_back(cur);
        return;
      }
      for (int v = -cnt[c]; v <= cnt[c]; v += 2) { cur[c] = (int8_t)v; go(c + 1, acc + v); }
    }
  } w{m, total, cnt, cur, out};
  w.go(0, 0);
  return out;
}

int sqn(const Tup &v) { int s = 0; for (int x : v) s += x * x; return s; }
int acorr(const Tup &v, int s) {
  int m = (int)v.size(), t = 0;
  for (int i = 0; i + s < m; i++) t += v[i] * v[i + s];
  return t;
}
// One side's eq-2.11b term (pair_auto, wz_match.cpp:277-281).
int side_auto(const Tup &x, const Tup &y, int s) {
  int m = (int)x.size();
  return acorr(x, s) + acorr(y, s) + acorr(x, m - s) + acorr(y, m - s);
}
// (T_1..T_3, norm) packed as in auto_key (wz_match.cpp:284-289).
uint64_t pack_auto(const int *T, int nT, int norm) {
  uint64_t k = (uint64_t)(uint16_t)norm;
  for (int i = 0; i < nT && i < 3; i++) k |= ((uint64_t)(uint16_t)(T[i] + 8192)) << (16 * (i + 1));
  return k;
}
Tup reduce63(const Tup &v) { return {(int8_t)(v[0] + v[3]), (int8_t)(v[1] + v[4]), (int8_t)(v[2] + v[5])}; }

// Thm 2.3 eq (18) (thm212_ok, wz_match.cpp:302-322) split into per-side residues:
// condition j reads (x[j-1] + x[c_j] + y[j-1] + y[c_j]) mod 4 == want_j, so each side
// contributes r_j(v) = (v[j-1] + v[c_j]) mod 4 and the pair passes iff
// r_j(x) + r_j(y) == want_j (mod 4) for every tested j.
struct Eq212 {
  int m, n;
  bool ab;
  int nj = 0, ja[8], jc[8], want[8];
  Eq212(int m_, int n_, bool ab_) : m(m_), n(n_), ab(ab_) {
    if (ab) {
      int c1 = n % m;
      ja[nj] = 0; jc[nj] = c1; want[nj] = (n % m) ? 2 : 0; nj++;
      for (int j = 2; j <= m; j++) {
        if (j == c1 + 1) continue;
        ja[nj] = j - 1; jc[nj] = ((n + 1 - j) % m + m) % m; want[nj] = 0; nj++;
      }
    } else {
      for (int j = 1; j <= m; j++) { ja[nj] = j - 1; jc[nj] = ((n - j) % m + m) % m; want[nj] = 0; nj++; }
    }
  }
  int residue_key(const Tup &v) const {      // 2 bits per tested condition
    int k = 0;
    for (int t = 0; t < nj; t++) k |= (((v[ja[t]] + v[jc[t]]) % 4 + 4) % 4) << (2 * t);
    return k;
  }
  int partner_key(const Tup &x) const {      // residue key y must have
    int k = 0;
    for (int t = 0; t < nj; t++) {
      int rx = ((x[ja[t]] + x[jc[t]]) % 4 + 4) % 4;
      k |= (((want[t] - rx) % 4 + 4) % 4) << (2 * t);
    }
    return k;
  }
  bool ok(const Tup &x, const Tup &y) const {
    for (int t = 0; t < nj; t++)
      if ((((x[ja[t]] + x[jc[t]] + y[ja[t]] + y[jc[t]]) % 4) + 4) % 4 != want[t]) return false;
    return true;
  }
};

struct NormBits {
  std::vector<uint8_t> has;
  void build(int L, int a, int b, int m, int tgt) {
    has.assign(tgt + 1, 0);
    auto K = class_tuples(L, a, m), R = class_tuples(L, b, m);
    std::vector<int> rn;
    for (auto &r : R) rn.push_back(sqn(r));
    std::sort(rn.begin(), rn.end());
    rn.erase(std::unique(rn.begin(), rn.end()), rn.end());
    for (auto &k : K) {
      int kn = sqn(k);
      if (kn > tgt) continue;
      for (int r : rn) { if (kn + r > tgt) break; has[kn + r] = 1; }
    }
  }
  bool ok(int v) const { return v >= 0 && v < (int)has.size() && has[v]; }
};

// The complement side's achievable (eq-2.11b tuple, norm) keys (PairAutoSet,
// wz_match.cpp:324-350), with the eq-2.12 witness filter on that side.
struct AutoKeys {
  std::unordered_set<uint64_t> keys;
  void build(int L, int a, int b, int m, int n, bool t212, bool abSide) {
    int tgt = 4 * n + 2, half = m / 2;
    auto K = class_tuples(L, a, m), R = class_tuples(L, b, m);
    Eq212 e(m, n, abSide);
    struct Side { int norm, key; int ca[3]; };
    auto prep = [&](const std::vector<Tup> &V) {
      std::vector<Side> out(V.size());
      for (size_t i = 0; i < V.size(); i++) {
        out[i].norm = sqn(V[i]);
        out[i].key = e.residue_key(V[i]);
        for (int s = 1; s <= half; s++) out[i].ca[s - 1] = acorr(V[i], s) + acorr(V[i], m - s);
      }
      return out;
    };
    auto KS = prep(K), RS = prep(R);
    std::unordered_map<int, std::vector<int>> rb;   // R indices per residue key
    for (size_t i = 0; i < R.size(); i++) rb[t212 ? RS[i].key : 0].push_back((int)i);
    keys.reserve(1 << 16);
    int T[3];
    for (size_t i = 0; i < K.size(); i++) {
      if (KS[i].norm > tgt) continue;
      auto it = rb.find(t212 ? e.partner_key(K[i]) : 0);
      if (it == rb.end()) continue;
      for (int j : it->second) {
        int nn = KS[i].norm + RS[j].norm;
        if (nn > tgt) continue;
        for (int s = 0; s < half; s++) T[s] = KS[i].ca[s] + RS[j].ca[s];
        keys.insert(pack_auto(T, half, nn));
      }
    }
  }
  bool has(const int *T, int nT, int need) const { return keys.count(pack_auto(T, nT, need)) != 0; }
};

// Integer cell key ordered exactly like the reference's string key "v,v,...,|v,v,...,"
// (wz_match.cpp:1665-1671) under std::string operator<. Every entry is the text "v," and
// those texts are prefix-free (the comma ends each), so comparing two keys is comparing their
// entry sequences lexicographically, each entry by the string order of its text. An entry's
// rank is its text's position among the texts of -31..31 (cells have |v| <= 11 for n <= 63;
// the quad switch Q maps a cell to entries of up to twice that, e.g. -16..16 at n = 44, where
// "-1," < "-10," < "-16," < "-2," < ... < "0," < "1," < "10," < ... < "9,"). Six bits per entry,
// 12 entries at m = 6: a 128-bit key.
struct ElemRank {
  uint8_t r[63];
  ElemRank() {
    std::vector<std::pair<std::string, int>> t;
    for (int v = -31; v <= 31; v++) t.push_back({std::to_string(v) + ",", v});
    std::sort(t.begin(), t.end());
    for (size_t i = 0; i < t.size(); i++) r[t[i].second + 31] = (uint8_t)i;
  }
};
const ElemRank kElemRank;
using CellKey = unsigned __int128;
struct CellKeyHash {
  size_t operator()(CellKey k) const {
    return std::hash<uint64_t>()((uint64_t)k ^ ((uint64_t)(k >> 64) * 0x9E3779B97F4A7C15ull));
  }
};
CellKey cell_rank_key(const int8_t *px, const int8_t *py, int m) {
  CellKey k = 0;
  for (int c = 0; c < m; c++) k = (k << 6) | kElemRank.r[px[c] + 31];
  for (int c = 0; c < m; c++) k = (k << 6) | kElemRank.r[py[c] + 31];
  return k;
}
std::string cell_string(const int8_t *px, const int8_t *py, int m) {
  std::string k;
  for (int c = 0; c < m; c++) { k += std::to_string((int)px[c]); k += ','; }
  k += '|';
  for (int c = 0; c < m; c++) { k += std::to_string((int)py[c]); k += ','; }
  return k;
}
long long prof_score(const HostCell &p, int m) {
  long long s = 0;
  for (int c = 0; c < m; c++) s += std::abs(p.px[c]) + std::abs(p.py[c]);
  return s;
}

}  // namespace

// survive_profiles (wz_match.cpp:439-487) for one side of the join: this side's m = 3
// profiles (px, py) that the complement side's norm sets admit, with the mod-6 tighten
// (some lift to m = 6 passes eq 2.12 and, with 2.11b, cancels its residue
// autocorrelations against the complement's achievable tuples), in px-major order.
// ab_side: the A,B side (length n+1, sigs a, b; complement C,D), else the C,D side.
int side_profiles3(const wzfh_config &c, int n, bool ab_side, std::vector<HostCell> &out) {
  const int tgt = 4 * n + 2;
  const bool t212 = c.thm212 != 0 || n >= 36, t211b = c.thm211b != 0 || n >= 36;
  const int L = ab_side ? n + 1 : n, Lc = ab_side ? n : n + 1;
  const int sx = ab_side ? c.a : c.c, sy = ab_side ? c.b : c.d;
  const int cx = ab_side ? c.c : c.a, cy = ab_side ? c.d : c.b;
  NormBits cmp3, cmp6;
  cmp3.build(Lc, cx, cy, 3, tgt);
  cmp6.build(Lc, cx, cy, 6, tgt);
  AutoKeys cmp6a;
  if (t211b) cmp6a.build(Lc, cx, cy, 6, n, t212, !ab_side);
  auto PX3 = class_tuples(L, sx, 3), PY3 = class_tuples(L, sy, 3);
  auto PX6 = class_tuples(L, sx, 6), PY6 = class_tuples(L, sy, 6);
  Eq212 e3(3, n, ab_side), e6(6, n, ab_side);
  int T[3];
  for (auto &px : PX3) {
    int npx = sqn(px);
    if (npx > tgt) continue;
    for (auto &py : PY3) {
      int need = tgt - npx - sqn(py);
      if (!cmp3.ok(need)) continue;
      if (t212 && !e3.ok(px, py)) continue;
      bool ok6 = false;
      for (auto &px6 : PX6) {
        if (reduce63(px6) != px) continue;
        int npx6 = sqn(px6);
        if (npx6 > tgt) continue;
        for (auto &py6 : PY6) {
          if (reduce63(py6) != py) continue;
          int need6 = tgt - npx6 - sqn(py6);
          if (need6 < 0) continue;
          if (t212 && !e6.ok(px6, py6)) continue;
          if (t211b) {
            for (int s = 1; s <= 3; s++) T[s - 1] = -side_auto(px6, py6, s);
            if (cmp6a.has(T, 3, need6)) { ok6 = true; break; }
          } else if (cmp6.ok(need6)) { ok6 = true; break; }
        }
        if (ok6) break;
      }
      if (!ok6) continue;
      HostCell h{};
      for (int q = 0; q < 3; q++) { h.px[q] = px[q]; h.py[q] = py[q]; }
      out.push_back(h);
      if (out.size() >= 20000000) return WZFH_ETRUNC;
    }
  }
  return WZFH_OK;
}

// Hall tables: cos / sin of i*j*pi/100 in FP64 from the host libm, exactly as
// init_hall_tables computes them (wz_match.cpp:106-114).
static void fill_hall_tables(wzfh_problem &P) {
  std::memset(P.hcos, 0, sizeof P.hcos);
  std::memset(P.hsin, 0, sizeof P.hsin);
  for (int j = 1; j <= 200; j++) {
    double th = j * M_PI / 100.0;
    for (int i = 0; i < P.L1 && i < kMaxL; i++) { P.hcos[j][i] = std::cos(i * th); P.hsin[j][i] = std::sin(i * th); }
  }
}

int build_problem(wzfh_problem &P) {
  const wzfh_config &c = P.cfg;
  const int n = P.n, L1 = P.L1, m = P.m, tgt = 4 * n + 2;
  const bool t212 = c.thm212 != 0, t211b = c.thm211b != 0;

  if (P.ab_side) {   // the A,B side of the join (WZ_JOIN22): its m = 3 profiles only
    std::vector<HostCell> ab;
    int rc = side_profiles3(c, n, true, ab);
    if (rc != WZFH_OK) return rc;
    P.cells.swap(ab);
    P.kept.assign(P.cells.size(), 1);
    P.cell_rows.assign(P.cells.size(), -1);
    fill_hall_tables(P);
    return WZFH_OK;
  }

  NormBits ab3, ab6;
  ab3.build(L1, c.a, c.b, 3, tgt);
  ab6.build(L1, c.a, c.b, 6, tgt);
  AutoKeys ab6a;
  if (t211b) ab6a.build(L1, c.a, c.b, 6, n, t212, true);

  std::vector<HostCell> cells;
  auto push = [&](const Tup &px, const Tup &py) {
    HostCell h{};
    for (int q = 0; q < (int)px.size(); q++) { h.px[q] = px[q]; h.py[q] = py[q]; }
    cells.push_back(h);
  };
  if (c.m6) {
    // survive_profiles6 (wz_match.cpp:792-837), C,D side, in the reference's
    // px-major / py-minor order.
    auto PX = class_tuples(n, c.c, 6), PY = class_tuples(n, c.d, 6);
    Eq212 e(6, n, false);
    struct YInfo { int norm, norm3; };
    std::vector<YInfo> yi(PY.size());
    std::unordered_map<int, std::vector<int>> yb;
    for (size_t j = 0; j < PY.size(); j++) {
      yi[j] = {sqn(PY[j]), sqn(reduce63(PY[j]))};
      yb[t212 ? e.residue_key(PY[j]) : 0].push_back((int)j);
    }
    int T[3];
    for (auto &px : PX) {
      int npx = sqn(px);
      if (npx > tgt) continue;
      int npx3 = sqn(reduce63(px));
      if (npx3 > tgt) continue;
      auto it = yb.find(t212 ? e.partner_key(px) : 0);
      if (it == yb.end()) continue;
      for (int j : it->second) {
        int need = tgt - npx - yi[j].norm;
        if (!ab6.ok(need)) continue;
        int need3 = tgt - npx3 - yi[j].norm3;
        if (!ab3.ok(need3)) continue;
        if (t211b) {
          for (int s = 1; s <= 3; s++) T[s - 1] = -side_auto(px, PY[j], s);
          if (!ab6a.has(T, 3, need)) continue;
        }
        push(px, PY[j]);
        if (cells.size() >= 20000000) return WZFH_ETRUNC;
      }
    }
    if (c.endpos) {   // ENDPOS cell parity filter (wz_match.cpp:1598-1622)
      std::vector<HostCell> keep;
      for (auto &p : cells) {
        int s = 0;
        for (int q = 0; q < 6; q++) s += p.px[q] + p.py[q];
        int mid = (n % 2 == 1) ? p.px[(n / 2) % 6] + p.py[(n / 2) % 6] : 0;
        if ((((s - mid) % 4) + 4) % 4 == 0) keep.push_back(p);
      }
      cells.swap(keep);
    }
  } else {
    int rc = side_profiles3(c, n, false, cells);
    if (rc != WZFH_OK) return rc;
  }

  // WZ_FH_PROF_ORDER 1/2 (wz_match.cpp:1633-1638): std::sort on the same input
  // sequence with the same comparator gives the same permutation as the reference.
  if (c.order == 1)
    std::sort(cells.begin(), cells.end(), [m](const HostCell &x, const HostCell &y) { return prof_score(x, m) < prof_score(y, m); });
  else if (c.order == 2)
    std::sort(cells.begin(), cells.end(), [m](const HostCell &x, const HostCell &y) { return prof_score(x, m) > prof_score(y, m); });

  P.kept.assign(cells.size(), 1);
  P.qp_active = false;
  if (c.canon) {
    // Orbit canonicalization (wz_match.cpp:1712-1808): the kept cell of an orbit is its member
    // with the smallest key, unless the Q-closure prune proves the orbit empty.
    bool q_safe = c.endpos || ((n % 2 == 0) && (((c.c + c.d) % 4 + 4) % 4 == 0));
    bool orbit_q = (c.orbit_q == 2) || (c.orbit_q == 1 && q_safe);
    int nvar = orbit_q ? 64 : 32;
    auto revp = [&](const int8_t *p, int8_t *q) { for (int k = 0; k < m; k++) q[((n - 1 - k) % m + m) % m] = p[k]; };
    std::vector<CellKey> own(cells.size()), oid(cells.size());
    std::unordered_map<CellKey, CellKey, CellKeyHash> orbit_min;
    orbit_min.reserve(cells.size());
    for (size_t i = 0; i < cells.size(); i++) {
      own[i] = cell_rank_key(cells[i].px, cells[i].py, m);
      CellKey best = ~(CellKey)0;
      for (int var = 0; var < nvar; var++) {
        int8_t x[8], y[8], t[8];
        std::memcpy(x, cells[i].px, 8); std::memcpy(y, cells[i].py, 8);
        if (var & 32) {   // quad switch Q (wz_match.cpp:1704-1711)
          int8_t rx[8], ry[8], nx[8], ny[8];
          revp(x, rx); revp(y, ry);
          for (int k = 0; k < m; k++) { nx[k] = (int8_t)((x[k] + y[k] + rx[k] - ry[k]) / 2); ny[k] = (int8_t)((x[k] + y[k] - rx[k] + ry[k]) / 2); }
          std::memcpy(x, nx, 8); std::memcpy(y, ny, 8);
        }
        if (var & 1) for (int k = 0; k < m; k++) x[k] = (int8_t)-x[k];
        if (var & 2) for (int k = 0; k < m; k++) y[k] = (int8_t)-y[k];
        if (var & 4) { revp(x, t); std::memcpy(x, t, 8); }
        if (var & 8) { revp(y, t); std::memcpy(y, t, 8); }
        const CellKey k = (var & 16) ? cell_rank_key(y, x, m) : cell_rank_key(x, y, m);
        best = std::min(best, k);
      }
      oid[i] = best;
      auto it = orbit_min.find(best);
      if (it == orbit_min.end() || own[i] < it->second) orbit_min[best] = own[i];
    }
    // Q-closure orbit elimination (WZ_FH_ORBIT_QPRUNE, wz_match.cpp:1735-1800): the list holds
    // every cell that can hold a solution, and Q maps solutions to solutions, so a cell whose
    // Q image is missing from the list holds none -- and neither does any member of its orbit --
    // when the image is certified impossible: not integral, a class sum out of range or of the
    // wrong parity, or failing the C,D-side eq 2.12 (with Thm 2.12 on). Any other missing image
    // proves nothing.
    std::unordered_set<CellKey, CellKeyHash> qp_dead;
    if (orbit_q && c.orbit_qprune && m == 6) {
      P.qp_active = true;
      std::unordered_set<CellKey, CellKeyHash> listed(own.begin(), own.end());
      int cnt[8];
      for (int k = 0; k < m; k++) cnt[k] = class_count(n, k, m);
      auto realizable = [&](const int8_t *p) {
        for (int k = 0; k < m; k++)
          if (std::abs(p[k]) > cnt[k] || ((p[k] + cnt[k]) & 1)) return false;
        return true;
      };
      const Eq212 eq212cd(m, n, false);
      long long *st = P.qp_stats;
      for (int k = 0; k < 7; k++) st[k] = 0;
      for (size_t i = 0; i < cells.size(); i++) {
        const int8_t *x = cells[i].px, *y = cells[i].py;
        int8_t rx[8], ry[8], qx[8] = {0}, qy[8] = {0};
        revp(x, rx); revp(y, ry);
        bool integral = true;
        for (int k = 0; k < m; k++) {
          if ((x[k] + y[k] + rx[k] - ry[k]) & 1) integral = false;
          qx[k] = (int8_t)((x[k] + y[k] + rx[k] - ry[k]) / 2);
          qy[k] = (int8_t)((x[k] + y[k] - rx[k] + ry[k]) / 2);
        }
        if (integral && listed.count(cell_rank_key(qx, qy, m))) continue;
        st[0]++;
        if (!integral) st[3]++;
        else if (!realizable(qx) || !realizable(qy)) st[1]++;
        else if (c.thm212 && !eq212cd.ok(Tup(qx, qx + m), Tup(qy, qy + m))) st[2]++;
        else { st[4]++; continue; }
        qp_dead.insert(oid[i]);
      }
      st[5] = (long long)qp_dead.size();
      for (size_t i = 0; i < cells.size(); i++) st[6] += qp_dead.count(oid[i]);
    }
    auto keep_of = [&](size_t i) { return orbit_min[oid[i]] == own[i] && !qp_dead.count(oid[i]); };
    if (c.order == 3) {   // Pass H total order (wz_match.cpp:1816-1856)
      std::unordered_map<CellKey, long long, CellKeyHash> omin;
      for (size_t i = 0; i < cells.size(); i++) {
        long long sc = prof_score(cells[i], m);
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
      std::vector<HostCell> sc(cells.size());
      std::vector<CellKey> own2(cells.size()), oid2(cells.size());
      for (size_t i = 0; i < perm.size(); i++) { sc[i] = cells[perm[i]]; own2[i] = own[perm[i]]; oid2[i] = oid[perm[i]]; }
      cells.swap(sc); own.swap(own2); oid.swap(oid2);
      std::unordered_set<CellKey, CellKeyHash> uniq(own.begin(), own.end());
      if (uniq.size() != own.size()) { set_error("duplicate cell keys in the raw list"); return WZFH_EINVAL; }
      uint64_t dl = 1469598103934665603ULL, dk = 1469598103934665603ULL;
      auto fnv = [](uint64_t h, const std::string &k) { for (unsigned char ch : k) { h ^= ch; h *= 1099511628211ULL; } return h; };
      for (size_t i = 0; i < cells.size(); i++) {
        dl = fnv(dl, cell_string(cells[i].px, cells[i].py, m));
        dk = fnv(dk, keep_of(i) ? "1" : "0");
      }
      char buf[64];
      std::snprintf(buf, sizeof buf, "%016llx:%016llx", (unsigned long long)dl, (unsigned long long)dk);
      P.digest = buf;
    }
    for (size_t i = 0; i < cells.size(); i++) P.kept[i] = keep_of(i);
  } else if (c.order == 3) {
    set_error("PROF_ORDER=3 requires WZ_FH_ORBIT_CANON=1");
    return WZFH_EINVAL;
  }
  P.cells.swap(cells);

  // A,B profile rows per cell key (wz_match.cpp:2318-2378).
  P.cell_rows.assign(P.cells.size(), -1);
  if (P.ab_prof) {
    const int half_m = m / 2;
    std::unordered_map<uint64_t, int> keyidx;
    std::vector<uint64_t> ckey(P.cells.size());
    std::vector<std::vector<AbRow>> lists;
    int T[3];
    for (size_t i = 0; i < P.cells.size(); i++) {
      Tup px(P.cells[i].px, P.cells[i].px + m), py(P.cells[i].py, P.cells[i].py + m);
      int need = tgt - sqn(px) - sqn(py);
      for (int s = 1; s <= half_m; s++) T[s - 1] = -side_auto(px, py, s);
      ckey[i] = pack_auto(T, half_m, need);
      if (P.kept[i] && !keyidx.count(ckey[i])) { keyidx.emplace(ckey[i], (int)lists.size()); lists.emplace_back(); }
    }
    struct ST { Tup v; int norm; int ca[3]; };
    auto tups = [&](int sig) {
      std::vector<ST> out;
      for (int sgn = 1; sgn >= (sig ? -1 : 1); sgn -= 2)
        for (auto &v : class_tuples(L1, sgn * sig, m)) {
          ST t; t.v = v; t.norm = sqn(v);
          if (t.norm > tgt) continue;
          for (int s = 1; s <= half_m; s++) t.ca[s - 1] = acorr(v, s) + acorr(v, m - s);
          out.push_back(std::move(t));
        }
      return out;
    };
    auto Kt = tups(c.a), Rt = tups(c.b);
    Eq212 e(m, n, true);
    std::unordered_map<int, std::vector<int>> rb;
    for (size_t j = 0; j < Rt.size(); j++) rb[e.residue_key(Rt[j].v)].push_back((int)j);
    // The reference visits (kt, rt) in Kt-major order and appends matching rows; the
    // bucketed visit keeps rt ascending inside each kt, so each list keeps its order.
    for (auto &kt : Kt) {
      auto it = rb.find(e.partner_key(kt.v));
      if (it == rb.end()) continue;
      for (int j : it->second) {
        const ST &rt = Rt[j];
        int nn = kt.norm + rt.norm;
        if (nn > tgt) continue;
        for (int s = 0; s < half_m; s++) T[s] = kt.ca[s] + rt.ca[s];
        auto f = keyidx.find(pack_auto(T, half_m, nn));
        if (f == keyidx.end()) continue;
        AbRow row{};
        for (int q = 0; q < m; q++) { row.k[q] = kt.v[q]; row.r[q] = rt.v[q]; }
        lists[f->second].push_back(row);
      }
    }
    for (auto &l : lists) {
      P.row_off.push_back((int)P.rows.size());
      P.row_len.push_back((int)l.size());
      P.rows.insert(P.rows.end(), l.begin(), l.end());
    }
    for (size_t i = 0; i < P.cells.size(); i++) {
      auto it = keyidx.find(ckey[i]);
      P.cell_rows[i] = (it == keyidx.end() || P.row_len[it->second] == 0) ? -1 : it->second;
    }
  }

  fill_hall_tables(P);
  return WZFH_OK;
}

}  // namespace wzfh
