// wzfh_api.hip -- host runtime and C ABI of libwzfh.so.
//
// Owns the device copies of a problem (cells, rows, hall tables, A1 tables), the
// per-call work buffers and the HIP stream, launches the stream and completer
// kernels, and turns per-item counts into the reference's exact global order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "wzfh_internal.h"
#include "wzfh_kernels.h"

namespace wzfh {

hipError_t launch_stream(const StreamArgs &a, int blocks, hipStream_t st);
hipError_t launch_stream_finalize(const StreamArgs &a, uint64_t *cell_tot, unsigned long long ncands, KeptCand *kept,
                                  unsigned long long *nkept, hipStream_t st);
hipError_t launch_complete(const CompleteArgs &a, const uint8_t *kab, int blocks, bool lds, size_t lds_bytes,
                           hipStream_t st);
size_t complete_lds_bytes(int L, int words);
hipError_t launch_complete_reduce(const CompleteArgs &a, int nent, int *fv, long long *fn, uint64_t *fA,
                                  uint64_t *fB, int *fj, long long *fcc, long long *fcur, hipStream_t st);
hipError_t complete_set_lds_limit(size_t bytes);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);

// ---------------------------------------------------------------------------
// Allocation-free completer (wzfh_complete_noalloc): every per-call device buffer of that entry
// point is carved from this file-scope array. The loader gives each GPU its own copy, held
// until the process ends; its address on a device comes from hipGetSymbolAddress (in host code
// the name `scratch` is only a host-side stand-in, never a device address).
static __device__ __align__(256) unsigned char scratch[WZFH_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[256];   // per device ordinal: the array's device address

// Looks up (once per device, with that device current) the device address of `scratch`.
static int scratch_address(int device, void **out) {
  if (device < 0 || device >= 256) return WZFH_EINVAL;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return WZFH_EDEVICE;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return WZFH_OK;
}

static thread_local std::string g_err;
std::string set_error(const std::string &msg) { g_err = msg; return msg; }

#define WZ_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                    \
      return WZFH_EDEVICE;                                                             \
    }                                                                                  \
  } while (0)

// A device buffer that only grows. It grows by at least half its size (rounded up to 2 MiB above
// 2 MiB), so a server whose launches vary in size reallocates rarely, and the buffer it outgrows
// is kept until release(): hipFree waits for the device to finish its queued work, which would
// stall every other stream's pipeline in the middle of serving (the kept buffers add at most
// twice the final size).
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  std::vector<void *> retired;
  int ensure(size_t want) {
    if (want <= bytes) return WZFH_OK;
    if (p) retired.push_back(p);
    p = nullptr;
    size_t b = std::max<size_t>(want, 256);
    if (bytes > 0) b = std::max(b, bytes + bytes / 2);
    if (b > (2u << 20)) b = (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) { p = nullptr; set_error("hipMalloc failed"); return WZFH_ENOMEM; }
    bytes = b;
    return WZFH_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    for (void *q : retired) (void)hipFree(q);
    retired.clear();
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return (T *)p; }
};

struct DeviceState {
  int device = 0, num_cu = 0;
  size_t complete_lds_max = 0;     // dynamic LDS bytes a completer block may use
  wzfh_tuning tune{16, 1, 4, 4, 64, 2, 0, 14, 0.0, 0.0};   // launch settings (speed only)
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // problem constants
  DBuf hcos, hsin, order, rej, outer2, outer3, rows, kab, hall_frag;
  StreamArgs base{};                       // tables + constants of the stream kernel
  std::vector<uint8_t> angle_order;        // current evaluation order (angle j-1)
  std::vector<unsigned long long> angle_hist;
  std::vector<uint8_t> kab_host;
  // per-call buffers
  DBuf cell_res, cell_rootok, cands, ctr, cell_tot, kept;
  // work-item records (static items, then donated ones) and per-cell walk state
  DBuf it_nodes, it_noff, it_loff, it_coff, it_leaves, it_cands, it_state, it_next, it_key, it_meta, it_cell, it_tail;
  DBuf dq, cell_lock, cell_crossed, cell_wp, cell_pn, cell_pl, cell_pc, cell_open, cell_next, cell_sp;
  uint32_t epoch = 0;              // launch epoch tagging record states
  DBuf c_cd, c_roff, c_rlen, c_verdict, c_nodes, c_A, c_B, c_charges, c_icharged, c_ebase, c_epath, c_ebud, c_eskip,
       c_fv, c_fn, c_fA, c_fB, c_fj, c_fcc, c_fcur, c_scratch, c_depth;
  double last_ms[2] = {0, 0};
  long long last_units[2] = {0, 0};
  // running totals since the last reset: device time of the stream kernel (kind 0) and of the
  // completer kernel without its reduce (kind 1), their units and launch counts
  std::mutex tot_mu;
  double tot_ms[2] = {0, 0};
  long long tot_units[2] = {0, 0}, tot_launches[2] = {0, 0};
  void *noalloc = nullptr;         // this device's copy of `scratch` (allocation-free completer)
  DBuf telem;                      // stream-kernel counters (WZFH_TELEMETRY=1)
  unsigned long long last_telem[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  long long walked = 0;            // completer nodes walked by all items of a run (WZFH_TELEMETRY=1)
  long long last_walked = 0;
  double hall_eps = 0.0;           // margin of the matrix-core leaf test
  // the last stream result, reused when the caller asks again for the same cells and
  // options (count first, then fetch) so the kernel runs once
  std::vector<int> memo_cells;
  wzfh_stream_opts memo_opts{};
  std::vector<wzfh_cell_stats> memo_stats;
  std::vector<uint64_t> memo_cd;
  std::vector<long long> memo_offs;
  bool memo_valid = false;
};

static int quad_bits(const int *v) {   // sign bits x1,y1,x2,y2 -> bit3..bit0 (set = -1)
  return ((v[0] < 0) << 3) | ((v[1] < 0) << 2) | ((v[2] < 0) << 1) | (v[3] < 0);
}

static int class_count(int L, int c, int m) { return c < L ? (L - c + m - 1) / m : 0; }

// Stream-kernel constants of a problem: child tables in DFS digit order (with
// WZ_FH_STREAM_REV and the C,D root pins folded in), per-depth class info.
static void build_stream_tables(const wzfh_problem &P, StreamArgs &a) {
  const int L = P.side_len(), m = P.m, half = L / 2;
  const bool rev = P.cfg.rev != 0, endpos = P.cfg.endpos != 0;
  a.L = L; a.half = half; a.m = m;
  a.odd = L & 1;
  a.cm_mid = half % m;
  a.nch0 = (endpos || P.ab_side) ? 8 : 16;
  int tot[8];
  for (int c = 0; c < m; c++) tot[c] = class_count(L, c, m);
  for (int d = 0; d < 32; d++) {
    if (d >= half) { a.depth_info[d] = 0; continue; }
    const int c1 = d % m, c2 = (L - 1 - d) % m;
    auto rem_after = [&](int c) {   // remaining positions of class c once depths 0..d are placed
      int placed = 0;
      for (int i = c; i < L; i += m) if (i <= d || i >= L - 1 - d) placed++;
      return tot[c] - placed;
    };
    a.depth_info[d] = (uint32_t)c1 | ((uint32_t)c2 << 8) | ((uint32_t)rem_after(c1) << 16) | ((uint32_t)rem_after(c2) << 24);
  }
  // root quads: P22_16 on the C,D side (P22_POS with ENDPOS), P22_NEG on the A,B side
  // (count_pairs22, wz_match.cpp:766-767)
  auto quad_of = [&](int d, int b) -> const int * {
    if (d == 0) {
      const int qi = rev ? a.nch0 - 1 - b : b;
      if (P.ab_side) return kQuads.neg[qi];
      return endpos ? kQuads.pos[qi] : kQuads.all[qi];
    }
    return kQuads.pos[rev ? 7 - b : b];
  };
  for (int b = 0; b < 8; b++) a.quad8[b] = (uint8_t)quad_bits(quad_of(1, b));
  a.root_pin_mask = 0;
  for (int b = 0; b < 16; b++) {
    if (b >= a.nch0) { a.quad16[b] = 0; a.root_pin_mask |= 1 << b; continue; }
    const int *v = quad_of(0, b);
    a.quad16[b] = (uint8_t)quad_bits(v);
    const bool pinX = P.ab_side ? P.cfg.a == 0 : P.pinC, pinY = P.ab_side ? P.cfg.b == 0 : P.pinD;
    if ((pinX && v[0] != 1) || (pinY && v[1] != 1)) a.root_pin_mask |= 1 << b;
  }
  auto pm_bit = [](int v) { return v > 0 ? 0 : 1; };
  auto sum_bit = [](int s) { return s == 2 ? 0 : (s == 0 ? 1 : 2); };
  for (int idx = 0; idx < 256; idx++) {
    const int mx1 = idx & 3, my1 = (idx >> 2) & 3, mx2 = (idx >> 4) & 3, my2 = (idx >> 6) & 3;
    unsigned m8 = 0, m16 = 0;
    for (int b = 0; b < 16; b++) {
      for (int depth = 0; depth < 2; depth++) {
        if (depth == 1 && b >= 8) continue;
        if (depth == 0 && b >= a.nch0) continue;
        const int *v = quad_of(depth == 0 ? 0 : 1, b);
        bool ok = ((mx1 >> pm_bit(v[0])) & 1) && ((my1 >> pm_bit(v[1])) & 1) && ((mx2 >> pm_bit(v[2])) & 1) &&
                  ((my2 >> pm_bit(v[3])) & 1);
        if (!ok) continue;
        if (depth == 1) m8 |= 1u << b;
        else if (!((a.root_pin_mask >> b) & 1)) m16 |= 1u << b;
      }
    }
    a.tab_ne[idx] = (uint8_t)m8;
    a.tab0_ne[idx] = (uint16_t)m16;
  }
  for (int idx = 0; idx < 64; idx++) {
    const int sx = idx & 7, sy = (idx >> 3) & 7;
    unsigned m8 = 0, m16 = 0;
    for (int b = 0; b < 16; b++) {
      for (int depth = 0; depth < 2; depth++) {
        if (depth == 1 && b >= 8) continue;
        if (depth == 0 && b >= a.nch0) continue;
        const int *v = quad_of(depth == 0 ? 0 : 1, b);
        bool ok = ((sx >> sum_bit(v[0] + v[2])) & 1) && ((sy >> sum_bit(v[1] + v[3])) & 1);
        if (!ok) continue;
        if (depth == 1) m8 |= 1u << b;
        else if (!((a.root_pin_mask >> b) & 1)) m16 |= 1u << b;
      }
    }
    a.tab_eq[idx] = (uint8_t)m8;
    a.tab0_eq[idx] = (uint16_t)m16;
  }
  a.lim = 4.0 * P.n + 2.0 + 0.5;
  a.hall_nkb = (a.L + 31) / 32;
}

// MFMA A-operand fragments of the leaf-test angle table (hall_mfma in wzfh_stream.hip):
// entry (((g*2 + t)*2 + kb)*2 + part)*64 + lane holds 8 fp16 values of row a = 16g + lane%16
// (angle j = a+1 for a < 100, j = 200 for a = 100, zero rows above), positions
// k = 32kb + 8*(lane/16) + e, of cos (t = 0) or sin (t = 1): part 0 = fp16(v), part 1 =
// fp16(v - part 0). Returns max |v - hi - lo| over the table (exact in FP64).
static double build_hall_frag(const wzfh_problem &P, std::vector<_Float16> &out) {
  constexpr int kGroups = 7;
  out.assign((size_t)kGroups * 2 * 2 * 2 * 64 * 8, (_Float16)0.0f);
  double worst = 0.0;
  for (int g = 0; g < kGroups; g++)
    for (int t = 0; t < 2; t++)
      for (int kb = 0; kb < 2; kb++)
        for (int lane = 0; lane < 64; lane++)
          for (int e = 0; e < 8; e++) {
            const int row = 16 * g + (lane & 15), k = 32 * kb + 8 * (lane >> 4) + e;
            double v = 0.0;
            if (row <= 100) {
              const int j = row < 100 ? row + 1 : 200;
              v = t == 0 ? P.hcos[j][k] : P.hsin[j][k];
            }
            const _Float16 hi = (_Float16)v;
            const _Float16 lo = (_Float16)(v - (double)hi);
            worst = std::max(worst, std::fabs((v - (double)hi) - (double)lo));
            out[((((size_t)(g * 2 + t) * 2 + kb) * 2 + 0) * 64 + lane) * 8 + e] = hi;
            out[((((size_t)(g * 2 + t) * 2 + kb) * 2 + 1) * 64 + lane) * 8 + e] = lo;
          }
  return worst;
}

// Margin of the matrix-core leaf test. A lane's sum r = sum_k s_k (hi_k + lo_k) differs from
// the exact sum over the FP64 table by at most delta = L*split + acc, where acc bounds the
// FP32 accumulation of 2L exact products (every partial sum is at most L(1 + 2^-11) in
// magnitude; 2^-23 per addition also covers truncating adders). The reference's own FP64
// sums are within 2L^2 * 2^-53 of exact. For the four sums of one (leaf, angle),
// |S' - S| <= 2*delta*sum|r_t| + 4*delta^2 <= 4*delta*sqrt(S) + 4*delta^2 (Cauchy-Schwarz), and
// forming S in FP32 adds at most 8 * S * 2^-23; only S up to lim + 1 matters for either
// certain decision. The margin used is twice that bound.
static void hall_thresholds(int L, double lim, double split_err, float &hi, float &lo, double &eps) {
  const double u23 = std::ldexp(1.0, -23);
  const double acc = 2.0 * L * (L * (1.0 + std::ldexp(1.0, -11))) * u23;
  const double ref = 2.0 * L * L * std::ldexp(1.0, -53);
  const double delta = L * split_err + acc + ref;
  const double smax = lim + 1.0;
  const double bound = 4.0 * delta * std::sqrt(smax) + 4.0 * delta * delta + 8.0 * smax * u23;
  eps = 2.0 * bound;
  hi = std::nextafter((float)(lim + eps), INFINITY);
  lo = std::nextafter((float)(lim - eps), -INFINITY);
}

// Kab(D, s): 2 * number of shift-s pairs of an A,B of length L not yet fully placed
// when D mirror quads are placed (the reference's incremental Kab, wz_match.cpp:1313-1314,
// 1115-1138, in closed form).
static std::vector<uint8_t> build_kab(int L) {
  std::vector<uint8_t> t(32 * 64, 0);
  for (int D = 0; D < 32; D++)
    for (int s = 1; s < 64 && s < L; s++) {
      int both = 0;
      for (int i = 0; i + s < L; i++) {
        const bool pi = i < D || i >= L - D, pj = (i + s) < D || (i + s) >= L - D;
        if (pi && pj) both++;
      }
      t[D * 64 + s] = (uint8_t)(2 * (L - s - both));
    }
  return t;
}

int device_init(wzfh_problem &P) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device: libwzfh has no CPU path");
    return WZFH_ENODEV;
  }
  DeviceState *S = new DeviceState();
  P.dev = S;
  WZ_HIP(hipGetDevice(&S->device));
  hipDeviceProp_t prop;
  WZ_HIP(hipGetDeviceProperties(&prop, S->device));
  S->num_cu = prop.multiProcessorCount;
  // dynamic LDS of the completer: the device's per-block LDS less the kernel's static tables
  S->complete_lds_max = prop.sharedMemPerBlock > 8192 ? prop.sharedMemPerBlock - 8192 : 0;
  if (S->complete_lds_max > 0) WZ_HIP(complete_set_lds_limit(S->complete_lds_max));
  WZ_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  for (auto &e : S->ev) WZ_HIP(hipEventCreate(&e));
  if (scratch_address(S->device, &S->noalloc) != WZFH_OK) {
    set_error("hipGetSymbolAddress of the allocation-free scratch array failed");
    return WZFH_EDEVICE;
  }
  const size_t tb = 201 * 64 * sizeof(double);
  if (S->hcos.ensure(tb) || S->hsin.ensure(tb)) return WZFH_ENOMEM;
  WZ_HIP(hipMemcpy(S->hcos.p, P.hcos, tb, hipMemcpyHostToDevice));
  WZ_HIP(hipMemcpy(S->hsin.p, P.hsin, tb, hipMemcpyHostToDevice));
  // initial angle order: j = 1..100 spread (golden-ratio stride), then 101..200; the
  // order is refined from the first-rejection histogram after every launch
  S->angle_order.clear();
  std::vector<int> seen(201, 0);
  for (int t = 0, j = 0; t < 100; t++) {
    j = (j + 61) % 100;
    while (seen[j + 1]) j = (j + 1) % 100;
    seen[j + 1] = 1;
    S->angle_order.push_back((uint8_t)j);
  }
  for (int j = 101; j <= 200; j++) S->angle_order.push_back((uint8_t)(j - 1));
  S->angle_hist.assign(200, 0);
  if (S->order.ensure(200) || S->rej.ensure(200 * sizeof(unsigned))) return WZFH_ENOMEM;
  if (S->outer2.ensure(sizeof kOuter.t2) || S->outer3.ensure(sizeof kOuter.t3)) return WZFH_ENOMEM;
  WZ_HIP(hipMemcpy(S->outer2.p, kOuter.t2, sizeof kOuter.t2, hipMemcpyHostToDevice));
  WZ_HIP(hipMemcpy(S->outer3.p, kOuter.t3, sizeof kOuter.t3, hipMemcpyHostToDevice));
  if (!P.rows.empty()) {
    if (S->rows.ensure(P.rows.size() * sizeof(AbRow))) return WZFH_ENOMEM;
    WZ_HIP(hipMemcpy(S->rows.p, P.rows.data(), P.rows.size() * sizeof(AbRow), hipMemcpyHostToDevice));
  }
  S->kab_host = build_kab(P.L1);
  if (S->kab.ensure(S->kab_host.size())) return WZFH_ENOMEM;
  WZ_HIP(hipMemcpy(S->kab.p, S->kab_host.data(), S->kab_host.size(), hipMemcpyHostToDevice));
  build_stream_tables(P, S->base);
  {
    std::vector<_Float16> frag;
    const double split_err = build_hall_frag(P, frag);
    if (S->hall_frag.ensure(frag.size() * sizeof(_Float16))) return WZFH_ENOMEM;
    WZ_HIP(hipMemcpy(S->hall_frag.p, frag.data(), frag.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    hall_thresholds(S->base.L, S->base.lim, split_err, S->base.hall_hi, S->base.hall_lo, S->hall_eps);
  }
  return WZFH_OK;
}

static void free_state(DeviceState *S) {
  if (!S) return;
  for (DBuf *b : {&S->telem, &S->hall_frag, &S->hcos, &S->hsin, &S->order, &S->rej, &S->outer2, &S->outer3, &S->rows, &S->kab,
                  &S->cell_res, &S->cell_rootok, &S->cands, &S->ctr, &S->cell_tot, &S->kept,
                  &S->it_nodes, &S->it_noff, &S->it_loff, &S->it_coff, &S->it_leaves, &S->it_cands, &S->it_state,
                  &S->it_next, &S->it_key, &S->it_meta, &S->it_cell, &S->it_tail, &S->dq, &S->cell_lock,
                  &S->cell_crossed, &S->cell_wp, &S->cell_pn, &S->cell_pl, &S->cell_pc, &S->cell_open,
                  &S->cell_next, &S->cell_sp, &S->c_cd, &S->c_roff, &S->c_rlen, &S->c_verdict, &S->c_nodes, &S->c_A,
                  &S->c_B, &S->c_charges, &S->c_icharged, &S->c_ebase, &S->c_epath, &S->c_ebud, &S->c_eskip,
                  &S->c_fv, &S->c_fn, &S->c_fA, &S->c_fB, &S->c_fj, &S->c_fcc, &S->c_fcur, &S->c_scratch,
                  &S->c_depth})
    b->release();
  for (auto &e : S->ev) if (e) (void)hipEventDestroy(e);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
}

void device_free(wzfh_problem &P) {
  free_state(P.dev);
  P.dev = nullptr;
}

// A device state for completer runs only, with its own HIP stream and buffers and copies of P's
// completer tables (profile rows, the A,B bound table, the A1 outer tables): a serving loop runs
// several completion runs of one problem side by side on such states.
static int complete_state_init(const wzfh_problem &P, DeviceState *S) {
  const DeviceState *D = P.dev;
  S->device = D->device;
  S->num_cu = D->num_cu;
  S->complete_lds_max = D->complete_lds_max;
  S->tune = D->tune;
  S->kab_host = D->kab_host;
  WZ_HIP(hipSetDevice(S->device));
  WZ_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  for (auto &e : S->ev) WZ_HIP(hipEventCreate(&e));
  if (S->outer2.ensure(sizeof kOuter.t2) || S->outer3.ensure(sizeof kOuter.t3)) return WZFH_ENOMEM;
  WZ_HIP(hipMemcpy(S->outer2.p, kOuter.t2, sizeof kOuter.t2, hipMemcpyHostToDevice));
  WZ_HIP(hipMemcpy(S->outer3.p, kOuter.t3, sizeof kOuter.t3, hipMemcpyHostToDevice));
  if (!P.rows.empty()) {
    if (S->rows.ensure(P.rows.size() * sizeof(AbRow))) return WZFH_ENOMEM;
    WZ_HIP(hipMemcpy(S->rows.p, P.rows.data(), P.rows.size() * sizeof(AbRow), hipMemcpyHostToDevice));
  }
  if (S->kab.ensure(S->kab_host.size())) return WZFH_ENOMEM;
  WZ_HIP(hipMemcpy(S->kab.p, S->kab_host.data(), S->kab_host.size(), hipMemcpyHostToDevice));
  return WZFH_OK;
}

int complete_state(const wzfh_problem &P, DeviceState **out) {
  *out = nullptr;
  if (!P.dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  DeviceState *S = new DeviceState();
  const int rc = complete_state_init(P, S);
  if (rc != WZFH_OK) { free_state(S); return rc; }
  *out = S;
  return WZFH_OK;
}

void complete_state_free(DeviceState *S) { free_state(S); }

void complete_state_sync_tuning(const wzfh_problem &P, DeviceState *S) {
  if (P.dev && S) S->tune = P.dev->tune;
}

int state_kernel_totals(DeviceState *S, wzfh_kernel_sums *out, int reset) {
  std::lock_guard<std::mutex> lk(S->tot_mu);
  for (int k = 0; k < 2; k++) {
    if (out) { out->ms[k] = S->tot_ms[k]; out->units[k] = S->tot_units[k]; out->launches[k] = S->tot_launches[k]; }
    if (reset) { S->tot_ms[k] = 0; S->tot_units[k] = 0; S->tot_launches[k] = 0; }
  }
  return WZFH_OK;
}

// ---------------------------------------------------------------------------
// Stream: launch + exact post-processing.

static int auto_split(const wzfh_problem &P, const wzfh_stream_opts &o, int ncells) {
  const int half = P.side_len() / 2;
  int d0 = o.split_depth;
  if (d0 <= 0) {
    // static items only seed the launch: running items donate their untried subtrees to
    // idle lanes, so a few hundred paths per cell (fewer for big batches) suffice
    d0 = ncells >= 256 ? 2 : 3;
    d0 = std::max(d0, half - 21);   // an item's digit word holds 21 levels below its base
  }
  return std::max(1, std::min(d0, std::max(1, half)));
}

int run_stream(wzfh_problem &P, const int *cells, int ncells, const wzfh_stream_opts &o,
                      wzfh_cell_stats *stats, StreamOut &out, const wzfh_problem *const *cell_prob, int cell_blocks) {
  DeviceState *S = P.dev;
  WZ_HIP(hipSetDevice(S->device));   // callers may be any host thread (serving workers)
  StreamArgs a = S->base;
  const int d0 = auto_split(P, o, ncells);
  if (P.side_len() / 2 - d0 > 21) { set_error("split depth too shallow: an item spans at most 21 levels"); return WZFH_EINVAL; }
  if (a.nch0 * (double)(1u << std::min(30, 3 * (d0 - 1))) * ncells > 2.0e9) {
    set_error("split depth too deep for this many cells");
    return WZFH_EINVAL;
  }
  const int npaths = a.nch0 * (1 << (3 * (d0 - 1)));
  int window = o.window > 0 ? o.window : ((o.cand_cap > 0 || o.node_budget > 0) ? std::min(npaths, 16) : npaths);
  window = std::max(1, std::min(window, npaths));
  a.d0 = d0; a.npaths = npaths; a.window = window; a.ncells = ncells;
  a.node_budget = o.node_budget; a.cand_cap = o.cand_cap;
  // per-cell residual words and root check
  std::vector<uint64_t> res(2 * (size_t)ncells);
  std::vector<int> rootok(ncells);
  const int L = P.side_len(), m = P.m;
  for (int i = 0; i < ncells; i++) {
    const HostCell &c = (cell_prob ? cell_prob[i] : &P)->cells[cells[i]];
    uint64_t rx = 0, ry = 0;
    bool ok = true;
    for (int q = 0; q < m; q++) {
      const int tot = class_count(L, q, m);
      rx |= (uint64_t)(c.px[q] + 16) << (5 * q);
      ry |= (uint64_t)(c.py[q] + 16) << (5 * q);
      if (std::abs(c.px[q]) > tot || std::abs(c.py[q]) > tot || ((c.px[q] + tot) & 1) || ((c.py[q] + tot) & 1)) ok = false;
    }
    res[2 * i] = rx; res[2 * i + 1] = ry; rootok[i] = ok;
  }
  const size_t nstatic = (size_t)ncells * npaths;
  // donated records: cleared at every launch, so sized to the batch
  const size_t ndonate = std::min<size_t>(1u << 24, std::max<size_t>(1u << 18, 16384 * (size_t)ncells));
  const size_t nrec = nstatic + ndonate;
  if (nrec >= 0xFFFFFFF0ull) { set_error("too many work items for one launch"); return WZFH_ECAPACITY; }
  if (S->cell_res.ensure(res.size() * 8) || S->cell_rootok.ensure(ncells * 4) || S->cell_tot.ensure(ncells * 24) ||
      S->ctr.ensure(64) || S->dq.ensure(64) || S->it_nodes.ensure(nrec * 8) || S->it_noff.ensure(nrec * 8) ||
      S->it_loff.ensure(nrec * 8) || S->it_coff.ensure(nrec * 8) || S->it_leaves.ensure(nrec * 4) ||
      S->it_cands.ensure(nrec * 4) || S->it_state.ensure(nrec * 4) || S->it_next.ensure(nrec * 4) ||
      S->it_key.ensure(nrec * 8) || S->it_meta.ensure(nrec * 4) || S->it_cell.ensure(nrec * 4) ||
      S->it_tail.ensure(nrec * 4) || S->cell_lock.ensure(ncells * 4) || S->cell_crossed.ensure(ncells * 4) ||
      S->cell_wp.ensure(ncells * 4) || S->cell_pn.ensure(ncells * 8) || S->cell_pl.ensure(ncells * 8) ||
      S->cell_pc.ensure(ncells * 8) || S->cell_open.ensure(ncells * 4) || S->cell_next.ensure(ncells * 4) ||
      S->cell_sp.ensure(ncells * 4))
    return WZFH_ENOMEM;
  // candidate buffer: what the caps allow, at most 4M records up front, grown on overflow
  unsigned long long want = o.cand_cap > 0 ? (unsigned long long)ncells * (unsigned long long)(o.cand_cap + 512)
                                           : (1ull << 20);
  unsigned long long cap = std::max<unsigned long long>(std::min<unsigned long long>(want, 1ull << 22),
                                                        S->cands.bytes / sizeof(CandRec));
  for (int attempt = 0; attempt < 8; attempt++) {
    if (S->cands.ensure(cap * sizeof(CandRec)) || S->kept.ensure(cap * sizeof(KeptCand))) return WZFH_ENOMEM;
    WZ_HIP(hipMemcpyAsync(S->cell_res.p, res.data(), res.size() * 8, hipMemcpyHostToDevice, S->stream));
    WZ_HIP(hipMemcpyAsync(S->cell_rootok.p, rootok.data(), ncells * 4, hipMemcpyHostToDevice, S->stream));
    WZ_HIP(hipMemcpyAsync(S->order.p, S->angle_order.data(), 200, hipMemcpyHostToDevice, S->stream));
    WZ_HIP(hipMemsetAsync(S->ctr.p, 0, 64, S->stream));
    WZ_HIP(hipMemsetAsync(S->rej.p, 0, 200 * sizeof(unsigned), S->stream));
    a.cell_res = S->cell_res.as<uint64_t>();
    a.cell_rootok = S->cell_rootok.as<int>();
    a.hcos = S->hcos.as<double>(); a.hsin = S->hsin.as<double>();
    a.hall_frag = S->hall_frag.p;
    a.hall_verify = (getenv("WZFH_HALL_VERIFY") && atoi(getenv("WZFH_HALL_VERIFY"))) ? 1 : 0;
    a.angle_order = S->order.as<uint8_t>();
    a.work_ctr = S->ctr.as<unsigned long long>();
    a.ncands = S->ctr.as<unsigned long long>() + 1;
    a.rec_cap = (uint32_t)nrec;
    a.epoch = ++S->epoch;
    if (S->epoch >= (1u << 29)) S->epoch = 0;   // tags use 30 bits; fresh buffers hold 0
    a.it_nodes = S->it_nodes.as<uint64_t>();
    a.it_noff = S->it_noff.as<uint64_t>();
    a.it_loff = S->it_loff.as<uint64_t>();
    a.it_coff = S->it_coff.as<uint64_t>();
    a.it_leaves = S->it_leaves.as<uint32_t>();
    a.it_cands = S->it_cands.as<uint32_t>();
    a.it_state = S->it_state.as<uint32_t>();
    a.it_next = S->it_next.as<uint32_t>();
    a.it_key = S->it_key.as<uint64_t>();
    a.it_meta = S->it_meta.as<uint32_t>();
    a.it_cell = S->it_cell.as<uint32_t>();
    a.it_tail = S->it_tail.as<uint32_t>();
    a.dq = S->dq.as<unsigned long long>();
    a.cell_lock = S->cell_lock.as<int>();
    a.cell_crossed = S->cell_crossed.as<int>();
    a.cell_wp = S->cell_wp.as<uint32_t>();
    a.cell_pn = S->cell_pn.as<unsigned long long>();
    a.cell_pl = S->cell_pl.as<unsigned long long>();
    a.cell_pc = S->cell_pc.as<unsigned long long>();
    a.cell_open = S->cell_open.as<int>();
    a.cell_next = S->cell_next.as<uint32_t>();
    a.cell_sp = S->cell_sp.as<uint32_t>();
    a.static_win = (uint32_t)S->tune.static_window;
    a.stall_ticks = 500000000ull;   // 5 s of the 100 MHz constant clock

    if (const char *e = getenv("WZFH_STATIC_WINDOW")) a.static_win = (uint32_t)std::max(1, atoi(e));
    a.cands = S->cands.as<CandRec>();
    a.cand_capacity = cap;
    a.angle_rej = S->rej.as<unsigned>();
    a.telem = nullptr;
    if (getenv("WZFH_TELEMETRY") && atoi(getenv("WZFH_TELEMETRY"))) {
      if (S->telem.ensure(80)) return WZFH_ENOMEM;
      WZ_HIP(hipMemsetAsync(S->telem.p, 0, 80, S->stream));
      a.telem = S->telem.as<unsigned long long>();
    }
    // persistent grid: as many waves as can be resident (4 per SIMD at the kernel's
    // register budget); donation keeps them busy once the static items are gone
    // small launches (a few cells) walk each cell with more waves: 4 blocks per CU measured
    // best for one cell (71-80 ms against 90-130 ms at 2); batches take the tuned value
    int bpc = ncells <= 4 ? 4 : S->tune.blocks_per_cu;
    if (const char *e = getenv("WZFH_BLOCKS_PER_CU")) bpc = atoi(e);
    int blocks = std::max(1, S->num_cu * std::max(1, std::min(4, bpc)));
    // at most cell_blocks blocks per cell (WZFH_CELL_BLOCKS overrides), so launches of a few
    // cells leave the rest of the GPU to concurrent launches (serving many n at once)
    long long kcell = cell_blocks;
    if (const char *e = getenv("WZFH_CELL_BLOCKS")) kcell = atoll(e);
    if (kcell > 0) blocks = (int)std::max(1ll, std::min<long long>(blocks, kcell * (long long)ncells));
    // idle polling backoff: waves per cell, so a few cells are not probed by thousands of
    // waves every iteration (the counters are device-scope atomics)
    a.scan_max = std::max(8, std::min(1024, (blocks * 4) / std::max(1, ncells)));
    a.don_moves = 1024;
    a.don_moves_front = 64;
    a.don_hops = (uint32_t)S->tune.donate_hops;
    // a launch of a few cells has the GPU to itself: one more generation of frontier
    // donations spreads each cell over more lanes (one cell: 61-78 ms against 83-104 ms)
    if (ncells <= 4) a.don_hops = std::max(a.don_hops, 2u);
    a.inner_moves = S->tune.inner_moves;
    a.idle_sleep = S->tune.idle_sleep;
    if (const char *e = getenv("WZFH_IDLE_SLEEP")) a.idle_sleep = std::max(0, std::min(256, atoi(e)));
    if (const char *e = getenv("WZFH_INNER_MOVES")) a.inner_moves = std::max(1, std::min(256, atoi(e)));
    a.don_low = (uint32_t)(blocks * 256 / 4);
    if (const char *e = getenv("WZFH_DONATE_MOVES")) a.don_moves = a.don_moves_front = (uint32_t)std::max(1, atoi(e));
    if (const char *e = getenv("WZFH_DONATE_FRONT")) a.don_moves_front = (uint32_t)std::max(1, atoi(e));
    if (const char *e = getenv("WZFH_DONATE_HOPS")) a.don_hops = (uint32_t)std::min(15, std::max(0, atoi(e)));
    a.don_any = (getenv("WZFH_DONATE_ANY") && atoi(getenv("WZFH_DONATE_ANY"))) ? 1 : 0;
    if (const char *e = getenv("WZFH_DONATE_LOW")) a.don_low = (uint32_t)std::max(0, atoi(e));
    a.don_max_level = std::min(21, std::max(0, a.half - 3));
    a.don_keep = 0;
    if (const char *e = getenv("WZFH_DONATE_KEEP")) a.don_keep = std::max(0, std::min(31, atoi(e)));
    WZ_HIP(hipEventRecord(S->ev[0], S->stream));
    WZ_HIP(launch_stream(a, blocks, S->stream));
    WZ_HIP(hipEventRecord(S->ev[1], S->stream));
    unsigned long long ctr[2], dqh[6];
    WZ_HIP(hipMemcpyAsync(ctr, S->ctr.p, 16, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(dqh, S->dq.p, sizeof dqh, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
    if (dqh[5]) {
      set_error(dqh[5] == 2 ? "stream kernel met an invalid work record (internal error)"
                            : "stream kernel made no progress for 5 s and stopped (internal error)");
      return WZFH_EKERNEL;
    }
    float ms = 0;
    WZ_HIP(hipEventElapsedTime(&ms, S->ev[0], S->ev[1]));
    S->last_ms[0] = ms;
    {
      std::lock_guard<std::mutex> lk(S->tot_mu);
      S->tot_ms[0] += ms;
      S->tot_launches[0]++;
    }
    if (a.telem) WZ_HIP(hipMemcpy(S->last_telem, S->telem.p, 80, hipMemcpyDeviceToHost));
    if (ctr[1] > cap) { cap = ctr[1] + (ctr[1] >> 2) + 1024; continue; }
    // ---- exact global order on the device; only the cell totals and kept candidates return
    unsigned long long *nkept = S->ctr.as<unsigned long long>() + 2;
    WZ_HIP(launch_stream_finalize(a, S->cell_tot.as<uint64_t>(), ctr[1], S->kept.as<KeptCand>(), nkept, S->stream));
    unsigned long long nk = 0;
    std::vector<uint64_t> tot(3 * (size_t)ncells);
    std::vector<unsigned> rej(200);
    WZ_HIP(hipMemcpyAsync(&nk, nkept, 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(tot.data(), S->cell_tot.p, tot.size() * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(rej.data(), S->rej.p, 200 * sizeof(unsigned), hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
    for (int cs = 0; cs < ncells; cs++)
      if (tot[3 * cs] == kUnsetOff) {
        set_error("stream: an uncut cell has an unfinished work item");
        return WZFH_EKERNEL;
      }
    std::vector<KeptCand> kc(nk);
    if (nk) WZ_HIP(hipMemcpy(kc.data(), S->kept.p, nk * sizeof(KeptCand), hipMemcpyDeviceToHost));
    // refine the angle order for the next launch (decision-invariant)
    {
      for (int t = 0; t < 200; t++) S->angle_hist[S->angle_order[t]] += rej[t];
      std::vector<uint8_t> ord(S->angle_order);
      std::stable_sort(ord.begin(), ord.end(), [&](uint8_t x, uint8_t y) { return S->angle_hist[x] > S->angle_hist[y]; });
      S->angle_order = ord;
    }
    std::sort(kc.begin(), kc.end(), [](const KeptCand &x, const KeptCand &y) {
      if (x.cell != y.cell) return x.cell < y.cell;
      return x.G < y.G;
    });
    out.cands.clear();
    out.cell_begin.assign(ncells + 1, 0);
    long long units = 0;
    size_t pos = 0;
    for (int cs = 0; cs < ncells; cs++) {
      out.cell_begin[cs] = (long long)out.cands.size();
      const uint64_t total_nodes = tot[3 * cs], total_leaves = tot[3 * cs + 1];
      long long taken = 0;
      uint64_t last_g = 0, last_leaf = 0;
      while (pos < kc.size() && kc[pos].cell == (uint64_t)cs) {
        if (!(o.cand_cap > 0 && taken >= o.cand_cap)) {
          CandRec r;
          r.X = kc[pos].X; r.Y = kc[pos].Y; r.ell = kc[pos].G; r.item = 0; r.leaf = 0;
          out.cands.push_back(r);
          taken++;
          last_g = kc[pos].G; last_leaf = kc[pos].leaf;
        }
        pos++;
      }
      wzfh_cell_stats &stt = stats[cs];
      stt.cand = taken;
      stt.hall_ok = taken;
      stt.reserved = 0;
      if (o.cand_cap > 0 && taken >= o.cand_cap) {
        // the reference stops right after the cap-th candidate's leaf (rec() entry check)
        stt.capped = 1;
        stt.dfs_nodes = (long long)last_g + 1;
        stt.leaves = (long long)last_leaf + 1;
      } else if (o.node_budget > 0 && total_nodes >= (uint64_t)o.node_budget) {
        stt.capped = total_nodes > (uint64_t)o.node_budget ? 1 : 0;
        stt.dfs_nodes = o.node_budget;
        stt.leaves = stt.capped ? -1 : (long long)total_leaves;
      } else {
        stt.capped = 0;
        stt.dfs_nodes = (long long)total_nodes;
        stt.leaves = (long long)total_leaves;
      }
      units += stt.dfs_nodes;
    }
    out.cell_begin[ncells] = (long long)out.cands.size();
    S->last_units[0] = units;
    {
      std::lock_guard<std::mutex> lk(S->tot_mu);
      S->tot_units[0] += units;
    }
    return WZFH_OK;
  }
  set_error("candidate buffer kept overflowing");
  return WZFH_ECAPACITY;
}

// ---------------------------------------------------------------------------
// One completer launch over nent entries (candidates, or subtrees of candidates below a
// quad path; inputs already resolved to row ranges). ext > 0 gives every entry 8^ext work
// items (its quad paths ext levels further down) and reduces them per entry: verdict 0/1/2/4
// is final, -1 marks an entry whose budget runs out inside item j (charges cc and counted
// nodes nd before it, its own charges cj). ext = 0 runs every entry as one item whose budget
// abort is counted exactly. Adds the kernels' time to *ms.
// Result of one work item of a split launch (verdict 0..5, counted nodes, charges, A, B).
struct ItemRes { int v; long long n, c; uint64_t A, B; };

struct CompleteEntry {
  int cand;              // candidate index of the caller's arrays
  long long path;        // quad path above the entry's subtree (3 bits per depth)
  int base;              // depth of the subtree root
  long long budget;      // charges left to the candidate's budget at the subtree start
  long long pre_nodes;   // counted nodes of the candidate before the subtree
  int skip = 0;          // items below this were walked by the capped item the subtree continues
};

// Device buffers of one completer launch: per entry (inputs, reduce outputs), per work item
// (results, charges) and the per-wave row scratch. They come either from the handle's growable
// buffers or, for the allocation-free entry point, from a carve of this device's `scratch`.
struct CompleteBufs {
  uint64_t *cd = nullptr; int *roff = nullptr, *rlen = nullptr, *ebase = nullptr, *eskip = nullptr;
  long long *epath = nullptr, *ebud = nullptr;
  int *verdict = nullptr; long long *nodes = nullptr; uint64_t *A = nullptr, *B = nullptr;
  long long *charges = nullptr, *icharged = nullptr;
  uint64_t *row_scratch = nullptr; unsigned long long *ctr = nullptr;
  int *fv = nullptr; long long *fn = nullptr; uint64_t *fA = nullptr, *fB = nullptr; int *fj = nullptr;
  long long *fcc = nullptr, *fcur = nullptr;
};

// A bump carve of a fixed device range (256-byte aligned pieces); nullptr once it runs out.
struct Carve {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  template <typename T> T *take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at + bytes > cap) return nullptr;
    used = at + bytes;
    return (T *)(base + at);
  }
};

// Per-item and per-entry bytes of a launch in a carve (with the alignment of each piece).
constexpr size_t kItemBytes = 4 + 8 + 8 + 8 + 8 + 8;
constexpr size_t kEntryBytes = 16 + 4 + 4 + 4 + 4 + 8 + 8 + 4 + 8 + 8 + 8 + 4 + 8 + 8;
constexpr size_t kCarveSlack = 25 * 256 + 64;

static int complete_buffers(DeviceState *S, Carve *carve, int nent, size_t nw, size_t row_scratch,
                            CompleteBufs &b) {
  if (!carve) {
    if (S->c_cd.ensure((size_t)nent * 16) || S->c_roff.ensure(nent * 4) || S->c_rlen.ensure(nent * 4) ||
        S->c_ebase.ensure(nent * 4) || S->c_epath.ensure(nent * 8) || S->c_ebud.ensure(nent * 8) ||
        S->c_eskip.ensure(nent * 4) ||
        S->c_verdict.ensure(nw * 4) || S->c_nodes.ensure(nw * 8) || S->c_A.ensure(nw * 8) || S->c_B.ensure(nw * 8) ||
        S->c_charges.ensure(nw * 8) || S->c_icharged.ensure(nw * 8) || S->c_scratch.ensure(row_scratch) ||
        S->ctr.ensure(64) || S->c_fv.ensure(nent * 4) || S->c_fn.ensure(nent * 8) || S->c_fA.ensure(nent * 8) ||
        S->c_fB.ensure(nent * 8) || S->c_fj.ensure(nent * 4) || S->c_fcc.ensure(nent * 8) ||
        S->c_fcur.ensure(nent * 8))
      return WZFH_ENOMEM;
    b.cd = S->c_cd.as<uint64_t>(); b.roff = S->c_roff.as<int>(); b.rlen = S->c_rlen.as<int>();
    b.ebase = S->c_ebase.as<int>(); b.epath = S->c_epath.as<long long>(); b.ebud = S->c_ebud.as<long long>();
    b.eskip = S->c_eskip.as<int>();
    b.verdict = S->c_verdict.as<int>(); b.nodes = S->c_nodes.as<long long>();
    b.A = S->c_A.as<uint64_t>(); b.B = S->c_B.as<uint64_t>();
    b.charges = S->c_charges.as<long long>(); b.icharged = S->c_icharged.as<long long>();
    b.row_scratch = S->c_scratch.as<uint64_t>(); b.ctr = S->ctr.as<unsigned long long>();
    b.fv = S->c_fv.as<int>(); b.fn = S->c_fn.as<long long>(); b.fA = S->c_fA.as<uint64_t>();
    b.fB = S->c_fB.as<uint64_t>(); b.fj = S->c_fj.as<int>(); b.fcc = S->c_fcc.as<long long>();
    b.fcur = S->c_fcur.as<long long>();
    return WZFH_OK;
  }
  carve->used = 0;
  b.ctr = carve->take<unsigned long long>(64);
  b.cd = carve->take<uint64_t>((size_t)nent * 16); b.roff = carve->take<int>(nent * 4);
  b.rlen = carve->take<int>(nent * 4); b.ebase = carve->take<int>(nent * 4);
  b.epath = carve->take<long long>(nent * 8); b.ebud = carve->take<long long>(nent * 8);
  b.eskip = carve->take<int>(nent * 4);
  b.verdict = carve->take<int>(nw * 4); b.nodes = carve->take<long long>(nw * 8);
  b.A = carve->take<uint64_t>(nw * 8); b.B = carve->take<uint64_t>(nw * 8);
  b.charges = carve->take<long long>(nw * 8); b.icharged = carve->take<long long>(nw * 8);
  b.fv = carve->take<int>(nent * 4); b.fn = carve->take<long long>(nent * 8); b.fA = carve->take<uint64_t>(nent * 8);
  b.fB = carve->take<uint64_t>(nent * 8); b.fj = carve->take<int>(nent * 4); b.fcc = carve->take<long long>(nent * 8);
  b.fcur = carve->take<long long>(nent * 8);
  b.row_scratch = carve->take<uint64_t>(row_scratch);
  if (!b.ctr || !b.cd || !b.roff || !b.rlen || !b.ebase || !b.eskip || !b.epath || !b.ebud || !b.verdict || !b.nodes || !b.A ||
      !b.B || !b.charges || !b.icharged || !b.fv || !b.fn || !b.fA || !b.fB || !b.fj || !b.fcc || !b.fcur ||
      !b.row_scratch) {
    set_error("allocation-free completer: a launch does not fit the static scratch array");
    return WZFH_ECAPACITY;
  }
  return WZFH_OK;
}

static int complete_launch(wzfh_problem &P, DeviceState *S, const uint64_t *cd, const int *roff, const int *rlen,
                           const std::vector<CompleteEntry> &ents, int maxrows, long long budget, int ext,
                           long long item_cap, int *v,
                           long long *nd, uint64_t *A, uint64_t *B, int *fj, long long *fcc, long long *fcur,
                           double *ms, Carve *carve, std::vector<std::vector<ItemRes>> *items,
                           std::vector<long long> *depth = nullptr) {
  const int nent = (int)ents.size();
  const int words = std::max(1, (maxrows + 63) / 64);
  // first-round launches under the root canon enumerate depth-0 digits 6 and 7 only
  bool root_pair = ext > 0 && !P.cfg.no_canon_ab;
  for (const CompleteEntry &x : ents) root_pair = root_pair && x.base == 0;
  const int items_log2 = 3 * ext - (root_pair ? 2 : 0);
  const long long nwork = (long long)nent << items_log2;
  int bpc = 8;
  if (const char *e = getenv("WZFH_COMPLETE_BPC")) bpc = std::max(1, std::min(32, atoi(e)));
  int blocks = (int)std::max(1ll, std::min((long long)S->num_cu * bpc, nwork));
  const size_t lds_bytes = complete_lds_bytes(P.L1, words);
  bool lds = lds_bytes <= S->complete_lds_max;
  if (const char *e = getenv("WZFH_COMPLETE_LDS")) lds = lds && atoi(e) != 0;
  const size_t nw = (size_t)nwork;
  if (carve && !lds) {   // row bitsets in global memory: as many waves as the carve holds
    const size_t fixed = kCarveSlack + (size_t)nent * kEntryBytes + nw * kItemBytes;
    const size_t room = carve->cap > fixed ? carve->cap - fixed : 0;
    blocks = (int)std::min<size_t>((size_t)blocks, room / lds_bytes);
    if (blocks < 1) { set_error("allocation-free completer: row lists too large for the static scratch array"); return WZFH_ECAPACITY; }
  }
  const size_t row_scratch = lds ? 64 : (size_t)blocks * lds_bytes;
  std::vector<uint64_t> ecd(2 * (size_t)nent);
  std::vector<int> ero(nent), erl(nent), ebase(nent);
  std::vector<long long> epath(nent), ebud(nent);
  std::vector<int> eskip(nent);
  bool any_skip = false;
  for (int e = 0; e < nent; e++) {
    const CompleteEntry &x = ents[e];
    ecd[2 * e] = cd[2 * (size_t)x.cand]; ecd[2 * e + 1] = cd[2 * (size_t)x.cand + 1];
    ero[e] = roff[x.cand]; erl[e] = rlen[x.cand];
    ebase[e] = x.base; epath[e] = x.path; ebud[e] = x.budget;
    eskip[e] = x.skip;
    any_skip = any_skip || x.skip > 0;
  }
  CompleteBufs bf;
  if (int rc = complete_buffers(S, carve, nent, nw, row_scratch, bf)) return rc;
  WZ_HIP(hipMemcpyAsync(bf.cd, ecd.data(), (size_t)nent * 16, hipMemcpyHostToDevice, S->stream));
  WZ_HIP(hipMemcpyAsync(bf.roff, ero.data(), nent * 4, hipMemcpyHostToDevice, S->stream));
  WZ_HIP(hipMemcpyAsync(bf.rlen, erl.data(), nent * 4, hipMemcpyHostToDevice, S->stream));
  WZ_HIP(hipMemcpyAsync(bf.ebase, ebase.data(), nent * 4, hipMemcpyHostToDevice, S->stream));
  WZ_HIP(hipMemcpyAsync(bf.epath, epath.data(), nent * 8, hipMemcpyHostToDevice, S->stream));
  WZ_HIP(hipMemcpyAsync(bf.ebud, ebud.data(), nent * 8, hipMemcpyHostToDevice, S->stream));
  if (any_skip) WZ_HIP(hipMemcpyAsync(bf.eskip, eskip.data(), nent * 4, hipMemcpyHostToDevice, S->stream));
  // counters start at zero (a fill kernel, not a memset node: callers of the allocation-free
  // entry point may record this stream into a graph)
  WZ_HIP(launch_zero(bf.ctr, 64, S->stream));
  if (ext > 0) WZ_HIP(launch_zero(bf.icharged, nw * 8, S->stream));
  CompleteArgs a{};
  a.n = P.n; a.L = P.L1; a.m = P.m;
  a.absA = std::abs(P.cfg.a); a.absB = std::abs(P.cfg.b);
  a.outer_k = P.cfg.outer_k; a.canon = !P.cfg.no_canon_ab;
  a.ncand = nent; a.budget = budget;
  a.split_ext = ext; a.nwork = nwork; a.root_pair = root_pair ? 1 : 0; a.items_log2 = items_log2;
  a.item_cap = ext > 0 ? item_cap : 0;
  a.ent_path = bf.epath; a.ent_base = bf.ebase; a.ent_budget = bf.ebud;
  a.ent_skip = any_skip ? bf.eskip : nullptr;
  a.snap_ext = (ext > 0 && item_cap > 0) ? ext : 0;   // capped items report where they stopped
  a.cd = bf.cd;
  a.row_off = bf.roff; a.row_len = bf.rlen;
  a.rows = S->rows.p ? S->rows.as<uint64_t>() : bf.cd;
  for (int d = 0; d < 32; d++)
    for (int c = 0; c < 8; c++) {
      int r = c < P.m ? class_count(P.L1, c, P.m) : 0;
      for (int i = c; c < P.m && i < P.L1; i += P.m)
        if (i <= d || i >= P.L1 - 1 - d) r--;
      a.rcap[d][c] = (uint8_t)std::max(0, r);
    }
  if (P.cfg.a2 == 2) {   // init_a2_tables (wz_match.cpp:1001-1016)
    int blk_of[8];
    for (int c = 0; c < 8; c++) blk_of[c] = -1;
    a.a2 = 1;
    for (int c = 0; c < P.m; c++) {
      if (blk_of[c] >= 0) continue;
      const int sr = ((P.L1 - 1 - c) % P.m + P.m) % P.m;
      a.a2_r[a.a2_nblk] = (int8_t)c; a.a2_s[a.a2_nblk] = (int8_t)sr; a.a2_mid[a.a2_nblk] = 0;
      blk_of[c] = blk_of[sr] = a.a2_nblk++;
    }
    const int half = P.L1 / 2;
    if (P.L1 % 2 == 1) a.a2_mid[blk_of[half % P.m]] = 1;
    for (int p = 0; p <= half && p < 33; p++)
      for (int i = p; i < half; i++) a.a2_cnt[p][blk_of[i % P.m]]++;
  }
  a.outer2 = S->outer2.as<uint8_t>(); a.outer3 = S->outer3.as<uint8_t>();
  a.pos_pack = 0;
  for (int sb = 0; sb < 16; sb++) a.pos_pack |= (uint64_t)(kOuter.pos_index[sb] + 1) << (4 * sb);
  a.scratch = bf.row_scratch; a.scratch_words = words;
  a.work_ctr = bf.ctr;
  a.verdict = bf.verdict; a.nodes = bf.nodes;
  a.outA = bf.A; a.outB = bf.B;
  a.charges = bf.charges;
  a.item_charged = bf.icharged;
  a.depth_out = nullptr;
  if (depth && ext == 0 && !carve) {   // per-depth node counts of whole walks (FH_TELEM)
    if (S->c_depth.ensure(nw * 64 * 8)) return WZFH_ENOMEM;
    a.depth_out = S->c_depth.as<long long>();
  }
  WZ_HIP(hipEventRecord(S->ev[2], S->stream));
  WZ_HIP(launch_complete(a, S->kab.as<uint8_t>(), blocks, lds, lds ? lds_bytes : 0, S->stream));
  WZ_HIP(hipEventRecord(S->ev[4], S->stream));
  // per-entry results: on the device for split launches (one wave per entry), from the
  // items themselves for unsplit ones; WZFH_CHECK_REDUCE=1 also scans the items on the host
  // and compares
  if (ext > 0)
    WZ_HIP(launch_complete_reduce(a, nent, bf.fv, bf.fn, bf.fA, bf.fB, bf.fj, bf.fcc, bf.fcur, S->stream));
  WZ_HIP(hipEventRecord(S->ev[3], S->stream));
  const bool check = ext > 0 && getenv("WZFH_CHECK_REDUCE") && atoi(getenv("WZFH_CHECK_REDUCE"));
  std::vector<int> hv(nent), hj(nent);
  std::vector<long long> hn(nent), hcc(nent), hcj(nent);
  std::vector<uint64_t> hA(nent), hB(nent);
  std::vector<int> iv;
  std::vector<long long> in, ic;
  std::vector<uint64_t> iA, iB;
  auto fetch_items = [&]() -> int {
    if (!iv.empty()) return WZFH_OK;
    iv.resize(nw); in.resize(nw); ic.resize(nw); iA.resize(nw); iB.resize(nw);
    WZ_HIP(hipMemcpyAsync(iv.data(), a.verdict, nw * 4, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(in.data(), a.nodes, nw * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(ic.data(), a.charges, nw * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(iA.data(), a.outA, nw * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(iB.data(), a.outB, nw * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
    return WZFH_OK;
  };
  if (getenv("WZFH_TELEMETRY") && atoi(getenv("WZFH_TELEMETRY"))) {   // every item's walk, kept or not
    std::vector<long long> wn(nw);
    WZ_HIP(hipMemcpyAsync(wn.data(), a.nodes, nw * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
    for (long long x : wn) S->walked += x;
  }
  if (a.depth_out) {
    depth->assign(nw * 64, 0);
    WZ_HIP(hipMemcpyAsync(depth->data(), a.depth_out, nw * 64 * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
  }
  if (ext == 0 || check) {
    if (int rc = fetch_items()) return rc;
    // the same scan as the device reduce, on the host
    const long long np = 1ll << items_log2;
    for (int e = 0; e < nent; e++) {
      long long cum_n = 0, cum_c = 0, cj = 0;
      int vo = 2, jo = -1;
      uint64_t Ao = 0, Bo = 0;
      for (long long p = 0; p < np; p++) {
        const size_t u = (size_t)((long long)e * np + p);
        if (iv[u] == 1) { vo = 1; cum_n = 0; break; }
        if (ext > 0 && (iv[u] == 3 || (budget > 0 && cum_c + ic[u] > ents[e].budget))) {
          vo = -1; jo = (int)p; cj = ic[u];
          break;
        }
        if (iv[u] == 5) { vo = -2; break; }   // capped before a decision
        cum_n += in[u];
        if (iv[u] == 0 || iv[u] == 3 || iv[u] == 4) {   // a hit, or (unsplit) an exactly counted abort
          vo = iv[u]; Ao = iA[u]; Bo = iB[u];
          if (ext > 0) cum_c += ic[u];   // a split hit reports the charges through it
          break;
        }
        cum_c += ic[u];
      }
      hv[e] = vo; hn[e] = cum_n; hA[e] = Ao; hB[e] = Bo; hj[e] = jo; hcc[e] = cum_c; hcj[e] = cj;
    }
  }
  if (ext > 0) {
    WZ_HIP(hipMemcpyAsync(v, bf.fv, nent * 4, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(nd, bf.fn, nent * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(A, bf.fA, nent * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(B, bf.fB, nent * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(fj, bf.fj, nent * 4, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(fcc, bf.fcc, nent * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipMemcpyAsync(fcur, bf.fcur, nent * 8, hipMemcpyDeviceToHost, S->stream));
    WZ_HIP(hipStreamSynchronize(S->stream));
    if (check)
      for (int e = 0; e < nent; e++) {
        const bool same = v[e] == hv[e] && nd[e] == hn[e] && A[e] == hA[e] && B[e] == hB[e] &&
                          (v[e] >= 0 || (fj[e] == hj[e] && fcc[e] == hcc[e] && fcur[e] == hcj[e]));
        if (!same) {
          set_error("completer: device reduce differs from the host scan");
          return WZFH_EDEVICE;
        }
      }
  } else {
    for (int e = 0; e < nent; e++) {
      v[e] = hv[e]; nd[e] = hn[e]; A[e] = hA[e]; B[e] = hB[e]; fj[e] = hj[e]; fcc[e] = hcc[e]; fcur[e] = hcj[e];
    }
  }
  if (items) {   // entries whose scan met a capped item first: their items, for the host
    items->assign(nent, {});
    const long long np = 1ll << items_log2;
    for (int e = 0; e < nent; e++)
      if (v[e] == -2) {
        if (int rc = fetch_items()) return rc;
        std::vector<ItemRes> &dst = (*items)[e];
        dst.resize(np);
        for (long long p = 0; p < np; p++) {
          const size_t u = (size_t)((long long)e * np + p);
          dst[p] = ItemRes{iv[u], in[u], ic[u], iA[u], iB[u]};
        }
      }
  }
  if (root_pair)   // item index -> quad path digits below the entry
    for (int e = 0; e < nent; e++)
      if (fj[e] >= 0) {
        const int lo = 3 * (ext - 1);
        fj[e] = ((6 | (fj[e] >> lo)) << lo) | (fj[e] & ((1 << lo) - 1));
      }
  float t = 0, tk = 0;
  WZ_HIP(hipEventElapsedTime(&t, S->ev[2], S->ev[3]));
  WZ_HIP(hipEventElapsedTime(&tk, S->ev[2], S->ev[4]));
  *ms += t;
  {
    std::lock_guard<std::mutex> lk(S->tot_mu);
    S->tot_ms[1] += tk;
    S->tot_launches[1]++;
  }
  return WZFH_OK;
}

// Split of the completer: every candidate's quad paths of depths 0..2 become separate work
// items (8^3 per candidate) so one candidate's tree runs on many waves; the subtree that holds
// a budget abort is split the same way again (3 more levels) until it is small, then walked as
// one item that counts the abort exactly.
static int complete_split_ext(const wzfh_problem &P, const DeviceState *S) {
  int ext = (P.L1 / 2 >= 6) ? S->tune.complete_split : 0;
  if (const char *e = getenv("WZFH_COMPLETE_SPLIT")) ext = std::max(0, std::min(6, atoi(e)));
  if (ext >= P.L1 / 2) ext = 0;
  return ext;
}
// Split items stop after 2^complete_cap_log2 charges (the tuning; WZFH_COMPLETE_ITEM_CAP = C
// overrides with C charges, 0 = no cap); the subtree of a capped item is split again in the next
// round, so no item walks much more than the cap.
static long long item_charge_cap(const DeviceState *S) {
  if (const char *e = getenv("WZFH_COMPLETE_ITEM_CAP")) return std::max(0ll, atoll(e));
  const int lg = S->tune.complete_cap_log2;
  return lg > 0 ? 1ll << std::min(lg, 40) : 0;
}

static long long exact_item_charges() {
  const char *e = getenv("WZFH_COMPLETE_EXACT_ITEM");
  return e ? std::max(0ll, atoll(e)) : 2048;   // a whole walk of 2048 charges is a few ms on one wave
}

// Cells of one search batch that fit in a 1/(2 nshare) share of the device's free memory (read
// at every call): per cell the stream's work records (16 K donated records of 68 B, the static
// ones), its candidate records (32 B raw + 32 B kept, up to the cap plus a margin, or 1,024
// without a cap), and the completer's split items (top_k candidates x items x 44 B).
long long serve_cell_slots(wzfh_problem &P, int nshare, const wzfh_search_opts &o) {
  size_t free_b = 0, total_b = 0;
  if (!P.dev || hipSetDevice(P.dev->device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
    return 1ll << 30;
  const int ext = complete_split_ext(P, P.dev);
  const double items = ext > 0 ? (double)(1ll << (3 * ext)) : 1.0;
  const double cands = o.stream.cand_cap > 0 ? (double)o.stream.cand_cap + 512.0 : 1024.0;
  const double per_cell = 16384.0 * 72.0 + cands * 64.0 + (double)std::max(1, o.top_k) * items * 44.0 + 65536.0;
  return std::max(1ll, (long long)((double)free_b / (2.0 * std::max(1, nshare)) / per_cell));
}

// Host side of the completer: rounds of launches over jobs until every candidate is decided.
//
// A job is a subtree of one candidate (the whole tree at first) with the charges left to the
// candidate's budget at its start. A split launch runs a job's quad paths `ext` levels further
// down as work items, every item stopping after `cap` charges; the device reduce scans the items
// in path order (the reference's DFS order) and reports the decision, or the item in which the
// budget runs out (that item's subtree becomes an exact-abort job), or that a capped item comes
// before any decision. In that last case the job becomes a frame: its items' results come to
// the host, every capped item's subtree becomes a job of the next round (all launched together,
// each with the charges left as if the capped items before it had charged only what they did
// before stopping -- an upper bound), and the frame is scanned in path order as its capped items
// resolve. A capped item's result is used as is when its decision cannot depend on the budget;
// when the budget runs out inside it, its subtree is run once more with the exact charges left.
// So every heavy subtree is split again and again until its pieces are small, and the whole
// candidate's verdict, node count and first A,B are the reference's.
namespace {
struct Result { int v; long long n, c, bud; uint64_t A, B; bool done; };
struct Job {
  int cand; long long path; int base;
  long long budget;          // charges left at the subtree's start
  long long pre_n, pre_c;    // nodes / charges of the target before this job
  int owner, slot;           // target: frame `owner`'s capped-item slot, or the candidate (-1)
  bool whole;                // walk as one item (an exactly counted abort)
  int skip = 0;              // split items below this were walked by the capped item it continues
};
struct Frame {
  int cand; long long path; int base, ext; bool root_pair;
  long long budget, pre_n, pre_c;
  int owner, slot;
  std::vector<ItemRes> items;
  std::vector<int> slot_of;  // item -> its capped-item slot (-1: decided by the launch)
  std::vector<Result> subs;
  size_t next = 0;
  long long cum_n = 0, cum_c = 0;
  bool closed = false;
};
}  // namespace

// The rounds of one completer run over the candidates cdv / cellv (packed (C, D) pairs and cell
// indices, -1 = no profile rows). With a feed, candidates also join between rounds (feed->more
// appends them) and each is reported through feed->done as soon as it is decided; fin holds every
// candidate's result at the end.
static int complete_rounds(wzfh_problem &P, std::vector<uint64_t> &cdv, std::vector<int> &cellv, long long budget,
                           Carve *carve, long long *depth_nodes, DeviceState *S, CompleteFeed *feed,
                           std::vector<Result> &fin) {
  WZ_HIP(hipSetDevice(S->device));
  std::vector<int> roff, rlen;
  int maxrows = 0;
  const bool limited = budget > 0;
  fin.clear();
  S->walked = 0;
  double ms = 0;
  // per-depth counts need whole walks: one item per candidate
  const int ext0 = depth_nodes ? 0 : complete_split_ext(P, S);
  const int max_base = std::min(P.L1 / 2 - 1, 21);   // paths are 3 bits per depth in 64 bits
  const long long exact_item = exact_item_charges();
  const long long cap = ext0 > 0 ? item_charge_cap(S) : 0;
  std::vector<Frame> frames;
  std::vector<Job> jobs, next_jobs;
  // candidates [from, cellv.size()) join the run: their rows and their first-round jobs
  auto add = [&](size_t from) -> int {
    for (size_t i = from; i < cellv.size(); i++) {
      int ro = -1, rl = 0;
      const int c = cellv[i];
      if (c >= 0 && P.ab_prof) {
        if (c >= (int)P.cell_rows.size()) { set_error("cell index out of range"); return WZFH_EINVAL; }
        const int li = P.cell_rows[c];
        if (li < 0) { ro = 0; rl = 0; }   // proven-dead cell: no row can complete
        else if ((long long)P.row_len[li] <= P.cfg.abp_cap) { ro = P.row_off[li]; rl = P.row_len[li]; }
        // else: oversized list, unconstrained
      }
      roff.push_back(ro);
      rlen.push_back(rl);
      maxrows = std::max(maxrows, rl);
      fin.push_back(Result{2, 0, 0, 0, 0, 0, false});
      jobs.push_back(Job{(int)i, 0, 0, budget, 0, 0, -1, (int)i, ext0 == 0});
    }
    return WZFH_OK;
  };
  if (int rc = add(0)) return rc;

  std::function<void(int, int, Result)> deliver;
  // scans frame f from its next item; returns when it waits for a capped item or is closed
  auto resume = [&](int f) {
    Frame &F = frames[f];
    auto finish = [&](Result r) {
      F.closed = true;
      r.n += F.pre_n; r.c += F.pre_c; r.bud = F.pre_c + F.budget;
      deliver(F.owner, F.slot, r);
    };
    auto abort_in = [&](size_t i, bool small) {   // the budget runs out inside item i: exact rerun
      const int lo = 3 * (F.ext - 1);
      const long long dig = F.root_pair ? (((6 | ((long long)i >> lo)) << lo) | ((long long)i & ((1ll << lo) - 1)))
                                        : (long long)i;
      Job y{F.cand, (F.path << (3 * F.ext)) | dig, F.base + F.ext, F.budget - F.cum_c, F.pre_n + F.cum_n,
            F.pre_c + F.cum_c, F.owner, F.slot, false};
      y.whole = small || y.base + ext0 > max_base;
      F.closed = true;
      next_jobs.push_back(y);
    };
    while (!F.closed && F.next < F.items.size()) {
      const size_t i = F.next;
      const ItemRes &it = F.items[i];
      if (it.v == 1) { finish(Result{1, 0, 0, 0, 0, 0, true}); return; }   // pre-rejected candidate
      if (it.v == 5) {
        // a capped item launched no subtree job only when the charges before it plus its own
        // partial charges already pass the budget: the budget runs out inside it
        if (F.slot_of[i] < 0) { abort_in(i, false); return; }
        const Result &r = F.subs[F.slot_of[i]];
        if (!r.done) return;   // waits for the capped item's subtree
        const long long left = F.budget - F.cum_c;
        if (limited && (r.v == 3 || r.c > left)) {
          if (r.v == 3 && r.bud == left) { finish(Result{3, F.cum_n + r.n, F.cum_c + r.c, 0, 0, 0, true}); return; }
          abort_in(i, false);
          return;
        }
        if (r.v == 0 || r.v == 4) { finish(Result{r.v, F.cum_n + r.n, F.cum_c + r.c, 0, r.A, r.B, true}); return; }
        F.cum_n += r.n; F.cum_c += r.c; F.next++;
        continue;
      }
      if (it.v == 3 || (limited && F.cum_c + it.c > F.budget)) { abort_in(i, it.c <= exact_item); return; }
      if (it.v == 0 || it.v == 4) { finish(Result{it.v, F.cum_n + it.n, F.cum_c + it.c, 0, it.A, it.B, true}); return; }
      F.cum_n += it.n; F.cum_c += it.c; F.next++;
    }
    if (!F.closed) finish(Result{2, F.cum_n, F.cum_c, 0, 0, 0, true});
  };
  deliver = [&](int owner, int slot, Result r) {
    r.done = true;
    if (owner < 0) {
      fin[slot] = r;
      {
        std::lock_guard<std::mutex> lk(S->tot_mu);
        S->tot_units[1] += r.n;
      }
      if (feed) {
        wzfh_completion c{};
        c.verdict = r.v;
        c.nodes = r.n;
        c.A = r.v == 0 || r.v == 4 ? r.A : 0;
        c.B = r.v == 0 || r.v == 4 ? r.B : 0;
        feed->done(slot, c);
      }
      return;
    }
    frames[owner].subs[slot] = r;
    resume(owner);
  };

  int round = 0;
  while (true) {
    if (feed) {   // candidates that arrived since the last round join this one
      const size_t before = cellv.size();
      if (feed->more(cdv, cellv) > 0) {
        if (int rc = add(before)) return rc;
        round = 0;
      }
    }
    if (jobs.empty()) break;
    if (++round > 4096) { set_error("completer: rounds did not converge (internal error)"); return WZFH_EKERNEL; }
    next_jobs.clear();
    // launches: exact walks, first-round splits (root canon pair), deeper splits
    // A run that candidates join (feed) launches first-round and deeper split jobs together: a
    // launch then enumerates every depth-0 digit, and under the root canon the six digits the
    // canon rejects end at once with no charge, so the results are the same and a round waits
    // for one split launch instead of two.
    const bool merge_splits = feed != nullptr;
    for (int kind = 0; kind < 3; kind++) {
      if (merge_splits && kind == 2) break;
      std::vector<Job> group;
      for (const Job &x : jobs)
        if ((kind == 0) == x.whole && (kind == 0 || merge_splits || (kind == 1) == (x.base == 0))) group.push_back(x);
      if (group.empty()) continue;
      const int ext = kind == 0 ? 0 : ext0;
      const long long icap = kind == 0 ? 0 : cap;
      // at most 4M items per launch (what the static scratch holds for the allocation-free path)
      long long max_items = 1ll << 22;
      if (carve) max_items = std::min<long long>(max_items, (long long)((carve->cap - kCarveSlack) / (kItemBytes + kEntryBytes)));
      const int chunk = (int)std::max(1ll, max_items >> (3 * ext));
      for (size_t c0 = 0; c0 < group.size(); c0 += chunk) {
        std::vector<CompleteEntry> part;
        for (size_t e = c0; e < std::min(group.size(), c0 + chunk); e++)
          part.push_back(CompleteEntry{group[e].cand, group[e].path, group[e].base, group[e].budget, 0, group[e].skip});
        const int ne = (int)part.size();
        std::vector<int> v(ne), j(ne);
        std::vector<long long> nd(ne), cc(ne), cj(ne);
        std::vector<uint64_t> A(ne), B(ne);
        std::vector<std::vector<ItemRes>> items;
        std::vector<long long> dep;
        int rc = complete_launch(P, S, cdv.data(), roff.data(), rlen.data(), part, maxrows, budget, ext, icap, v.data(),
                                 nd.data(), A.data(), B.data(), j.data(), cc.data(), cj.data(), &ms, carve, &items,
                                 depth_nodes ? &dep : nullptr);
        if (rc != WZFH_OK) return rc;
        if (depth_nodes && !dep.empty())
          for (int e = 0; e < ne; e++)
            std::copy(dep.begin() + (size_t)e * 64, dep.begin() + (size_t)e * 64 + 64,
                      depth_nodes + (size_t)group[c0 + e].cand * 64);
        // as complete_launch decides it: only a launch of root entries alone takes the root pair
        bool root_pair = ext > 0 && !P.cfg.no_canon_ab;
        for (const CompleteEntry &x : part) root_pair = root_pair && x.base == 0;
        for (int e = 0; e < ne; e++) {
          const Job &x = group[c0 + e];
          if (v[e] >= 0) {   // decided: hit, clean, pre-rejected, or an exactly counted abort
            deliver(x.owner, x.slot, Result{v[e], x.pre_n + nd[e], x.pre_c + cc[e], x.pre_c + x.budget, A[e], B[e], true});
            continue;
          }
          if (v[e] == -1) {   // the budget runs out inside item j (already mapped to path digits)
            Job y{x.cand, (x.path << (3 * ext)) | j[e], x.base + ext, x.budget - cc[e], x.pre_n + nd[e],
                  x.pre_c + cc[e], x.owner, x.slot, false};
            y.whole = cj[e] <= exact_item || y.base + ext0 > max_base;
            next_jobs.push_back(y);
            continue;
          }
          // a capped item comes before any decision: the job becomes a frame
          Frame F;
          F.cand = x.cand; F.path = x.path; F.base = x.base; F.ext = ext; F.root_pair = root_pair;
          F.budget = x.budget; F.pre_n = x.pre_n; F.pre_c = x.pre_c; F.owner = x.owner; F.slot = x.slot;
          F.items = std::move(items[e]);
          F.slot_of.assign(F.items.size(), -1);
          const int f = (int)frames.size();
          long long lower = 0;   // charges before each item, capped items counted as far as they got
          for (size_t i = 0; i < F.items.size(); i++) {
            const ItemRes &it = F.items[i];
            if (it.v == 1 || it.v == 0 || it.v == 4 || it.v == 3) break;   // a decision: later items never matter
            if (limited && lower + it.c > F.budget) break;                 // the budget runs out by here
            if (it.v == 5) {
              const int lo = 3 * (ext - 1);
              const long long dig = root_pair ? (((6 | ((long long)i >> lo)) << lo) | ((long long)i & ((1ll << lo) - 1)))
                                              : (long long)i;
              Job y{x.cand, (x.path << (3 * ext)) | dig, x.base + ext, x.budget - lower, 0, 0, f, (int)F.subs.size(), false};
              y.whole = y.base + ext0 > max_base;
              if (!y.whole && ext0 == ext) {
                // the capped item's walk covered its subtree's sub-paths below the one it stopped
                // in: the next round splits the subtree from that sub-path on, and the charges and
                // nodes counted before it are carried as the job's start
                const long long sc = (long long)(it.A & ((1ull << 44) - 1));
                const int sj = (int)(it.A >> 44);
                y.skip = sj;
                y.pre_c = sc;
                y.pre_n = (long long)it.B;
                y.budget -= sc;
              }
              F.slot_of[i] = (int)F.subs.size();
              F.subs.push_back(Result{0, 0, 0, 0, 0, 0, false});
              next_jobs.push_back(y);
            }
            lower += it.c;
          }
          frames.push_back(std::move(F));
          resume(f);
        }
      }
    }
    jobs.swap(next_jobs);
  }
  S->last_ms[1] = ms;
  S->last_walked = S->walked;
  long long units = 0;
  for (const Result &r : fin) {
    if (!r.done) { set_error("completer: a candidate was left undecided (internal error)"); return WZFH_EKERNEL; }
    units += r.n;
  }
  S->last_units[1] = units;
  return WZFH_OK;
}

int run_complete(wzfh_problem &P, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                 wzfh_completion *out, Carve *carve, long long *depth_nodes, DeviceState *dev) {
  DeviceState *S = dev ? dev : P.dev;
  if (ncand <= 0) return WZFH_OK;
  std::vector<uint64_t> cdv(cd, cd + 2 * (size_t)ncand);
  std::vector<int> cellv(ncand, -1);
  if (cell_of) std::copy(cell_of, cell_of + ncand, cellv.begin());
  std::vector<Result> fin;
  if (int rc = complete_rounds(P, cdv, cellv, budget, carve, depth_nodes, S, nullptr, fin)) return rc;
  bool internal = false;
  for (int i = 0; i < ncand; i++) {
    out[i].verdict = fin[i].v;
    out[i].reserved = 0;
    out[i].nodes = fin[i].n;
    out[i].A = fin[i].v == 0 || fin[i].v == 4 ? fin[i].A : 0;
    out[i].B = fin[i].v == 0 || fin[i].v == 4 ? fin[i].B : 0;
    if (fin[i].v == 4) internal = true;
  }
  if (internal) { set_error("completer FOUND failed the exact NPAF recheck"); return WZFH_EINTERNAL; }
  return WZFH_OK;
}

int run_complete_feed(wzfh_problem &P, long long budget, DeviceState *dev, CompleteFeed &feed) {
  DeviceState *S = dev ? dev : P.dev;
  std::vector<uint64_t> cdv;
  std::vector<int> cellv;
  std::vector<Result> fin;
  return complete_rounds(P, cdv, cellv, budget, nullptr, nullptr, S, &feed, fin);
}

// Per cell of a stream batch: stable-sort its candidates by flat score and take the first K
// (the drain's stable_sort, wz_match.cpp:2654-2655, front-only top-K). Appends the chosen
// candidates' packed words, owning cell slot, sorted rank and in-cell stream index.
void select_topk(int n, const StreamOut &so, int ncells, int K, TopK &t) {
  for (int i = 0; i < ncells; i++) {
    const long long b = so.cell_begin[i], e = so.cell_begin[i + 1];
    std::vector<std::pair<long long, long long>> key;
    key.reserve(e - b);
    for (long long j = b; j < e; j++) key.push_back({wzfh_flat_score(n, so.cands[j].X, so.cands[j].Y), j - b});
    std::stable_sort(key.begin(), key.end(), [](auto &x, auto &y) { return x.first < y.first; });
    const long long take = std::min<long long>(K, (long long)key.size());
    for (long long r = 0; r < take; r++) {
      const CandRec &c = so.cands[b + key[r].second];
      t.cd.push_back(c.X); t.cd.push_back(c.Y);
      t.owner.push_back(i); t.rank.push_back(r); t.sidx.push_back(key[r].second);
    }
  }
}

// The drain's counters per cell (wz_match.cpp:2686-2714): the sorted ranks are completed in
// order and the drain stops at the first hit (or internal error, :2591-2595); ranks past it
// are not counted. A cell's completions sit in rank order in the arrays.
void aggregate_completions(const TopK &t, const wzfh_completion *comp, wzfh_search_result *res) {
  for (size_t k = 0; k < t.owner.size(); k++) {
    wzfh_search_result &r = res[t.owner[k]];
    if (r.hit || r.internal) continue;
    r.completed++;
    r.ab_nodes += comp[k].nodes;
    if (comp[k].verdict == 2) r.clean_no++;
    if (comp[k].verdict == 3) r.aborted++;
    if (comp[k].verdict == 4) r.internal = 1;
    if (comp[k].verdict == 0) {
      r.hit = 1; r.hit_rank = t.rank[k]; r.hit_idx = t.sidx[k];
      r.A = comp[k].A; r.B = comp[k].B; r.C = t.cd[2 * k]; r.D = t.cd[2 * k + 1];
    }
  }
}

std::string stream_key(const wzfh_problem &P) {
  return std::to_string(P.n) + "/" + std::to_string(P.m) + "/" + std::to_string(P.cfg.rev != 0) + "/" +
         std::to_string(P.cfg.endpos != 0) + "/" + std::to_string(P.ab_side) + "/" + std::to_string(P.pinC) +
         std::to_string(P.pinD);
}

int stream_handle(const wzfh_problem &P, wzfh_problem **out) {
  *out = nullptr;
  wzfh_problem *Q = new wzfh_problem();
  Q->cfg = P.cfg;
  Q->n = P.n; Q->L1 = P.L1; Q->m = P.m;
  Q->pinC = P.pinC; Q->pinD = P.pinD; Q->pins_dropped = P.pins_dropped;
  Q->ab_prof = false; Q->ab_side = P.ab_side;
  std::memcpy(Q->hcos, P.hcos, sizeof Q->hcos);
  std::memcpy(Q->hsin, P.hsin, sizeof Q->hsin);
  int rc = device_init(*Q);
  if (rc == WZFH_OK && P.dev) Q->dev->tune = P.dev->tune;
  if (rc != WZFH_OK) { device_free(*Q); delete Q; return rc; }
  *out = Q;
  return WZFH_OK;
}

// A problem handle: setup of one side's cell list and, with a device, its device state.
// ab_side builds the A,B side of the join instead of the C,D first-hit cells.
int problem_init(const wzfh_config &cfg_in, bool ab_side, bool with_device, wzfh_problem **out) {
  *out = nullptr;
  const wzfh_config *cfg = &cfg_in;
  const int n = cfg->n;
  if (n < 4) { set_error("n must be >= 4"); return WZFH_EINVAL; }
  if (n + 1 > kMaxL - 1) { set_error("n too large for one 64-bit word per sequence"); return WZFH_EINVAL; }
  if (cfg->a * cfg->a + cfg->b * cfg->b + cfg->c * cfg->c + cfg->d * cfg->d != 4 * n + 2) {
    set_error("sig norm != 4n+2");
    return WZFH_EINVAL;
  }
  if (cfg->outer_k != 0 && cfg->outer_k != 2 && cfg->outer_k != 3) { set_error("outer_k must be 0, 2 or 3"); return WZFH_EINVAL; }
  if (cfg->a2 != 0 && cfg->a2 != 2) { set_error("a2 must be 0 (off) or 2 (prune); the shadow mode 1 is not provided"); return WZFH_EINVAL; }
  wzfh_problem *P = new wzfh_problem();
  P->cfg = *cfg;
  P->ab_side = ab_side;
  if (P->cfg.abp_cap <= 0) P->cfg.abp_cap = 20000;
  P->n = n; P->L1 = n + 1; P->m = cfg->m6 ? 6 : 3;
  if (n >= 36) { P->cfg.thm211b = 1; P->cfg.thm212 = 1; }   // wz_match.cpp:1388
  P->pinC = cfg->c == 0; P->pinD = cfg->d == 0;
  if (cfg->canon && (P->pinC || P->pinD)) { P->pinC = P->pinD = false; P->pins_dropped = true; }   // :1661-1662
  P->ab_prof = cfg->ab_prof < 0 ? (cfg->m6 != 0) : (cfg->ab_prof != 0);
  int rc = build_problem(*P);
  if (rc == WZFH_OK && with_device) rc = device_init(*P);
  if (rc != WZFH_OK) { device_free(*P); delete P; return rc; }
  *out = P;
  return WZFH_OK;
}

}  // namespace wzfh

using namespace wzfh;

extern "C" {

const char *wzfh_last_error(void) { return g_err.c_str(); }

int wzfh_config_default(wzfh_config *cfg, int n, int a, int b, int c, int d) {
  if (!cfg) return WZFH_EINVAL;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->n = n; cfg->a = a; cfg->b = b; cfg->c = c; cfg->d = d;
  cfg->m6 = n >= 36;
  cfg->thm211b = n >= 36;
  cfg->thm212 = n >= 36;
  cfg->ab_prof = -1;
  cfg->outer_k = 3;
  cfg->abp_cap = 20000;
  return WZFH_OK;
}

static int init_impl(const wzfh_config *cfg, wzfh_problem **out, bool with_device) {
  if (!cfg || !out) return WZFH_EINVAL;
  return problem_init(*cfg, false, with_device, out);
}

int wzfh_init(const wzfh_config *cfg, wzfh_problem **out) { return init_impl(cfg, out, true); }

int wzfh_init_host(const wzfh_config *cfg, wzfh_problem **out) { return init_impl(cfg, out, false); }

void wzfh_destroy(wzfh_problem *p) {
  if (!p) return;
  device_free(*p);
  delete p;
}

int wzfh_num_cells(const wzfh_problem *p) { return p ? (int)p->cells.size() : 0; }

int wzfh_cell_info(const wzfh_problem *p, int pi, int *px, int *py, int *kept, int *rows) {
  if (!p || pi < 0 || pi >= (int)p->cells.size()) return WZFH_EINVAL;
  for (int c = 0; c < p->m; c++) { if (px) px[c] = p->cells[pi].px[c]; if (py) py[c] = p->cells[pi].py[c]; }
  if (kept) *kept = p->kept[pi];
  if (rows) *rows = !p->ab_prof ? -2 : (p->cell_rows[pi] < 0 ? -1 : p->row_len[p->cell_rows[pi]]);
  return WZFH_OK;
}

int wzfh_live_cells(const wzfh_problem *p, int *out, int cap) {
  if (!p) return WZFH_EINVAL;
  int cnt = 0;
  for (size_t i = 0; i < p->cells.size(); i++) {
    if (!p->kept[i]) continue;
    if (p->ab_prof && p->cell_rows[i] < 0) continue;
    if (out && cnt < cap) out[cnt] = (int)i;
    cnt++;
  }
  return cnt;
}

int wzfh_qprune_stats(const wzfh_problem *p, long long *stats) {
  if (!p) return WZFH_EINVAL;
  if (stats) for (int k = 0; k < 7; k++) stats[k] = p->qp_stats[k];
  return p->qp_active ? 1 : 0;
}

int wzfh_order_digest(const wzfh_problem *p, char *buf, int buflen) {
  if (!p || !buf || buflen < 34) return WZFH_EINVAL;
  std::snprintf(buf, buflen, "%s", p->digest.c_str());
  return WZFH_OK;
}

int wzfh_stream_cells(wzfh_problem *p, const int *cells, int ncells, const wzfh_stream_opts *opts,
                      wzfh_cell_stats *stats, uint64_t *cd, long long cd_capacity, long long *offsets) {
  if (!p || !cells || ncells <= 0 || !stats || !offsets) return WZFH_EINVAL;
  for (int i = 0; i < ncells; i++)
    if (cells[i] < 0 || cells[i] >= (int)p->cells.size()) { set_error("cell index out of range"); return WZFH_EINVAL; }
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  wzfh_stream_opts o{};
  if (opts) o = *opts;
  DeviceState *S = p->dev;
  const bool same = S->memo_valid && S->memo_cells.size() == (size_t)ncells &&
                    std::equal(cells, cells + ncells, S->memo_cells.begin()) &&
                    std::memcmp(&S->memo_opts, &o, sizeof o) == 0;
  if (!same) {
    StreamOut so;
    S->memo_valid = false;
    S->memo_stats.assign(ncells, wzfh_cell_stats{});
    int rc = run_stream(*p, cells, ncells, o, S->memo_stats.data(), so);
    if (rc != WZFH_OK) return rc;
    S->memo_cells.assign(cells, cells + ncells);
    S->memo_opts = o;
    S->memo_offs = so.cell_begin;
    S->memo_cd.resize(2 * so.cands.size());
    for (size_t j = 0; j < so.cands.size(); j++) { S->memo_cd[2 * j] = so.cands[j].X; S->memo_cd[2 * j + 1] = so.cands[j].Y; }
    S->memo_valid = true;
  }
  std::copy(S->memo_stats.begin(), S->memo_stats.end(), stats);
  for (int i = 0; i <= ncells; i++) offsets[i] = S->memo_offs[i];
  const long long total = S->memo_offs[ncells];
  if (cd) {
    if (total > cd_capacity) return WZFH_ECAPACITY;
    std::copy(S->memo_cd.begin(), S->memo_cd.begin() + 2 * total, cd);
  }
  return WZFH_OK;
}

int wzfh_complete(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                  wzfh_completion *out) {
  if (!p || (ncand > 0 && (!cd || !out)) || ncand < 0 || budget < 0) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  return run_complete(*p, cd, cell_of, ncand, budget, out);
}

int wzfh_complete_depth(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                        wzfh_completion *out, long long *depth_nodes) {
  if (!p || (ncand > 0 && (!cd || !out || !depth_nodes)) || ncand < 0 || budget < 0) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  std::fill(depth_nodes, depth_nodes + (size_t)ncand * 64, 0ll);
  return run_complete(*p, cd, cell_of, ncand, budget, out, nullptr, depth_nodes);
}

int wzfh_complete_noalloc(wzfh_problem *p, const uint64_t *cd, const int *cell_of, int ncand, long long budget,
                          wzfh_completion *out) {
  if (!p || !cd || ncand < 0 || !out) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  if (ncand > WZFH_NOALLOC_MAX_CAND) {
    set_error("wzfh_complete_noalloc: at most " + std::to_string(WZFH_NOALLOC_MAX_CAND) + " candidates per call");
    return WZFH_EINVAL;
  }
  if (!p->dev->noalloc) { set_error("allocation-free scratch not set up on this device"); return WZFH_EDEVICE; }
  Carve carve;
  carve.base = (char *)p->dev->noalloc;
  carve.cap = WZFH_NOALLOC_SCRATCH_BYTES;
  return run_complete(*p, cd, cell_of, ncand, budget, out, &carve);
}

long long wzfh_flat_score(int n, uint64_t C, uint64_t D) {
  long long sc = 0;
  for (int s = 1; s < n; s++) {
    const uint64_t mk = (1ull << (n - s)) - 1;
    const int dis = __builtin_popcountll((C ^ (C >> s)) & mk) + __builtin_popcountll((D ^ (D >> s)) & mk);
    sc += std::abs(2 * (n - s) - 2 * dis);
  }
  return sc;
}

int wzfh_npaf_check(int n, const int8_t *A, const int8_t *B, const int8_t *C, const int8_t *D, int *max_abs) {
  if (!A || !B || !C || !D || n < 1) return WZFH_EINVAL;
  int mx = 0;
  const int n1 = n + 1;
  for (int s = 1; s <= n; s++) {
    int v = 0;
    for (int i = 0; i + s < n1; i++) v += A[i] * A[i + s] + B[i] * B[i + s];
    for (int i = 0; i + s < n; i++) v += C[i] * C[i + s] + D[i] * D[i + s];
    mx = std::max(mx, std::abs(v));
  }
  if (max_abs) *max_abs = mx;
  return WZFH_OK;
}

int wzfh_last_kernel_ms(const wzfh_problem *p, int kind, double *ms, long long *units) {
  if (!p || !p->dev || kind < 0 || kind > 13) return WZFH_EINVAL;
  if (kind == 13) {   // completer nodes walked by the last run's items (WZFH_TELEMETRY=1)
    if (ms) *ms = (double)p->dev->last_walked;
    if (units) *units = p->dev->last_walked;
    return WZFH_OK;
  }
  if (kind == 10) {   // margin of the matrix-core leaf test
    if (ms) *ms = p->dev->hall_eps;
    if (units) *units = 0;
    return WZFH_OK;
  }
  if (kind >= 2) {   // stream-kernel counters of the last launch (WZFH_TELEMETRY=1)
    const int slot = kind <= 9 ? kind - 2 : kind - 3;
    if (ms) *ms = (double)p->dev->last_telem[slot];
    if (units) *units = (long long)p->dev->last_telem[slot];
    return WZFH_OK;
  }
  if (ms) *ms = p->dev->last_ms[kind];
  if (units) *units = p->dev->last_units[kind];
  return WZFH_OK;
}

int wzfh_kernel_totals(wzfh_problem *p, wzfh_kernel_sums *out, int reset) {
  if (!p) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  return state_kernel_totals(p->dev, out, reset);
}

int wzfh_search_cells(wzfh_problem *p, const int *cells, int ncells, const wzfh_search_opts *opts,
                      wzfh_search_result *out) {
  if (!p || !cells || ncells <= 0 || !opts || !out) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  for (int i = 0; i < ncells; i++)
    if (cells[i] < 0 || cells[i] >= (int)p->cells.size()) { set_error("cell index out of range"); return WZFH_EINVAL; }
  const int K = opts->top_k;
  auto one_batch = [&](const int *cl, int nc, wzfh_search_result *res) -> int {
    std::vector<wzfh_cell_stats> st(nc);
    StreamOut so;
    int rc = run_stream(*p, cl, nc, opts->stream, st.data(), so);
    if (rc != WZFH_OK) return rc;
    TopK t;
    select_topk(p->n, so, nc, K, t);
    for (int i = 0; i < nc; i++) {
      res[i] = wzfh_search_result{};
      res[i].stream = st[i];
      res[i].hit_rank = -1; res[i].hit_idx = -1;
    }
    const int nk = (int)t.owner.size();
    std::vector<int> cellof(nk);
    for (int k = 0; k < nk; k++) cellof[k] = cl[t.owner[k]];
    std::vector<wzfh_completion> comp(nk);
    const double stream_ms = p->dev->last_ms[0];
    rc = run_complete(*p, t.cd.data(), cellof.data(), nk, opts->budget, comp.data());
    p->dev->last_ms[0] = stream_ms;
    if (rc != WZFH_OK && rc != WZFH_EINTERNAL) return rc;
    aggregate_completions(t, comp.data(), res);
    return rc;
  };
  if (opts->batched) return one_batch(cells, ncells, out);
  // plain path: one launch per cell; every cell is searched, and an error other than a
  // failed recheck (EINTERNAL, flagged per cell) fails the whole call
  double sms = 0, cms = 0;
  long long su = 0, cu = 0;
  int status = WZFH_OK;
  std::string internal_msg;
  for (int i = 0; i < ncells; i++) {
    int rc = one_batch(cells + i, 1, out + i);
    if (rc == WZFH_EINTERNAL) { status = rc; internal_msg = wzfh_last_error(); }
    else if (rc != WZFH_OK) {
      const std::string msg = wzfh_last_error();
      for (int j = 0; j < ncells; j++) out[j] = wzfh_search_result{};
      set_error(msg);
      return rc;
    }
    sms += p->dev->last_ms[0]; cms += p->dev->last_ms[1];
    su += p->dev->last_units[0]; cu += p->dev->last_units[1];
  }
  p->dev->last_ms[0] = sms; p->dev->last_ms[1] = cms;
  p->dev->last_units[0] = su; p->dev->last_units[1] = cu;
  if (status != WZFH_OK) set_error(internal_msg);
  return status;
}


int wzfh_get_tuning(const wzfh_problem *p, wzfh_tuning *out) {
  if (!p || !out) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  *out = p->dev->tune;
  return WZFH_OK;
}

int wzfh_set_tuning(wzfh_problem *p, const wzfh_tuning *t) {
  if (!p || !t) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  if (t->inner_moves < 1 || t->inner_moves > 256 || t->donate_hops < 0 || t->donate_hops > 15 ||
      t->static_window < 1 || t->complete_split < 0 || t->complete_split > 6 || t->idle_sleep < 0 ||
      t->idle_sleep > 256 || t->blocks_per_cu < 1 || t->blocks_per_cu > 4 || t->complete_cap_log2 < 0 ||
      t->complete_cap_log2 > 40) {
    set_error("tuning value out of range");
    return WZFH_EINVAL;
  }
  p->dev->tune = *t;
  return WZFH_OK;
}

// Coordinate descent over the launch settings on a sample of cells. Every trial must give
// the same per-cell results as the first: the settings only move work between lanes.
int wzfh_autotune(wzfh_problem *p, const int *cells, int ncells, const wzfh_search_opts *opts, wzfh_tuning *out) {
  if (!p || !cells || ncells <= 0 || !opts) return WZFH_EINVAL;
  if (!p->dev) { set_error("problem was set up without a device"); return WZFH_ENODEV; }
  for (int i = 0; i < ncells; i++)
    if (cells[i] < 0 || cells[i] >= (int)p->cells.size()) { set_error("cell index out of range"); return WZFH_EINVAL; }
  DeviceState *S = p->dev;
  wzfh_tuning best = S->tune;
  int trials = 0;
  std::vector<wzfh_cell_stats> ref_st;
  auto stream_trial = [&](const wzfh_tuning &t, double *ms) -> int {
    S->tune = t;
    std::vector<wzfh_cell_stats> st(ncells);
    for (int rep = 0; rep < 2; rep++) {   // best of two: launch times vary run to run
      StreamOut so;
      int rc = run_stream(*p, cells, ncells, opts->stream, st.data(), so);
      if (rc != WZFH_OK) return rc;
      trials++;
      *ms = rep == 0 ? S->last_ms[0] : std::min(*ms, S->last_ms[0]);
    }
    if (ref_st.empty()) ref_st = st;
    for (int i = 0; i < ncells; i++)
      if (st[i].cand != ref_st[i].cand || st[i].dfs_nodes != ref_st[i].dfs_nodes || st[i].leaves != ref_st[i].leaves ||
          st[i].hall_ok != ref_st[i].hall_ok) {
        set_error("autotune: stream results differ between launch settings");
        return WZFH_EKERNEL;
      }
    return WZFH_OK;
  };
  double best_ms = 0;
  int rc = stream_trial(best, &best_ms);
  if (rc != WZFH_OK) { S->tune = best; return rc; }
  const std::vector<int> moves = {4, 8, 16, 32, 64}, hops = {0, 1, 2}, wins = {2, 4, 8, 16}, sleeps = {0, 16, 64, 128, 256}, bpcs = {1, 2, 3, 4};
  for (int knob = 0; knob < 5; knob++) {
    const std::vector<int> &vals = knob == 0 ? moves : knob == 1 ? hops : knob == 2 ? wins : knob == 3 ? sleeps : bpcs;
    for (int v : vals) {
      wzfh_tuning t = best;
      int &field = knob == 0 ? t.inner_moves : knob == 1 ? t.donate_hops : knob == 2 ? t.static_window
                 : knob == 3 ? t.idle_sleep : t.blocks_per_cu;
      if (field == v) continue;
      field = v;
      double ms = 0;
      rc = stream_trial(t, &ms);
      if (rc != WZFH_OK) { S->tune = best; return rc; }
      if (ms < best_ms) { best_ms = ms; best = t; }
    }
  }
  // completer split depth, then the charge cap of split items (the longest item a launch waits
  // for: a lower cap means shorter launches and more rounds), on the sample's own top-K candidates
  wzfh_search_opts so = *opts;
  so.batched = 1;
  std::vector<wzfh_search_result> ref_res;
  double best_c = -1;
  std::vector<std::pair<int, int>> ctrials;   // (knob 0 split / 1 cap, value)
  for (int ext : {best.complete_split, 3, 4, 5}) ctrials.push_back({0, ext});
  for (int lg : {10, 12, 16, 0}) ctrials.push_back({1, lg});
  for (size_t ti = 0; ti < ctrials.size(); ti++) {
    const int knob = ctrials[ti].first, val = ctrials[ti].second;
    if (ti > 0 && ((knob == 0 && val == best.complete_split) || (knob == 1 && val == best.complete_cap_log2))) continue;
    wzfh_tuning t = best;
    if (knob == 0) t.complete_split = val;
    else t.complete_cap_log2 = val;
    S->tune = t;
    std::vector<wzfh_search_result> res(ncells);
    double ms = 0;
    for (int rep = 0; rep < 2; rep++) {   // best of two: completer times vary run to run
      rc = wzfh_search_cells(p, cells, ncells, &so, res.data());
      if (rc != WZFH_OK && rc != WZFH_EINTERNAL) { S->tune = best; return rc; }
      trials++;
      ms = rep == 0 ? S->last_ms[1] : std::min(ms, S->last_ms[1]);
    }
    if (ref_res.empty()) ref_res = res;
    for (int i = 0; i < ncells; i++)
      if (res[i].hit != ref_res[i].hit || res[i].ab_nodes != ref_res[i].ab_nodes || res[i].A != ref_res[i].A ||
          res[i].B != ref_res[i].B || res[i].clean_no != ref_res[i].clean_no || res[i].aborted != ref_res[i].aborted) {
        S->tune = best;
        set_error("autotune: completer results differ between launch settings");
        return WZFH_EKERNEL;
      }
    if (best_c < 0 || ms < best_c) { best_c = ms; best = t; }
  }
  best.stream_ms = best_ms;
  best.complete_ms = best_c;
  best.trials = trials;
  S->tune = best;
  if (out) *out = best;
  return WZFH_OK;
}

}  // extern "C"
