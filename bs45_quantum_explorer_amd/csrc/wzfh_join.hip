// wzfh_join.hip -- the WZ_JOIN22 hash join (wz_match.cpp:2905-3104) on one MI355X.
//
// BUILD: the C,D stream of every C,D profile (count_pairs22 on the C,D side with the leaf
// test, the same stream kernel as the first-hit path) goes into an open-addressed table in
// HBM keyed by the exact autocorrelation vector v(s) = -(N_C(s) + N_D(s)), s = 1..n. The
// reference keys a 64-bit FNV-1a of that vector and re-enumerates C,D to resolve hits
// (:855-871, :3039-3090); here the vector itself is stored, so equal keys are equal
// vectors. One entry per distinct vector, holding the earliest C,D in stream order
// (profile order, then DFS order).
// STREAM: every A,B candidate (count_pairs22 on the A,B side: length n+1, P22_NEG roots)
// probes with N_A(s) + N_B(s); a match means NPAF(s) = 0 for s = 1..n, i.e. a solution.
//
// Table layout (structure of arrays): tag u64 (a 64-bit hash of the vector, 0 = empty),
// order u64, C,D u64 x 2, vector int16 x vs. Inserts are lock-free and never wait on
// another thread: a claim kernel takes empty slots with one CAS (or meets an equal tag),
// the claimers write their entries, a verify kernel compares equal-tag meetings with the
// stored vector (a true 64-bit collision probes on in another round), and a final kernel
// lets the earliest duplicate write its C,D. Probing is HBM-latency / bandwidth bound:
// per A,B candidate one 8 B tag read per probed slot plus 2n B on a tag match.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "wzfh_internal.h"

namespace wzfh {
namespace {

enum : uint8_t { K_TODO = 0, K_NEWOWNER = 1, K_SAMETAG = 2, K_DUP = 3, K_OWNER = 4, K_FULL = 5 };

__device__ __forceinline__ int npaf1(uint64_t x, int len, int s) {
  if (s >= len) return 0;
  const int w = len - s;
  const uint64_t m = w >= 64 ? ~0ull : ((1ull << w) - 1ull);
  return w - 2 * __popcll((x ^ (x >> s)) & m);
}
// Key entry s of a pair: sign * (N_X(s) + N_Y(s)) over length len.
__device__ __forceinline__ int keyv(uint64_t X, uint64_t Y, int len, int sign, int s) {
  return sign * (npaf1(X, len, s) + npaf1(Y, len, s));
}
__device__ uint64_t key_hash(uint64_t X, uint64_t Y, int len, int sign, int n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (int s = 1; s <= n; s++) {
    h ^= (uint64_t)(uint16_t)(int16_t)keyv(X, Y, len, sign, s);
    h *= 0x100000001b3ull;
  }
  h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
  return h ? h : 1ull;
}

struct Table {
  unsigned long long *tag, *order;
  uint64_t *cd;
  int16_t *vec;
  unsigned long long mask;
  int vs, n;
};

__device__ bool vec_equal(const Table &t, unsigned long long slot, uint64_t X, uint64_t Y, int len, int sign) {
  const int16_t *v = t.vec + slot * (unsigned long long)t.vs;
  for (int s = 1; s <= t.n; s++)
    if (v[s - 1] != (int16_t)keyv(X, Y, len, sign, s)) return false;
  return true;
}

// C,D candidates i < N that still need a slot take one: an empty slot by CAS (new owner),
// or the first slot whose tag equals theirs (same tag: compared in join_verify).
__global__ void join_claim(Table t, const uint64_t *cdc, unsigned long long N, uint8_t *kind,
                           unsigned long long *slot_of, unsigned long long *owners) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || kind[i] != K_TODO) return;
  const uint64_t X = cdc[2 * i], Y = cdc[2 * i + 1];
  const unsigned long long h = key_hash(X, Y, t.n, -1, t.n);
  unsigned long long slot = slot_of[i] == ~0ull ? (h & t.mask) : slot_of[i];
  for (unsigned long long probe = 0; probe <= t.mask; probe++) {
    unsigned long long cur = __hip_atomic_load(&t.tag[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0) {
      cur = atomicCAS(&t.tag[slot], 0ull, h);
      if (cur == 0) {
        kind[i] = K_NEWOWNER;
        slot_of[i] = slot;
        atomicAdd(owners, 1ull);
        return;
      }
    }
    if (cur == h) {
      kind[i] = K_SAMETAG;
      slot_of[i] = slot;
      return;
    }
    slot = (slot + 1) & t.mask;
  }
  kind[i] = K_FULL;
}

__global__ void join_write(Table t, const uint64_t *cdc, const unsigned long long *ord, unsigned long long N,
                           uint8_t *kind, const unsigned long long *slot_of) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || kind[i] != K_NEWOWNER) return;
  const uint64_t X = cdc[2 * i], Y = cdc[2 * i + 1];
  const unsigned long long slot = slot_of[i];
  int16_t *v = t.vec + slot * (unsigned long long)t.vs;
  for (int s = 1; s <= t.n; s++) v[s - 1] = (int16_t)keyv(X, Y, t.n, -1, s);
  t.order[slot] = ord[i];
  t.cd[2 * slot] = X;
  t.cd[2 * slot + 1] = Y;
  kind[i] = K_OWNER;
}

// Same tag: an equal vector is a duplicate (the earliest order wins the entry); a different
// one (a true 64-bit collision) probes on from the next slot in another claim round.
__global__ void join_verify(Table t, const uint64_t *cdc, const unsigned long long *ord, unsigned long long N,
                            uint8_t *kind, unsigned long long *slot_of, unsigned long long *reprobe) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || kind[i] != K_SAMETAG) return;
  const uint64_t X = cdc[2 * i], Y = cdc[2 * i + 1];
  const unsigned long long slot = slot_of[i];
  if (vec_equal(t, slot, X, Y, t.n, -1)) {
    atomicMin(&t.order[slot], ord[i]);
    kind[i] = K_DUP;
  } else {
    slot_of[i] = (slot + 1) & t.mask;
    kind[i] = K_TODO;
    atomicAdd(reprobe, 1ull);
  }
}

// The entry's C,D becomes the earliest candidate with that vector (orders are unique).
__global__ void join_rep(Table t, const uint64_t *cdc, const unsigned long long *ord, unsigned long long N,
                         const uint8_t *kind, const unsigned long long *slot_of) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || kind[i] != K_DUP) return;
  const unsigned long long slot = slot_of[i];
  if (t.order[slot] == ord[i]) {
    t.cd[2 * slot] = cdc[2 * i];
    t.cd[2 * slot + 1] = cdc[2 * i + 1];
  }
}

struct Hit {
  unsigned long long ab;     // index of the A,B candidate in its window
  unsigned long long slot;
};

__global__ void join_probe(Table t, const uint64_t *abc, unsigned long long N, int len, Hit *hits,
                           unsigned long long cap, unsigned long long *nhits) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint64_t X = abc[2 * i], Y = abc[2 * i + 1];
  const unsigned long long h = key_hash(X, Y, len, 1, t.n);
  unsigned long long slot = h & t.mask;
  for (unsigned long long probe = 0; probe <= t.mask; probe++) {
    const unsigned long long cur = t.tag[slot];
    if (cur == 0) return;
    if (cur == h && vec_equal(t, slot, X, Y, len, 1)) {
      const unsigned long long k = atomicAdd(nhits, 1ull);
      if (k < cap) hits[k] = Hit{i, slot};
      return;
    }
    slot = (slot + 1) & t.mask;
  }
}

// Member lists: every C,D candidate of the build is linked into the list of its table slot
// (the slot of its exact vector), so a key's C,D are found without re-streaming.
__global__ void join_chain(const unsigned long long *slot_of, unsigned long long N, unsigned long long *head,
                           unsigned long long *next) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  next[i] = atomicExch(&head[slot_of[i]], i);
}

// Solutions of each A,B hit: every C,D on its slot's list. Pass 1 counts, pass 2 writes
// (A, B, C, D) from the hit's offset.
__global__ void join_sol_count(const Hit *hits, unsigned long long nh, const unsigned long long *head,
                               const unsigned long long *next, unsigned long long *cnt) {
  const unsigned long long h = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  unsigned long long c = 0;
  for (unsigned long long j = head[hits[h].slot]; j != ~0ull; j = next[j]) c++;
  cnt[h] = c;
}

__global__ void join_sol_emit(const Hit *hits, unsigned long long nh, const unsigned long long *head,
                              const unsigned long long *next, const unsigned long long *offs, const uint64_t *abxy,
                              const uint64_t *cdxy, uint64_t *out) {
  const unsigned long long h = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  unsigned long long o = offs[h];
  const uint64_t A = abxy[2 * hits[h].ab], B = abxy[2 * hits[h].ab + 1];
  for (unsigned long long j = head[hits[h].slot]; j != ~0ull; j = next[j], o++) {
    out[4 * o] = A;
    out[4 * o + 1] = B;
    out[4 * o + 2] = cdxy[2 * j];
    out[4 * o + 3] = cdxy[2 * j + 1];
  }
}

template <typename T> int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc failed (join)");
    return WZFH_ENOMEM;
  }
  return WZFH_OK;
}
template <typename T> void dfree(T *&p) { if (p) (void)hipFree(p); p = nullptr; }

unsigned blocks_for(unsigned long long n) { return (unsigned)std::max<unsigned long long>(1, (n + 255) / 256); }

}  // namespace

// The whole stream of one side in profile order: candidates (X, Y) and their order keys
// (profile index << 40 | DFS index of the leaf).
static int stream_side(wzfh_problem &P, std::vector<uint64_t> &xy, std::vector<unsigned long long> &ord,
                       long long &count) {
  const int ncells = (int)P.cells.size();
  const int window = 4096;
  wzfh_stream_opts o{};
  std::vector<int> cl;
  std::vector<wzfh_cell_stats> st;
  for (int w0 = 0; w0 < ncells; w0 += window) {
    const int nc = std::min(window, ncells - w0);
    cl.resize(nc);
    st.resize(nc);
    for (int i = 0; i < nc; i++) cl[i] = w0 + i;
    StreamOut out;
    int rc = run_stream(P, cl.data(), nc, o, st.data(), out);
    if (rc != WZFH_OK) return rc;
    for (int cs = 0; cs < nc; cs++)
      for (long long j = out.cell_begin[cs]; j < out.cell_begin[cs + 1]; j++) {
        xy.push_back(out.cands[j].X);
        xy.push_back(out.cands[j].Y);
        ord.push_back(((unsigned long long)(w0 + cs) << 40) | (unsigned long long)out.cands[j].ell);
      }
  }
  count = (long long)ord.size();
  return WZFH_OK;
}

}  // namespace wzfh

using namespace wzfh;

struct wzfh_join {
  wzfh_problem *cd = nullptr, *ab = nullptr;
  int n = 0;
  hipStream_t stream = nullptr;
  Table t{};
  unsigned long long slots = 0;
  unsigned long long *ctr = nullptr;   // [0] owners, [1] reprobe, [2] hits
  bool built = false, probed = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double build_ms = 0, probe_ms = 0;   // device time of the table kernels (last build / probe)
  // kept for wzfh_join_solutions: the C,D candidates and their slots (build), the A,B
  // candidates and the hits (probe), the slots' member lists (built on first use)
  uint64_t *cdxy = nullptr, *abxy = nullptr;
  unsigned long long *cd_slot = nullptr, *head = nullptr, *next = nullptr;
  Hit *hits = nullptr;
  unsigned long long ncd = 0, nhits = 0;
  void drop_build() { dfree(cdxy); dfree(cd_slot); dfree(head); dfree(next); ncd = 0; }
  void drop_probe() { dfree(abxy); dfree(hits); nhits = 0; probed = false; }
};

extern "C" {

int wzfh_join_open(const wzfh_config *cfg, wzfh_join **out) {
  if (!cfg || !out) return WZFH_EINVAL;
  *out = nullptr;
  wzfh_config c = *cfg;
  c.m6 = 0; c.canon = 0; c.order = 0; c.endpos = 0; c.ab_prof = 0; c.orbit_q = 0;   // JOIN22 uses cdProfs / abProfs
  wzfh_join *J = new wzfh_join();
  J->n = c.n;
  int rc = problem_init(c, false, true, &J->cd);
  if (rc == WZFH_OK) rc = problem_init(c, true, true, &J->ab);
  if (rc == WZFH_OK && hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed (join)");
    rc = WZFH_EDEVICE;
  }
  if (rc == WZFH_OK) rc = dalloc(&J->ctr, 4);
  if (rc == WZFH_OK && (hipEventCreate(&J->ev[0]) != hipSuccess || hipEventCreate(&J->ev[1]) != hipSuccess)) {
    set_error("hipEventCreate failed (join)");
    rc = WZFH_EDEVICE;
  }
  if (rc != WZFH_OK) { wzfh_join_close(J); return rc; }
  *out = J;
  return WZFH_OK;
}

}  // extern "C"

// The table from N C,D pairs xy (2 words each) with their order keys (the earliest order of a
// vector keeps the entry): claim / write / verify rounds until no equal-tag meeting probes on,
// then the earliest duplicate writes its C,D. Device time of the table kernels -> J->build_ms.
static int build_table(wzfh_join *J, const uint64_t *xy, const unsigned long long *ord, long long N, int slots_log2,
                       wzfh_join_result *r) {
  int rc;
  unsigned long long slots = 1024;
  if (slots_log2 > 0) {
    if (slots_log2 > 40) { set_error("slots_log2 too large"); return WZFH_EINVAL; }
    slots = 1ull << slots_log2;
  } else {
    while (slots < 2ull * (unsigned long long)N) slots <<= 1;
  }
  if ((unsigned long long)N > slots) { set_error("join table smaller than the C,D stream"); return WZFH_ECAPACITY; }
  const int vs = (J->n + 3) & ~3;
  Table &t = J->t;
  dfree(t.tag); dfree(t.order); dfree(t.cd); dfree(t.vec);
  if ((rc = dalloc(&t.tag, slots)) || (rc = dalloc(&t.order, slots)) || (rc = dalloc(&t.cd, 2 * slots)) ||
      (rc = dalloc(&t.vec, slots * (unsigned long long)vs)))
    return rc;
  t.mask = slots - 1;
  t.vs = vs;
  t.n = J->n;
  J->slots = slots;
  J->drop_build();
  J->drop_probe();
  J->built = false;
  uint64_t *dxy = nullptr;
  unsigned long long *dord = nullptr, *slot_of = nullptr;
  uint8_t *kind = nullptr;
  auto cleanup = [&]() { dfree(dord); dfree(kind); };
  if ((rc = dalloc(&dxy, 2 * (size_t)N)) || (rc = dalloc(&dord, (size_t)N)) || (rc = dalloc(&slot_of, (size_t)N)) ||
      (rc = dalloc(&kind, (size_t)N))) {
    cleanup();
    dfree(dxy);
    dfree(slot_of);
    return rc;
  }
  J->cdxy = dxy;
  J->cd_slot = slot_of;
  J->ncd = (unsigned long long)N;
  hipStream_t s = J->stream;
  bool ok = hipMemsetAsync(t.tag, 0, slots * 8, s) == hipSuccess &&
            hipMemsetAsync(J->ctr, 0, 32, s) == hipSuccess &&
            hipMemsetAsync(kind, 0, (size_t)N, s) == hipSuccess &&
            hipMemsetAsync(slot_of, 0xFF, (size_t)N * 8, s) == hipSuccess &&
            (N == 0 || (hipMemcpyAsync(dxy, xy, (size_t)N * 16, hipMemcpyHostToDevice, s) == hipSuccess &&
                        hipMemcpyAsync(dord, ord, (size_t)N * 8, hipMemcpyHostToDevice, s) == hipSuccess));
  const unsigned nb = blocks_for((unsigned long long)N);
  J->build_ms = 0;
  float ems = 0;
  for (int round = 0; ok && N > 0 && round < 16; round++) {
    ok = hipEventRecord(J->ev[0], s) == hipSuccess;
    hipLaunchKernelGGL(join_claim, dim3(nb), dim3(256), 0, s, t, dxy, (unsigned long long)N, kind, slot_of, J->ctr);
    hipLaunchKernelGGL(join_write, dim3(nb), dim3(256), 0, s, t, dxy, dord, (unsigned long long)N, kind, slot_of);
    ok = hipMemsetAsync(J->ctr + 1, 0, 8, s) == hipSuccess;
    hipLaunchKernelGGL(join_verify, dim3(nb), dim3(256), 0, s, t, dxy, dord, (unsigned long long)N, kind, slot_of,
                       J->ctr + 1);
    ok = ok && hipEventRecord(J->ev[1], s) == hipSuccess;
    unsigned long long re = 0;
    ok = ok && hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(&re, J->ctr + 1, 8, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess && hipEventElapsedTime(&ems, J->ev[0], J->ev[1]) == hipSuccess;
    J->build_ms += ems;
    if (re == 0) break;
  }
  if (ok && N > 0) {
    ok = hipEventRecord(J->ev[0], s) == hipSuccess;
    hipLaunchKernelGGL(join_rep, dim3(nb), dim3(256), 0, s, t, dxy, dord, (unsigned long long)N, kind, slot_of);
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(J->ev[1], s) == hipSuccess &&
         hipEventSynchronize(J->ev[1]) == hipSuccess && hipEventElapsedTime(&ems, J->ev[0], J->ev[1]) == hipSuccess;
    J->build_ms += ems;
  }
  unsigned long long owners = 0;
  std::vector<uint8_t> kh((size_t)N);
  ok = ok && hipMemcpyAsync(&owners, J->ctr, 8, hipMemcpyDeviceToHost, s) == hipSuccess &&
       (N == 0 || hipMemcpyAsync(kh.data(), kind, (size_t)N, hipMemcpyDeviceToHost, s) == hipSuccess) &&
       hipStreamSynchronize(s) == hipSuccess;
  cleanup();
  if (!ok) { set_error("join build: device call failed"); return WZFH_EDEVICE; }
  for (uint8_t k : kh)
    if (k != K_OWNER && k != K_DUP) { set_error("join build: table full or unresolved insert"); return WZFH_ECAPACITY; }
  r->distinct_keys = (long long)owners;
  r->slots = (long long)slots;
  J->built = true;
  return WZFH_OK;
}

extern "C" {

int wzfh_join_build(wzfh_join *J, int slots_log2, wzfh_join_result *r) {
  if (!J || !r) return WZFH_EINVAL;
  *r = wzfh_join_result{};
  r->cd_profiles = (long long)J->cd->cells.size();
  r->ab_profiles = (long long)J->ab->cells.size();
  std::vector<uint64_t> xy;
  std::vector<unsigned long long> ord;
  long long N = 0;
  int rc = stream_side(*J->cd, xy, ord, N);
  if (rc != WZFH_OK) return rc;
  r->cd_stream = N;
  return build_table(J, xy.data(), ord.data(), N, slots_log2, r);
}

int wzfh_join_build_from(wzfh_join *J, const uint64_t *cd, long long ncd, int slots_log2, wzfh_join_result *r) {
  if (!J || !r || ncd < 0 || (ncd > 0 && !cd)) return WZFH_EINVAL;
  *r = wzfh_join_result{};
  r->cd_stream = ncd;
  std::vector<unsigned long long> ord((size_t)ncd);
  for (long long i = 0; i < ncd; i++) ord[(size_t)i] = (unsigned long long)i;
  return build_table(J, cd, ord.data(), ncd, slots_log2, r);
}

}  // extern "C"

// Probes the table with N A,B pairs xy (order keys ord decide which hit is reported).
static int probe_table(wzfh_join *J, const uint64_t *xy, const unsigned long long *ord, long long N,
                       wzfh_join_result *r) {
  int rc;
  r->ab_stream = N;
  r->hits = 0;
  r->found = 0;
  J->drop_probe();
  if (N == 0) { J->probed = true; return WZFH_OK; }
  hipStream_t s = J->stream;
  J->drop_probe();
  uint64_t *dxy = nullptr;
  Hit *hits = nullptr;
  const unsigned long long cap = (unsigned long long)N;
  if ((rc = dalloc(&dxy, 2 * (size_t)N)) || (rc = dalloc(&hits, (size_t)cap))) { dfree(dxy); dfree(hits); return rc; }
  J->abxy = dxy;
  J->hits = hits;
  unsigned long long nh = 0;
  bool ok = hipMemsetAsync(J->ctr + 2, 0, 8, s) == hipSuccess &&
            hipMemcpyAsync(dxy, xy, (size_t)N * 16, hipMemcpyHostToDevice, s) == hipSuccess;
  J->probe_ms = 0;
  if (ok) {
    float ems = 0;
    ok = hipEventRecord(J->ev[0], s) == hipSuccess;
    hipLaunchKernelGGL(join_probe, dim3(blocks_for((unsigned long long)N)), dim3(256), 0, s, J->t, dxy,
                       (unsigned long long)N, J->n + 1, hits, cap, J->ctr + 2);
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(J->ev[1], s) == hipSuccess &&
         hipMemcpyAsync(&nh, J->ctr + 2, 8, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess && hipEventElapsedTime(&ems, J->ev[0], J->ev[1]) == hipSuccess;
    J->probe_ms = ems;
  }
  std::vector<Hit> hh(std::min(nh, cap));
  ok = ok && (hh.empty() || hipMemcpy(hh.data(), hits, hh.size() * sizeof(Hit), hipMemcpyDeviceToHost) == hipSuccess);
  if (!ok) { J->drop_probe(); set_error("join probe: device call failed"); return WZFH_EDEVICE; }
  J->nhits = std::min(nh, cap);
  J->probed = true;
  r->hits = (long long)nh;
  if (hh.empty()) return WZFH_OK;
  // The reported solution is the one the reference's single-threaded resolve prints: it
  // re-enumerates C,D in stream order and stops at the first C,D whose key has hits
  // (:3044-3090), i.e. the entry with the smallest order; among that key's hits it takes
  // the first of hitmap.equal_range (:3057-3060), and libstdc++'s unordered_multimap puts
  // a new element in front of its equal keys, so that is the hit streamed last.
  unsigned long long best_ord = ~0ull, best_slot = 0, best_ab = ~0ull;
  uint64_t A = 0, B = 0;
  for (const Hit &h : hh) {
    unsigned long long o = 0;
    if (hipMemcpy(&o, J->t.order + h.slot, 8, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("join probe: device call failed");
      return WZFH_EDEVICE;
    }
    const unsigned long long oab = ord[h.ab];
    if (o < best_ord || (o == best_ord && oab > best_ab)) {
      best_ord = o; best_slot = h.slot; best_ab = oab;
      A = xy[2 * h.ab]; B = xy[2 * h.ab + 1];
    }
  }
  uint64_t cdw[2];
  if (hipMemcpy(cdw, J->t.cd + 2 * best_slot, 16, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("join probe: device call failed");
    return WZFH_EDEVICE;
  }
  r->found = 1;
  r->A = A; r->B = B; r->C = cdw[0]; r->D = cdw[1];
  return WZFH_OK;
}

extern "C" {

int wzfh_join_probe(wzfh_join *J, wzfh_join_result *r) {
  if (!J || !r) return WZFH_EINVAL;
  if (!J->built) { set_error("join probe before build"); return WZFH_EINVAL; }
  std::vector<uint64_t> xy;
  std::vector<unsigned long long> ord;
  long long N = 0;
  int rc = stream_side(*J->ab, xy, ord, N);
  if (rc != WZFH_OK) return rc;
  return probe_table(J, xy.data(), ord.data(), N, r);
}

int wzfh_join_probe_from(wzfh_join *J, const uint64_t *ab, long long nab, wzfh_join_result *r) {
  if (!J || !r || nab < 0 || (nab > 0 && !ab)) return WZFH_EINVAL;
  if (!J->built) { set_error("join probe before build"); return WZFH_EINVAL; }
  std::vector<unsigned long long> ord((size_t)nab);
  for (long long i = 0; i < nab; i++) ord[(size_t)i] = (unsigned long long)i;
  return probe_table(J, ab, ord.data(), nab, r);
}

int wzfh_join_kernel_ms(const wzfh_join *J, double *build_ms, double *probe_ms) {
  if (!J) return WZFH_EINVAL;
  if (build_ms) *build_ms = J->build_ms;
  if (probe_ms) *probe_ms = J->probe_ms;
  return WZFH_OK;
}

int wzfh_join_solutions(wzfh_join *J, uint64_t *out, long long cap, long long *count) {
  if (!J || !count || cap < 0) return WZFH_EINVAL;
  *count = 0;
  if (!J->probed) { set_error("join solutions before probe"); return WZFH_EINVAL; }
  if (J->nhits == 0) return WZFH_OK;
  hipStream_t s = J->stream;
  int rc;
  if (!J->head) {
    if ((rc = dalloc(&J->head, (size_t)J->slots)) || (rc = dalloc(&J->next, (size_t)std::max(1ull, J->ncd)))) {
      dfree(J->head); dfree(J->next);
      return rc;
    }
    bool ok = hipMemsetAsync(J->head, 0xFF, J->slots * 8, s) == hipSuccess;
    if (ok && J->ncd) {
      hipLaunchKernelGGL(join_chain, dim3(blocks_for(J->ncd)), dim3(256), 0, s, J->cd_slot, J->ncd, J->head, J->next);
      ok = hipGetLastError() == hipSuccess;
    }
    if (!ok) { dfree(J->head); dfree(J->next); set_error("join solutions: device call failed"); return WZFH_EDEVICE; }
  }
  const unsigned long long nh = J->nhits;
  unsigned long long *cnt = nullptr;
  if ((rc = dalloc(&cnt, (size_t)nh))) return rc;
  std::vector<unsigned long long> offs(nh);
  hipLaunchKernelGGL(join_sol_count, dim3(blocks_for(nh)), dim3(256), 0, s, J->hits, nh, J->head, J->next, cnt);
  bool ok = hipGetLastError() == hipSuccess &&
            hipMemcpyAsync(offs.data(), cnt, nh * 8, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess;
  if (!ok) { dfree(cnt); set_error("join solutions: device call failed"); return WZFH_EDEVICE; }
  unsigned long long total = 0;
  for (unsigned long long h = 0; h < nh; h++) { const unsigned long long c = offs[h]; offs[h] = total; total += c; }
  *count = (long long)total;
  if (!out) { dfree(cnt); return WZFH_OK; }
  if ((long long)total > cap) { dfree(cnt); set_error("join solutions: output buffer too small"); return WZFH_ECAPACITY; }
  uint64_t *dout = nullptr;
  if ((rc = dalloc(&dout, 4 * (size_t)total))) { dfree(cnt); return rc; }
  ok = hipMemcpyAsync(cnt, offs.data(), nh * 8, hipMemcpyHostToDevice, s) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(join_sol_emit, dim3(blocks_for(nh)), dim3(256), 0, s, J->hits, nh, J->head, J->next, cnt,
                       J->abxy, J->cdxy, dout);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(out, dout, 32 * (size_t)total, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
  }
  dfree(dout);
  dfree(cnt);
  if (!ok) { set_error("join solutions: device call failed"); return WZFH_EDEVICE; }
  // ascending (A, B, C, D): the order does not depend on how the device scheduled the lists
  struct Q { uint64_t w[4]; };
  Q *q = (Q *)out;
  std::sort(q, q + total, [](const Q &x, const Q &y) {
    return std::lexicographical_compare(x.w, x.w + 4, y.w, y.w + 4);
  });
  return WZFH_OK;
}

void wzfh_join_close(wzfh_join *J) {
  if (!J) return;
  J->drop_build();
  J->drop_probe();
  for (auto &e : J->ev) if (e) (void)hipEventDestroy(e);
  dfree(J->t.tag); dfree(J->t.order); dfree(J->t.cd); dfree(J->t.vec);
  dfree(J->ctr);
  if (J->stream) (void)hipStreamDestroy(J->stream);
  if (J->cd) wzfh_destroy(J->cd);
  if (J->ab) wzfh_destroy(J->ab);
  delete J;
}

}  // extern "C"
