// CPU emulation of the HIP runtime and the wave intrinsics libwzfh's kernels use, for tests
// only: each block runs on host threads (one per lane), blocks one after another, and every
// wave-level operation (ballot, shuffle, readlane, MFMA) synchronizes the wave's 64 threads
// through a barrier. It lets the kernels' control logic (work splitting, reductions, budget
// recursion) be checked against the golden fixtures on a machine without a GPU. Not a model
// of memory ordering or of concurrency between blocks.
#pragma once
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <thread>
#include <vector>
struct dim3 { unsigned x = 1, y = 1, z = 1; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct EmuTid { unsigned x, y, z; };
inline thread_local EmuTid tl_tid{0, 0, 0};
inline thread_local EmuTid tl_bid{0, 0, 0};
inline EmuTid g_bdim{256, 1, 1}, g_gdim{1, 1, 1};
#define threadIdx tl_tid
#define blockIdx tl_bid
#define blockDim g_bdim
#define gridDim g_gdim
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __constant__ static
#define __shared__ static thread_local_shared
#define thread_local_shared
#define __launch_bounds__(...)
typedef int hipError_t;
typedef void *hipStream_t;
typedef double *hipEvent_t;
enum { hipSuccess = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipStreamNonBlocking = 1 };
struct hipDeviceProp_t { int multiProcessorCount; size_t sharedMemPerBlock; char name[64]; };
struct ulonglong2 { unsigned long long x, y; };
#define HIP_DYNAMIC_SHARED(type, var) static type var[1 << 15];
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
inline hipError_t hipFuncSetAttribute(const void *, int, int) { return 0; }
// ---- wave machinery: 4 waves of 64 threads per block, blocks run one after another
inline std::barrier<> *g_wbar[16];
inline std::barrier<> *g_bbar;
inline uint64_t g_wslot[16][64];
inline int wv_id() { return tl_tid.x >> 6; }
inline void wsync() { g_wbar[wv_id()]->arrive_and_wait(); }
inline unsigned long long __ballot(int p) {
  int w = wv_id();
  g_wslot[w][tl_tid.x & 63] = p ? 1 : 0; wsync();
  unsigned long long r = 0; for (int i = 0; i < 64; i++) r |= g_wslot[w][i] << i;
  wsync(); return r;
}
inline int __any(int p) { return __ballot(p) != 0; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffs(int x) { return __builtin_ffs(x); }
template <class T> T __shfl(T v, int src) {
  int w = wv_id(); uint64_t u = 0; memcpy(&u, &v, sizeof(T)); g_wslot[w][tl_tid.x & 63] = u; wsync();
  uint64_t r = g_wslot[w][src]; wsync(); T o; memcpy(&o, &r, sizeof(T)); return o;
}
inline int __builtin_amdgcn_readfirstlane(int v) { return __shfl(v, 0); }
template <class T> T __shfl_up(T v, unsigned d) { int l = (int)(tl_tid.x & 63); return __shfl(v, l >= (int)d ? l - (int)d : l); }
template <class T> T __shfl_xor(T v, int m) { return __shfl(v, (int)((tl_tid.x & 63) ^ (unsigned)m)); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
inline unsigned long long __builtin_amdgcn_s_memrealtime() { return (unsigned long long)(std::chrono::steady_clock::now().time_since_epoch().count() / 10); }
// MFMA 16x16x32 f16 (layout verified on gfx950): A lane l holds A[l%16][8(l/16)+e],
// B lane l holds B[8(l/16)+e][l%16], D lane l reg r holds D[4(l/16)+r][l%16].
typedef _Float16 emu_half8 __attribute__((ext_vector_type(8)));
typedef float emu_float4 __attribute__((ext_vector_type(4)));
inline float g_mfA[16][64][8], g_mfB[16][64][8];
inline emu_float4 __builtin_amdgcn_mfma_f32_16x16x32_f16(emu_half8 A, emu_half8 B, emu_float4 C, int, int, int) {
  const int w = wv_id(), l = tl_tid.x & 63;
  for (int e = 0; e < 8; e++) { g_mfA[w][l][e] = (float)A[e]; g_mfB[w][l][e] = (float)B[e]; }
  wsync();
  emu_float4 D = C;
  for (int r = 0; r < 4; r++) {
    const int row = 4 * (l / 16) + r, col = l % 16;
    float acc = C[r];
    for (int k = 0; k < 32; k++) acc += g_mfA[w][row + 16 * (k / 8)][k % 8] * g_mfB[w][col + 16 * (k / 8)][k % 8];
    D[r] = acc;
  }
  wsync();
  return D;
}
inline int __builtin_amdgcn_readlane(int v, int l) {
  if (l < 0 || l > 63) { fprintf(stderr, "readlane index %d\n", l); abort(); }
  return __shfl(v, l);
}
#define __builtin_amdgcn_fence(a, b) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() wsync()
inline void __syncthreads() { g_bbar->arrive_and_wait(); }
template <class T, class U> T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> T atomicSub(T *p, U v) { return __atomic_fetch_sub(p, (T)v, __ATOMIC_SEQ_CST); }
using std::min;
using std::max;
template <class T> T atomicMin(T *p, T v) { T o = __atomic_load_n(p, __ATOMIC_SEQ_CST); while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {} return o; }
inline int atomicMin_int_unused(int *p, int v) { int o = __atomic_load_n(p, __ATOMIC_SEQ_CST); while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {} return o; }
#define __hip_atomic_load(p, a, b) __atomic_load_n(p, __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, a, b) __atomic_store_n(p, v, __ATOMIC_SEQ_CST)
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
template <class T> T atomicCAS(T *p, T cmp, T val) { T o = cmp; __atomic_compare_exchange_n(p, &o, val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST); return o; }
template <class T> T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long clock64() { return (unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count(); }
inline double __dadd_rn(double a, double b) { volatile double r = a + b; return r; }
inline double __dmul_rn(double a, double b) { volatile double r = a * b; return r; }
// ---- runtime
inline const char *hipGetErrorString(hipError_t) { return "emu"; }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return 0; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return 0; }
inline hipError_t hipSetDevice(int) { return 0; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) { p->multiProcessorCount = getenv("EMU_CU") ? atoi(getenv("EMU_CU")) : 1; p->sharedMemPerBlock = getenv("EMU_LDS") ? atoi(getenv("EMU_LDS")) : 163840; return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, int) { *s = (void *)1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline double emu_now() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new double(0); return 0; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return 0; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { *e = emu_now(); return 0; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(*b - *a); return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipMalloc(void **p, size_t b) {
  *p = calloc(1, b);
  if (*p && getenv("EMU_STALE_FILL")) {   // reused memory: fill with a stale 32-bit value
    const unsigned v = (unsigned)atoi(getenv("EMU_STALE_FILL"));
    for (size_t i = 0; i + 4 <= b; i += 4) memcpy((char *)*p + i, &v, 4);
  }
  return *p ? 0 : 1;
}
inline hipError_t hipFree(void *p) { free(p); return 0; }
#ifndef __align__
#define __align__(n) __attribute__((aligned(n)))
#endif
#define HIP_SYMBOL(X) X
template <typename T> hipError_t hipGetSymbolAddress(void **p, const T &sym) { *p = (void *)&sym; return 0; }
inline hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b) { *free_b = size_t(8) << 30; *total_b = size_t(16) << 30; return 0; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t b, int) { memcpy(d, s, b); return 0; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t b, int, hipStream_t) { memcpy(d, s, b); return 0; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t b, hipStream_t) { memset(d, v, b); return 0; }
inline hipError_t hipMemset(void *d, int v, size_t b) { memset(d, v, b); return 0; }
inline void emu_launch(dim3 grid, dim3 block, const std::function<void()> &body) {
  g_bdim = {block.x, 1, 1}; g_gdim = {grid.x, 1, 1};
  int nw = (block.x + 63) / 64;
  for (unsigned b = 0; b < grid.x; b++) {
    std::barrier<> bb(block.x); g_bbar = &bb;
    std::vector<std::barrier<> *> wb;
    for (int w = 0; w < nw; w++) { g_wbar[w] = new std::barrier<>(64); }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < block.x; t++) th.emplace_back([&, t, b] { tl_tid = {t, 0, 0}; tl_bid = {b, 0, 0}; body(); });
    for (auto &x : th) x.join();
    for (int w = 0; w < nw; w++) delete g_wbar[w];
  }
}
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) emu_launch(G, B, [&] { K(__VA_ARGS__); })
