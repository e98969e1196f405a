// CPU emulation of the HIP device API subset the selection kernels use, for
// checking kernel logic without a GPU (tests/test_kernel_emulation.py):
//   * one workgroup at a time, one std::thread per work-item;
//   * __syncthreads is a barrier that also checks every work-item arrives at
//     the same call site, and that none has left the kernel while others wait;
//   * "one-lane waves": a lane's ballot is its own predicate, so the kernels'
//     wave-aggregated LDS atomics take their per-lane path;
//   * wave-wide scans and shuffles (called by full waves only) are emulated
//     across the 64 threads of a wave;
//   * __shared__ variables become statics (one workgroup runs at a time).
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
#include <cmath>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __shared__ static
#define __HIPCC__ 1

struct dim3 { uint32_t x = 1, y = 1, z = 1; dim3() {} dim3(uint32_t a, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {} };
struct uint4 { uint32_t x, y, z, w; };
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
extern thread_local dim3 threadIdx;
extern dim3 blockIdx, blockDim, gridDim;

struct EmuBarrier {
  std::mutex m; std::condition_variable cv;
  uint32_t n = 0, arrived = 0, gen = 0, exited = 0;
  const char* site_file = nullptr; int site_line = -1; bool bad = false;
  void reset(uint32_t nt) { n = nt; arrived = 0; gen = 0; exited = 0; site_line = -1; bad = false; }
  void wait(const char* f, int line) {
    std::unique_lock<std::mutex> lk(m);
    if (arrived == 0) { site_file = f; site_line = line; }
    else if (line != site_line) {
      fprintf(stderr, "EMU: barrier mismatch: thread %u at %s:%d, others at %s:%d (block %u)\n", threadIdx.x, f, line,
              site_file, site_line, blockIdx.x);
      bad = true;
    }
    if (exited) { fprintf(stderr, "EMU: barrier at %s:%d after %u threads exited (block %u)\n", f, line, exited, blockIdx.x); std::abort(); }
    uint32_t g = gen;
    if (++arrived == n) { arrived = 0; gen++; cv.notify_all(); return; }
    if (!cv.wait_for(lk, std::chrono::seconds(20), [&] { return gen != g; })) {
      fprintf(stderr, "EMU: barrier timeout at %s:%d (%u of %u arrived, exited %u, block %u)\n", f, line, arrived, n,
              exited, blockIdx.x);
      std::abort();
    }
  }
  void exit_thread() {
    std::unique_lock<std::mutex> lk(m);
    exited++;
    if (arrived) { fprintf(stderr, "EMU: thread %u exits while %u threads wait at %s:%d (block %u)\n", threadIdx.x, arrived, site_file, site_line, blockIdx.x); std::abort(); }
  }
};
extern EmuBarrier g_bar;
#define __syncthreads() g_bar.wait(__FILE__, __LINE__)
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }

template <class T> inline T atomicAdd(T* p, T v) { return std::atomic_ref<T>(*p).fetch_add(v); }
inline float atomicAdd(float* p, float v) { return std::atomic_ref<float>(*p).fetch_add(v); }
template <class T> inline T atomicOr(T* p, T v) { return std::atomic_ref<T>(*p).fetch_or(v); }
template <class T> inline T atomicExch(T* p, T v) { return std::atomic_ref<T>(*p).exchange(v); }
template <class T> inline T atomicMin(T* p, T v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (v < cur && !a.compare_exchange_weak(cur, v)) {
  }
  return cur;
}
template <class T> inline T atomicMax(T* p, T v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (v > cur && !a.compare_exchange_weak(cur, v)) {
  }
  return cur;
}
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline uint64_t __ballot(int p) { return p ? 1ull : 0ull; }
inline uint32_t __lane_id() { return threadIdx.x & 63; }
inline int __builtin_amdgcn_readlane(int v, int) { return v; }
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t, uint32_t x) { return x; }
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t, uint32_t x) { return x; }
inline float __expf(float x) { return std::exp(x); }
float __shfl_xor(float v, int mask, int width);
// integer lane exchange inside a wave (all 64 threads of the wave must call)
struct EmuWaveX { std::mutex m; std::condition_variable cv; uint32_t arrived = 0, gen = 0; int v[64]; int out[64]; };
inline EmuWaveX g_emu_wx[16];
inline int __shfl_xor(int v, int mask, int width) {
  (void)width;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  EmuWaveX& s = g_emu_wx[w];
  std::unique_lock<std::mutex> lk(s.m);
  const uint32_t g = s.gen;
  s.v[l] = v;
  if (++s.arrived == 64) {
    for (int i = 0; i < 64; ++i) s.out[i] = s.v[i ^ (mask & 63)];
    s.arrived = 0;
    s.gen++;
    s.cv.notify_all();
  } else if (!s.cv.wait_for(lk, std::chrono::seconds(20), [&] { return s.gen != g; })) {
    fprintf(stderr, "EMU: shuffle timeout (thread %u)\n", threadIdx.x);
    std::abort();
  }
  return s.out[l];
}
inline unsigned long long wall_clock64() { return 0; }
extern unsigned char* g_dyn_smem;

// wave-wide inclusive scan among the 64 threads of a wave (all must call)
uint32_t emu_wave_incl_scan(uint32_t v);

// wave-level ordering point: a barrier among the 64 threads of the caller's wave
#define __builtin_amdgcn_fence(order, scope) ((void)0)
inline EmuWaveX g_emu_wb[16];
inline void __builtin_amdgcn_wave_barrier() {
  EmuWaveX& s = g_emu_wb[threadIdx.x >> 6];
  std::unique_lock<std::mutex> lk(s.m);
  const uint32_t g = s.gen;
  if (++s.arrived == 64) {
    s.arrived = 0;
    s.gen++;
    s.cv.notify_all();
  } else if (!s.cv.wait_for(lk, std::chrono::seconds(20), [&] { return s.gen != g; })) {
    fprintf(stderr, "EMU: wave barrier timeout (thread %u)\n", threadIdx.x);
    std::abort();
  }
}
