// CPU emulation of the HIP device API subset the selection kernels use, for
// checking kernel logic without a GPU (tests/test_kernel_emulation.py):
//   * one std::thread per work-item; workgroups run one at a time, or (the
//     batched kernel) all of a grid's workgroups at once, each with its own
//     LDS (EmuWG);
//   * __syncthreads is a barrier of the caller's workgroup that also checks
//     every work-item arrives at the same call site, and that none has left
//     the kernel while others wait;
//   * "one-lane waves": a lane's ballot is its own predicate (its own bit),
//     so the kernels' wave-aggregated LDS atomics take their per-lane path;
//   * wave-wide scans and shuffles (called by full waves only) are emulated
//     across the 64 threads of a wave;
//   * __shared__ variables become per-workgroup slots (emu_shared) that a
//     driver can poison before a launch, as LDS holds garbage on the GPU.
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
inline thread_local dim3 threadIdx;
inline thread_local dim3 blockIdx;
inline dim3 blockDim, gridDim;

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
// integer lane exchange / wave barrier / wave scan rendezvous of one wave
struct EmuWaveX { std::mutex m; std::condition_variable cv; uint32_t arrived = 0, gen = 0; int v[64]; int out[64]; };
struct EmuScan { std::mutex m; std::condition_variable cv; uint32_t arrived = 0, gen = 0; uint32_t v[64]; uint32_t out[64]; };
// one workgroup's emulation state
struct EmuWG {
  uint32_t id = 0;
  EmuBarrier bar;
  unsigned char* dyn_smem = nullptr;
  EmuWaveX wx[16], wb[16];
  EmuScan scan[16];
};
inline EmuWG emu_default_wg;
inline thread_local EmuWG* emu_wg = &emu_default_wg;
#define g_bar (emu_wg->bar)
#define g_dyn_smem (emu_wg->dyn_smem)
#define __syncthreads() g_bar.wait(__FILE__, __LINE__)

// __shared__ variables: one slot per workgroup (at most EMU_MAX_WG run at
// once); every slot of every declaration site reached so far is registered so
// that a driver can fill a workgroup's slots with garbage before it starts
constexpr uint32_t EMU_MAX_WG = 16;
struct EmuSite { unsigned char* base; size_t stride; size_t size; };
inline std::mutex emu_sites_mu;
inline std::vector<EmuSite> emu_sites;
template <class T, int ID> inline T& emu_shared() {
  struct alignas(64) Slot { T v; };
  static Slot slots[EMU_MAX_WG];
  static std::once_flag once;
  std::call_once(once, [] {
    std::lock_guard<std::mutex> lk(emu_sites_mu);
    emu_sites.push_back({reinterpret_cast<unsigned char*>(slots), sizeof(Slot), sizeof(T)});
  });
  if (emu_wg->id >= EMU_MAX_WG) { fprintf(stderr, "EMU: workgroup slot %u out of range\n", emu_wg->id); std::abort(); }
  return slots[emu_wg->id].v;
}
inline void emu_poison_wg(uint32_t id, unsigned char byte) {
  std::lock_guard<std::mutex> lk(emu_sites_mu);
  for (auto& e : emu_sites) memset(e.base + e.stride * id, byte, e.size);
}
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_system() { std::atomic_thread_fence(std::memory_order_seq_cst); }

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
inline uint64_t __ballot(int p) { return p ? 1ull << (threadIdx.x & 63) : 0ull; }  // the lane's own bit
inline uint32_t __lane_id() { return threadIdx.x & 63; }
inline int __builtin_amdgcn_readlane(int v, int) { return v; }
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t, uint32_t x) { return x; }
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t, uint32_t x) { return x; }
inline float __expf(float x) { return std::exp(x); }
float __shfl_xor(float v, int mask, int width);
// integer lane exchange inside a wave (all 64 threads of the wave must call)
inline int __shfl_xor(int v, int mask, int width) {
  (void)width;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  EmuWaveX& s = emu_wg->wx[w];
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

// wave-wide inclusive scan among the 64 threads of a wave (all must call)
inline uint32_t emu_wave_incl_scan(uint32_t v) {
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  EmuScan& s = emu_wg->scan[w];
  std::unique_lock<std::mutex> lk(s.m);
  const uint32_t g = s.gen;
  s.v[l] = v;
  if (++s.arrived == 64) {
    uint32_t acc = 0;
    for (int i = 0; i < 64; ++i) { acc += s.v[i]; s.out[i] = acc; }
    s.arrived = 0; s.gen++; s.cv.notify_all();
  } else if (!s.cv.wait_for(lk, std::chrono::seconds(20), [&] { return s.gen != g; })) {
    fprintf(stderr, "EMU: wave scan timeout (thread %u)\n", threadIdx.x); std::abort();
  }
  return s.out[l];
}

// wave-level ordering point: a barrier among the 64 threads of the caller's wave
#define __builtin_amdgcn_fence(order, scope) ((void)0)
inline void __builtin_amdgcn_wave_barrier() {
  EmuWaveX& s = emu_wg->wb[threadIdx.x >> 6];
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
