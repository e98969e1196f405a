// CPU emulation of the HIP device API subset the selection kernels use, for
// checking kernel logic without a GPU (tests/test_kernel_emulation.py):
//   * one std::thread per work-item; workgroups run one at a time, or (the
//     batched kernel) all of a grid's workgroups at once, each with its own
//     LDS (EmuWG);
//   * __syncthreads is a barrier of the caller's workgroup that also checks
//     every work-item arrives at the same call site, and that none has left
//     the kernel while others wait;
//   * 64-lane waves: ballots, readlanes, shuffles and scans are rendezvous of
//     the wave's active lanes (those not at the workgroup barrier or gone),
//     so the kernels' wave-aggregated and peeled LDS atomics take the paths
//     they take on the GPU (see EmuWave);
//   * __shared__ variables become per-workgroup slots (emu_shared) that a
//     driver can poison before a launch, as LDS holds garbage on the GPU.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
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

// One wave of 64 lanes.  A wave operation (ballot, readlane, readfirstlane,
// shuffle, scan, wave barrier) completes once every lane of the wave is
// blocked somewhere -- at a wave operation, at the workgroup barrier, or gone
// from the kernel -- and the lanes waiting at that operation are its active
// lanes, as on the GPU, where the lanes that took the other branch or left a
// loop early are masked off.  Lanes waiting at different operations (a branch
// with an operation on each side, or a loop that some lanes have left) are
// released one call site at a time, lowest code address first: the body of a
// branch or loop lies before the code after it, so the lanes that skipped it
// wait there for the others, as they do at the reconvergence point on the GPU.
enum EmuOp : int { EMU_BALLOT, EMU_READLANE, EMU_READFIRST, EMU_SHFL_XOR, EMU_SCAN, EMU_WBAR };
struct EmuWave {
  std::mutex m; std::condition_variable cv;
  uint32_t lanes = 0;               // lanes of this wave in the launch
  uint32_t n_bar = 0, n_exit = 0;   // lanes at the workgroup barrier / gone
  uint64_t waiting = 0, ready = 0;  // lanes at a wave operation / released with their result
  const void* site[64]; int op[64]; uint64_t val[64]; int arg[64]; uint64_t out[64];
};
// counts over all waves: operations, those with more than one active lane,
// those with some lanes of the wave inactive, releases that left lanes waiting
// at another site, readlanes from an inactive lane
struct EmuWaveStats { std::atomic<uint64_t> ops{0}, multi{0}, partial{0}, split{0}, inactive_src{0}; };
inline EmuWaveStats emu_wave_stats;
inline void emu_print_wave_stats() {
  printf("waves: %llu operations, %llu with several active lanes, %llu with some lanes inactive, %llu split releases, "
         "%llu inactive readlane sources\n", (unsigned long long)emu_wave_stats.ops.load(),
         (unsigned long long)emu_wave_stats.multi.load(), (unsigned long long)emu_wave_stats.partial.load(),
         (unsigned long long)emu_wave_stats.split.load(), (unsigned long long)emu_wave_stats.inactive_src.load());
}
// caller holds s.m
inline void emu_try_release(EmuWave& s) {
  if (!s.waiting || (uint32_t)__builtin_popcountll(s.waiting) + s.n_bar + s.n_exit < s.lanes) return;
  const void* site = nullptr;
  for (uint64_t w = s.waiting; w; w &= w - 1) {
    const int l = __builtin_ctzll(w);
    if (!site || std::less<const void*>()(s.site[l], site)) site = s.site[l];
  }
  uint64_t grp = 0;
  for (uint64_t w = s.waiting; w; w &= w - 1) {
    const int l = __builtin_ctzll(w);
    if (s.site[l] == site) grp |= 1ull << l;
  }
  const uint64_t full = s.lanes >= 64 ? ~0ull : ((1ull << s.lanes) - 1);
  const int first = __builtin_ctzll(grp);
  const int op = s.op[first];
  for (uint64_t w = grp; w; w &= w - 1)
    if (s.op[__builtin_ctzll(w)] != op) { fprintf(stderr, "EMU: different wave operations at one site\n"); std::abort(); }
  if ((op == EMU_SHFL_XOR || op == EMU_SCAN) && grp != full) {
    fprintf(stderr, "EMU: wave shuffle/scan with lanes %016llx of %016llx active\n", (unsigned long long)grp,
            (unsigned long long)full);
    std::abort();
  }
  emu_wave_stats.ops++;
  if (grp & (grp - 1)) emu_wave_stats.multi++;
  if (grp != full) emu_wave_stats.partial++;
  if (grp != s.waiting) emu_wave_stats.split++;
  uint64_t bal = 0, acc = 0;
  for (uint64_t w = grp; w; w &= w - 1) {
    const int l = __builtin_ctzll(w);
    if (s.val[l]) bal |= 1ull << l;
  }
  for (uint64_t w = grp; w; w &= w - 1) {
    const int l = __builtin_ctzll(w);
    switch (op) {
      case EMU_BALLOT: s.out[l] = bal; break;
      case EMU_READLANE: {
        const int src = s.arg[l] & 63;
        const bool act = (grp >> src) & 1;
        if (!act) emu_wave_stats.inactive_src++;
        s.out[l] = act ? s.val[src] : 0xDEADBEEFu;  // an inactive lane's register holds a stale value
        break;
      }
      case EMU_READFIRST: s.out[l] = s.val[first]; break;
      case EMU_SHFL_XOR: s.out[l] = s.val[(l ^ s.arg[l]) & 63]; break;
      case EMU_SCAN: acc = (uint32_t)(acc + s.val[l]); s.out[l] = acc; break;
      default: s.out[l] = 0; break;
    }
  }
  s.waiting &= ~grp;
  s.ready |= grp;
  s.cv.notify_all();
}
// a thread outside any launch (a driver calling device code itself) is a wave of one lane
inline thread_local bool emu_solo = false;

struct EmuBarrier {
  std::mutex m; std::condition_variable cv;
  uint32_t n = 0, arrived = 0, gen = 0, exited = 0;
  const char* site_file = nullptr; int site_line = -1; bool bad = false;
  EmuWave wave[16];
  void reset(uint32_t nt) {
    n = nt; arrived = 0; gen = 0; exited = 0; site_line = -1; bad = false;
    for (uint32_t w = 0; w < 16; ++w) {
      EmuWave& v = wave[w];
      v.lanes = 64 * w < nt ? std::min<uint32_t>(64, nt - 64 * w) : 0;
      v.n_bar = v.n_exit = 0;
      v.waiting = v.ready = 0;
    }
  }
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
    if (++arrived == n) {
      // every thread is here: none stays blocked at the barrier
      for (uint32_t w = 0; w < 16; ++w) {
        std::lock_guard<std::mutex> wl(wave[w].m);
        wave[w].n_bar = 0;
      }
      arrived = 0; gen++; cv.notify_all(); return;
    }
    {
      EmuWave& v = wave[threadIdx.x >> 6];
      std::lock_guard<std::mutex> wl(v.m);
      v.n_bar++;
      emu_try_release(v);
    }
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
    EmuWave& v = wave[threadIdx.x >> 6];
    std::lock_guard<std::mutex> wl(v.m);
    v.n_exit++;
    emu_try_release(v);
  }
};
// one workgroup's emulation state
struct EmuWG {
  uint32_t id = 0;
  EmuBarrier bar;
  unsigned char* dyn_smem = nullptr;
};
inline EmuWG emu_default_wg;
inline thread_local EmuWG* emu_wg = &emu_default_wg;
#define g_bar (emu_wg->bar)
#define g_dyn_smem (emu_wg->dyn_smem)
#define __syncthreads() g_bar.wait(__FILE__, __LINE__)

// a wave operation of the calling lane at `site` (its caller's code address)
inline uint64_t emu_wave_op(const void* site, int op, uint64_t v, int arg) {
  const uint32_t l = threadIdx.x & 63;
  if (emu_solo) {
    switch (op) {
      case EMU_BALLOT: return v ? 1ull << l : 0ull;
      case EMU_SHFL_XOR:
        if ((arg & 63) == 0) return v;
        fprintf(stderr, "EMU: shuffle in a one-lane wave\n");
        std::abort();
      case EMU_WBAR: return 0;
      default: return v;
    }
  }
  EmuWave& s = g_bar.wave[threadIdx.x >> 6];
  std::unique_lock<std::mutex> lk(s.m);
  s.site[l] = site; s.op[l] = op; s.val[l] = v; s.arg[l] = arg;
  s.waiting |= 1ull << l;
  emu_try_release(s);
  if (!s.cv.wait_for(lk, std::chrono::seconds(20), [&] { return ((s.ready >> l) & 1) != 0; })) {
    fprintf(stderr, "EMU: wave operation timeout (thread %u, block %u): waiting %016llx, %u at the barrier, %u gone, %u lanes\n",
            threadIdx.x, blockIdx.x, (unsigned long long)s.waiting, s.n_bar, s.n_exit, s.lanes);
    std::abort();
  }
  s.ready &= ~(1ull << l);
  return s.out[l];
}
#define EMU_SITE __builtin_return_address(0)

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
inline uint32_t __lane_id() { return threadIdx.x & 63; }
__attribute__((noinline)) inline uint64_t __ballot(int p) { return emu_wave_op(EMU_SITE, EMU_BALLOT, p != 0, 0); }
__attribute__((noinline)) inline int __builtin_amdgcn_readlane(int v, int src) {
  return (int)(uint32_t)emu_wave_op(EMU_SITE, EMU_READLANE, (uint32_t)v, src);
}
__attribute__((noinline)) inline int __builtin_amdgcn_readfirstlane(int v) {
  return (int)(uint32_t)emu_wave_op(EMU_SITE, EMU_READFIRST, (uint32_t)v, 0);
}
// lanes below the caller's with their bit set in the mask, plus x
inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t lo, uint32_t x) {
  const uint32_t l = threadIdx.x & 63;
  const uint32_t below = l >= 32 ? 0xFFFFFFFFu : ((1u << l) - 1);
  return x + (uint32_t)__builtin_popcount(lo & below);
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t hi, uint32_t x) {
  const uint32_t l = threadIdx.x & 63;
  const uint32_t below = l < 32 ? 0u : ((1u << (l - 32)) - 1);
  return x + (uint32_t)__builtin_popcount(hi & below);
}
inline float __expf(float x) { return std::exp(x); }
// lane exchange inside a wave (every lane of the wave calls; mask < width)
__attribute__((noinline)) inline int __shfl_xor(int v, int mask, int width) {
  (void)width;
  return (int)(uint32_t)emu_wave_op(EMU_SITE, EMU_SHFL_XOR, (uint32_t)v, mask);
}
__attribute__((noinline)) inline float __shfl_xor(float v, int mask, int width) {
  (void)width;
  uint32_t b;
  memcpy(&b, &v, 4);
  b = (uint32_t)emu_wave_op(EMU_SITE, EMU_SHFL_XOR, b, mask);
  float r;
  memcpy(&r, &b, 4);
  return r;
}
inline unsigned long long wall_clock64() { return 0; }

// wave-wide inclusive scan among the 64 lanes of a wave (every lane calls)
__attribute__((noinline)) inline uint32_t emu_wave_incl_scan(uint32_t v) {
  return (uint32_t)emu_wave_op(EMU_SITE, EMU_SCAN, v, 0);
}

// wave-level ordering point: a rendezvous of the wave's active lanes
#define __builtin_amdgcn_fence(order, scope) ((void)0)
__attribute__((noinline)) inline void __builtin_amdgcn_wave_barrier() { emu_wave_op(EMU_SITE, EMU_WBAR, 0, 0); }
