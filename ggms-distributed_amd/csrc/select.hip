// Host driver and C ABI of the selection engine.
//
// The reference drives every iteration from the host with ~6 synchronous
// copies per pass (DistributedSMOS/iterativeSMOS.cu:1348-1512).  Here the
// whole pass (count, scan, findKBuckets, re-ranking, next splitter,
// compaction, finishing) is decided on the device; the host reads back one
// 128-byte control block per pass to size the next launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ggms.h"
#include "ggms_internal.hpp"
#include "splitter.hpp"

using namespace ggms;

namespace ggms {

const float* learned_weights_device(int device);  // mlp.hip

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) {                             \
      last_hip_error() = e_;                            \
      return GGMS_EDEVICE;                              \
    }                                                   \
  } while (0)

hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

constexpr int kMaxBuffers = 48;  // capacity of Layout::off; every workspace buffer needs an entry
struct Layout {
  size_t off[kMaxBuffers];
  size_t total;
};

// Resolved tunables of one call.
struct Config {
  uint32_t B;          // requested buckets per pass
  uint32_t B_cap;      // cap for passes >= 1
  uint32_t B0_bound;   // upper bound of first-pass buckets
  uint32_t R0_bound;   // upper bound of first-pass ranges
  uint32_t S;          // sample size
  uint32_t P;          // pivots (linear) or pieces (learned / sketch)
  uint32_t E;          // edge buckets
  uint32_t mode;
  uint32_t NT;         // streaming workgroup size
  uint32_t G_max;
  uint32_t G_max_next;  // workgroups of the passes after the first
  uint32_t min_tile;
  uint32_t max_passes;
  uint32_t fin_cap;
  uint64_t seed;
  uint32_t force_finish;
  bool one_pass;       // bucketMultiselect: one bucketing pass, every k-bucket finished
  bool seg_first;      // first pass compacts every k-bucket into its own segment
  bool device_schedule;  // every launch enqueued up front, no host reads between kernels
};

static uint32_t pow2ceil_host(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

template <class K>
static Config resolve_config(ggms_alg alg, uint64_t n, uint32_t nk, const ggms_opts& o, int cus) {
  Config c{};
  c.mode = kModeLinear;
  if (alg == GGMS_ALG_DISTRIBUTED_1GPU || alg == GGMS_ALG_DISTRIBUTED_BUCKET_1GPU) c.mode = kModeStrided;
  c.one_pass = alg == GGMS_ALG_BUCKET || alg == GGMS_ALG_DISTRIBUTED_BUCKET_1GPU;
  // The iterative forms carry every first-pass k-bucket into one merged second
  // pass, which re-buckets them all under one bucket budget: right for few ks,
  // where the k-buckets are few and large.  With many ks (automatic: at least
  // one k per four first-pass buckets) the k-buckets are many and most of them
  // fit the finish, so the first pass writes each into its own segment, those
  // within the finish capacity are finished at once and only the larger ones
  // are re-bucketed, each with its own share of the budget.  Measured at 2^28
  // fp32 N(0,1), merged / segmented: 2048 ks 0.89 / 1.17 ms, 3072 ks
  // 1.17 / 1.23, 3584 ks 1.28 / 1.27, 4096 ks 1.37 / 1.32, 6144 ks 2.44 / 1.49
  // (above ~5000 ks the pass-1 range parameters no longer fit LDS).
  const uint32_t B_auto = o.num_buckets ? o.num_buckets : 14336u;
  if (c.one_pass || o.first_pass == GGMS_FIRST_PASS_SEGMENTED) c.seg_first = true;
  else if (o.first_pass == GGMS_FIRST_PASS_MERGED) c.seg_first = false;
  else c.seg_first = (uint64_t)nk * 4 >= B_auto;
  if (o.splitter == GGMS_SPLIT_LEARNED) c.mode = kModeLearned;
  if (o.splitter == GGMS_SPLIT_SKETCH) c.mode = kModeSketch;
  // 0 = automatic: 14336 for every splitter (the reference uses 8192).  More
  // first-pass buckets leave fewer survivors for pass 1 and the finish; the
  // first-pass count keeps its speed at one workgroup per CU, and above ~14K
  // the planning kernel and the column sums grow faster than pass 1 shrinks.
  // Measured at 2^28 fp32, 1001 ks, linear splitter: 8192 -> 0.802 ms,
  // 12288 -> 0.757, 14336 -> 0.750, 16384 -> 0.780.
  // With the segmented first pass on a large input, 16384: fewer survivors
  // to move twice and to finish outweigh the larger plan and column sums
  // (8192 random ks, 14336 -> 16384: 2^28 fp32 mixture 2.20 -> 2.11 ms linear,
  // 1.739 -> 1.730 sketch; 2^27 fp64 N(0,1) 1.82 -> 1.73 ms; at 2^26,
  // |N(0,1)| with 8192 even ks, 0.661 -> 0.672 ms, so not below 2^27 keys).
  // The same calls default to 128 quantile pieces and one first-pass tile per
  // CU (fewer histogram rows to sum; the grouped first pass runs one workgroup
  // per CU anyway): mixture 2.09 -> 2.06 ms linear, 1.719 -> 1.658 sketch,
  // 1.730 -> 1.670 learned; 2^27 fp32 N(0,1) 0.980 -> 0.945 ms, fp64 1.732 ->
  // 1.701 (one tile per CU makes the headline, 2^28 with 1001 ks, 0.68 -> 0.84
  // ms and |N(0,1)| at 2^26 0.655 -> 0.679).
  const bool big_seg = c.seg_first && !c.one_pass && n >= (1ull << 27);
  c.B = o.num_buckets ? o.num_buckets : (big_seg ? 16384u : 14336u);
  c.B = std::max<uint32_t>(16, std::min<uint32_t>(c.B, kMaxBuckets));
  c.B_cap = kMaxBuckets;
  if (c.mode == kModeLinear || c.mode == kModeStrided) {
    c.P = o.num_pivots ? o.num_pivots : 17;
    c.P = std::max<uint32_t>(2, std::min<uint32_t>(c.P, 1024));
    // the reference's 1024-key sample; a large call with few ks and 4-byte
    // keys draws 2048 (2^28 fp32 N(0,1), 1001 ks: 0.684 -> 0.672 ms, its
    // pieces closer to the quantiles; the many-k calls keep 1024: 2048 is
    // 3 % faster on the 2^28 mixture but 2 % slower on 2^26 uint32 keys, and
    // with 8-byte keys the larger sort costs more than it saves)
    const bool big_few = c.mode == kModeLinear && !c.seg_first && sizeof(K) == 4 && n >= (1ull << 27);
    c.S = o.sample_size ? o.sample_size : (big_few ? 2048u : 1024u);
  } else {
    c.P = o.learned_pieces ? o.learned_pieces : (big_seg ? 128u : 64u);
    c.P = std::max<uint32_t>(2, std::min<uint32_t>(c.P, kMaxPieces));
    c.S = o.sample_size ? o.sample_size : 8192;
    c.S = std::max<uint32_t>(c.S, 2 * kMlpWin);
  }
  c.S = std::max<uint32_t>(8, std::min<uint32_t>(c.S, sizeof(K) == 4 ? 16384u : 8192u));
  uint32_t pieces = (c.mode == kModeLinear || c.mode == kModeStrided) ? std::max<uint32_t>(1, c.P - 1) : c.P;
  c.E = std::max<uint32_t>(1, c.B / pieces / 8);
  uint32_t Q = (c.mode == kModeLinear || c.mode == kModeStrided) ? c.P : c.P + 1;
  c.R0_bound = 2 * Q + 6;
  c.B0_bound = c.B + 2 * c.E + 4 + Q;
  c.NT = o.block_threads ? o.block_threads : 1024;
  if (c.NT % 64 || c.NT < 64 || c.NT > 1024) c.NT = 1024;
  uint32_t bpc = o.blocks_per_cu ? o.blocks_per_cu : (big_seg ? 1u : 2u);
  c.G_max = std::max<uint32_t>(1, std::min<uint32_t>(bpc * (uint32_t)cus, kMaxWorkgroups));
  // passes >= 1: one workgroup per CU (their LDS tables and scattered writes
  // favour it: pass-1 count 0.059 -> 0.052 ms, compaction 0.085 -> 0.076 ms)
  c.G_max_next = o.blocks_per_cu ? c.G_max : std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)cus, kMaxWorkgroups));
  c.min_tile = kMinTile;
  c.device_schedule = o.schedule == GGMS_SCHEDULE_DEVICE;
  c.max_passes = o.max_passes ? std::min<uint32_t>(o.max_passes, GGMS_MAX_PASS_STATS) : 64;
  // the device-driven schedule launches every pass it may need: 3 by default
  // (2^28 fp32 with 1001 ks needs 2, 8192 ks 3); the last finishes whatever is left
  if (c.device_schedule) c.max_passes = o.device_passes ? std::min<uint32_t>(o.device_passes, 8) : 3;
  if (c.max_passes < 2) c.max_passes = 2;  // the first pass always hands its k-buckets to a second
  c.fin_cap = sizeof(K) == 4 ? 32768 : 16384;  // 128 KB of LDS in the finishing kernel
  c.seed = o.seed;
  c.force_finish = 0;
  (void)n;
  return c;
}

// Tiling of the first pass: at most G_max workgroups, each owning a tile of
// at least min_tile elements (a multiple of 64), the last tile ragged.
static void first_pass_tiling(uint64_t n, uint32_t G_max, uint32_t min_tile, uint32_t* tile, uint32_t* G) {
  uint64_t g = std::min<uint64_t>(G_max, (n + min_tile - 1) / min_tile);
  if (g < 1) g = 1;
  const uint64_t t = ((n + g - 1) / g + 63) & ~63ull;
  *tile = (uint32_t)t;
  *G = (uint32_t)((n + t - 1) / t);
}

struct Workspace {
  void* base = nullptr;
  size_t bytes = 0;
};

}  // namespace ggms

struct ggms_ctx {
  int device = 0;
  int cus = 256;
  size_t max_lds = 65536;
  Workspace ws;
  void* pinned = nullptr;  // host staging (ctl blocks, ks, outputs)
  void* pinned_dev = nullptr;  // the same memory as the device addresses it
  size_t pinned_bytes = 0;
  uint32_t mirror_seq = 0;       // last sequence number a plan kernel was asked to publish
  ggms_stats stats{};
  std::vector<hipEvent_t> events;  // timing events (opts.profile)
  std::mutex mu;
};

namespace ggms {

static ggms_status ensure_pinned(ggms_ctx* c, size_t bytes) {
  if (c->pinned_bytes >= bytes) return GGMS_OK;
  if (c->pinned) (void)hipHostFree(c->pinned);
  c->pinned = nullptr;
  c->pinned_bytes = 0;
  size_t b = std::max<size_t>(bytes, 1 << 16);
  // coherent: the plan kernel writes control blocks here, then a sequence
  // word the host polls, so the host reads them while the next kernels run
  if (hipHostMalloc(&c->pinned, b, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return GGMS_ENOMEM;
  if (hipHostGetDevicePointer(&c->pinned_dev, c->pinned, 0) != hipSuccess) {
    (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    return GGMS_EDEVICE;
  }
  c->pinned_bytes = b;
  return GGMS_OK;
}

static ggms_status ensure_ws(ggms_ctx* c, size_t bytes) {
  if (c->ws.bytes >= bytes) return GGMS_OK;
  if (c->ws.base) (void)hipFree(c->ws.base);
  c->ws.base = nullptr;
  c->ws.bytes = 0;
  if (hipMalloc(&c->ws.base, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return GGMS_ENOMEM;
  }
  c->ws.bytes = bytes;
  return GGMS_OK;
}

// Optional per-kernel timing: a start/stop event pair around every launch, on
// the launch stream, read after the call's final synchronisation.
struct Prof {
  ggms_ctx* c;
  hipStream_t st;
  bool on;
  struct Rec {
    int kind, pass;
    size_t e0;
  };
  std::vector<Rec> recs;
  size_t used = 0;
  Prof(ggms_ctx* c_, hipStream_t st_, bool on_) : c(c_), st(st_), on(on_) {}
  void begin(int kind, int pass) {
    if (!on) return;
    while (c->events.size() < used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        on = false;
        return;
      }
      c->events.push_back(e);
    }
    (void)hipEventRecord(c->events[used], st);
    recs.push_back({kind, pass, used});
    used += 2;
  }
  void end() {
    if (!on || recs.empty()) return;
    (void)hipEventRecord(c->events[recs.back().e0 + 1], st);
  }
  void collect(ggms_stats& s) {
    if (!on) return;
    for (const Rec& r : recs) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->events[r.e0], c->events[r.e0 + 1]) != hipSuccess) continue;
      s.kernel_ms[r.kind] += ms;
      s.kernel_launches[r.kind] += 1;
      if (r.pass >= 0 && r.pass < GGMS_MAX_PASS_STATS) {
        if (r.kind == GGMS_KERNEL_COUNT0 || r.kind == GGMS_KERNEL_COUNTS) s.count_ms[r.pass] += ms;
        if (r.kind == GGMS_KERNEL_COMPACT0 || r.kind == GGMS_KERNEL_COMPACTS) s.compact_ms[r.pass] += ms;
      }
    }
  }
};

enum Buf {
  bBufA, bBufB, bRows, bTotals, bRanges0, bRanges1, bRangesFirst, bCtl, bKRank, bKState, bKAux, bKList,
  bSlotOfBucket, bSlotDst, bSlotBucket, bActBucket, bActCls, bActDst, bActSlot, bActKlo, bActKhi,
  bSmallStart, bSmallLen, bSmallKStart, bWgFirst, bLut, bSmallKlo, bSmallKhi, bParams1, bActCnt, bRank0, bSmallOrder,
  bGrpCursor, bGrpStart, bGrpIds, bCount
};
static_assert(bCount <= kMaxBuffers, "workspace layout table too small");

template <class T>
static Layout make_layout(uint64_t n, uint32_t nk, const Config& c) {
  using K = typename KeyTraits<T>::K;
  size_t sz[bCount];
  size_t Bm = std::max<size_t>(c.B0_bound, c.B_cap) + 64;
  size_t Rm = std::max<size_t>(c.R0_bound, Bm) + 64;
  size_t nkp = std::max<uint32_t>(nk, 1);
  const size_t nbuf = n;
  sz[bBufA] = nbuf * sizeof(K);
  sz[bBufB] = nbuf * sizeof(K);  // also the spare buffer of the finish's global-memory sort
  // one histogram row per workgroup: the first pass's tiles, or one per CU after it
  uint32_t tile0 = 0, G0 = 1;
  if (n > 0 && n <= kMaxN) first_pass_tiling(n, c.G_max, c.min_tile, &tile0, &G0);
  sz[bRows] = (size_t)std::max(G0, c.G_max_next) * Bm * 4;
  sz[bTotals] = Bm * 4;
  sz[bRanges0] = Rm * sizeof(Range<K>);
  sz[bRanges1] = Rm * sizeof(Range<K>);
  sz[bRangesFirst] = ((size_t)c.R0_bound + 64) * sizeof(Range<K>);
  sz[bCtl] = (size_t)(c.max_passes + 2) * sizeof(PassCtl);
  sz[bKRank] = nkp * 4;
  sz[bKState] = nkp * 4;
  sz[bKAux] = nkp * 4;
  sz[bKList] = nkp * 4;
  sz[bSlotOfBucket] = Bm * 2 * 2;  // double-buffered by pass parity
  sz[bSlotDst] = Bm * 4;
  sz[bSlotBucket] = Bm * 4;
  sz[bActBucket] = Bm * 4;
  sz[bActCls] = Bm * 4;
  sz[bActCnt] = Bm * 4;
  sz[bRank0] = (size_t)rank_word_count((uint32_t)Bm) * 8;
  sz[bActDst] = Bm * 4;
  sz[bActSlot] = Bm * 4;
  sz[bActKlo] = Bm * sizeof(K);
  sz[bActKhi] = Bm * sizeof(K);
  sz[bSmallStart] = Bm * 4;
  sz[bSmallLen] = Bm * 4;
  sz[bSmallKStart] = (Bm + 1) * 4;
  sz[bSmallOrder] = Bm * 4;
  sz[bWgFirst] = (size_t)c.G_max * 4 * 2;  // double-buffered by pass parity
  sz[bLut] = LutGeom<K>::bytes();
  sz[bParams1] = Bm * sizeof(MapParam<K>);
  sz[bSmallKlo] = Bm * sizeof(K);
  sz[bSmallKhi] = Bm * sizeof(K);
  sz[bGrpCursor] = (size_t)kMaxGroups * 4;
  sz[bGrpStart] = (size_t)(kMaxGroups + 1) * 4;
  // slot index inside its group of every key of the grouped compaction's level 1
  sz[bGrpIds] = c.seg_first ? nbuf : 0;
  Layout L{};
  size_t o = 0;
  for (int i = 0; i < bCount; ++i) {
    L.off[i] = o;
    o += align256(sz[i]);
  }
  L.total = o;
  return L;
}

// Dynamic LDS of the first-pass count: map parameters, then the histogram (the
// cell table is the kernel's static LDS array).
template <class T>
static size_t count_lds0(const Config& c) {
  using K = typename KeyTraits<T>::K;
  return (size_t)c.R0_bound * sizeof(MapParam<K>) + (size_t)c.B0_bound * 4;
}
template <class T>
static size_t count_lds0_total(const Config& c) {
  using K = typename KeyTraits<T>::K;
  return count_lds0<T>(c) + LutGeom<K>::bytes();
}

// The one-pass bucketMultiselect finishes after its bucketing pass every
// k-bucket of up to this many times the LDS capacity (the larger ones from
// global memory); only k-buckets above that are re-bucketed first.
constexpr uint32_t kOnePassFinishCap = 4;

// at most this many ks: the setup runs inside the first-pass planning kernel
constexpr uint32_t kSetupInPlan0MaxKs = kSetupKsPerThread * 1024;

// LDS ceiling of one workgroup on gfx950 (160 KB per CU; static LDS of the
// kernels stays under 2 KB).
constexpr size_t kLdsLimit = 158 * 1024;
// dynamic LDS that leaves room for only one 1024-thread workgroup per CU
constexpr size_t kLdsOneWgPerCu = 81 * 1024;

// The plan's LDS: its tables for B buckets (prefix, active index, rank
// words), then room for one word per small segment (at most one per k) when
// that fits, where it keeps the segments' class information for its last
// phases instead of reading it back from global memory.
static size_t plan_lds_bytes(uint32_t B, uint32_t nk) {
  const size_t base = (size_t)B * 8 + (size_t)rank_word_count(B) * 8;
  const size_t want = base + ((size_t)std::min(nk, B) + 1) * 4;
  return want <= kLdsLimit ? want : base;
}

// Raises a kernel's dynamic-LDS limit on the current device when a launch
// needs more than the default 64 KB; each (device, kernel) pair is raised once
// to what it has needed so far, so repeated calls (and graph capture) make no
// runtime call.
struct LdsAttr {
  int dev;
  const void* fn;
  size_t bytes;
};
static void set_lds_attr(const void* fn, size_t bytes) {
  if (bytes <= 65536) return;
  static std::mutex mu;
  static LdsAttr done[512];  // fixed table: no heap use on this path
  static uint32_t used = 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  for (uint32_t i = 0; i < used; ++i)
    if (done[i].dev == dev && done[i].fn == fn) {
      if (done[i].bytes >= bytes) return;
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess)
        done[i].bytes = bytes;
      return;
    }
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess && used < 512)
    done[used++] = {dev, fn, bytes};
}

#define LDS_CHECK(bytes)                              \
  do {                                                \
    if ((bytes) > kLdsLimit) return GGMS_ENOTSUP;     \
  } while (0)

// Arguments of the first-pass planning kernel (sampling, pivots, range table;
// optionally the setup step).
template <class T>
static Plan0Args plan0_args(const Config& c, int device, uint32_t nk, uint32_t ctl_words, bool setup_in_plan0,
                            const uint32_t* d_ks, uint32_t* k_rank, uint32_t* k_state, uint32_t* dbad,
                            const PassCtl& c0) {
  Plan0Args pa{};
  pa.mode = c.mode;
  pa.S = c.S;
  pa.P = c.P;
  pa.B = c.B;
  pa.E = c.E;
  pa.strideS = 8192;
  pa.seed = c.seed;
  pa.mlp = c.mode == kModeLearned ? learned_weights_device(device) : nullptr;
  pa.setup = setup_in_plan0 ? 1u : 0u;
  pa.nk = nk;
  pa.ctl_words = ctl_words;
  pa.ks = d_ks;
  pa.k_rank = k_rank;
  pa.k_state = k_state;
  pa.bad = dbad;
  pa.c0 = c0;
  return pa;
}

// Plan-kernel arguments that do not change from pass to pass.
template <class K>
static PlanArgs<K> plan_args_common(const Config& c, const Layout& L, char* W, uint32_t nk, void* d_out) {
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  PlanArgs<K> pl{};
  pl.totals = (uint32_t*)P(bTotals);
  pl.nk = nk;
  pl.k_rank = (uint32_t*)P(bKRank);
  pl.k_state = (uint32_t*)P(bKState);
  pl.k_aux = (uint32_t*)P(bKAux);
  pl.slot_dst = (uint32_t*)P(bSlotDst);
  pl.slot_bucket = (uint32_t*)P(bSlotBucket);
  pl.act_bucket = (uint32_t*)P(bActBucket);
  pl.act_cls = (uint32_t*)P(bActCls);
  pl.act_cnt = (uint32_t*)P(bActCnt);
  pl.act_dst = (uint32_t*)P(bActDst);
  pl.act_slot = (uint32_t*)P(bActSlot);
  pl.act_klo = (K*)P(bActKlo);
  pl.act_khi = (K*)P(bActKhi);
  pl.small_start = (uint32_t*)P(bSmallStart);
  pl.small_len = (uint32_t*)P(bSmallLen);
  pl.small_klo = (K*)P(bSmallKlo);
  pl.small_khi = (K*)P(bSmallKhi);
  pl.small_kstart = (uint32_t*)P(bSmallKStart);
  pl.small_order = (uint32_t*)P(bSmallOrder);
  pl.klist = (uint32_t*)P(bKList);
  pl.out = d_out;
  pl.fin_cap = c.fin_cap;
  pl.finish_cap = c.fin_cap;
  pl.B_req = c.B;
  pl.B_cap = c.B_cap;
  pl.G_max = c.G_max_next;
  pl.min_tile = c.min_tile;
  return pl;
}

// The pass-1 kernels hold the pass-1 range parameters in LDS whenever their
// tables with them fit this (they run one workgroup per CU either way);
// beyond it every key reads its range's parameters from the caches.  At 2^28
// N(0,1) with 2048 ks (1900 pass-1 ranges) the parameters took the pass-1
// kernels past the former 112 KB limit: compaction 0.27 ms with them in
// global memory.
constexpr size_t kPass1ParamsLds = kLdsLimit;

// Dynamic LDS the pass-1 kernels are given in a schedule fixed before the
// first pass has run (bounds instead of the first pass's exact table sizes):
// their tables without the pass-1 range parameters, and at least the
// kPass1ParamsLds up to which the parameters join them (the kernels decide
// that on the device).
template <class K>
static size_t pass1_lds_fixed(const Config& c, size_t own) {
  const size_t base = own + pass1_map_bytes<K>(c.B0_bound, c.R0_bound, LutGeom<K>::L2TABLES, 0);
  return std::max<size_t>(base, kPass1ParamsLds);
}

// The whole call enqueued on `st` with no host reads between kernels (a
// device-driven schedule): every pass's kernels are launched with shapes and
// LDS sized for the largest pass the call can produce; a kernel whose pass
// turned out empty (its control block still zero) or that has nothing to do
// in it returns at once.  The last of c.max_passes passes finishes every
// k-bucket left, whatever its size.  flags: device-visible words
// [0] finish error, [1] a k outside [1, n], [2] NaNs in the input -- the
// caller reads them after the stream has completed.  Capturable in a hipGraph
// (no host synchronisation, no allocation, no memset).
template <class T>
static ggms_status launch_fixed(const Config& c, const Layout& L, char* W, uint32_t* flags, const uint32_t* d_ks,
                                T* d_out, int device, int cus, const T* d_vec, uint32_t n, uint32_t nk,
                                hipStream_t st, Prof& prof) {
  using K = typename KeyTraits<T>::K;
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  K* bufA = (K*)P(bBufA);
  K* bufB = (K*)P(bBufB);
  uint32_t* rows = (uint32_t*)P(bRows);
  uint32_t* totals = (uint32_t*)P(bTotals);
  Range<K>* rng[2] = {(Range<K>*)P(bRanges0), (Range<K>*)P(bRanges1)};
  Range<K>* rng_first = (Range<K>*)P(bRangesFirst);
  PassCtl* ctl = (PassCtl*)P(bCtl);
  uint32_t* k_rank = (uint32_t*)P(bKRank);
  uint32_t* k_state = (uint32_t*)P(bKState);
  uint32_t* klist = (uint32_t*)P(bKList);
  const size_t ctlb = sizeof(PassCtl);

  PassCtl c0{};
  c0.n = n;
  uint32_t tile0 = 0, G0 = 0;
  first_pass_tiling(n, c.G_max, c.min_tile, &tile0, &G0);
  c0.tile = tile0;
  c0.G = G0;
  c0.nan_lo_bucket = kNone32;
  c0.nan_hi_bucket = kNone32;
  c0.hot_bucket = kNone32;
  // the status words start clear (a fill kernel: capturable, unlike a memset)
  k_fill_u32<<<1, 64, 0, st>>>(flags, 3, 0u);
  HIP_TRY(hipGetLastError());
  const uint32_t ctl_words = (uint32_t)((c.max_passes + 2) * ctlb / 4);
  const bool setup_in_plan0 = n > c.fin_cap && nk <= kSetupInPlan0MaxKs;
  if (!setup_in_plan0) {
    prof.begin(GGMS_KERNEL_OTHER, -1);
    const uint32_t threads = std::max(ctl_words, nk);
    k_setup<<<(threads + 255) / 256, 256, 0, st>>>(d_ks, nk, n, k_rank, k_state, ctl, ctl_words, c0, flags + 1);
    prof.end();
    HIP_TRY(hipGetLastError());
  }
  const uint32_t fin_grid = std::max<uint32_t>(1, std::min<uint32_t>(4 * (uint32_t)cus, 65535));
  const size_t fin_lds = (size_t)(c.fin_cap / 2) * sizeof(K);
  set_lds_attr((const void*)k_finish<T>, fin_lds);
  if (n <= c.fin_cap) {
    const bool wide = finish_wide_segment<K>(n, (K)0, KeyMax<K>::v, c.fin_cap);
    prof.begin(GGMS_KERNEL_OTHER, -1);
    k_tiny_setup<T><<<1, 1024, 0, st>>>(d_vec, n, nk, bufA, k_rank, k_state, (uint32_t*)P(bSmallStart),
                                        (uint32_t*)P(bSmallLen), (K*)P(bSmallKlo), (K*)P(bSmallKhi),
                                        (uint32_t*)P(bSmallKStart), klist, ctl, wide ? 1u : 0u);
    prof.end();
    HIP_TRY(hipGetLastError());
    prof.begin(GGMS_KERNEL_FINISH, 0);
    k_finish<T><<<1, 1024, fin_lds, st>>>(
        bufA, ctl, (uint32_t*)P(bSmallStart), (uint32_t*)P(bSmallLen), (K*)P(bSmallKlo), (K*)P(bSmallKhi),
        (uint32_t*)P(bSmallKStart), nullptr, klist, k_rank, k_state, d_out, c.fin_cap, flags);
    if (wide)
      k_finish_wide<T><<<1, 1024, 0, st>>>(bufA, bufB, ctl, (uint32_t*)P(bSmallStart), (uint32_t*)P(bSmallLen),
                                           (K*)P(bSmallKlo), (K*)P(bSmallKhi), (uint32_t*)P(bSmallKStart), nullptr,
                                           klist, k_rank, k_state, d_out, c.fin_cap);
    prof.end();
    HIP_TRY(hipGetLastError());
    return GGMS_OK;
  }
  const Plan0Args pa = plan0_args<T>(c, device, nk, ctl_words, setup_in_plan0, d_ks, k_rank, k_state, flags + 1, c0);
  const size_t lds_p0 = plan0_lds_bytes<K>(pow2ceil_host(c.S));
  LDS_CHECK(lds_p0);
  set_lds_attr((const void*)k_plan0<T>, lds_p0);
  prof.begin(GGMS_KERNEL_PLAN0, 0);
  k_plan0<T><<<1, 1024, lds_p0, st>>>(d_vec, pa, rng_first, (K*)P(bLut), ctl);
  prof.end();
  HIP_TRY(hipGetLastError());

  PlanArgs<K> pl = plan_args_common<K>(c, L, W, nk, d_out);
  uint16_t* slot_tab[2] = {(uint16_t*)P(bSlotOfBucket),
                           (uint16_t*)P(bSlotOfBucket) + (std::max<size_t>(c.B0_bound, c.B_cap) + 64)};
  const uint32_t NT = c.NT;
  const uint32_t Gn = c.G_max_next;
  const uint32_t U_slot_bound = std::min<uint32_t>(nk, std::max(c.B0_bound, c.B_cap));
  const uint32_t passes = c.max_passes;
  for (uint32_t pass = 0; pass < passes; ++pass) {
    PassCtl* cp = ctl + pass;
    Range<K>* rc = pass == 0 ? rng_first : rng[pass & 1];
    K* src = (pass & 1) ? bufA : bufB;
    K* dst = (pass & 1) ? bufB : bufA;
    uint32_t* wgf_cur = (uint32_t*)P(bWgFirst) + (size_t)(pass & 1) * c.G_max;
    const uint32_t B_bound = pass == 0 ? c.B0_bound : c.B_cap;
    // count
    if (pass == 0) {
      size_t lds = count_lds0<T>(c);
      LDS_CHECK(count_lds0_total<T>(c));
      const size_t hist2 = (size_t)c.B0_bound * 4;
      const uint32_t tpw =
          (count_lds0_total<T>(c) > kLdsOneWgPerCu && count_lds0_total<T>(c) + hist2 <= kLdsLimit) ? 2u : 1u;
      if (tpw == 2) lds += hist2;
      set_lds_attr((const void*)k_count0<T>, lds);
      prof.begin(GGMS_KERNEL_COUNT0, (int)pass);
      k_count0<T><<<(G0 + tpw - 1) / tpw, NT, lds, st>>>(d_vec, rc, (K*)P(bLut), cp, rows, tpw);
      prof.end();
    } else if (pass == 1 && !c.seg_first) {
      const size_t lds = pass1_lds_fixed<K>(c, ((size_t)B_bound * 4 + 15) / 16 * 16);
      LDS_CHECK(lds);
      set_lds_attr((const void*)k_count1<K>, lds);
      prof.begin(GGMS_KERNEL_COUNTS, (int)pass);
      k_count1<K><<<Gn, NT, lds, st>>>(src, ctl, (K*)P(bLut), rng_first, (uint64_t*)P(bRank0),
                                        (MapParam<K>*)P(bParams1), (uint32_t)kPass1ParamsLds, cp, rows);
      prof.end();
    } else {
      const size_t lds = (size_t)B_bound * 4;
      set_lds_attr((const void*)k_countS<K>, lds);
      prof.begin(GGMS_KERNEL_COUNTS, (int)pass);
      k_countS<K><<<Gn, NT, lds, st>>>(src, rc, wgf_cur, cp, rows);
      prof.end();
    }
    HIP_TRY(hipGetLastError());
    prof.begin(GGMS_KERNEL_COLSUM, (int)pass);
    k_colsum<<<(B_bound + 63) / 64, 1024, 0, st>>>(rows, cp, totals);
    prof.end();
    HIP_TRY(hipGetLastError());
    // plan
    pl.ranges = rc;
    pl.ctl = cp;
    pl.ctl_next = ctl + pass + 1;
    pl.next_ranges = rng[(pass + 1) & 1];
    pl.slot_of_bucket = slot_tab[pass & 1];
    pl.wg_first_next = (uint32_t*)P(bWgFirst) + (size_t)((pass + 1) & 1) * c.G_max;
    pl.force_finish = (pass + 1 >= passes) ? 1u : 0u;
    pl.finish_cap = (pass == 0 && c.one_pass) ? kOnePassFinishCap * c.fin_cap : c.fin_cap;
    pl.merge_small = (pass == 0 && !c.seg_first) ? 1u : 0u;
    pl.next_params = (pass == 0 && !c.seg_first) ? (MapParam<K>*)P(bParams1) : nullptr;
    pl.rank_words = (pass == 0 && !c.seg_first) ? (uint64_t*)P(bRank0) : nullptr;
    pl.mirror = nullptr;
    pl.nan_out = pass == 0 ? flags + 2 : nullptr;
    plan_groups<K>(pl, L, W, pass == 0 && c.seg_first);
    const size_t lds_plan = plan_lds_bytes(B_bound, nk);
    pl.lds_words = (uint32_t)(lds_plan / 4);
    LDS_CHECK(lds_plan);
    set_lds_attr((const void*)k_plan<T>, lds_plan);
    prof.begin(GGMS_KERNEL_PLAN, (int)pass);
    k_plan<T><<<1, 1024, lds_plan, st>>>(pl);
    prof.end();
    HIP_TRY(hipGetLastError());
    // compaction
    if (pass == 0 && c.seg_first) {
      ggms_status s1 = launch_segmented_first_pass<T>(c, L, W, d_vec, n, G0, cp, pl, src, dst, st, prof);
      if (s1) return s1;
    } else if (pass == 0) {
      const size_t lds = (size_t)c.R0_bound * sizeof(MapParam<K>) + (size_t)B_bound * 2;
      LDS_CHECK(lds + LutGeom<K>::bytes());
      set_lds_attr((const void*)k_compact0<T>, lds);
      prof.begin(GGMS_KERNEL_COMPACT0, (int)pass);
      k_compact0<T><<<G0, NT, lds, st>>>(d_vec, rc, (K*)P(bLut), cp, rows, pl.slot_of_bucket, pl.slot_bucket, dst);
      prof.end();
    } else if (pass == 1 && !c.seg_first) {
      const size_t lds = pass1_lds_fixed<K>(c, compact_lds_bytes(B_bound, U_slot_bound + 4));
      LDS_CHECK(lds);
      set_lds_attr((const void*)k_compact1<K>, lds);
      prof.begin(GGMS_KERNEL_COMPACTS, (int)pass);
      k_compact1<K><<<Gn, NT, lds, st>>>(src, ctl, (K*)P(bLut), rng_first, (uint64_t*)P(bRank0),
                                          (MapParam<K>*)P(bParams1), (uint32_t)kPass1ParamsLds, cp, rows, pl.slot_of_bucket,
                                          pl.slot_dst, pl.slot_bucket, dst);
      prof.end();
    } else {
      const size_t lds = compact_lds_bytes(B_bound, U_slot_bound + 4);
      LDS_CHECK(lds);
      set_lds_attr((const void*)k_compactS<K>, lds);
      prof.begin(GGMS_KERNEL_COMPACTS, (int)pass);
      k_compactS<K><<<Gn, NT, lds, st>>>(src, rc, wgf_cur, cp, rows, pl.slot_of_bucket, pl.slot_dst,
                                          pl.slot_bucket, dst);
      prof.end();
    }
    HIP_TRY(hipGetLastError());
    // finish (the iterative first pass hands every k-bucket to pass 1)
    if (pass > 0 || c.seg_first) {
      prof.begin(GGMS_KERNEL_FINISH, (int)pass);
      k_finish<T><<<fin_grid, 1024, fin_lds, st>>>(dst, cp, pl.small_start, pl.small_len, pl.small_klo,
                                                    pl.small_khi, pl.small_kstart, pl.small_order, klist, k_rank,
                                                    k_state, d_out, c.fin_cap, flags);
      k_finish_wide<T><<<(uint32_t)cus, 1024, 0, st>>>(dst, src, cp, pl.small_start, pl.small_len, pl.small_klo,
                                                        pl.small_khi, pl.small_kstart, pl.small_order, klist, k_rank,
                                                        k_state, d_out, c.fin_cap);
      prof.end();
      HIP_TRY(hipGetLastError());
    }
  }
  return GGMS_OK;
}

// Segmented first pass in two levels (k_group_pass0, then k_group_split):
// level 1 moves every survivor into its group's span of `scratch` (with its
// slot index beside it in the id buffer), level 2 moves it to its segment of
// `dst`.  Level 2 runs one workgroup per chunk of an upper bound of the group
// chunks; those past the last chunk return at once.
template <class T>
static ggms_status launch_segmented_first_pass(const Config& c, const Layout& L, char* W, const T* d_vec,
                                               uint32_t n, uint32_t G0, PassCtl* cp, PlanArgs<typename KeyTraits<T>::K>& pl,
                                               typename KeyTraits<T>::K* scratch, typename KeyTraits<T>::K* dst,
                                               hipStream_t st, Prof& prof) {
  using K = typename KeyTraits<T>::K;
  using S = GroupStage<K>;
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  const Range<K>* rc = (const Range<K>*)P(bRangesFirst);
  const K* lut = (const K*)P(bLut);
  const size_t lds1 = Pass0Map<K>::param_bytes(c.R0_bound) + S::pass0_bytes(c.B0_bound);
  LDS_CHECK(lds1 + LutGeom<K>::bytes() + 5 * 1024);
  set_lds_attr((const void*)k_group_pass0<T>, lds1);
  const size_t lds_sp = Pass0Map<K>::param_bytes(c.R0_bound) + ((size_t)c.B0_bound * 2 + 15) / 16 * 16;
  set_lds_attr((const void*)k_group_scatter<T>, lds_sp);
  prof.begin(GGMS_KERNEL_COMPACT0, 0);
  // the plan picks one of the two (ctl->grp_sparse); the other returns at once
  k_group_scatter<T><<<G0, 1024, lds_sp, st>>>(d_vec, rc, lut, cp, (const uint32_t*)P(bRows), pl.slot_of_bucket,
                                               pl.slot_bucket, (uint32_t*)P(bGrpCursor), scratch,
                                               (uint8_t*)P(bGrpIds));
  k_group_pass0<T><<<G0, S::kThreads, lds1, st>>>(d_vec, rc, lut, cp, (const uint32_t*)P(bRows), pl.slot_of_bucket,
                                                  pl.slot_bucket, (uint32_t*)P(bGrpCursor), scratch,
                                                  (uint8_t*)P(bGrpIds));
  HIP_TRY(hipGetLastError());
  const size_t lds2 = S::split_bytes();
  set_lds_attr((const void*)k_group_split<K>, lds2);
  // every group has at most one chunk beyond len / CH2 (its first starts up to
  // N - 1 keys before the span): the chunks total at most n / CH2 + groups + 1
  const uint32_t chunks = (uint32_t)(((uint64_t)n + S::CH2 - 1) / S::CH2) + kMaxGroups + 1;
  k_group_split<K><<<chunks, S::kThreads2, lds2, st>>>(scratch, (const uint8_t*)P(bGrpIds), cp,
                                                      (const uint32_t*)P(bGrpStart), pl.slot_dst, dst);
  prof.end();
  HIP_TRY(hipGetLastError());
  return GGMS_OK;
}

// Plan arguments of the grouped compaction (a segmented first pass).
template <class K>
static void plan_groups(PlanArgs<K>& pl, const Layout& L, char* W, bool on) {
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  pl.grp_cursor = on ? (uint32_t*)P(bGrpCursor) : nullptr;
  pl.grp_start = on ? (uint32_t*)P(bGrpStart) : nullptr;
}

// Status of a device-driven call from its status words.
static ggms_status fixed_status(const uint32_t* f) {
  if (f[1]) return GGMS_EINVAL;
  if (f[2]) return GGMS_ENAN;
  if (f[0]) return GGMS_EDEVICE;
  return GGMS_OK;
}

// Waits until the plan kernel has published sequence number `want` (after
// its mirror of the control blocks, with a system-scope fence between).  Every
// 65536 polls it asks whether the stream has drained: a drained stream without
// the word means the plan never published (a device failure), not a wait.
static ggms_status wait_mirror(volatile uint32_t* seq, uint32_t want, hipStream_t st) {
  for (uint64_t spins = 1; *seq != want; ++spins) {
    if ((spins & 0xFFFF) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipSuccess) return *seq == want ? GGMS_OK : GGMS_EDEVICE;
      if (q != hipErrorNotReady) return GGMS_EDEVICE;
    }
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#endif
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return GGMS_OK;
}

template <class T>
static ggms_status run_select(ggms_ctx* ctx, ggms_alg alg, const T* d_vec, uint64_t n64, const uint32_t* ks,
                              bool ks_on_device, uint32_t nk, void* out, bool out_on_device, const ggms_opts& o) {
  using K = typename KeyTraits<T>::K;
  if (n64 == 0 || n64 > kMaxN) return GGMS_EINVAL;
  if (nk == 0) return GGMS_OK;
  if (!d_vec || !ks || !out) return GGMS_EINVAL;
  const uint32_t n = (uint32_t)n64;
  hipStream_t st = (hipStream_t)o.stream;
  Config c = resolve_config<K>(alg, n64, nk, o, ctx->cus);
  if (c.mode == kModeLearned && !learned_weights_device(ctx->device)) return GGMS_ENOTSUP;
  if (!ks_on_device)
    for (uint32_t i = 0; i < nk; ++i)
      if (ks[i] < 1 || ks[i] > n) return GGMS_EINVAL;

  Layout L = make_layout<T>(n64, nk, c);
  const size_t ctlb = sizeof(PassCtl);
  const size_t ks_bytes = ((size_t)nk * 4 + 255) / 256 * 256;
  const size_t pinned_need = 6 * ctlb + 256 + ks_bytes + (size_t)nk * sizeof(T) + 4096;
  // under hipGraph capture only the asynchronous device-driven form can run,
  // and it must find its memory ready (ggms_ctx_reserve): nothing allocates
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
  if (cap != hipStreamCaptureStatusNone) {
    if (!c.device_schedule || !ks_on_device || !out_on_device || o.profile) return GGMS_EINVAL;
    if (ctx->ws.bytes < L.total || ctx->pinned_bytes < pinned_need) return GGMS_ENOMEM;
  }
  ggms_status s = ensure_ws(ctx, L.total);
  if (s) return s;
  s = ensure_pinned(ctx, pinned_need);
  if (s) return s;
  char* W = (char*)ctx->ws.base;
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  K* bufA = (K*)P(bBufA);
  K* bufB = (K*)P(bBufB);
  uint32_t* rows = (uint32_t*)P(bRows);
  uint32_t* totals = (uint32_t*)P(bTotals);
  Range<K>* rng[2] = {(Range<K>*)P(bRanges0), (Range<K>*)P(bRanges1)};  // passes >= 1, by parity
  Range<K>* rng_first = (Range<K>*)P(bRangesFirst);                      // pass 0, kept for pass 1
  PassCtl* ctl = (PassCtl*)P(bCtl);

  uint32_t* k_rank = (uint32_t*)P(bKRank);
  uint32_t* k_state = (uint32_t*)P(bKState);
  uint32_t* k_aux = (uint32_t*)P(bKAux);
  uint32_t* klist = (uint32_t*)P(bKList);

  // pinned host memory (the kernels read and write it directly): ctl blocks
  // [0, 4) for copies, the plan's mirror of the current and next control
  // blocks [4, 6), the finish's error flag and the bad-k flag, then the ks
  // (read by k_setup) and the outputs (written by the plan and the finish)
  char* H = (char*)ctx->pinned;
  char* Hd = (char*)ctx->pinned_dev;
  PassCtl* hctl = (PassCtl*)H;
  PassCtl* hmirror = hctl + 4;
  PassCtl* dmirror = (PassCtl*)(Hd + 4 * ctlb);
  volatile uint32_t* hflag = (volatile uint32_t*)(H + 6 * ctlb);
  uint32_t* dflag = (uint32_t*)(Hd + 6 * ctlb);
  volatile uint32_t* hbad = hflag + 1;
  uint32_t* dbad = dflag + 1;
  volatile uint32_t* hseq = hflag + 3;  // the plan's mirror sequence word (written after the mirror)
  uint32_t* dseq = dflag + 3;
  char* hks = H + 6 * ctlb + 256;
  const uint32_t* d_ks = ks_on_device ? ks : (const uint32_t*)(Hd + 6 * ctlb + 256);
  char* hout = hks + ks_bytes;
  T* d_out = out_on_device ? (T*)out : (T*)(Hd + 6 * ctlb + 256 + ks_bytes);
  *hflag = 0;
  *hbad = 0;
  if (!ks_on_device) memcpy(hks, ks, (size_t)nk * 4);
  ggms_stats& stt = ctx->stats;
  memset(&stt, 0, sizeof(stt));
  Prof prof(ctx, st, o.profile != 0);

  if (c.device_schedule) {
    s = launch_fixed<T>(c, L, W, dflag, d_ks, d_out, ctx->device, ctx->cus, d_vec, n, nk, st, prof);
    if (s) return s;
    if (out_on_device && ks_on_device) return GGMS_OK;  // asynchronous: ggms_ctx_device_status after the stream
    HIP_TRY(hipStreamSynchronize(st));
    if (!out_on_device) memcpy(out, hout, (size_t)nk * sizeof(T));
    prof.collect(stt);
    return fixed_status((const uint32_t*)hflag);
  }

  // control block 0
  PassCtl c0{};
  c0.n = n;
  uint32_t tile0 = 0, G0 = 0;
  first_pass_tiling(n64, c.G_max, c.min_tile, &tile0, &G0);
  c0.tile = tile0;
  c0.G = G0;
  c0.nan_lo_bucket = kNone32;
  c0.nan_hi_bucket = kNone32;
  c0.hot_bucket = kNone32;
  // setup (control blocks, ks -> ranks): its own kernel, or folded into the
  // first-pass planning kernel when the ks are few enough for one workgroup
  const uint32_t ctl_words = (uint32_t)((c.max_passes + 2) * ctlb / 4);
  const bool setup_in_plan0 = n > c.fin_cap && nk <= kSetupInPlan0MaxKs;
  if (!setup_in_plan0) {
    prof.begin(GGMS_KERNEL_OTHER, -1);
    const uint32_t threads = std::max(ctl_words, nk);
    k_setup<<<(threads + 255) / 256, 256, 0, st>>>(d_ks, nk, n, k_rank, k_state, ctl, ctl_words, c0, dbad);
    prof.end();
    HIP_TRY(hipGetLastError());
  }

  const uint32_t fin_grid = std::max<uint32_t>(1, std::min<uint32_t>(4 * (uint32_t)ctx->cus, 65535));
  const size_t fin_lds = (size_t)(c.fin_cap / 2) * sizeof(K);  // two workgroups per CU
  set_lds_attr((const void*)k_finish<T>, fin_lds);

  if (n <= c.fin_cap) {
    const bool wide = finish_wide_segment<K>(n, (K)0, KeyMax<K>::v, c.fin_cap);
    prof.begin(GGMS_KERNEL_OTHER, -1);
    k_tiny_setup<T><<<1, 1024, 0, st>>>(d_vec, n, nk, bufA, k_rank, k_state, (uint32_t*)P(bSmallStart),
                                        (uint32_t*)P(bSmallLen), (K*)P(bSmallKlo), (K*)P(bSmallKhi),
                                        (uint32_t*)P(bSmallKStart), klist, ctl, wide ? 1u : 0u);
    prof.end();
    HIP_TRY(hipGetLastError());
    prof.begin(GGMS_KERNEL_FINISH, 0);
    k_finish<T><<<1, 1024, fin_lds, st>>>(
        bufA, ctl, (uint32_t*)P(bSmallStart), (uint32_t*)P(bSmallLen), (K*)P(bSmallKlo), (K*)P(bSmallKhi),
        (uint32_t*)P(bSmallKStart), nullptr, klist, k_rank, k_state, d_out, c.fin_cap, dflag);
    if (wide)
      k_finish_wide<T><<<1, 1024, 0, st>>>(bufA, bufB, ctl, (uint32_t*)P(bSmallStart), (uint32_t*)P(bSmallLen),
                                           (K*)P(bSmallKlo), (K*)P(bSmallKhi), (uint32_t*)P(bSmallKStart), nullptr,
                                           klist, k_rank, k_state, d_out, c.fin_cap);
    prof.end();
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    if (*hbad) return GGMS_EINVAL;
    if (*hflag) return GGMS_EDEVICE;
    stt.passes = 0;
    stt.finished_segments = 1;
  } else {
    // first-pass splitter
    const Plan0Args pa = plan0_args<T>(c, ctx->device, nk, ctl_words, setup_in_plan0, d_ks, k_rank, k_state, dbad, c0);
    size_t lds_p0 = plan0_lds_bytes<K>(pow2ceil_host(c.S));
    LDS_CHECK(lds_p0);
    set_lds_attr((const void*)k_plan0<T>, lds_p0);
    prof.begin(GGMS_KERNEL_PLAN0, 0);
    k_plan0<T><<<1, 1024, lds_p0, st>>>(d_vec, pa, rng_first, (K*)P(bLut), ctl);
    prof.end();
    HIP_TRY(hipGetLastError());

    uint32_t G = G0, B_bound = c.B0_bound, R_bound = c.R0_bound;
    uint32_t B0 = c.B0_bound, R0 = c.R0_bound, L2 = LutGeom<K>::L2TABLES;  // exact after pass 0
    uint16_t* slot_tab[2] = {(uint16_t*)P(bSlotOfBucket),
                             (uint16_t*)P(bSlotOfBucket) + (std::max<size_t>(c.B0_bound, c.B_cap) + 64)};
    uint32_t U_slot_bound = std::min<uint32_t>(nk, B_bound);
    for (uint32_t pass = 0;; ++pass) {
      PassCtl* cp = ctl + pass;
      PassCtl* cn = ctl + pass + 1;
      Range<K>* rc = pass == 0 ? rng_first : rng[pass & 1];
      Range<K>* rn = rng[(pass + 1) & 1];
      K* src = (pass & 1) ? bufA : bufB;  // input of pass >= 1
      K* dst = (pass & 1) ? bufB : bufA;  // output of this pass's compaction
      // tile -> first range table: read by this pass's streaming kernels, written
      // by this pass's plan for the next pass (so it must not alias)
      uint32_t* wgf_cur = (uint32_t*)P(bWgFirst) + (size_t)(pass & 1) * c.G_max;
      uint32_t* wgf_next = (uint32_t*)P(bWgFirst) + (size_t)((pass + 1) & 1) * c.G_max;
      const uint32_t NT = c.NT;
      // pass-1 range parameters live in LDS unless that would push either
      // pass-1 kernel past kPass1ParamsLds (then they are read from global memory)
      uint32_t p1_lds = 0;
      if (pass == 1 && !c.seg_first) {
        size_t cnt = ((size_t)B_bound * 4 + 15) / 16 * 16 + pass1_map_bytes<K>(B0, R0, L2, R_bound);
        size_t cmp = compact_lds_bytes(B_bound, U_slot_bound + 4) + pass1_map_bytes<K>(B0, R0, L2, R_bound);
        p1_lds = std::max(cnt, cmp) <= kPass1ParamsLds ? 1u : 0u;
      }
      // count
      if (pass == 0) {
        size_t lds = count_lds0<T>(c);
        LDS_CHECK(count_lds0_total<T>(c));
        // when the tables already allow only one workgroup per CU and a second
        // histogram still fits, each workgroup counts two consecutive tiles
        // (one row each) without a barrier between them, so its loads do not
        // drain and restart at a tile boundary (2^28 fp32: 0.185 -> 0.175 ms)
        const size_t hist2 = (size_t)c.B0_bound * 4;
        const uint32_t tpw =
            (count_lds0_total<T>(c) > kLdsOneWgPerCu && count_lds0_total<T>(c) + hist2 <= kLdsLimit) ? 2u : 1u;
        if (tpw == 2) lds += hist2;
        set_lds_attr((const void*)k_count0<T>, lds);
        prof.begin(GGMS_KERNEL_COUNT0, (int)pass);
        k_count0<T><<<(G + tpw - 1) / tpw, NT, lds, st>>>(d_vec, rc, (K*)P(bLut), cp, rows, tpw);
        prof.end();
      } else if (pass == 1 && !c.seg_first) {
        size_t lds = ((size_t)B_bound * 4 + 15) / 16 * 16 + pass1_map_bytes<K>(B0, R0, L2, p1_lds ? R_bound : 0);
        LDS_CHECK(lds);
        set_lds_attr((const void*)k_count1<K>, lds);
        prof.begin(GGMS_KERNEL_COUNTS, (int)pass);
        k_count1<K><<<G, NT, lds, st>>>(src, ctl, (K*)P(bLut), rng_first, (uint64_t*)P(bRank0),
                                        (MapParam<K>*)P(bParams1), p1_lds ? (uint32_t)lds : 0u, cp, rows);
        prof.end();
      } else {
        size_t lds = (size_t)B_bound * 4;
        LDS_CHECK(lds);
        set_lds_attr((const void*)k_countS<K>, lds);
        prof.begin(GGMS_KERNEL_COUNTS, (int)pass);
        k_countS<K><<<G, NT, lds, st>>>(src, rc, wgf_cur, cp, rows);
        prof.end();
      }
      HIP_TRY(hipGetLastError());
      prof.begin(GGMS_KERNEL_COLSUM, (int)pass);
      k_colsum<<<(B_bound + 63) / 64, 1024, 0, st>>>(rows, cp, totals);
      prof.end();
      HIP_TRY(hipGetLastError());
      // plan
      PlanArgs<K> pl = plan_args_common<K>(c, L, W, nk, d_out);
      pl.ranges = rc;
      pl.ctl = cp;
      pl.ctl_next = cn;
      pl.next_ranges = rn;
      pl.slot_of_bucket = slot_tab[pass & 1];
      pl.wg_first_next = wgf_next;
      pl.force_finish = (pass + 1 >= c.max_passes) ? 1u : c.force_finish;
      pl.finish_cap = (pass == 0 && c.one_pass) ? kOnePassFinishCap * c.fin_cap : c.fin_cap;
      pl.merge_small = (pass == 0 && !c.seg_first) ? 1u : 0u;
      pl.next_params = (pass == 0 && !c.seg_first) ? (MapParam<K>*)P(bParams1) : nullptr;
      pl.rank_words = (pass == 0 && !c.seg_first) ? (uint64_t*)P(bRank0) : nullptr;
      pl.mirror = dmirror;
      pl.mirror_seq = dseq;
      if (++ctx->mirror_seq == 0) ++ctx->mirror_seq;
      pl.seq = ctx->mirror_seq;
      pl.nan_out = nullptr;
      plan_groups<K>(pl, L, W, pass == 0 && c.seg_first);
      size_t lds_plan = plan_lds_bytes(B_bound, nk);
      pl.lds_words = (uint32_t)(lds_plan / 4);
      LDS_CHECK(lds_plan);
      set_lds_attr((const void*)k_plan<T>, lds_plan);
      prof.begin(GGMS_KERNEL_PLAN, (int)pass);
      k_plan<T><<<1, 1024, lds_plan, st>>>(pl);
      prof.end();
      HIP_TRY(hipGetLastError());
      // compaction
      if (pass == 0 && c.seg_first) {
        // every k-bucket into its own segment (the finish reads segments)
        s = launch_segmented_first_pass<T>(c, L, W, d_vec, n, G, cp, pl, src, dst, st, prof);
        if (s) return s;
      } else if (pass == 0) {
        // dynamic part: map parameters then the slot table (cell table is static LDS)
        size_t lds = (size_t)R_bound * sizeof(MapParam<K>) + (size_t)B_bound * 2;
        LDS_CHECK(lds + LutGeom<K>::bytes());
        set_lds_attr((const void*)k_compact0<T>, lds);
        prof.begin(GGMS_KERNEL_COMPACT0, (int)pass);
        k_compact0<T><<<G, NT, lds, st>>>(d_vec, rc, (K*)P(bLut), cp, rows, pl.slot_of_bucket, pl.slot_bucket,
                                          dst);
        prof.end();
      } else if (pass == 1 && !c.seg_first) {
        size_t lds = compact_lds_bytes(B_bound, U_slot_bound + 4) + pass1_map_bytes<K>(B0, R0, L2, p1_lds ? R_bound : 0);
        // one workgroup per CU: its writes scatter over every k-bucket segment,
        // and twice the workgroups in flight keep twice the partial HBM lines
        // open (measured 0.197 ms with two per CU, 0.154 ms with one)
        lds = std::max<size_t>(lds, kLdsOneWgPerCu);
        LDS_CHECK(lds);
        set_lds_attr((const void*)k_compact1<K>, lds);
        prof.begin(GGMS_KERNEL_COMPACTS, (int)pass);
        k_compact1<K><<<G, NT, lds, st>>>(src, ctl, (K*)P(bLut), rng_first, (uint64_t*)P(bRank0),
                                          (MapParam<K>*)P(bParams1), p1_lds ? (uint32_t)lds : 0u, cp, rows,
                                          pl.slot_of_bucket, pl.slot_dst, pl.slot_bucket, dst);
        prof.end();
      } else {
        size_t lds = compact_lds_bytes(B_bound, U_slot_bound + 4);
        LDS_CHECK(lds);
        set_lds_attr((const void*)k_compactS<K>, lds);
        prof.begin(GGMS_KERNEL_COMPACTS, (int)pass);
        k_compactS<K><<<G, NT, lds, st>>>(src, rc, wgf_cur, cp, rows, pl.slot_of_bucket,
                                          pl.slot_dst, pl.slot_bucket, dst);
        prof.end();
      }
      HIP_TRY(hipGetLastError());
      // the next pass (and this pass's finish) is planned from the control
      // blocks the plan kernel wrote to host memory, while the compaction
      // runs.  The host polls the sequence word the plan writes after them:
      // an event recorded between the plan and the compaction would make the
      // GPU drain and publish at that point (~6 us idle before each compaction).
      s = wait_mirror(hseq, pl.seq, st);
      if (s) return s;
      const PassCtl h = hmirror[0];
      const PassCtl hn = hmirror[1];
      // the iterative first pass hands every k-bucket to pass 1
      if ((pass > 0 || c.seg_first) && h.U_small > 0) {
        // every segment within a quarter of the capacity: half-size workgroups
        // (512 threads, half the LDS) so twice as many run at once
        const bool half = h.small_max <= c.fin_cap / 4;
        // many small segments beside a few large ones: the large ones first
        // (full workgroups), then the small ones in half-size workgroups
        const uint32_t U_grp = h.U_small - h.U_wave - h.U_wide;
        const bool split = !half && c.fin_cap / 4 >= kFinishSplitLen &&
                           U_grp - std::min(h.U_fin_big, U_grp) >= kFinishSplitMinSegs;
        prof.begin(GGMS_KERNEL_FINISH, (int)pass);
        if (split) {
          k_finish<T><<<fin_grid, 1024, fin_lds, st>>>(dst, cp, pl.small_start, pl.small_len, pl.small_klo,
                                                       pl.small_khi, pl.small_kstart, pl.small_order, klist, k_rank,
                                                       k_state, d_out, c.fin_cap, dflag, 1u);
          k_finish<T><<<fin_grid, 512, fin_lds / 2, st>>>(dst, cp, pl.small_start, pl.small_len, pl.small_klo,
                                                          pl.small_khi, pl.small_kstart, pl.small_order, klist, k_rank,
                                                          k_state, d_out, c.fin_cap / 2, dflag, 2u);
        } else {
          k_finish<T><<<fin_grid, half ? 512 : 1024, half ? fin_lds / 2 : fin_lds, st>>>(
              dst, cp, pl.small_start, pl.small_len, pl.small_klo, pl.small_khi, pl.small_kstart, pl.small_order,
              klist, k_rank, k_state, d_out, half ? c.fin_cap / 2 : c.fin_cap, dflag);
        }
        if (h.U_wide > 0)
          k_finish_wide<T><<<std::min<uint32_t>(h.U_wide, (uint32_t)ctx->cus), 1024, 0, st>>>(
              dst, src, cp, pl.small_start, pl.small_len, pl.small_klo, pl.small_khi, pl.small_kstart, pl.small_order,
              klist, k_rank, k_state, d_out, c.fin_cap);
        prof.end();
        HIP_TRY(hipGetLastError());
      }
      if (pass == 0 && (*hbad || h.nan_count)) {
        HIP_TRY(hipStreamSynchronize(st));
        return *hbad ? GGMS_EINVAL : GGMS_ENAN;
      }
      if (pass == 0) {
        B0 = h.B;
        R0 = h.R;
        L2 = h.l2_used;
      }
      if (pass < GGMS_MAX_PASS_STATS) {
        stt.n_in[pass] = h.n;
        stt.active[pass] = h.U;
        stt.n_kept[pass] = (uint64_t)h.N_big + h.N_small;
        stt.buckets[pass] = h.B;
      }
      stt.passes = pass + 1;
      stt.finished_segments += h.U_small;
      if (h.N_big == 0 || h.U_big == 0) break;
      if (pass + 1 >= c.max_passes) {  // cannot happen: the last pass forces finishing
        HIP_TRY(hipStreamSynchronize(st));
        return GGMS_EDEVICE;
      }
      G = hn.G;
      B_bound = std::max<uint32_t>(hn.B, 1);
      R_bound = hn.R;
      U_slot_bound = std::min<uint32_t>(nk, B_bound);
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  if (!out_on_device) memcpy(out, hout, (size_t)nk * sizeof(T));
  if (*hflag) return GGMS_EDEVICE;  // a segment over the finish capacity (never expected)
  prof.collect(stt);
  return GGMS_OK;
}

template <class T>
static ggms_status run_first_pass(ggms_ctx* ctx, const T* d_vec, uint64_t n64, const ggms_opts& o,
                                  ggms_range_u64* ranges_host, uint32_t* num_ranges, uint32_t* hist_host,
                                  uint32_t* num_buckets) {
  using K = typename KeyTraits<T>::K;
  if (n64 == 0 || n64 > kMaxN || !d_vec || !num_ranges || !num_buckets) return GGMS_EINVAL;
  const uint32_t n = (uint32_t)n64;
  hipStream_t st = (hipStream_t)o.stream;
  Config c = resolve_config<K>(GGMS_ALG_ITERATIVE, n64, 1, o, ctx->cus);
  if (c.mode == kModeLearned && !learned_weights_device(ctx->device)) return GGMS_ENOTSUP;
  Layout L = make_layout<T>(n64, 1, c);
  ggms_status s = ensure_ws(ctx, L.total);
  if (s) return s;
  const size_t ctlb = sizeof(PassCtl);
  size_t rbytes = (size_t)c.R0_bound * sizeof(Range<K>);
  s = ensure_pinned(ctx, 4 * ctlb + rbytes + (size_t)c.B0_bound * 4 + 4096);
  if (s) return s;
  char* W = (char*)ctx->ws.base;
  auto P = [&](int b) { return (void*)(W + L.off[b]); };
  uint32_t* rows = (uint32_t*)P(bRows);
  uint32_t* totals = (uint32_t*)P(bTotals);
  Range<K>* rng = (Range<K>*)P(bRanges0);
  PassCtl* ctl = (PassCtl*)P(bCtl);
  char* H = (char*)ctx->pinned;
  PassCtl* hctl = (PassCtl*)H;
  char* hdata = H + 4 * ctlb;
  PassCtl c0{};
  c0.n = n;
  uint32_t tile0 = 0, G0 = 0;
  first_pass_tiling(n64, c.G_max, c.min_tile, &tile0, &G0);
  c0.tile = tile0;
  c0.G = G0;
  c0.nan_lo_bucket = kNone32;
  c0.nan_hi_bucket = kNone32;
  c0.hot_bucket = kNone32;
  memcpy(&hctl[2], &c0, ctlb);
  HIP_TRY(hipMemcpyAsync(ctl, &hctl[2], ctlb, hipMemcpyHostToDevice, st));
  Plan0Args pa{};
  pa.mode = c.mode;
  pa.S = c.S;
  pa.P = c.P;
  pa.B = c.B;
  pa.E = c.E;
  pa.strideS = 8192;
  pa.seed = c.seed;
  pa.mlp = c.mode == kModeLearned ? learned_weights_device(ctx->device) : nullptr;
  size_t lds_p0 = plan0_lds_bytes<K>(pow2ceil_host(c.S));
  LDS_CHECK(lds_p0);
  set_lds_attr((const void*)k_plan0<T>, lds_p0);
  k_plan0<T><<<1, 1024, lds_p0, st>>>(d_vec, pa, rng, (K*)P(bLut), ctl);
  HIP_TRY(hipGetLastError());
  size_t lds = count_lds0<T>(c);
  LDS_CHECK(count_lds0_total<T>(c));
  set_lds_attr((const void*)k_count0<T>, lds);
  k_count0<T><<<G0, c.NT, lds, st>>>(d_vec, rng, (K*)P(bLut), ctl, rows, 1u);
  HIP_TRY(hipGetLastError());
  k_colsum<<<(c.B0_bound + 63) / 64, 1024, 0, st>>>(rows, ctl, totals);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&hctl[0], ctl, ctlb, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const uint32_t R = hctl[0].R, B = hctl[0].B;
  if (*num_ranges < R || *num_buckets < B) {
    *num_ranges = R;
    *num_buckets = B;
    return GGMS_EINVAL;
  }
  Range<K>* hr = (Range<K>*)hdata;
  uint32_t* hh = (uint32_t*)(hdata + rbytes);
  HIP_TRY(hipMemcpyAsync(hr, rng, (size_t)R * sizeof(Range<K>), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hh, totals, (size_t)B * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (uint32_t i = 0; i < R; ++i) {
    ranges_host[i].lo = hr[i].lo;
    ranges_host[i].hi = hr[i].hi;
    ranges_host[i].base = hr[i].base;
    ranges_host[i].nb = hr[i].nb;
    ranges_host[i].sh = hr[i].sh;
    ranges_host[i].mult = hr[i].mult;
  }
  if (hist_host) memcpy(hist_host, hh, (size_t)B * 4);
  *num_ranges = R;
  *num_buckets = B;
  return GGMS_OK;
}

// Default contexts, one per device, for the reference-shaped wrappers.
static std::mutex g_default_mu;
static std::vector<ggms_ctx*> g_default;

static ggms_ctx* default_ctx() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_default_mu);
  if ((int)g_default.size() <= dev) g_default.resize(dev + 1, nullptr);
  if (!g_default[dev]) {
    ggms_ctx* c = nullptr;
    if (ggms_ctx_create(dev, &c) != GGMS_OK) return nullptr;
    g_default[dev] = c;
  }
  return g_default[dev];
}

// Status of the calling thread's last reference-shaped wrapper call.
static ggms_status& wrapper_status() {
  static thread_local ggms_status s = GGMS_OK;
  return s;
}

template <class T>
static int wrapper(ggms_alg alg, ggms_dtype dt, T* d_vector, int length, const uint32_t* kVals_ori, int numKs,
                   T* outputs, int blocks, int threads) {
  (void)blocks;
  ggms_status s = GGMS_OK;
  ggms_ctx* c = nullptr;
  if (length <= 0 || numKs < 0) s = GGMS_EINVAL;
  else if (!(c = default_ctx())) s = GGMS_EDEVICE;
  else {
    ggms_opts o;
    ggms_opts_default(&o);
    if (threads > 0 && threads % 64 == 0 && threads <= 1024) o.block_threads = (uint32_t)threads;
    s = ggms_select(c, alg, dt, d_vector, (uint64_t)length, kVals_ori, (uint32_t)numKs, outputs, &o);
  }
  wrapper_status() = s;
  return s;
}

template <class T>
static ggms_status reserve_t(ggms_ctx* ctx, uint64_t n, uint32_t nk, const ggms_opts& o) {
  using K = typename KeyTraits<T>::K;
  if (n == 0 || n > kMaxN) return GGMS_EINVAL;
  const Config c = resolve_config<K>(GGMS_ALG_ITERATIVE, n, nk, o, ctx->cus);
  const Config cb = resolve_config<K>(GGMS_ALG_BUCKET, n, nk, o, ctx->cus);
  const Layout L = make_layout<T>(n, nk, c), Lb = make_layout<T>(n, nk, cb);
  ggms_status s = ensure_ws(ctx, std::max(L.total, Lb.total));
  if (s) return s;
  const size_t ks_bytes = ((size_t)nk * 4 + 255) / 256 * 256;
  return ensure_pinned(ctx, 6 * sizeof(PassCtl) + 256 + ks_bytes + (size_t)nk * sizeof(T) + 4096);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (ggms_select_noalloc): one fixed
// array per device, whose device address is looked up once per device with
// hipGetSymbolAddress (in host code the array's name is not a device address).
// Every ggms_select_noalloc call on a device uses that device's array, so calls
// that overlap on one device race; the array keeps the last call's contents,
// and every call rebuilds what it reads (status words cleared by a fill
// kernel, control blocks by the setup step).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)GGMS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices];

static ggms_status noalloc_scratch(int device, char** out) {
  if (device < 0 || device >= kMaxDevices) return GGMS_EDEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return GGMS_EDEVICE;
    }
    g_scratch_ptr[device] = (char*)p;
  }
  *out = g_scratch_ptr[device];
  return GGMS_OK;
}

template <class T>
static ggms_status run_noalloc(ggms_alg alg, const T* d_vec, uint64_t n64, const uint32_t* d_ks, uint32_t nk, T* d_out,
                               const ggms_opts& o) {
  using K = typename KeyTraits<T>::K;
  if (n64 == 0 || n64 > kMaxN) return GGMS_EINVAL;
  if (nk == 0) return GGMS_OK;
  if (!d_vec || !d_ks || !d_out) return GGMS_EINVAL;
  int device = 0, cus = 0;
  if (hipGetDevice(&device) != hipSuccess) return GGMS_EDEVICE;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (cus <= 0) cus = 256;
  ggms_opts od = o;
  od.schedule = GGMS_SCHEDULE_DEVICE;
  const Config c = resolve_config<K>(alg, n64, nk, od, cus);
  if (c.mode == kModeLearned && !learned_weights_device(device)) return GGMS_ENOTSUP;
  const Layout L = make_layout<T>(n64, nk, c);
  // the status words sit behind the workspace
  if (L.total + 256 > kNoAllocScratchBytes) return GGMS_ENOMEM;
  char* base = nullptr;
  ggms_status s = noalloc_scratch(device, &base);
  if (s) return s;
  uint32_t* flags = (uint32_t*)(base + L.total);
  hipStream_t st = (hipStream_t)o.stream;
  Prof prof(nullptr, st, false);
  s = launch_fixed<T>(c, L, base, flags, d_ks, d_out, device, cus, d_vec, (uint32_t)n64, nk, st, prof);
  if (s) return s;
  HIP_TRY(hipStreamSynchronize(st));
  uint32_t words[3] = {0, 0, 0};
  HIP_TRY(hipMemcpy(words, flags, sizeof(words), hipMemcpyDeviceToHost));
  return fixed_status(words);
}

}  // namespace ggms

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void ggms_opts_default(ggms_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->num_buckets = 0;
  o->num_pivots = 17;
  o->sample_size = 0;
  o->block_threads = 1024;
  o->blocks_per_cu = 0;
  o->splitter = GGMS_SPLIT_LINEAR;
  o->learned_pieces = 0;  // automatic (see resolve_config)
  o->max_passes = 64;
  o->seed = 0x5EEDull;
  o->stream = nullptr;
}

ggms_status ggms_last_wrapper_status(void) { return wrapper_status(); }

const char* ggms_status_str(ggms_status s) {
  switch (s) {
    case GGMS_OK: return "GGMS_OK";
    case GGMS_EINVAL: return "GGMS_EINVAL";
    case GGMS_ENAN: return "GGMS_ENAN";
    case GGMS_ENOMEM: return "GGMS_ENOMEM";
    case GGMS_EDEVICE: return "GGMS_EDEVICE";
    case GGMS_ENOTSUP: return "GGMS_ENOTSUP";
  }
  return "GGMS_UNKNOWN";
}

ggms_status ggms_ctx_create(int device, ggms_ctx** out) {
  if (!out) return GGMS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return GGMS_EDEVICE;
  ggms_ctx* c = new (std::nothrow) ggms_ctx();
  if (!c) return GGMS_ENOMEM;
  c->device = device;
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return GGMS_EDEVICE;
  }
  int cus = 0, lds = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  c->cus = cus > 0 ? cus : 256;
  c->max_lds = lds > 0 ? (size_t)lds : 65536;
  (void)hipSetDevice(cur);
  *out = c;
  return GGMS_OK;
}

void ggms_ctx_destroy(ggms_ctx* c) {
  if (!c) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(c->device);
  if (c->ws.base) (void)hipFree(c->ws.base);
  if (c->pinned) (void)hipHostFree(c->pinned);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  (void)hipSetDevice(cur);
  delete c;
}

ggms_status ggms_ctx_reserve(ggms_ctx* ctx, ggms_dtype t, uint64_t n, uint32_t nk, const ggms_opts* opts) {
  if (!ctx) return GGMS_EINVAL;
  ggms_opts o;
  if (opts) o = *opts; else ggms_opts_default(&o);
  std::lock_guard<std::mutex> lk(ctx->mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != ctx->device) (void)hipSetDevice(ctx->device);
  ggms_status s = GGMS_EINVAL;
  switch (t) {
    case GGMS_F32: s = reserve_t<float>(ctx, n, nk, o); break;
    case GGMS_F64: s = reserve_t<double>(ctx, n, nk, o); break;
    case GGMS_U32: s = reserve_t<uint32_t>(ctx, n, nk, o); break;
    case GGMS_I32: s = reserve_t<int32_t>(ctx, n, nk, o); break;
  }
  if (cur != ctx->device) (void)hipSetDevice(cur);
  return s;
}

ggms_status ggms_ctx_device_status(const ggms_ctx* ctx) {
  if (!ctx || !ctx->pinned) return GGMS_EINVAL;
  return fixed_status((const uint32_t*)((const char*)ctx->pinned + 6 * sizeof(PassCtl)));
}

ggms_status ggms_ctx_stats(const ggms_ctx* c, ggms_stats* out) {
  if (!c || !out) return GGMS_EINVAL;
  *out = c->stats;
  return GGMS_OK;
}

ggms_status ggms_first_pass_tiling(uint64_t n, uint32_t workgroups, uint32_t* tile, uint32_t* groups) {
  if (n == 0 || n > kMaxN || workgroups == 0 || !tile || !groups) return GGMS_EINVAL;
  first_pass_tiling(n, std::min<uint32_t>(workgroups, kMaxWorkgroups), kMinTile, tile, groups);
  return GGMS_OK;
}

uint64_t ggms_workspace_bytes(ggms_dtype t, uint64_t n, uint32_t nk, const ggms_opts* opts) {
  ggms_opts o;
  if (opts) o = *opts; else ggms_opts_default(&o);
  int cus = 256;
  switch (t) {
    case GGMS_F32: return make_layout<float>(n, nk, resolve_config<uint32_t>(GGMS_ALG_ITERATIVE, n, nk, o, cus)).total;
    case GGMS_F64: return make_layout<double>(n, nk, resolve_config<uint64_t>(GGMS_ALG_ITERATIVE, n, nk, o, cus)).total;
    case GGMS_U32: return make_layout<uint32_t>(n, nk, resolve_config<uint32_t>(GGMS_ALG_ITERATIVE, n, nk, o, cus)).total;
    case GGMS_I32: return make_layout<int32_t>(n, nk, resolve_config<uint32_t>(GGMS_ALG_ITERATIVE, n, nk, o, cus)).total;
  }
  return 0;
}

static ggms_status select_common(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n,
                                 const uint32_t* ks, bool ks_dev, uint32_t nk, void* out, bool out_dev,
                                 const ggms_opts* opts) {
  if (!ctx) return GGMS_EINVAL;
  if ((unsigned)alg > 3u) return GGMS_EINVAL;
  ggms_opts o;
  if (opts) o = *opts; else ggms_opts_default(&o);
  std::lock_guard<std::mutex> lk(ctx->mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != ctx->device) (void)hipSetDevice(ctx->device);
  ggms_status s;
  switch (t) {
    case GGMS_F32: s = run_select<float>(ctx, alg, (const float*)d_vec, n, ks, ks_dev, nk, out, out_dev, o); break;
    case GGMS_F64: s = run_select<double>(ctx, alg, (const double*)d_vec, n, ks, ks_dev, nk, out, out_dev, o); break;
    case GGMS_U32: s = run_select<uint32_t>(ctx, alg, (const uint32_t*)d_vec, n, ks, ks_dev, nk, out, out_dev, o); break;
    case GGMS_I32: s = run_select<int32_t>(ctx, alg, (const int32_t*)d_vec, n, ks, ks_dev, nk, out, out_dev, o); break;
    default: s = GGMS_EINVAL;
  }
  if (cur != ctx->device) (void)hipSetDevice(cur);
  return s;
}

ggms_status ggms_select(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n,
                        const uint32_t* ks_host, uint32_t nk, void* out_host, const ggms_opts* opts) {
  return select_common(ctx, alg, t, d_vec, n, ks_host, false, nk, out_host, false, opts);
}

ggms_status ggms_select_noalloc(ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n, const uint32_t* d_ks,
                                uint32_t nk, void* d_out, const ggms_opts* opts) {
  if ((unsigned)alg > 3u) return GGMS_EINVAL;
  ggms_opts o;
  if (opts) o = *opts; else ggms_opts_default(&o);
  switch (t) {
    case GGMS_F32: return run_noalloc<float>(alg, (const float*)d_vec, n, d_ks, nk, (float*)d_out, o);
    case GGMS_F64: return run_noalloc<double>(alg, (const double*)d_vec, n, d_ks, nk, (double*)d_out, o);
    case GGMS_U32: return run_noalloc<uint32_t>(alg, (const uint32_t*)d_vec, n, d_ks, nk, (uint32_t*)d_out, o);
    case GGMS_I32: return run_noalloc<int32_t>(alg, (const int32_t*)d_vec, n, d_ks, nk, (int32_t*)d_out, o);
  }
  return GGMS_EINVAL;
}

ggms_status ggms_select_device(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n,
                               const uint32_t* d_ks, uint32_t nk, void* d_out, const ggms_opts* opts) {
  return select_common(ctx, alg, t, d_vec, n, d_ks, true, nk, d_out, true, opts);
}

ggms_status ggms_select_topk(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n, uint32_t K,
                             void* out_host, const ggms_opts* opts) {
  if (!ctx || !d_vec || n == 0 || n > kMaxN) return GGMS_EINVAL;
  if (K == 0) return GGMS_OK;
  if (K > n || !out_host) return GGMS_EINVAL;
  std::vector<uint32_t> ks;
  try {
    ks.resize(K);
  } catch (const std::bad_alloc&) {
    return GGMS_ENOMEM;
  }
  for (uint32_t i = 0; i < K; ++i) ks[i] = i + 1;  // k = 1 .. K: the K largest, largest first
  return select_common(ctx, alg, t, d_vec, n, ks.data(), false, K, out_host, false, opts);
}

ggms_status ggms_debug_first_pass(ggms_ctx* ctx, ggms_dtype t, const void* d_vec, uint64_t n, const ggms_opts* opts,
                                  ggms_range_u64* ranges_host, uint32_t* num_ranges, uint32_t* hist_host,
                                  uint32_t* num_buckets) {
  if (!ctx) return GGMS_EINVAL;
  ggms_opts o;
  if (opts) o = *opts; else ggms_opts_default(&o);
  std::lock_guard<std::mutex> lk(ctx->mu);
  switch (t) {
    case GGMS_F32: return run_first_pass<float>(ctx, (const float*)d_vec, n, o, ranges_host, num_ranges, hist_host, num_buckets);
    case GGMS_F64: return run_first_pass<double>(ctx, (const double*)d_vec, n, o, ranges_host, num_ranges, hist_host, num_buckets);
    case GGMS_U32: return run_first_pass<uint32_t>(ctx, (const uint32_t*)d_vec, n, o, ranges_host, num_ranges, hist_host, num_buckets);
    case GGMS_I32: return run_first_pass<int32_t>(ctx, (const int32_t*)d_vec, n, o, ranges_host, num_ranges, hist_host, num_buckets);
  }
  return GGMS_EINVAL;
}

#define GGMS_WRAPPERS(SUF, TY, DT)                                                                              \
  int ggms_iterativeSMOSWrapper_##SUF(TY* d_vector, int length, const uint32_t* kVals_ori, int numKs,         \
                                      TY* outputs, int blocks, int threads) {                                   \
    return wrapper<TY>(GGMS_ALG_ITERATIVE, DT, d_vector, length, kVals_ori, numKs, outputs, blocks, threads);   \
  }                                                                                                             \
  int ggms_bucketMultiselectWrapper_##SUF(TY* d_vector, int length, const uint32_t* kVals_ori, int numKs,     \
                                          TY* outputs, int blocks, int threads) {                               \
    return wrapper<TY>(GGMS_ALG_BUCKET, DT, d_vector, length, kVals_ori, numKs, outputs, blocks, threads);      \
  }                                                                                                             \
  int ggms_distributedSMOSWrapper_##SUF(TY* d_vector, int length, const uint32_t* kVals_ori, int numKs,       \
                                        TY* outputs, int blocks, int threads, int rank) {                       \
    (void)rank;                                                                                                 \
    return wrapper<TY>(GGMS_ALG_DISTRIBUTED_1GPU, DT, d_vector, length, kVals_ori, numKs, outputs, blocks,      \
                       threads);                                                                                \
  }                                                                                                             \
  int ggms_distributedBucketMultiselectWrapper_##SUF(TY* d_vector, int length, const uint32_t* kVals_ori,       \
                                                     int numKs, TY* outputs, int blocks, int threads, int rank) {\
    (void)rank;                                                                                                 \
    return wrapper<TY>(GGMS_ALG_DISTRIBUTED_BUCKET_1GPU, DT, d_vector, length, kVals_ori, numKs, outputs,       \
                       blocks, threads);                                                                        \
  }

GGMS_WRAPPERS(f32, float, GGMS_F32)
GGMS_WRAPPERS(f64, double, GGMS_F64)
GGMS_WRAPPERS(u32, uint32_t, GGMS_U32)
GGMS_WRAPPERS(i32, int32_t, GGMS_I32)

}  // extern "C"
