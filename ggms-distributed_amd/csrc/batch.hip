// Batched selection: thousands of independent multiselect problems of uneven
// sizes in ONE persistent launch.
//
// Every workgroup loops: take the next problem from an atomic queue (problems
// are queued largest first, so the tail of the batch is made of small ones),
// solve it alone, repeat.  A problem is solved entirely by one workgroup with
// the same method as the multi-workgroup engine, scaled down:
//   * its ks are sorted by rank in LDS, so the ks of any bucket form one
//     contiguous run (no per-k bucket tables);
//   * problems that fit the LDS finish buffer go straight to the finish;
//   * otherwise a first bucketing pass: Philox sample, LDS sort, equal-quantile
//     pivots, linear-slope ranges, the two-level key-prefix table in LDS;
//     count into an LDS histogram, prefix, locate every k, resolve single-key
//     buckets, compact the other k-buckets into the problem's own scratch
//     region (per-bucket LDS cursors);
//   * k-buckets larger than the finish buffer go on a work stack and are
//     re-bucketed with one linear range (ping-pong between two scratch
//     buffers, each item keeping its own region);
//   * finish: radix select (few ks) or bitonic sort (many ks) in LDS.
// The last workgroup to leave resets the queue, so the launch can be captured
// once in a hipGraph and replayed.  Per-size-class parameters (buckets,
// sample size, pieces) come from a table the autotuner fills.
//
// Reference counterpart: none -- the reference solves one problem per call
// (DistributedSMOS/iterativeSMOS.cu:1069-1577).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ggms.h"
#include "ggms_internal.hpp"
#include "select_kernels.hpp"

namespace ggms {

constexpr uint32_t kBatchMaxKs = GGMS_BATCH_MAX_KS;  // ks per problem
constexpr uint32_t kBatchNB = 8192;        // LDS histogram capacity
constexpr uint32_t kBatchMaxPieces = 64;
constexpr uint32_t kBatchMaxRanges = 2 * (kBatchMaxPieces + 1) + 6;
constexpr uint32_t kBatchClasses = GGMS_BATCH_CLASSES;
// a run of first-pass children whose every member has at most this many ks is
// answered by per-wave radix selects instead of one sort of the whole run
constexpr uint32_t kWaveSelectMaxKs = 8;

template <class K> struct BatchItem {
  K klo, khi;
  uint32_t region;  // offset of the item's keys inside the problem's scratch region
  uint32_t len;
  uint32_t j0, j1;  // run of sorted ks
  uint32_t buf;     // 0: scratch A, 1: scratch B
  uint32_t pad;
};

struct BatchArgs {
  const void* values;
  const uint64_t* off;      // [nprob + 1]
  const uint32_t* ks;
  const uint32_t* koff;     // [nprob + 1]
  void* out;
  const uint32_t* order;    // queue order (largest first)
  uint32_t* queue;          // [2]: next problem, workgroups done
  uint32_t* status;         // per problem
  void* scratchA;
  void* scratchB;
  void* stacks;             // per workgroup item stack
  uint32_t nprob;
  uint32_t stack_cap;
  uint64_t seed;
  const ggms_batch_class* cfg;  // [GGMS_BATCH_CLASSES], device memory
};

template <class K> __host__ __device__ constexpr uint32_t batch_fin() { return sizeof(K) == 4 ? 16384u : 8192u; }

template <class K> struct BatchLds {
  K* A;             // finish buffer / sample / cell table + slot table + cursors
  uint32_t* H;      // histogram, then its inclusive prefix
  uint32_t* Kr;     // sorted ks: rank inside the current item (1-based)
  uint32_t* Ko;     // sorted ks: output index
  K* starts;        // range starts of the first pass
  uint32_t* req;    // requested buckets per range
  MapParam<K>* prm; // map parameters per range (the first-pass map's parameter arrays)
  uint32_t* rbase;  // first bucket per range
  uint32_t* cursor; // per child item: next write position
  uint32_t* runb;   // per k: its bucket

  static __host__ __device__ constexpr size_t bytes() {
    return (size_t)batch_fin<K>() * sizeof(K) + kBatchNB * 4 + 4 * kBatchMaxKs * 4 +
           kBatchMaxRanges * sizeof(MapParam<K>) + kBatchMaxRanges * sizeof(K) + 2 * kBatchMaxRanges * 4;
  }
  __device__ void carve(unsigned char* p) {
    size_t o = 0;
    A = reinterpret_cast<K*>(p + o);
    o += (size_t)batch_fin<K>() * sizeof(K);
    H = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchNB * 4;
    Kr = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchMaxKs * 4;
    Ko = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchMaxKs * 4;
    prm = reinterpret_cast<MapParam<K>*>(p + o);
    o += kBatchMaxRanges * sizeof(MapParam<K>);
    starts = reinterpret_cast<K*>(p + o);
    o += kBatchMaxRanges * sizeof(K);
    req = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchMaxRanges * 4;
    rbase = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchMaxRanges * 4;
    cursor = reinterpret_cast<uint32_t*>(p + o);
    o += kBatchMaxKs * 4;
    runb = reinterpret_cast<uint32_t*>(p + o);
  }
};

__device__ __forceinline__ uint32_t class_of(uint32_t n) {
  uint32_t l = n > 1 ? 31u - (uint32_t)__builtin_clz(n - 1) + 1u : 0u;  // ceil(log2 n)
  return l < kBatchClasses ? l : kBatchClasses - 1;
}

// Answers the ks [j0, j1) of an item whose len keys are in LDS buffer A.
template <class T>
__device__ __forceinline__ void batch_finish(BatchLds<typename KeyTraits<T>::K>& L, uint32_t len, typename KeyTraits<T>::K klo,
                             typename KeyTraits<T>::K khi, uint32_t j0, uint32_t j1, T* out, uint32_t* hist256,
                             uint32_t* bcast) {
  using K = typename KeyTraits<T>::K;
  if (j1 - j0 <= kSelectMaxKs) {
    for (uint32_t j = j0; j < j1; ++j) {
      K key = radix_select_lds<K>(L.A, len, klo, khi, L.Kr[j], hist256, bcast);
      if (threadIdx.x == 0) out[L.Ko[j]] = KeyTraits<T>::dec(key);
    }
  } else {
    uint32_t M = pow2ceil_u32(len);
    for (uint32_t i = len + threadIdx.x; i < M; i += blockDim.x) L.A[i] = KeyMax<K>::v;
    __syncthreads();
    bitonic_sort_lds<K, true, false>(L.A, M);
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) out[L.Ko[j]] = KeyTraits<T>::dec(L.A[L.Kr[j] - 1]);
  }
  __syncthreads();
}

// After the histogram L.H of an item (nb buckets, inclusive prefix done) and
// its ks [j0, j1) with item-relative ranks: turn every run of ks that share a
// bucket into an output (single-key bucket) or a child item.  owner(b) gives
// the bucket's key interval.  Returns child items in `kids` (global), count
// in *nkids; children are laid out contiguously from `region`.
template <class T, class IntervalF>
__device__ __forceinline__ void batch_plan(BatchLds<typename KeyTraits<T>::K>& L, uint32_t nb, uint32_t j0, uint32_t j1,
                           uint32_t region, uint32_t buf, IntervalF&& interval, T* out,
                           BatchItem<typename KeyTraits<T>::K>* kids, uint32_t* nkids, uint16_t* slot,
                           uint32_t* cursor, uint32_t* run_bucket, uint32_t* scan_s) {
  using K = typename KeyTraits<T>::K;
  // bucket of each k (ks are sorted by rank, so buckets are non-decreasing)
  for (uint32_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    uint32_t g = L.Kr[j];
    uint32_t lo = 0, hi = nb - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (L.H[mid] >= g) hi = mid; else lo = mid + 1;
    }
    run_bucket[j - j0] = lo;
  }
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) slot[b] = (uint16_t)kNone16;
  __syncthreads();
  // runs: j starts a run when its bucket differs from j-1's
  uint32_t n = j1 - j0;
  uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  uint32_t q0 = threadIdx.x * per, q1 = q0 + per < n ? q0 + per : n;
  uint32_t nrun = 0, clen = 0;
  for (uint32_t q = q0; q < q1; ++q) {
    if (q == 0 || run_bucket[q] != run_bucket[q - 1]) {
      uint32_t b = run_bucket[q];
      uint32_t c = L.H[b] - (b ? L.H[b - 1] : 0u);
      K klo, khi;
      interval(b, klo, khi);
      if (klo != khi) {
        nrun++;
        clen += c;
      }
    }
  }
  uint32_t totr, totc;
  uint32_t er = block_excl_scan_u32(nrun, scan_s, &totr);
  uint32_t ec = block_excl_scan_u32(clen, scan_s, &totc);
  for (uint32_t q = q0; q < q1; ++q) {
    if (q == 0 || run_bucket[q] != run_bucket[q - 1]) {
      uint32_t b = run_bucket[q];
      uint32_t below = b ? L.H[b - 1] : 0u;
      uint32_t c = L.H[b] - below;
      uint32_t e = q + 1;
      while (e < n && run_bucket[e] == b) ++e;
      K klo, khi;
      interval(b, klo, khi);
      if (klo == khi) {
        for (uint32_t j = j0 + q; j < j0 + e; ++j) out[L.Ko[j]] = KeyTraits<T>::dec(klo);
      } else {
        slot[b] = (uint16_t)er;
        cursor[er] = region + ec;
        BatchItem<K> it;
        it.klo = klo;
        it.khi = khi;
        it.region = region + ec;
        it.len = c;
        it.j0 = j0 + q;
        it.j1 = j0 + e;
        it.buf = buf;
        it.pad = 0;
        kids[er] = it;
        er++;
        ec += c;
      }
      for (uint32_t j = j0 + q; j < j0 + e; ++j) L.Kr[j] -= below;
    }
  }
  if (threadIdx.x == 0) *nkids = totr;
  __threadfence_block();
  __syncthreads();
}

template <class T>
__global__ void __launch_bounds__(1024) k_batch(BatchArgs a) {
  using K = typename KeyTraits<T>::K;
  using KT = KeyTraits<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_prob, s_n, s_flag, s_nkids, s_sp;
  __shared__ uint32_t hist256[256];
  __shared__ uint32_t bcast[2];
  __shared__ uint32_t scan_s[32];
  __shared__ uint32_t s_wbc[16][2];  // per-wave digit broadcast of wave_radix_select
  BatchLds<K> L;
  L.carve(smem);
  BatchItem<K>* stack = reinterpret_cast<BatchItem<K>*>(a.stacks) + (size_t)blockIdx.x * a.stack_cap;
  const T* values = reinterpret_cast<const T*>(a.values);
  T* out = reinterpret_cast<T*>(a.out);
  K* scr[2] = {reinterpret_cast<K*>(a.scratchA), reinterpret_cast<K*>(a.scratchB)};
  constexpr uint32_t FIN = batch_fin<K>();
  // the slot table (bucket -> child item) sits in the finish buffer: after
  // the cell table in the first pass, at its start in later passes
  static_assert(LutGeom<K>::bytes() + kBatchNB * 2 <= (size_t)batch_fin<K>() * sizeof(K), "slot table");

  for (;;) {
    if (threadIdx.x == 0) s_prob = atomicAdd(&a.queue[0], 1u);
    __syncthreads();
    const uint32_t qi = s_prob;
    __syncthreads();
    if (qi >= a.nprob) break;
    const uint32_t p = a.order[qi];
    if (threadIdx.x == 0) a.status[p] = GGMS_OK;
    const uint64_t base = a.off[p];
    const uint32_t n = (uint32_t)(a.off[p + 1] - base);
    const uint32_t kb = a.koff[p], nk = a.koff[p + 1] - kb;
    if (nk == 0) continue;
    if (nk > kBatchMaxKs || n == 0) {
      if (threadIdx.x == 0) a.status[p] = GGMS_EINVAL;
      continue;
    }
    // ks: rank (1-based, smallest first) paired with output index, sorted by rank
    uint64_t* kp = reinterpret_cast<uint64_t*>(L.A);
    uint32_t M = pow2ceil_u32(nk);
    if (threadIdx.x == 0) s_flag = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < M; j += blockDim.x) {
      uint64_t v = ~0ull;
      if (j < nk) {
        uint32_t k = a.ks[kb + j];
        if (k < 1 || k > n) {
          atomicOr(&s_flag, 1u);
          k = 1;
        }
        v = ((uint64_t)(n - k + 1) << 32) | j;
      }
      kp[j] = v;
    }
    __syncthreads();
    if (s_flag) {
      if (threadIdx.x == 0) a.status[p] = GGMS_EINVAL;
      continue;
    }
    bitonic_sort_lds<uint64_t, true, false>(kp, M);
    for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) {
      L.Kr[j] = (uint32_t)(kp[j] >> 32);
      L.Ko[j] = kb + (uint32_t)(kp[j] & 0xFFFFFFFFu);
    }
    __syncthreads();
    const T* x = values + base;
    K* sA = scr[0] + base;
    K* sB = scr[1] + base;

    if (n <= FIN) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) L.A[i] = KT::enc(x[i]);
      __syncthreads();
      bool nan = false;
      if (KT::kFloat)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) nan |= (L.A[i] < KT::kNegInf) | (L.A[i] > KT::kPosInf);
      if (nan) atomicOr(&s_flag, 1u);
      __syncthreads();
      if (s_flag) {
        if (threadIdx.x == 0) a.status[p] = GGMS_ENAN;
        continue;
      }
      batch_finish<T>(L, n, (K)0, KeyMax<K>::v, 0, nk, out, hist256, bcast);
      continue;
    }

    // ---- first pass: sample, pivots, ranges, cell table -------------------
    const ggms_batch_class cf = a.cfg[class_of(n)];
    const uint32_t S = cf.sample_size;
    const uint32_t P = cf.pieces;
    const uint32_t NB = cf.num_buckets;
    const uint32_t Sp = pow2ceil_u32(S);
    K* samp = L.A;
    const K lowest = KT::kFloat ? KT::kNegInf : (K)0;
    const K highest = KT::kFloat ? KT::kPosInf : KeyMax<K>::v;
    for (uint32_t i0 = threadIdx.x; i0 < Sp; i0 += 2 * blockDim.x) {  // S <= 2048: two loads in flight
      uint32_t pos[2];
      T v[2];
#pragma unroll
      for (uint32_t u = 0; u < 2; ++u) {
        uint32_t i = i0 + u * blockDim.x;
        pos[u] = i < S ? (uint32_t)(((uint64_t)philox_draw(a.seed ^ ((uint64_t)p << 20), kStreamSample, i).v[0] * n) >> 32)
                       : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 2; ++u) v[u] = x[pos[u]];
#pragma unroll
      for (uint32_t u = 0; u < 2; ++u) {
        uint32_t i = i0 + u * blockDim.x;
        if (i < Sp) {
          K key = KeyMax<K>::v;
          if (i < S) {
            key = KT::enc(v[u]);
            key = key < lowest ? lowest : key;
            key = key > highest ? highest : key;
          }
          samp[i] = key;
        }
      }
    }
    __syncthreads();
    bitonic_sort_lds<K, true, false>(samp, Sp);
    if (threadIdx.x == 0) {
      // equal-quantile pivots, NaN / edge / one-key ranges as in the large-problem splitter
      uint32_t R = 0;
      auto emit = [&](K lo, uint32_t rq) {
        L.starts[R] = lo;
        L.req[R] = rq;
        R++;
      };
      uint32_t share = NB / P > 0 ? NB / P : 1;
      if (KT::kFloat) emit((K)0, 1);
      K first = samp[0];
      if (first > lowest) emit(lowest, 1);
      K prev = first;
      uint32_t run = 1;
      bool open = false;
      for (uint32_t j = 1; j <= P; ++j) {
        uint32_t idx = (uint32_t)(((uint64_t)j * (S - 1)) / P);
        K v = samp[idx];
        if (v == prev) {
          run++;
          continue;
        }
        if (run > 1) {
          emit(prev, 1);  // heavy value: its own one-key range
          if ((K)(v - prev) >= 2) emit((K)(prev + 1), share);
        } else {
          emit(prev, share);
        }
        open = true;
        prev = v;
        run = 1;
      }
      (void)open;
      emit(prev, 1);
      if (prev < highest) emit((K)(prev + 1), 1);
      if (KT::kFloat) emit((K)(highest + 1), 1);
      s_n = R;
    }
    __syncthreads();
    const uint32_t R = s_n;
    Pass0Map<K> mp;
    mp.set_params(reinterpret_cast<unsigned char*>(L.prm), R);
    {
      uint32_t sum = 0;
      uint32_t per = (R + blockDim.x - 1) / blockDim.x;
      uint32_t r0 = threadIdx.x * per, r1 = r0 + per < R ? r0 + per : R;
      for (uint32_t r = r0; r < r1; ++r) {
        Range<K> rr;
        rr.lo = L.starts[r];
        rr.hi = r + 1 < R ? (K)(L.starts[r + 1] - 1) : KeyMax<K>::v;
        make_map<K>(rr, L.req[r]);
        L.req[r] = rr.nb;
        sum += rr.nb;
      }
      uint32_t tot;
      uint32_t o = block_excl_scan_u32(sum, scan_s, &tot);
      for (uint32_t r = r0; r < r1; ++r) {
        Range<K> rr;
        rr.lo = L.starts[r];
        rr.hi = r + 1 < R ? (K)(L.starts[r + 1] - 1) : KeyMax<K>::v;
        make_map<K>(rr, L.req[r]);
        rr.base = o;
        L.rbase[r] = o;
        mp.store_param(r, rr);
        o += rr.nb;
      }
      if (threadIdx.x == 0) s_n = tot;
    }
    __syncthreads();
    const uint32_t B0 = s_n;
    build_lut<K>(L.starts, R, L.A);  // cell table in LDS (sample no longer needed)
    __syncthreads();
    mp.lut = L.A;
    for (uint32_t i = threadIdx.x; i < B0; i += blockDim.x) L.H[i] = 0;
    __syncthreads();
    constexpr int NV = Vec16<T>::N;
    stream_tile_vec_any<T, false, true>(
        x, 0, n,
        [&](const T (&v)[NV]) {
          K k[NV];
#pragma unroll
          for (int j = 0; j < NV; ++j) k[j] = KT::enc(v[j]);
          uint32_t b[NV];
          mp.template buckets<NV>(k, b);
          lds_count_n<NV>(L.H, b);
        },
        [&](T v) { lds_count(L.H, mp.bucket(KT::enc(v))); });
    __syncthreads();
    if (KT::kFloat && (L.H[0] | L.H[B0 - 1])) {
      if (threadIdx.x == 0) a.status[p] = GGMS_ENAN;
      __syncthreads();
      continue;
    }
    block_incl_scan_lds(L.H, B0, scan_s);
    // plan: runs of ks per bucket; tables for the compaction live after the cell table
    uint16_t* slot = reinterpret_cast<uint16_t*>(reinterpret_cast<unsigned char*>(L.A) + LutGeom<K>::bytes());
    auto interval0 = [&](uint32_t b, K& klo, K& khi) {
      uint32_t lo = 0, hi = R - 1;
      while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (L.rbase[mid] <= b) lo = mid; else hi = mid - 1;
      }
      Range<K> rr;
      rr.lo = L.starts[lo];
      rr.hi = lo + 1 < R ? (K)(L.starts[lo + 1] - 1) : KeyMax<K>::v;
      make_map<K>(rr, L.req[lo]);
      bucket_interval<K>(rr, b - L.rbase[lo], klo, khi);
    };
    batch_plan<T>(L, B0, 0, nk, 0, 0, interval0, out, stack, &s_nkids, slot, L.cursor, L.runb, scan_s);
    uint32_t nkids = s_nkids;
    if (nkids) {
      stream_tile_vec_any<T, false, true>(
          x, 0, n,
          [&](const T (&v)[NV]) {
            K k[NV];
#pragma unroll
            for (int j = 0; j < NV; ++j) k[j] = KT::enc(v[j]);
            uint32_t b[NV];
            mp.template buckets<NV>(k, b);
            uint32_t s[NV];
            bool any = false;
#pragma unroll
            for (int j = 0; j < NV; ++j) {
              s[j] = slot[b[j]];
              any |= s[j] != kNone16;
            }
            // survivors are a few percent of the keys and rarely share a child
            // within a wave: one LDS atomic per survivor, nothing for a wave
            // without any
            if (wave_ballot(any)) {
#pragma unroll
              for (int j = 0; j < NV; ++j)
                if (s[j] != kNone16) sA[atomicAdd(&L.cursor[s[j]], 1u)] = k[j];
            }
          },
          [&](T v) {
            K key = KT::enc(v);
            uint32_t s = slot[mp.bucket(key)];
            bool act = s != kNone16;
            uint32_t pos = lds_claim(L.cursor, s, act);
            if (act) sA[pos] = key;
          });
      __threadfence_block();
      __syncthreads();
    }
    // ---- first-pass children.  They sit in sA in bucket order, so their key
    // intervals are disjoint and increasing: a run of consecutive small
    // children is loaded as one batch; each k is then either radix-selected
    // inside its child by one wave (every child of the run holds at most
    // kWaveSelectMaxKs ks: the waves work on different children with no
    // workgroup barrier) or read off one sort of the whole run at (child
    // offset in the batch + rank in child).  Big children move to the front
    // of the stack for re-bucketing.
    {
      // child tables in the histogram's space (nkids <= kBatchMaxKs: 16 bytes
      // per child always fit), key intervals behind them when they fit too;
      // the wave histograms take the compaction cursors' and the k-bucket
      // table's space (both free until the work stack)
      uint32_t* c_len = L.H;
      uint32_t* c_j0 = c_len + nkids;
      uint32_t* c_j1 = c_j0 + nkids;
      uint32_t* c_reg = c_j1 + nkids;
      K* c_klo = reinterpret_cast<K*>(c_reg + nkids);
      K* c_khi = c_klo + nkids;
      const bool kfit = (size_t)nkids * (16 + 2 * sizeof(K)) <= (size_t)kBatchNB * 4;
      uint32_t* whist = L.cursor;  // 256 bins per wave
      static_assert(16 * kBatchMaxKs <= kBatchNB * 4, "child tables");
      static_assert(16 * 256 <= 2 * kBatchMaxKs, "wave histograms: cursor + runb");
      for (uint32_t c = threadIdx.x; c < nkids; c += blockDim.x) {
        const BatchItem<K> it = stack[c];
        c_len[c] = it.len;
        c_j0[c] = it.j0;
        c_j1[c] = it.j1;
        c_reg[c] = it.region;
        if (kfit) {
          c_klo[c] = it.klo;
          c_khi[c] = it.khi;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t nbig = 0;
        for (uint32_t c = 0; c < nkids; ++c)
          if (c_len[c] > FIN) stack[nbig++] = stack[c];
        s_sp = nbig;
      }
      __threadfence_block();
      uint32_t c = 0;
      while (c < nkids) {  // every thread walks the same LDS table: uniform
        if (c_len[c] > FIN) {
          ++c;
          continue;
        }
        uint32_t e = c, tot = 0;
        bool few = kfit;
        while (e < nkids && c_len[e] <= FIN && tot + c_len[e] <= FIN) {
          few &= c_j1[e] - c_j0[e] <= kWaveSelectMaxKs;
          tot += c_len[e++];
        }
        const K* src = sA + c_reg[c];
        for (uint32_t i = threadIdx.x; i < tot; i += blockDim.x) L.A[i] = src[i];
        __syncthreads();
        if (e - c == 1) {
          batch_finish<T>(L, tot, (K)0, KeyMax<K>::v, c_j0[c], c_j1[c], out, hist256, bcast);
        } else if (few) {
          const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
          for (uint32_t q = c + w; q < e; q += nw) {
            const K* cs = L.A + (c_reg[q] - c_reg[c]);
            for (uint32_t j = c_j0[q]; j < c_j1[q]; ++j) {
              const K key = wave_radix_select<K>(cs, c_len[q], c_klo[q], c_khi[q], L.Kr[j], whist + w * 256, s_wbc[w]);
              if (lane_id() == 0) out[L.Ko[j]] = KeyTraits<T>::dec(key);
            }
          }
          __syncthreads();
        } else {
          const uint32_t M = pow2ceil_u32(tot);
          for (uint32_t i = tot + threadIdx.x; i < M; i += blockDim.x) L.A[i] = KeyMax<K>::v;
          __syncthreads();
          bitonic_sort_lds<K, true, false>(L.A, M);
          for (uint32_t q = c; q < e; ++q) {
            const uint32_t off = c_reg[q] - c_reg[c];
            for (uint32_t j = c_j0[q] + threadIdx.x; j < c_j1[q]; j += blockDim.x)
              out[L.Ko[j]] = KeyTraits<T>::dec(L.A[off + L.Kr[j] - 1]);
          }
          __syncthreads();
        }
        c = e;
      }
    }
    __syncthreads();
    // ---- work stack: finish small items, re-bucket large ones -----------------
    for (;;) {
      __syncthreads();
      const uint32_t sp = s_sp;
      if (sp == 0) break;
      const BatchItem<K> it = stack[sp - 1];
      __syncthreads();  // every thread has read s_sp and the item
      uint32_t next_sp = sp - 1;
      const K* src = (it.buf ? sB : sA) + it.region;
      if (it.len <= FIN) {
        for (uint32_t i = threadIdx.x; i < it.len; i += blockDim.x) L.A[i] = src[i];
        __syncthreads();
        batch_finish<T>(L, it.len, it.klo, it.khi, it.j0, it.j1, out, hist256, bcast);
      } else {
        // one linear range over the item's key interval; its children replace
        // it on the stack (at most one child per k, so the stack never holds
        // more items than the problem has ks)
        Range<K> rr;
        rr.lo = it.klo;
        rr.hi = it.khi;
        rr.base = 0;
        make_map<K>(rr, cf.rebucket);
        const uint32_t nb = rr.nb;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) L.H[i] = 0;
        __syncthreads();
        constexpr int NK = Vec16<K>::N;
        stream_tile_vec_any<K, false, true>(
            src, 0, it.len,
            [&](const K (&k)[NK]) {
              uint32_t b[NK];
#pragma unroll
              for (int j = 0; j < NK; ++j) b[j] = map_off<K>(k[j], rr.lo, rr.sh, rr.mult);
              lds_count_n<NK>(L.H, b);
            },
            [&](K key) { lds_count(L.H, map_off<K>(key, rr.lo, rr.sh, rr.mult)); });
        __syncthreads();
        block_incl_scan_lds(L.H, nb, scan_s);
        auto interval1 = [&](uint32_t b, K& klo, K& khi) { bucket_interval<K>(rr, b, klo, khi); };
        uint16_t* slot1 = reinterpret_cast<uint16_t*>(L.A);
        batch_plan<T>(L, nb, it.j0, it.j1, it.region, it.buf ^ 1u, interval1, out, stack + (sp - 1), &s_nkids, slot1,
                      L.cursor, L.runb, scan_s);
        const uint32_t nk2 = s_nkids;
        K* dst = (it.buf ? sA : sB);
        if (nk2) {
          auto claim = [&](K key) {
            uint32_t s = slot1[map_off<K>(key, rr.lo, rr.sh, rr.mult)];
            bool act = s != kNone16;
            uint32_t pos = lds_claim(L.cursor, s, act);
            if (act) dst[pos] = key;
          };
          stream_tile_vec_any<K, false, true>(
              src, 0, it.len,
              [&](const K (&k)[NK]) {
                uint32_t s[NK];
                bool any = false;
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                  s[j] = slot1[map_off<K>(k[j], rr.lo, rr.sh, rr.mult)];
                  any |= s[j] != kNone16;
                }
                if (wave_ballot(any)) {  // as in the first pass: one LDS atomic per survivor
#pragma unroll
                  for (int j = 0; j < NK; ++j)
                    if (s[j] != kNone16) dst[atomicAdd(&L.cursor[s[j]], 1u)] = k[j];
                }
              },
              claim);
          __threadfence_block();
        }
        next_sp += nk2;
      }
      if (threadIdx.x == 0) s_sp = next_sp;
    }
  }
  // last workgroup out resets the queue for the next replay
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&a.queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&a.queue[0], 0u);
      atomicExch(&a.queue[1], 0u);
    }
  }
}

}  // namespace ggms

using namespace ggms;

namespace {
ggms_batch_class default_class(uint32_t c) {
  ggms_batch_class k{};
  k.num_buckets = c >= 20 ? 8192u : (c >= 17 ? 4096u : 2048u);
  k.sample_size = c >= 20 ? 2048u : 1024u;
  k.pieces = c >= 20 ? 64u : 32u;
  k.rebucket = kBatchNB;
  return k;
}
}  // namespace

struct ggms_batch_plan {
  int device = 0;
  ggms_dtype t = GGMS_F32;
  uint32_t nprob = 0;
  uint32_t grid = 0;
  uint32_t threads = 1024;
  size_t lds = 0;
  void* dev = nullptr;  // one allocation: offsets, koffsets, order, queue, status, scratch A/B, stacks
  BatchArgs args{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint32_t> h_status;
};

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

template <class T>
static ggms_status batch_launch(ggms_batch_plan* bp, hipStream_t st) {
  hipLaunchKernelGGL(k_batch<T>, dim3(bp->grid), dim3(bp->threads), bp->lds, st, bp->args);
  return hipGetLastError() == hipSuccess ? GGMS_OK : GGMS_EDEVICE;
}

static ggms_status launch_plan(ggms_batch_plan* bp, hipStream_t st) {
  switch (bp->t) {
    case GGMS_F32: return batch_launch<float>(bp, st);
    case GGMS_F64: return batch_launch<double>(bp, st);
    case GGMS_U32: return batch_launch<uint32_t>(bp, st);
    case GGMS_I32: return batch_launch<int32_t>(bp, st);
  }
  return GGMS_EINVAL;
}

extern "C" {

void ggms_batch_default_classes(ggms_batch_class* classes) {
  if (!classes) return;
  for (uint32_t c = 0; c < GGMS_BATCH_CLASSES; ++c) classes[c] = default_class(c);
}


ggms_status ggms_batch_plan_create(int device, ggms_dtype t, const void* d_values, uint32_t num_problems,
                                   const uint64_t* h_offsets, const uint32_t* d_ks, const uint32_t* h_koffsets,
                                   void* d_out, const ggms_batch_opts* opts, ggms_batch_plan** out_plan) {
  if (!out_plan || !h_offsets || !h_koffsets || num_problems == 0) return GGMS_EINVAL;
  if ((unsigned)t > 3u) return GGMS_EINVAL;
  *out_plan = nullptr;
  for (uint32_t p = 0; p < num_problems; ++p) {
    if (h_offsets[p + 1] < h_offsets[p] || h_offsets[p + 1] - h_offsets[p] > kMaxN) return GGMS_EINVAL;
    if (h_koffsets[p + 1] < h_koffsets[p] || h_koffsets[p + 1] - h_koffsets[p] > kBatchMaxKs) return GGMS_EINVAL;
  }
  const uint64_t total = h_offsets[num_problems];
  if (total > 0 && !d_values) return GGMS_EINVAL;
  if (h_koffsets[num_problems] > 0 && (!d_ks || !d_out)) return GGMS_EINVAL;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return GGMS_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return GGMS_EDEVICE;
  // k ranks are checked here, once per plan: every k of problem p in [1, size_p]
  if (h_koffsets[num_problems] > 0) {
    std::vector<uint32_t> hk(h_koffsets[num_problems]);
    if (hipMemcpy(hk.data(), d_ks, hk.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipSetDevice(cur);
      return GGMS_EDEVICE;
    }
    for (uint32_t p = 0; p < num_problems; ++p) {
      const uint64_t n = h_offsets[p + 1] - h_offsets[p];
      for (uint32_t j = h_koffsets[p]; j < h_koffsets[p + 1]; ++j)
        if (hk[j] < 1 || hk[j] > n) {
          (void)hipSetDevice(cur);
          return GGMS_EINVAL;
        }
    }
  }
  ggms_batch_plan* bp = new (std::nothrow) ggms_batch_plan();
  if (!bp) return GGMS_ENOMEM;
  bp->device = device;
  bp->t = t;
  bp->nprob = num_problems;
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  uint32_t wgs = opts && opts->workgroups ? opts->workgroups : (uint32_t)(cus > 0 ? cus : 256);
  bp->grid = std::min<uint32_t>(wgs, num_problems);
  const uint32_t bt = opts && opts->block_threads ? opts->block_threads : 1024u;
  if (bt != 256 && bt != 512 && bt != 1024) {
    delete bp;
    (void)hipSetDevice(cur);
    return GGMS_EINVAL;
  }
  bp->threads = bt;
  const size_t ksz = (t == GGMS_F64) ? 8 : 4;
  bp->lds = ksz == 4 ? BatchLds<uint32_t>::bytes() : BatchLds<uint64_t>::bytes();
  const uint32_t stack_cap = kBatchMaxKs + 8;
  const size_t item = ksz == 4 ? sizeof(BatchItem<uint32_t>) : sizeof(BatchItem<uint64_t>);
  size_t o_off = 0;
  size_t o_koff = o_off + al256((size_t)(num_problems + 1) * 8);
  size_t o_order = o_koff + al256((size_t)(num_problems + 1) * 4);
  size_t o_queue = o_order + al256((size_t)num_problems * 4);
  size_t o_status = o_queue + 256;
  size_t o_sa = o_status + al256((size_t)num_problems * 4);
  size_t o_sb = o_sa + al256((size_t)total * ksz);
  size_t o_stack = o_sb + al256((size_t)total * ksz);
  size_t o_cfg = o_stack + al256((size_t)bp->grid * stack_cap * item);
  size_t bytes = o_cfg + al256(sizeof(ggms_batch_class) * GGMS_BATCH_CLASSES);
  if (hipMalloc(&bp->dev, bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete bp;
    (void)hipSetDevice(cur);
    return GGMS_ENOMEM;
  }
  char* D = (char*)bp->dev;
  // queue order: largest problems first
  std::vector<uint32_t> order(num_problems);
  for (uint32_t p = 0; p < num_problems; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return h_offsets[a + 1] - h_offsets[a] > h_offsets[b + 1] - h_offsets[b];
  });
  ggms_status s = GGMS_OK;
  if (hipMemcpy(D + o_off, h_offsets, (size_t)(num_problems + 1) * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(D + o_koff, h_koffsets, (size_t)(num_problems + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(D + o_order, order.data(), (size_t)num_problems * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(D + o_queue, 0, 256) != hipSuccess ||
      hipMemset(D + o_status, 0, (size_t)num_problems * 4) != hipSuccess)
    s = GGMS_EDEVICE;
  BatchArgs& a = bp->args;
  a.values = d_values;
  a.off = (const uint64_t*)(D + o_off);
  a.ks = d_ks;
  a.koff = (const uint32_t*)(D + o_koff);
  a.out = d_out;
  a.order = (const uint32_t*)(D + o_order);
  a.queue = (uint32_t*)(D + o_queue);
  a.status = (uint32_t*)(D + o_status);
  a.scratchA = D + o_sa;
  a.scratchB = D + o_sb;
  a.stacks = D + o_stack;
  a.nprob = num_problems;
  a.stack_cap = stack_cap;
  a.seed = opts ? opts->seed : 0x5EEDull;
  ggms_batch_class cfg[GGMS_BATCH_CLASSES];
  for (uint32_t c = 0; c < GGMS_BATCH_CLASSES; ++c) {
    ggms_batch_class k = (opts && opts->classes[c].num_buckets) ? opts->classes[c] : default_class(c);
    k.num_buckets = std::max<uint32_t>(16, std::min<uint32_t>(k.num_buckets, kBatchNB - 2 * kBatchMaxPieces - 16));
    k.pieces = std::max<uint32_t>(2, std::min<uint32_t>(k.pieces, kBatchMaxPieces));
    k.sample_size = std::max<uint32_t>(k.pieces + 1, std::min<uint32_t>(k.sample_size, 2048u));
    k.rebucket = k.rebucket ? std::max<uint32_t>(2, std::min<uint32_t>(k.rebucket, kBatchNB)) : kBatchNB;
    cfg[c] = k;
  }
  a.cfg = (const ggms_batch_class*)(D + o_cfg);
  if (s == GGMS_OK && hipMemcpy(D + o_cfg, cfg, sizeof(cfg), hipMemcpyHostToDevice) != hipSuccess) s = GGMS_EDEVICE;
  bp->h_status.assign(num_problems, 0);
  if (s == GGMS_OK && bp->lds > 65536) {
    const void* fn = nullptr;
    switch (t) {
      case GGMS_F32: fn = (const void*)k_batch<float>; break;
      case GGMS_F64: fn = (const void*)k_batch<double>; break;
      case GGMS_U32: fn = (const void*)k_batch<uint32_t>; break;
      default: fn = (const void*)k_batch<int32_t>; break;
    }
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bp->lds);
  }
  if (s == GGMS_OK && opts && opts->use_graph) {
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) s = GGMS_EDEVICE;
    else {
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) s = GGMS_EDEVICE;
      if (s == GGMS_OK) s = launch_plan(bp, cs);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(cs, &g) != hipSuccess) s = GGMS_EDEVICE;
      bp->graph = g;
      if (s == GGMS_OK && hipGraphInstantiate(&bp->exec, g, nullptr, nullptr, 0) != hipSuccess) s = GGMS_EDEVICE;
      (void)hipStreamDestroy(cs);
    }
  }
  (void)hipSetDevice(cur);
  if (s != GGMS_OK) {
    ggms_batch_plan_destroy(bp);
    return s;
  }
  *out_plan = bp;
  return GGMS_OK;
}

ggms_status ggms_batch_plan_launch(ggms_batch_plan* bp, void* stream) {
  if (!bp) return GGMS_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (bp->exec) return hipGraphLaunch(bp->exec, st) == hipSuccess ? GGMS_OK : GGMS_EDEVICE;
  return launch_plan(bp, st);
}

ggms_status ggms_batch_plan_status(ggms_batch_plan* bp, uint32_t* h_status) {
  if (!bp || !h_status) return GGMS_EINVAL;
  if (hipMemcpy(h_status, bp->args.status, (size_t)bp->nprob * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return GGMS_EDEVICE;
  for (uint32_t p = 0; p < bp->nprob; ++p)
    if (h_status[p]) return (ggms_status)h_status[p];
  return GGMS_OK;
}

void ggms_batch_plan_destroy(ggms_batch_plan* bp) {
  if (!bp) return;
  if (bp->exec) (void)hipGraphExecDestroy(bp->exec);
  if (bp->graph) (void)hipGraphDestroy(bp->graph);
  if (bp->dev) (void)hipFree(bp->dev);
  delete bp;
}

ggms_status ggms_select_batched(int device, ggms_dtype t, const void* d_values, uint32_t num_problems,
                                const uint64_t* h_offsets, const uint32_t* d_ks, const uint32_t* h_koffsets,
                                void* d_out, const ggms_batch_opts* opts) {
  ggms_batch_plan* bp = nullptr;
  ggms_batch_opts o;
  if (opts) o = *opts; else ggms_batch_opts_default(&o);
  o.use_graph = 0;
  ggms_status s = ggms_batch_plan_create(device, t, d_values, num_problems, h_offsets, d_ks, h_koffsets, d_out, &o, &bp);
  if (s) return s;
  s = ggms_batch_plan_launch(bp, nullptr);
  if (s == GGMS_OK && hipDeviceSynchronize() != hipSuccess) s = GGMS_EDEVICE;
  if (s == GGMS_OK) s = ggms_batch_plan_status(bp, bp->h_status.data());
  ggms_batch_plan_destroy(bp);
  return s;
}

void ggms_batch_opts_default(ggms_batch_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->seed = 0x5EEDull;
  ggms_batch_default_classes(o->classes);
}

}  // extern "C"
