// Selection kernels for gfx950 (wave64, 160 KB LDS per CU).
//
// One bucketing pass = count -> colsum -> plan -> compact -> finish:
//   k_plan0    one workgroup: sample, sort the sample in LDS, place pivots,
//              build the first pass's range table (linear-slope splitter).
//   k_count0 / k_countS
//              fused "assign bucket + count": 128-bit loads, per-workgroup LDS
//              histogram, wave-peeled atomics; writes one histogram row per
//              workgroup.  No per-element bucket ids are stored.
//   k_colsum   column totals of the rows (the pass histogram).
//   k_plan     one workgroup: inclusive scan of the histogram, findKBuckets for
//              every live k, dedupe into active buckets, classify each active
//              bucket (single key / small / big), lay out the next buffer,
//              re-rank ks, and build the next pass's ranges.
//   k_compact0 / k_compactS
//              recompute each element's bucket and copy active-bucket elements
//              to their segment, at offsets the workgroup claims with one
//              global atomic per segment (pass 0: one for the whole tile).
//   k_finish   one workgroup per small segment: bitonic sort in LDS and pick.
//
// Reference counterparts (DistributedSMOS/): assignSmartBucket_iterative
// iterativeSMOS.cu:488-571, sumCounts :580-586, findKBuckets :207-232,
// updatekVals_iterative :302-376, reindexCounts :593-605,
// copyElements_iterative :611-656, updatePivots_iterative :715-758,
// generateBucketsandSlopes_iterative :404-477, generatePivots :1011-1054.
#pragma once
#include <type_traits>

#include "ggms_device.hpp"
#include "philox.hpp"

namespace ggms {

constexpr uint32_t kNone16 = 0xFFFFu;
constexpr uint32_t kNone32 = 0xFFFFFFFFu;
// grouped (segmented) first-pass compaction: consecutive slots per group, and
// a bound on the groups (above (kMaxBuckets + edge and pivot buckets) / kGroupSlots)
constexpr uint32_t kGroupSlots = 128;
constexpr uint32_t kMaxGroups = 256;

enum KState : uint32_t { kLive = 0, kSmall = 1, kDone = 2, kSmallNew = 3 };
enum ActClass : uint32_t { kResolved = 0, kClsSmall = 1, kClsBig = 2 };
enum SplitMode : uint32_t { kModeLinear = 0, kModeLearned = 1, kModeSketch = 2, kModeStrided = 3 };

// Device-side control block of one pass.  The plan kernel of pass p fills the
// block of pass p+1; kernels read their own block, so the host only needs
// upper bounds for launch shapes.
struct PassCtl {
  uint32_t n;        // elements entering the pass (pass >= 1: big-segment elements)
  uint32_t tile;     // elements per workgroup
  uint32_t G;        // workgroups that own a tile
  uint32_t R;        // ranges of the pass
  uint32_t Rp;       // R rounded up to a power of two (pass-0 search)
  uint32_t B;        // buckets of the pass
  uint32_t U;        // active buckets found by the plan of this pass
  uint32_t U_big;    // big active buckets (ranges of the next pass)
  uint32_t U_small;  // small active buckets (finished by LDS sort)
  uint32_t U_slots;  // U_big + U_small
  uint32_t N_big;    // elements of big segments (next pass input)
  uint32_t N_small;  // elements of small segments
  uint32_t nan_count;
  uint32_t live_ks;
  uint32_t nan_lo_bucket;  // bucket holding keys below key(-inf), or kNone32
  uint32_t nan_hi_bucket;  // bucket holding keys above key(+inf), or kNone32
  uint32_t append;         // first-pass compaction: next free position of the survivor buffer
  uint32_t l2_used;        // first pass: level-2 tables of the cell table in use
  uint32_t U_wide;         // small segments selected from global memory (the last U_wide of the finish order)
  uint32_t U_wave;         // small segments finished one per wave (the U_wave before those)
  uint32_t small_max;      // longest small segment of the pass
  // first pass: window of the key-prefix cell table, [lut_base, lut_base +
  // (cells << lut_shift)); lut_shift == key bits - level-1 bits is the whole key space
  uint32_t lut_base_lo, lut_base_hi, lut_shift;
  // ranges of the keys outside that window: those below it lie in ranges
  // [0, lut_r_below], those above it in [lut_r_above, R - 1]
  uint32_t lut_r_below, lut_r_above;
  // first pass: bucket of a dominant value (a one-key range holding at least
  // a quarter of the sample), counted in registers by k_count0; else kNone32
  uint32_t hot_bucket;
  // segmented first pass: nonzero when fewer than a quarter of the keys
  // survive, so the unstaged grouped compaction runs instead of the staged one
  uint32_t grp_sparse;
  // finish order: the whole-workgroup segments of at least kFinishSplitLen
  // keys come first, in [0, U_fin_big)
  uint32_t U_fin_big;
  uint32_t pad[3];
};

constexpr uint32_t kSetupKsPerThread = 8;
struct Plan0Args {
  uint32_t mode;        // SplitMode
  uint32_t S;           // sample size (<= 8192)
  uint32_t P;           // linear: number of pivots; learned/sketch: pieces
  uint32_t B;           // bucket budget
  uint32_t E;           // buckets of each edge range (below sample min / above sample max)
  uint32_t strideS;     // strided mode: first-level strided sample size (8192)
  uint64_t seed;
  const float* mlp;     // learned mode: MLP parameters (see mlp layout in select.hip)
  // setup != 0: the kernel also does k_setup's work (nk <= kSetupKsPerThread
  // * blockDim: saves a launch)
  uint32_t setup;
  uint32_t nk;
  uint32_t ctl_words;
  const uint32_t* ks;
  uint32_t* k_rank;
  uint32_t* k_state;
  uint32_t* bad;
  PassCtl c0;
};

// ---------------------------------------------------------------------------
// workgroup utilities
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Key of rank `rank` (1-based) among a[0, len), all keys in [klo, khi], by
// ONE wave with no workgroup barrier: 8-bit digits over the bits where klo
// and khi differ, a 256-bin histogram in the wave's own LDS (hist), the digit
// found by a wave scan and handed to every lane through bc[0..1].
template <class K>
__device__ K wave_radix_select(const K* __restrict__ a, uint32_t len, K klo, K khi, uint32_t rank, uint32_t* hist,
                               uint32_t* bc) {
  constexpr uint32_t kb = sizeof(K) * 8;
  const uint32_t l = lane_id();
  const uint32_t vary = bitlen64((uint64_t)(klo ^ khi));
  K pre = vary >= kb ? (K)0 : (K)(klo & (K)(~(K)0 << vary));
  for (int shift = (int)((vary + 7) & ~7u) - 8; shift >= 0; shift -= 8) {
    const K hmask = (uint32_t)shift + 8 >= kb ? (K)0 : (K)(~(K)0 << (shift + 8));
    const K want = (K)(pre & hmask);
    for (uint32_t b = l; b < 256; b += 64) hist[b] = 0;
    wave_sync();
    auto count = [&](K key) {
      if ((K)(key & hmask) == want) atomicAdd(&hist[(uint32_t)(key >> shift) & 255u], 1u);
    };
    uint32_t i = l;
    for (; i + 192 < len; i += 256) {  // four reads in flight (a may be global memory)
      const K k0 = a[i], k1 = a[i + 64], k2 = a[i + 128], k3 = a[i + 192];
      count(k0);
      count(k1);
      count(k2);
      count(k3);
    }
    for (; i < len; i += 64) count(a[i]);
    wave_sync();
    const uint32_t c0 = hist[4 * l], c1 = hist[4 * l + 1], c2 = hist[4 * l + 2], c3 = hist[4 * l + 3];
    const uint32_t sum = c0 + c1 + c2 + c3;
    uint32_t cum = wave_incl_scan_u32(sum) - sum;
    const uint32_t c[4] = {c0, c1, c2, c3};
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
      if (cum < rank && cum + c[b] >= rank) {
        bc[0] = 4 * l + b;
        bc[1] = rank - cum;
      }
      cum += c[b];
    }
    wave_sync();
    pre = (K)(pre | ((K)bc[0] << shift));
    rank = bc[1];
    wave_sync();  // every lane has read bc and hist before the next round writes them
  }
  return pre;
}

// Exclusive scan of one value per thread over the whole workgroup.
// scratch: >= 16 u32 of LDS.  Returns exclusive prefix, total in *total.
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* scratch, uint32_t* total) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  uint32_t inc = wave_incl_scan_u32(v);
  __syncthreads();
  if (lane == 63) scratch[wv] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t s = scratch[w];
    if (w < wv) off += s;
    tot += s;
  }
  *total = tot;
  return off + inc - v;
}

// Four exclusive scans at once (one pair of barriers instead of four).
// scratch: >= 64 u32 of LDS.
__device__ __forceinline__ void block_excl_scan_u32x4(const uint32_t (&v)[4], uint32_t (&off)[4], uint32_t (&tot)[4],
                                                      uint32_t* scratch) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  uint32_t inc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) inc[i] = wave_incl_scan_u32(v[i]);
  __syncthreads();
  if (lane == 63) {
#pragma unroll
    for (int i = 0; i < 4; ++i) scratch[wv * 4 + i] = inc[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    off[i] = 0;
    tot[i] = 0;
  }
  for (uint32_t w = 0; w < nw; ++w) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t s = scratch[w * 4 + i];
      if (w < wv) off[i] += s;
      tot[i] += s;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) off[i] += inc[i] - v[i];
}

// In-place inclusive scan of an LDS array of length L.  scratch >= 16 u32.
__device__ void block_incl_scan_lds(uint32_t* a, uint32_t L, uint32_t* scratch) {
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  uint32_t per = (L + nt - 1) / nt;
  uint32_t b0 = t * per, b1 = b0 + per < L ? b0 + per : L;
  uint32_t s = 0;
  for (uint32_t i = b0; i < b1; ++i) s += a[i];
  uint32_t tot;
  uint32_t off = block_excl_scan_u32(s, scratch, &tot);
  for (uint32_t i = b0; i < b1; ++i) {
    off += a[i];
    a[i] = off;
  }
  __syncthreads();
}

// Bitonic sort (ascending) of M keys in LDS, M a power of two, by the whole
// workgroup.  When M = E * blockDim with E in {4, 8, 16}, thread t owns keys
// [t*E, t*E + E): every compare-exchange distance below E runs in registers
// (one load, all short-distance stages, one store), so only distances >= E
// cost a barrier each.
//
// SW: between the first and the last register phase the array is kept
// swizzled: inside every 128-byte block the eight 16-byte units are permuted
// by XOR with the block number (keys keep their order inside a unit).  A
// lane's E keys are E / (16 / sizeof(K)) whole units, read and written as
// 16-byte accesses; the eight lanes served together by one 16-byte LDS access
// then land on eight different units (plain, the 64-byte chunks of lanes t and
// t + 2 share banks: 4-way conflicts for 16 words per lane), and the stages
// whose lanes read consecutive keys touch each block's units once.  No extra
// space: the permutation stays inside each block.
template <class K>
__device__ __forceinline__ uint32_t bitonic_swz(uint32_t i) {
  constexpr uint32_t KU = 16 / sizeof(K), KB = 128 / sizeof(K);
  return (i & ~(KB - 1)) | ((((i / KU) ^ (i / KB)) & 7u) * KU) | (i & (KU - 1));
}

template <class K, bool SW = false>
__device__ __forceinline__ void bitonic_stage_lds(K* s, uint32_t M, uint32_t size, uint32_t stride) {
  for (uint32_t i = threadIdx.x; i < (M >> 1); i += blockDim.x) {
    uint32_t lo = 2 * i - (i & (stride - 1));
    uint32_t hi = lo + stride;
    bool asc = (lo & size) == 0;
    const uint32_t plo = SW ? bitonic_swz<K>(lo) : lo, phi = SW ? bitonic_swz<K>(hi) : hi;
    K a = s[plo], b = s[phi];
    if ((a > b) == asc) {
      s[plo] = b;
      s[phi] = a;
    }
  }
}

template <class K, int E>
__device__ __forceinline__ void bitonic_reg_merge(K (&r)[E], uint32_t base, uint32_t size, uint32_t stride_max) {
#pragma unroll
  for (int st = E / 2; st > 0; st >>= 1) {
    if ((uint32_t)st <= stride_max) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        if ((i & st) == 0) {
          bool asc = ((base + (uint32_t)i) & size) == 0;
          K a = r[i], b = r[i + st];
          bool sw = (a > b) == asc;
          r[i] = sw ? b : a;
          r[i + st] = sw ? a : b;
        }
      }
    }
  }
}

// A lane's E keys in the swizzled layout: whole 16-byte units where E keys
// fill at least one (the lane's first key then starts a unit), else key by key.
template <class K, int E>
__device__ __forceinline__ void swz_load(const K* s, uint32_t base, K (&r)[E]) {
  constexpr int KU = 16 / sizeof(K);
  if constexpr (E >= KU) {
#pragma unroll
    for (int j = 0; j < E / KU; ++j) {
      const uint4 v = *reinterpret_cast<const uint4*>(s + bitonic_swz<K>(base + (uint32_t)(j * KU)));
      K t[KU];
      __builtin_memcpy(t, &v, 16);
#pragma unroll
      for (int q = 0; q < KU; ++q) r[j * KU + q] = t[q];
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) r[i] = s[bitonic_swz<K>(base + (uint32_t)i)];
  }
}
template <class K, int E>
__device__ __forceinline__ void swz_store(K* s, uint32_t base, const K (&r)[E]) {
  constexpr int KU = 16 / sizeof(K);
  if constexpr (E >= KU) {
#pragma unroll
    for (int j = 0; j < E / KU; ++j) {
      K t[KU];
#pragma unroll
      for (int q = 0; q < KU; ++q) t[q] = r[j * KU + q];
      uint4 v;
      __builtin_memcpy(&v, t, 16);
      *reinterpret_cast<uint4*>(s + bitonic_swz<K>(base + (uint32_t)(j * KU))) = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) s[bitonic_swz<K>(base + (uint32_t)i)] = r[i];
  }
}

// Threads past M / E hold no keys (they only take part in the barriers and
// the LDS stages).  Input and output are in plain order.  SW needs s 16-byte
// aligned and M a multiple of 128 bytes of keys.
template <class K, int E, bool SW = false>
__device__ void bitonic_sort_lds_reg(K* s, uint32_t M) {
  const uint32_t base = threadIdx.x * E;
  const bool own = base < M;
  K r[E];
  if (own) {
#pragma unroll
    for (int i = 0; i < E; ++i) r[i] = s[base + i];
    for (uint32_t size = 2; size <= (uint32_t)E; size <<= 1) bitonic_reg_merge<K, E>(r, base, size, size >> 1);
  }
  if (SW) {
    __syncthreads();  // every plain read done before the swizzled writes
    if (own) swz_store<K, E>(s, base, r);
  } else if (own) {
#pragma unroll
    for (int i = 0; i < E; ++i) s[base + i] = r[i];
  }
  __syncthreads();
  for (uint32_t size = 2 * E; size <= M; size <<= 1) {
    for (uint32_t stride = size >> 1; stride >= (uint32_t)E; stride >>= 1) {
      bitonic_stage_lds<K, SW>(s, M, size, stride);
      __syncthreads();
    }
    if (own) {
      if (SW) {
        swz_load<K, E>(s, base, r);
      } else {
#pragma unroll
        for (int i = 0; i < E; ++i) r[i] = s[base + i];
      }
      bitonic_reg_merge<K, E>(r, base, size, E / 2);
    }
    if (SW && size == M) {  // last merge: back to plain order
      __syncthreads();
      if (own) {
#pragma unroll
        for (int i = 0; i < E; ++i) s[base + i] = r[i];
      }
    } else if (own) {
      if (SW) {
        swz_store<K, E>(s, base, r);
      } else {
#pragma unroll
        for (int i = 0; i < E; ++i) s[base + i] = r[i];
      }
    }
    __syncthreads();
  }
}

// Key exchange with lane (lane ^ j) of the same wave (j < 64).
template <class K>
__device__ __forceinline__ K shfl_xor_key(K v, int j) {
  if constexpr (sizeof(K) == 8) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, j, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), j, 64);
    return ((K)hi << 32) | (K)lo;
  } else {
    return (K)(uint32_t)__shfl_xor((int)(uint32_t)v, j, 64);
  }
}

// Bitonic sort with the keys in registers throughout: thread t holds keys
// [t*E, t*E + E) of M = E * T, T = M / E owning threads (a multiple of 64).  A
// compare-exchange at distance d pairs thread t with t ^ (d / E): d < E stays
// in registers, d / E < 64 is a lane shuffle inside the wave (no barrier),
// and only d >= 64 E goes through LDS (own keys out, partner's keys in, two
// barriers).  For M = 8192, E = 8 that is 10 barrier stages instead of 55.
template <class K, int E>
__device__ void bitonic_sort_shfl(K* s, uint32_t M) {
  const uint32_t t = threadIdx.x, base = t * E;
  const bool own = base < M;  // uniform per wave: M / E is a multiple of 64
  K r[E];
  if (own) {
#pragma unroll
    for (int i = 0; i < E; ++i) r[i] = s[base + i];
    for (uint32_t size = 2; size <= (uint32_t)E; size <<= 1) bitonic_reg_merge<K, E>(r, base, size, size >> 1);
  }
  for (uint32_t size = 2 * E; size <= M; size <<= 1) {
    const bool asc = (base & size) == 0;
    for (uint32_t stride = size >> 1; stride >= (uint32_t)E; stride >>= 1) {
      const uint32_t j = stride / E;
      const bool keep_min = ((t & j) == 0) == asc;
      if (j >= 64) {
        if (own) {
#pragma unroll
          for (int i = 0; i < E; ++i) s[base + i] = r[i];
        }
        __syncthreads();
        K p[E];
        if (own) {
#pragma unroll
          for (int i = 0; i < E; ++i) p[i] = s[(t ^ j) * E + i];
        }
        __syncthreads();  // every partner read done before the next stage writes
        if (own) {
#pragma unroll
          for (int i = 0; i < E; ++i) r[i] = keep_min ? (p[i] < r[i] ? p[i] : r[i]) : (p[i] > r[i] ? p[i] : r[i]);
        }
      } else if (own) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const K o = shfl_xor_key<K>(r[i], (int)j);
          r[i] = keep_min ? (o < r[i] ? o : r[i]) : (o > r[i] ? o : r[i]);
        }
      }
    }
    if (own) bitonic_reg_merge<K, E>(r, base, size, E / 2);
  }
  if (own) {
#pragma unroll
    for (int i = 0; i < E; ++i) s[base + i] = r[i];
  }
  __syncthreads();
}

// SHFL: the register-resident form (bitonic_sort_shfl) where it applies; it
// holds 2 E keys per thread at its LDS stages, so kernels held to 64 VGPRs
// pass false and keep the LDS-stage form.
template <class K, bool SW = false, bool SHFL = true>
__device__ void bitonic_sort_lds(K* s, uint32_t M) {
  const uint32_t e = M / blockDim.x;
  if (M == e * blockDim.x &&
      (e == 4 || e == 8 || (e == 16 && sizeof(K) <= 4) || (e == 32 && sizeof(K) == 2))) {
    if (SHFL && e == 4) bitonic_sort_shfl<K, 4>(s, M);
    else if (SHFL && e == 8) bitonic_sort_shfl<K, 8>(s, M);
    else if (SHFL && e == 16) bitonic_sort_shfl<K, 16>(s, M);
    else if (e == 4) bitonic_sort_lds_reg<K, 4, SW>(s, M);
    else if (e == 8) bitonic_sort_lds_reg<K, 8, SW>(s, M);
    else if (e == 16) bitonic_sort_lds_reg<K, 16, SW>(s, M);
    else bitonic_sort_lds_reg<K, 32, SW>(s, M);
    return;
  }
  if (e < 4 && M >= 256 && (M & (M - 1)) == 0) {  // fewer keys than 4 per thread: M / 4 threads hold 4 each
    if (SHFL) bitonic_sort_shfl<K, 4>(s, M);
    else bitonic_sort_lds_reg<K, 4, SW>(s, M);
    return;
  }
  for (uint32_t size = 2; size <= M; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      bitonic_stage_lds<K>(s, M, size, stride);
      __syncthreads();
    }
  }
}

// Half-width word of a key type: offsets from a segment's low key that fit
// one are stored in it.
template <class K> struct HalfKey;
template <> struct HalfKey<uint32_t> { using T = uint16_t; };
template <> struct HalfKey<uint64_t> { using T = uint32_t; };

// The finish holds a segment in fin_cap / 2 full-width keys of LDS (two
// workgroups per CU): as keys when it has at most fin_cap / 2 of them, or as
// half-width offsets from its low key when its interval is that narrow.  The
// rest (wide and large, rare) are selected straight from global memory.
template <class K>
__host__ __device__ __forceinline__ bool finish_fits_half(uint32_t len, K klo, K khi, uint32_t fin_cap) {
  using H = typename HalfKey<K>::T;
  return len <= fin_cap / 2 || (uint64_t)(K)(khi - klo) <= (uint64_t)(H)~(H)0;
}

// Size classes of small segments for the finish order (1024 keys per class,
// the largest first), then class 1: segments of at most kFinishWaveLen keys
// and kFinishWaveKs ks, finished one per wave, then class 0, ordered last, for
// the segments selected from global memory (too wide and too large for the
// LDS layouts, or larger than the finish capacity: the one-pass
// bucketMultiselect finishes every k-bucket whatever its size).  With 4-byte
// keys a segment of up to kFinishWaveLen4 keys and one k is one-wave too.  The
// one-wave class is used only when a pass leaves at least kFinishWaveMinSegs small segments: a wave
// sweeping its segment from L2 is slower per segment than a workgroup with
// the segment in LDS (2.6 K keys: ~35 vs ~12 us), and wins only once the
// segments outnumber the workgroups the GPU holds at once (8163 segments:
// finish 0.206 -> 0.092 ms; 1000 segments: 0.026 -> 0.041 ms).
constexpr uint32_t kOrderBins = 35;
constexpr uint32_t kFinishWaveLen = 4096;
constexpr uint32_t kFinishWaveKs = 4;
constexpr uint32_t kFinishWaveMinSegs = 2048;
// 4-byte keys: segments of up to kFinishWaveLen4 keys with one k too (|N(0,1)|
// 2^26, 8192 even ks, ~5.5 K keys a segment: finish 0.170 -> 0.089 ms; mixture
// 2^28 8192 ks: 0.318 -> 0.299 ms; with up to 2 ks, or up to 12288 keys, the
// mixture's finish is slower than without the longer form)
constexpr uint32_t kFinishWaveLen4 = 8192;
// With at least kFinishSplitMinSegs whole-workgroup segments below
// kFinishSplitLen keys, the host-driven schedule finishes those with 512-thread
// workgroups and half the LDS (three per CU instead of two) in a launch of
// their own, after the larger ones.
constexpr uint32_t kFinishSplitLen = 8192;
constexpr uint32_t kFinishSplitMinSegs = 2048;
template <class K>
__host__ __device__ __forceinline__ bool finish_wide_segment(uint32_t len, K klo, K khi, uint32_t fin_cap) {
  return len > fin_cap || !finish_fits_half<K>(len, klo, khi, fin_cap);
}
// The part of a segment's class that its length and key interval decide:
// low 8 bits the class without the one-wave option, bit 8 set when its
// length allows the one-wave class, bit 9 when the longer 4-byte-key form
// does; order_bin adds the ks' count.
template <class K>
__host__ __device__ __forceinline__ uint32_t order_bin_info(uint32_t len, K klo, K khi, uint32_t fin_cap) {
  if (finish_wide_segment<K>(len, klo, khi, fin_cap)) return 0;
  const uint32_t base = len >> 10 < kOrderBins - 3 ? 2 + (len >> 10) : kOrderBins - 1;
  uint32_t wave = len <= kFinishWaveLen ? 0x100u : 0u;
  if (sizeof(K) == 4 && len <= kFinishWaveLen4) wave |= 0x200u;
  return base | wave;
}
__host__ __device__ __forceinline__ uint32_t order_bin_of(uint32_t info, uint32_t nks, bool waves) {
  const bool wave = ((info & 0x100u) && nks <= kFinishWaveKs) || ((info & 0x200u) && nks == 1);
  return (waves && wave) ? 1u : (info & 0xFFu);
}
template <class K>
__host__ __device__ __forceinline__ uint32_t order_bin(uint32_t len, K klo, K khi, uint32_t fin_cap, uint32_t nks,
                                                       bool waves) {
  return order_bin_of(order_bin_info<K>(len, klo, khi, fin_cap), nks, waves);
}

// words of a slot rank table (32 buckets per 64-bit word)
__host__ __device__ __forceinline__ uint32_t rank_word_count(uint32_t B) { return (B + 31) / 32; }

__device__ __forceinline__ uint32_t pow2ceil_u32(uint32_t v) {
  if (v <= 1) return 1;
  return 1u << (32 - __builtin_clz(v - 1));
}

// ---------------------------------------------------------------------------
// vector loads: 16 bytes per lane
// ---------------------------------------------------------------------------
template <class V> struct Vec16;
template <> struct Vec16<float> {
  static constexpr int N = 4;
  __device__ static void unpack(const uint4& q, float* o) {
    o[0] = __builtin_bit_cast(float, q.x); o[1] = __builtin_bit_cast(float, q.y);
    o[2] = __builtin_bit_cast(float, q.z); o[3] = __builtin_bit_cast(float, q.w);
  }
};
template <> struct Vec16<uint32_t> {
  static constexpr int N = 4;
  __device__ static void unpack(const uint4& q, uint32_t* o) { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }
};
template <> struct Vec16<int32_t> {
  static constexpr int N = 4;
  __device__ static void unpack(const uint4& q, int32_t* o) {
    o[0] = (int32_t)q.x; o[1] = (int32_t)q.y; o[2] = (int32_t)q.z; o[3] = (int32_t)q.w;
  }
};
template <> struct Vec16<double> {
  static constexpr int N = 2;
  __device__ static void unpack(const uint4& q, double* o) {
    o[0] = __builtin_bit_cast(double, ((uint64_t)q.y << 32) | q.x);
    o[1] = __builtin_bit_cast(double, ((uint64_t)q.w << 32) | q.z);
  }
};
template <> struct Vec16<uint64_t> {
  static constexpr int N = 2;
  __device__ static void unpack(const uint4& q, uint64_t* o) {
    o[0] = ((uint64_t)q.y << 32) | q.x;
    o[1] = ((uint64_t)q.w << 32) | q.z;
  }
};

// 16-byte streaming load with the nontemporal hint, for data each element of
// which is read once (measured on one MI355X: a plain 1 GiB read sweep
// 6.6 TB/s, with the hint 7.0 TB/s).
__device__ __forceinline__ uint4 ld_stream(const uint4* p) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 t = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(t.x, t.y, t.z, t.w);
}

// Streams elements [t0, t1) of x (16-byte aligned base, t0 a multiple of the
// vector width) through f(value, position).  Four 16-byte loads in flight per
// lane.
template <class V, class F>
__device__ __forceinline__ void stream_tile(const V* __restrict__ x, uint32_t t0, uint32_t t1, F&& f) {
  constexpr int N = Vec16<V>::N;
  const uint32_t nt = blockDim.x;
  const uint4* xv = reinterpret_cast<const uint4*>(x);
  uint32_t v1 = t1 / N;
  uint32_t v = t0 / N + threadIdx.x;
  for (; v + 3 * nt < v1; v += 4 * nt) {
    uint4 q0 = xv[v];
    uint4 q1 = xv[v + nt];
    uint4 q2 = xv[v + 2 * nt];
    uint4 q3 = xv[v + 3 * nt];
    V e[4][N];
    Vec16<V>::unpack(q0, e[0]);
    Vec16<V>::unpack(q1, e[1]);
    Vec16<V>::unpack(q2, e[2]);
    Vec16<V>::unpack(q3, e[3]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < N; ++j) f(e[u][j], (v + u * nt) * N + j);
  }
  for (; v < v1; v += nt) {
    uint4 q0 = xv[v];
    V e[N];
    Vec16<V>::unpack(q0, e);
#pragma unroll
    for (int j = 0; j < N; ++j) f(e[j], v * N + j);
  }
  uint32_t tail0 = (v1 * N > t0) ? v1 * N : t0;
  for (uint32_t i = tail0 + threadIdx.x; i < t1; i += nt) f(x[i], i);
}

// Vector form: fv(const V (&e)[N]) receives the N values of each 16-byte load
// (so per-element table lookups of one load can be issued together), f1(V)
// the ragged tail.  Same alignment rules as stream_tile.  NT: nontemporal
// loads, for data read from HBM exactly once (the caller's vector in the first
// pass).  Later passes read what the previous compaction just wrote, much of
// it still in the last-level cache, and lose with the hint (pass-1 compaction
// 0.151 ms plain, 0.160 ms with it).  PIPE: the next four loads are issued
// before the current four are processed (16 more VGPRs), for a workgroup that
// streams alone on its CU and would otherwise leave HBM idle while it computes.
template <class V, bool NT = false, bool PIPE = false, class FV, class F1>
__device__ __forceinline__ void stream_tile_vec(const V* __restrict__ x, uint32_t t0, uint32_t t1, FV&& fv, F1&& f1) {
  constexpr int N = Vec16<V>::N;
  const uint32_t nt = blockDim.x;
  const uint4* xv = reinterpret_cast<const uint4*>(x);
  uint32_t v1 = t1 / N;
  uint32_t v = t0 / N + threadIdx.x;
  if constexpr (PIPE) {
    if (v + 3 * nt < v1) {
      uint4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = NT ? ld_stream(xv + v + u * nt) : xv[v + u * nt];
      for (;;) {
        const uint32_t vn = v + 4 * nt;
        const bool more = vn + 3 * nt < v1;
        uint4 r[4];
        if (more) {
#pragma unroll
          for (int u = 0; u < 4; ++u) r[u] = NT ? ld_stream(xv + vn + u * nt) : xv[vn + u * nt];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          V e[N];
          Vec16<V>::unpack(q[u], e);
          fv(e);
        }
        v = vn;
        if (!more) break;
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = r[u];
      }
    }
  }
  for (; v + 3 * nt < v1; v += 4 * nt) {
    uint4 q0 = NT ? ld_stream(xv + v) : xv[v];
    uint4 q1 = NT ? ld_stream(xv + v + nt) : xv[v + nt];
    uint4 q2 = NT ? ld_stream(xv + v + 2 * nt) : xv[v + 2 * nt];
    uint4 q3 = NT ? ld_stream(xv + v + 3 * nt) : xv[v + 3 * nt];
    V e[N];
    Vec16<V>::unpack(q0, e);
    fv(e);
    Vec16<V>::unpack(q1, e);
    fv(e);
    Vec16<V>::unpack(q2, e);
    fv(e);
    Vec16<V>::unpack(q3, e);
    fv(e);
  }
  for (; v < v1; v += nt) {
    uint4 q0 = NT ? ld_stream(xv + v) : xv[v];
    V e[N];
    Vec16<V>::unpack(q0, e);
    fv(e);
  }
  uint32_t tail0 = (v1 * N > t0) ? v1 * N : t0;
  for (uint32_t i = tail0 + threadIdx.x; i < t1; i += nt) f1(x[i]);
}

// Any start: the elements before the first 16-byte boundary go one per lane,
// the rest through the vector path.
template <class V, bool NT = false, bool PIPE = false, class FV, class F1>
__device__ __forceinline__ void stream_tile_vec_any(const V* __restrict__ x, uint32_t t0, uint32_t t1, FV&& fv,
                                                    F1&& f1) {
  if (t0 >= t1) return;
  const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(x + t0) & 15u);
  if (mis % sizeof(V) != 0) {  // not even element-aligned to a 16-byte grid: scalar only
    for (uint32_t i = t0 + threadIdx.x; i < t1; i += blockDim.x) f1(x[i]);
    return;
  }
  uint32_t head = mis ? (16u - mis) / (uint32_t)sizeof(V) : 0u;
  if (head > t1 - t0) head = t1 - t0;
  if (threadIdx.x < head) f1(x[t0 + threadIdx.x]);
  stream_tile_vec<V, NT, PIPE>(x + t0 + head, 0, t1 - t0 - head, fv, f1);
}

// hist[b[j]] += 1 for N buckets per lane (every lane of the wave active for
// all N).  One compare + ballot decides whether the group needs duplicate
// peeling; the common case is N plain LDS atomics.
template <int N>
__device__ __forceinline__ void lds_count_n(uint32_t* hist, const uint32_t (&b)[N]) {
  // a value dominant enough to need peeling shows up in the first key of the
  // group as well; the other keys of the group are not checked
  const uint32_t b0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)b[0]);
  const bool peel = __builtin_popcountll(wave_ballot(b[0] == b0)) >= kPeelMin;
  if (peel) {
#pragma unroll
    for (int j = 0; j < N; ++j) lds_count(hist, b[j]);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) atomicAdd(&hist[b[j]], 1u);
  }
}

// Caller vectors may start at any element; 16-byte loads need 16-byte
// alignment, so an unaligned base takes the one-element-per-lane path.
template <class V, class F>
__device__ __forceinline__ void stream_tile_any(const V* __restrict__ x, uint32_t t0, uint32_t t1, F&& f) {
  if ((reinterpret_cast<uintptr_t>(x) & 15u) == 0) {
    stream_tile<V>(x, t0, t1, f);
  } else {
    for (uint32_t i = t0 + threadIdx.x; i < t1; i += blockDim.x) f(x[i], i);
  }
}

// ---------------------------------------------------------------------------
// First-pass map in LDS: a two-level key-prefix table.
//
// Level 1 has 8192 cells (top 13 key bits) for 32-bit keys, 4096 (top 12)
// for 64-bit keys.  A cell whose key span contains at
// most one range start is a leaf: its first range r0 and the offset of that
// start inside the cell, so the range of a key is
//     r0 + ((key & cell_mask) > start_offset - 1)
// with no branch.  A cell holding two or more starts points to a level-2
// table of 64 sub-cells (next 6 key bits) in the same leaf format; when the
// level-2 budget is used up, or a sub-cell still holds two or more starts,
// the entry says "search" and carries the first/last candidate range for a
// short binary search.  Both rarer paths sit behind wave-uniform tests.
//
// Entry layout (E = K: 32-bit entries for u32 keys, 64-bit for u64 keys;
// KB = key bits, S1 = KB - level-1 bits, S2 = S1 - 6):
//   level-1 leaf:   bit KB-1 = 0 | r0 in bits [S1, S1+11) | start_offset-1 in bits [0, S1)
//   level-2 leaf:   bit KB-1 = 0 | r0 in bits [S2, S2+17) | start_offset-1 in bits [0, S2)
//   level-1 table:  bit KB-1 = 1, bit KB-2 = 1 | table index
//   search (1, 2):  bit KB-1 = 1, bit KB-2 = 0 | r0 in bits [0, 11) | r1 in bits [11, 22)
// The map parameters of a range are packed for one 16-byte (u32 keys) or
// 32-byte (u64 keys) LDS read.
// ---------------------------------------------------------------------------
template <class K> struct MapParam;
template <> struct __align__(16) MapParam<uint32_t> {
  uint32_t lo, base, mult, sh;
};
template <> struct __align__(16) MapParam<uint64_t> {
  uint64_t lo;
  uint32_t base, mult, sh, pad0, pad1, pad2;
};

template <class K> struct LutGeom {
  static constexpr uint32_t KB = sizeof(K) * 8;
  // 13 level-1 bits for u32 keys: with up to ~128 quantile ranges over a
  // normal-like distribution every cell holding real mass is a leaf (at 12
  // bits, 60 % of the mass of a 128-range list sits in two-start cells and
  // takes the level-2 path).  u64 keys keep 12 bits (the table is 8-byte).
  static constexpr uint32_t L1BITS = sizeof(K) == 4 ? 13u : 12u;
  static constexpr uint32_t L1CELLS = 1u << L1BITS;
  static constexpr uint32_t S1 = KB - L1BITS;
  static constexpr uint32_t S2 = S1 - 6;
  static constexpr uint32_t L2TABLES = 32u;  // 8 KB (u32) / 16 KB (u64) of level 2
  static constexpr K FLAG = (K)1 << (KB - 1);
  static constexpr K TABLE = (K)1 << (KB - 2);
  static constexpr K M1 = ((K)1 << S1) - 1;
  static constexpr K M2 = ((K)1 << S2) - 1;
  static __host__ __device__ constexpr size_t bytes() { return (size_t)(L1CELLS + L2TABLES * 64) * sizeof(K); }
  static __host__ __device__ constexpr uint32_t max_ranges() { return 2047; }
  static constexpr uint32_t kEntries = L1CELLS + L2TABLES * 64;
};

template <class K> __device__ __forceinline__ MapParam<K> to_param(const Range<K>& r) {
  MapParam<K> q{};
  q.lo = r.lo;
  q.base = r.base;
  q.mult = r.mult;
  q.sh = r.sh;
  return q;
}

// Branch-free: mult == 0 (identity map) gives mulhi == 0 and adds t instead,
// so a wave never splits on the map kind.
template <class K> __device__ __forceinline__ uint32_t param_bucket(const MapParam<K>& q, K key) {
  K d = key - q.lo;
  uint32_t t = sizeof(K) == 4 ? (uint32_t)d : (uint32_t)(d >> q.sh);
  uint32_t h = __umulhi(t, q.mult);
  uint32_t id = q.mult == 0 ? t : 0u;
  return q.base + h + id;
}

// The per-range map parameters are stored as separate arrays, so lanes of a
// wave that read different ranges hit different LDS banks (one packed record
// per range made ranges r and r + 8 collide).
//   32-bit keys: lo, mult, and c = base*2^32 + adj - lo*mult (mod 2^64); the
//     bucket is the high word of key*mult + c, one v_mad_u64_u32.  For a
//     non-identity map that is base + mulhi(key - lo, mult) (the 64-bit sum is
//     exact: (key - lo)*mult < nb*2^32).  An identity map (mult == 0: every key
//     its own bucket, key - lo < nb) is stored as mult = adj = 2^32 - 1:
//     (t + 1)(2^32 - 1) = t*2^32 + (2^32 - 1 - t) has high word t.
//   64-bit keys: lo, mult, base, sh, as param_bucket.
// lo also serves the binary search of the rare many-start cells.
template <class K> struct Pass0Map {
  using G = LutGeom<K>;
  K* lut;  // level 1 followed by the level-2 tables
  // cell table window (WIN lookups): cell c covers keys [wbase + (c << wsh),
  // wbase + ((c + 1) << wsh)); keys outside [wbase, wbase + wmax] take the
  // binary-search path.  Without WIN the cells are the top key bits (wsh = S1).
  K wbase = 0;
  K wmax = (K)~(K)0;
  uint32_t wsh = G::S1;
  uint32_t r_below = 0, r_above = 0;  // search spans of keys below / above the window
  __device__ void set_window(const PassCtl* c0) {
    wbase = sizeof(K) == 4 ? (K)c0->lut_base_lo : (K)(((uint64_t)c0->lut_base_hi << 32) | c0->lut_base_lo);
    wsh = c0->lut_shift;
    wmax = wsh >= G::S1 ? (K)~(K)0 : (K)((((K)G::L1CELLS) << wsh) - 1);
    r_below = c0->lut_r_below;
    r_above = c0->lut_r_above;
  }
  __device__ bool windowed() const { return wsh < G::S1; }
  K* plo;
  uint32_t* pmult;
  uint32_t* pbase;  // 64-bit keys
  uint32_t* psh;    // 64-bit keys
  uint64_t* pc;     // 32-bit keys
  uint32_t Rn = 0;  // ranges
  // (Per-field arrays rather than one record per range: random ranges across
  // a wave then spread over all LDS banks; 16-byte records put every lane's
  // first word on a quarter of them and measured slower in the pass-1 kernels.)

  // LDS bytes of the parameter arrays for R ranges (16 bytes per range for
  // 32-bit keys, 20 for 64-bit keys), rounded to 16
  static __host__ __device__ size_t param_bytes(uint32_t R) {
    return (((size_t)R * (sizeof(K) + 12)) + 15) / 16 * 16;
  }
  static __host__ __device__ size_t bytes(uint32_t R) { return G::bytes() + param_bytes(R); }
  // cell table holding only the first l2 level-2 tables
  static __host__ __device__ size_t lut_bytes(uint32_t l2) { return ((size_t)G::L1CELLS + (size_t)l2 * 64) * sizeof(K); }
  static __host__ __device__ size_t bytes(uint32_t R, uint32_t l2) { return lut_bytes(l2) + param_bytes(R); }
  // p must be 8-byte aligned
  __device__ void set_params(unsigned char* p, uint32_t R) {
    Rn = R;
    plo = reinterpret_cast<K*>(p);
    pmult = reinterpret_cast<uint32_t*>(p + (size_t)R * sizeof(K));
    if constexpr (sizeof(K) == 4) {
      pc = reinterpret_cast<uint64_t*>(p + (size_t)R * 8);
      pbase = nullptr;
      psh = nullptr;
    } else {
      pbase = pmult + R;
      psh = pbase + R;
      pc = nullptr;
    }
  }
  __device__ void carve(unsigned char* p, uint32_t R, uint32_t l2 = G::L2TABLES) {
    lut = reinterpret_cast<K*>(p);
    set_params(p + lut_bytes(l2), R);
  }
  // g: a Range<K> or a MapParam<K> (fields lo, mult, base, sh)
  template <class P> __device__ __forceinline__ void store_param(uint32_t r, const P& g) {
    plo[r] = g.lo;
    if constexpr (sizeof(K) == 4) {
      const uint32_t m = g.mult ? g.mult : 0xFFFFFFFFu;
      const uint64_t adj = g.mult ? 0ull : 0xFFFFFFFFull;
      pmult[r] = m;
      pc[r] = ((uint64_t)g.base << 32) + adj - (uint64_t)(uint32_t)g.lo * m;
    } else {
      pmult[r] = g.mult;
      pbase[r] = g.base;
      psh[r] = g.sh;
    }
  }
  __device__ void load_params(const Range<K>* __restrict__ g, uint32_t R) {
    for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) store_param(i, g[i]);
  }
  __device__ void load(const K* __restrict__ g_lut, const Range<K>* __restrict__ g, uint32_t R,
                       uint32_t l2 = G::L2TABLES) {
    for (uint32_t i = threadIdx.x; i < G::L1CELLS + l2 * 64; i += blockDim.x) lut[i] = g_lut[i];
    load_params(g, R);
  }
  // bucket of key in range r (branch-free, equal to param_bucket)
  __device__ __forceinline__ uint32_t range_bucket(uint32_t r, K key) const {
    if constexpr (sizeof(K) == 4) {
      return (uint32_t)(((uint64_t)(uint32_t)key * pmult[r] + pc[r]) >> 32);
    } else {
      K d = key - plo[r];
      uint32_t t = (uint32_t)(d >> psh[r]);
      uint32_t m = pmult[r];
      uint32_t h = __umulhi(t, m);
      uint32_t id = m == 0 ? t : 0u;
      return pbase[r] + h + id;
    }
  }
  // Range of a key in a level-1 leaf entry: r0 (the bits above S1; the flag
  // bit is 0 for a leaf) plus one when the key's in-cell offset passes the
  // stored start offset.  Splicing the in-cell offset under the entry's high
  // bits turns that test into one compare of whole words.  Meaningless for
  // non-leaf entries, which the callers overwrite.
  static __device__ __forceinline__ uint32_t leaf_range(K e, K off) {
    const K spliced = off | (e & (K)~G::M1);
    return (uint32_t)(e >> G::S1) + (spliced > e ? 1u : 0u);
  }
  // Level-1 cell of a key, its offset inside the cell and whether it lies
  // outside the window.
  template <bool WIN> __device__ __forceinline__ uint32_t cell_of(K key, K& off, bool& out) const {
    if constexpr (WIN) {
      const K d = (K)(key - wbase);
      out = d > wmax;
      off = d & (K)(((K)1 << wsh) - 1);
      return out ? 0u : (uint32_t)(d >> wsh);
    } else {
      out = false;
      off = key & G::M1;
      return (uint32_t)(key >> G::S1);
    }
  }
  // The rarer lookups, each behind its own wave-uniform test: a level-2 table
  // (one more LDS read), then a binary search over the range starts for a
  // multi-start (sub)cell or a key outside the window.  Split in two, a wave
  // whose keys need level-2 tables but no search (multi-modal inputs: a few
  // dense cells hold several starts) skips the search code entirely.
  // level2: the sub-cell entry of key, whose level-1 entry e points to a
  // table; r becomes its range when that entry is a leaf.
  template <bool WIN> __device__ __forceinline__ K level2(K key, K e, uint32_t& r) const {
    uint32_t sub;
    K soff;
    if constexpr (WIN) {
      const K d = (K)(key - wbase);
      sub = (uint32_t)(d >> (wsh - 6)) & 63u;
      soff = d & (K)(((K)1 << (wsh - 6)) - 1);
    } else {
      sub = (uint32_t)(key >> G::S2) & 63u;
      soff = key & G::M2;
    }
    const K e2 = lut[G::L1CELLS + (((uint32_t)(e & (G::TABLE - 1))) << 6) + sub];
    r = (uint32_t)((e2 >> G::S2) & 0x1FFFF) + (soff > (e2 & G::M2) ? 1u : 0u);
    return e2;
  }
  // search: the range of key by binary search over the candidates a search
  // entry e names, or over the ranges below / above the window for a key
  // outside it
  __device__ __forceinline__ uint32_t search(K key, K e, bool out) const {
    uint32_t lo, hi;
    if (out) {
      const bool below = key < wbase;
      lo = below ? 0u : r_above;
      hi = below ? r_below : (uint32_t)(Rn - 1);
    } else {
      lo = (uint32_t)(e & 0x7FF);
      hi = (uint32_t)((e >> 11) & 0x7FF);
    }
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (plo[mid] <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
  }
  template <bool WIN = false> __device__ __forceinline__ uint32_t range_of(K key) const {
    K off;
    bool out;
    K e = lut[cell_of<WIN>(key, off, out)];
    uint32_t r = leaf_range(e, off);
    bool nl = (e & G::FLAG) != 0 || out;
    if (wave_ballot(nl)) {
      if (!out && (e & G::TABLE)) e = level2<WIN>(key, e, r);
      const bool ns = out || (e & G::FLAG) != 0;
      if (wave_ballot(ns)) {
        if (ns) r = search(key, e, out);
      }
    }
    return r;
  }
  template <bool WIN = false> __device__ __forceinline__ uint32_t bucket(K key) const {
    return range_bucket(range_of<WIN>(key), key);
  }

  // Buckets of N keys: the N level-1 reads are issued together, one ballot
  // covers the rare non-leaf cells of all N, then the N parameter reads.
  template <int N, bool WIN = false> __device__ __forceinline__ void buckets(const K (&key)[N], uint32_t (&b)[N]) const {
    K e[N], off[N];
    bool out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) e[j] = lut[cell_of<WIN>(key[j], off[j], out[j])];
    uint32_t r[N];
    bool nl = false;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      r[j] = leaf_range(e[j], off[j]);
      nl |= (e[j] & G::FLAG) != 0 || out[j];
    }
    if (wave_ballot(nl)) {
      bool ns = false;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (!out[j] && (e[j] & G::TABLE)) e[j] = level2<WIN>(key[j], e[j], r[j]);
        ns |= out[j] || (e[j] & G::FLAG) != 0;
      }
      if (wave_ballot(ns)) {
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (out[j] || (e[j] & G::FLAG)) r[j] = search(key[j], e[j], out[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) b[j] = range_bucket(r[j], key[j]);
  }
};

// Range of the last start <= key among starts[lo..hi].
template <class K> __device__ __forceinline__ uint32_t last_start_le(const K* starts, uint32_t lo, uint32_t hi, K key) {
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Level-2 entry of a sub-cell with low key lo whose last range starting at
// or below lo is r0 and last starting at or below its high key r1: a leaf
// (at most one start inside) or a search span.
template <class K>
__device__ __forceinline__ K l2_entry(const K* starts, uint32_t r0, uint32_t r1, K lo) {
  using G = LutGeom<K>;
  if (r1 == r0) return ((K)r0 << G::S2) | G::M2;
  if (r1 == r0 + 1) return ((K)r0 << G::S2) | (K)(starts[r1] - lo - 1);
  return G::FLAG | (K)r0 | ((K)r1 << 11);
}

// Two-level table of a range list given its sorted starts (in LDS; the ranges
// partition the key space, starts[0] == 0, R <= 2047).  Called by a whole
// workgroup; writes level 1 and the level-2 tables to g_lut and returns the
// number of level-2 tables in use (same value in every thread).
template <class K>
__device__ uint32_t build_lut(const K* starts, uint32_t R, K* __restrict__ g_lut) {
  using G = LutGeom<K>;
  __shared__ uint32_t tab_cell[G::L2TABLES];
  __shared__ uint32_t scan_s[32];
  const uint32_t nt = blockDim.x;
  // level 1; multi-start cells are numbered in cell order for level-2 tables
  uint32_t per = (G::L1CELLS + nt - 1) / nt;
  uint32_t c0 = threadIdx.x * per, c1 = c0 + per < G::L1CELLS ? c0 + per : G::L1CELLS;
  uint32_t nmulti = 0;
  for (uint32_t c = c0; c < c1; ++c) {
    K lo = (K)((K)c << G::S1), hi = (K)(lo + G::M1);
    uint32_t r0 = last_start_le<K>(starts, 0, R - 1, lo);
    uint32_t r1 = last_start_le<K>(starts, r0, R - 1, hi);
    nmulti += r1 > r0 + 1;
  }
  uint32_t tot;
  uint32_t t = block_excl_scan_u32(nmulti, scan_s, &tot);
  for (uint32_t c = c0; c < c1; ++c) {
    K lo = (K)((K)c << G::S1), hi = (K)(lo + G::M1);
    uint32_t r0 = last_start_le<K>(starts, 0, R - 1, lo);
    uint32_t r1 = last_start_le<K>(starts, r0, R - 1, hi);
    K e;
    if (r1 == r0) {
      e = ((K)r0 << G::S1) | G::M1;
    } else if (r1 == r0 + 1) {
      e = ((K)r0 << G::S1) | (K)(starts[r1] - lo - 1);
    } else if (t < G::L2TABLES) {
      tab_cell[t] = c;
      e = G::FLAG | G::TABLE | (K)t;
      t++;
    } else {
      e = G::FLAG | (K)r0 | ((K)r1 << 11);
    }
    g_lut[c] = e;
  }
  __syncthreads();
  // level 2: 64 sub-cells of each tabled cell
  uint32_t used = tot < G::L2TABLES ? tot : G::L2TABLES;
  for (uint32_t i = threadIdx.x; i < used * 64; i += nt) {
    uint32_t tb = i >> 6, sub = i & 63;
    uint32_t c = tab_cell[tb];
    K lo = (K)(((K)c << G::S1) | ((K)sub << G::S2)), hi = (K)(lo + G::M2);
    uint32_t r0 = last_start_le<K>(starts, 0, R - 1, lo);
    uint32_t r1 = last_start_le<K>(starts, r0, R - 1, hi);
    g_lut[G::L1CELLS + i] = l2_entry<K>(starts, r0, r1, lo);
  }
  return used;
}

// Same table as build_lut, with every level-1 cell decided in O(1) from
// counts instead of two binary searches, over a window of the key space:
// cell c covers [wbase + (c << wsh), wbase + ((c + 1) << wsh)) (wbase = 0,
// wsh = S1: the top key bits, the whole key space).  Each range start inside
// the window is counted in its cell (16-bit counters, two per word) or, when
// it sits exactly on the cell's low edge, flagged in a bitmap; starts below
// the window are counted once.  An exclusive scan of the per-cell totals then
// gives, for cell c with P starts before it (below the window or in earlier
// cells),
//   r0 = P + edge(c) - 1 (last start <= the cell's low key),
//   r1 = r0 + inside(c)  (last start <= the cell's high key).
// A cell with two or more starts gets a level-2 table of 64 sub-cells when
// its width allows (wsh >= 6) and tables are left, else a search entry.
// cw: LDS scratch of lut_count_words<K>() words.  Returns the level-2 tables
// in use (same value in every thread).
template <class K> __host__ __device__ constexpr uint32_t lut_count_words() {
  return LutGeom<K>::L1CELLS / 2 + LutGeom<K>::L1CELLS / 32;
}
template <class K>
__device__ uint32_t build_lut_counted(const K* starts, uint32_t R, K* __restrict__ g_lut, uint32_t* cw,
                                      K wbase = 0, uint32_t wsh = LutGeom<K>::S1) {
  using G = LutGeom<K>;
  constexpr uint32_t C = G::L1CELLS;
  uint32_t* edge = cw + C / 2;
  __shared__ uint32_t tab_cell[G::L2TABLES];
  __shared__ uint32_t scan_s[32];
  __shared__ uint32_t below_s;
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  const K wmask = wsh >= G::S1 ? G::M1 : (K)(((K)1 << wsh) - 1);
  const K wmax = wsh >= G::S1 ? (K)~(K)0 : (K)((((K)C) << wsh) - 1);
  for (uint32_t i = t; i < lut_count_words<K>(); i += nt) cw[i] = 0;
  if (t == 0) below_s = 0;
  __syncthreads();
  for (uint32_t r = t; r < R; r += nt) {
    const K st = starts[r];
    if (st < wbase) {
      atomicAdd(&below_s, 1u);
      continue;
    }
    const K d = (K)(st - wbase);
    if (d > wmax) continue;  // above the window: in no cell
    const uint32_t c = (uint32_t)(d >> wsh);
    if (d & wmask) atomicAdd(&cw[c >> 1], 1u << (16 * (c & 1)));
    else atomicOr(&edge[c >> 5], 1u << (c & 31));
  }
  __syncthreads();
  auto inside = [&](uint32_t c) { return (cw[c >> 1] >> (16 * (c & 1))) & 0xFFFFu; };
  auto on_edge = [&](uint32_t c) { return (edge[c >> 5] >> (c & 31)) & 1u; };
  const uint32_t per = (C + nt - 1) / nt;
  const uint32_t c0 = t * per, c1 = c0 + per < C ? c0 + per : C;
  uint32_t nst = 0, nmulti = 0;
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t k = inside(c);
    nst += k + on_edge(c);
    nmulti += k >= 2;
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan_u32((nmulti << 16) | nst, scan_s, &tot);  // both fit 16 bits
  uint32_t P = below_s + (ex & 0xFFFFu), tix = ex >> 16;
  const bool tables = wsh >= 6;
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t k = inside(c), a = on_edge(c);
    const uint32_t r0 = P + a - 1, r1 = r0 + k;
    const K lo = (K)(wbase + ((K)c << wsh));
    K e;
    if (k == 0) {
      e = ((K)r0 << G::S1) | G::M1;
    } else if (k == 1) {
      e = ((K)r0 << G::S1) | (K)(starts[r1] - lo - 1);
    } else if (tables && tix < G::L2TABLES) {
      tab_cell[tix] = c;
      e = G::FLAG | G::TABLE | (K)tix;
      tix++;
    } else {
      e = G::FLAG | (K)r0 | ((K)r1 << 11);
    }
    g_lut[c] = e;
    P += a + k;
  }
  __syncthreads();
  // level 2: 64 sub-cells of each tabled cell (few: binary search)
  const uint32_t nm = tables ? tot >> 16 : 0u;
  const uint32_t used = nm < G::L2TABLES ? nm : G::L2TABLES;
  const uint32_t ssh = wsh - 6;
  for (uint32_t i = t; i < used * 64; i += nt) {
    uint32_t tb = i >> 6, sub = i & 63;
    uint32_t c = tab_cell[tb];
    K lo = (K)(wbase + ((K)c << wsh) + ((K)sub << ssh)), hi = (K)(lo + (((K)1 << ssh) - 1));
    uint32_t r0 = last_start_le<K>(starts, 0, R - 1, lo);
    uint32_t r1 = last_start_le<K>(starts, r0, R - 1, hi);
    g_lut[G::L1CELLS + i] = l2_entry<K>(starts, r0, r1, lo);
  }
  return used;
}

// Plain fill (used instead of hipMemsetAsync so the sequence stays capturable).
static __global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t words, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) p[i] = v;
}

// ---------------------------------------------------------------------------
// k_setup: the first kernel of a call.  Zeroes the control blocks and writes
// block 0 from the host's values; turns every k (k-th largest, 1-based) into a
// k-th smallest rank.  A k outside [1, n] sets *bad (a flag word in
// host-visible memory) and is never selected.  ks may be host-visible memory.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void setup_item(uint32_t i, const uint32_t* __restrict__ ks, uint32_t nk, uint32_t n,
                                           uint32_t* __restrict__ k_rank, uint32_t* __restrict__ k_state,
                                           PassCtl* __restrict__ ctl, uint32_t ctl_words, const PassCtl& c0,
                                           uint32_t* __restrict__ bad) {
  constexpr uint32_t kCtlWords = sizeof(PassCtl) / 4;
  if (i == 0) ctl[0] = c0;
  else if (i >= kCtlWords && i < ctl_words) reinterpret_cast<uint32_t*>(ctl)[i] = 0u;
  if (i >= nk) return;
  uint32_t k = ks[i];
  if (k < 1 || k > n) {
    *bad = 1u;
    k_rank[i] = 1;
    k_state[i] = kDone;
    return;
  }
  k_rank[i] = n - k + 1;
  k_state[i] = kLive;
}

static __global__ void k_setup(const uint32_t* __restrict__ ks, uint32_t nk, uint32_t n, uint32_t* __restrict__ k_rank,
                               uint32_t* __restrict__ k_state, PassCtl* __restrict__ ctl, uint32_t ctl_words,
                               PassCtl c0, uint32_t* __restrict__ bad) {
  setup_item(blockIdx.x * blockDim.x + threadIdx.x, ks, nk, n, k_rank, k_state, ctl, ctl_words, c0, bad);
}

// ---------------------------------------------------------------------------
// k_count0: first pass over the caller's vector (values of type T).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(1024) k_count0(const T* __restrict__ x, const Range<typename KeyTraits<T>::K>* __restrict__ ranges,
                                                 const typename KeyTraits<T>::K* __restrict__ lut,
                                                 const PassCtl* __restrict__ ctl, uint32_t* __restrict__ rows,
                                                 uint32_t tiles_per_wg) {
  using K = typename KeyTraits<T>::K;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, G = ctl->G;
  const uint32_t w0 = blockIdx.x * tiles_per_wg;
  if (w0 >= G) return;
  // the cell table is a static LDS array: its address is a compile-time
  // constant folded into every cell read (a dynamic-LDS base costs an add per key)
  __shared__ __align__(16) K s_lut[LutGeom<K>::kEntries];
  Pass0Map<K> mp;
  mp.lut = s_lut;
  mp.set_params(smem, ctl->R);
  // tiles_per_wg (1 or 2) consecutive tiles, each counted into its own
  // histogram with no barrier between them, so the loads keep streaming
  // across the tile boundary; one histogram row per tile
  uint32_t* hist0 = reinterpret_cast<uint32_t*>(smem + Pass0Map<K>::param_bytes(ctl->R));
  for (uint32_t i = threadIdx.x; i < B * tiles_per_wg; i += blockDim.x) hist0[i] = 0;
  mp.load(lut, ranges, ctl->R);
  mp.set_window(ctl);
  __syncthreads();
  // With a dominant value (its bucket hb), a key of it only adds to a
  // register counter: no LDS atomic and no wave-wide peeling for most keys;
  // each wave adds its lanes' counts to the histogram once per tile.
  const uint32_t hb = ctl->hot_bucket;
  auto run = [&](auto win, auto hot) {
    constexpr bool W = decltype(win)::value;
    constexpr bool H = decltype(hot)::value;
    for (uint32_t h = 0; h < tiles_per_wg && w0 + h < G; ++h) {
      uint32_t* hist = hist0 + h * B;
      uint32_t t0 = (w0 + h) * tile;
      uint32_t t1 = tile_end(t0, tile, n);
      constexpr int N = Vec16<T>::N;
      uint32_t hot_cnt = 0;
      stream_tile_vec_any<T, true>(
          x, t0, t1,
          [&](const T (&v)[N]) {
            K k[N];
#pragma unroll
            for (int j = 0; j < N; ++j) k[j] = KeyTraits<T>::enc(v[j]);
            uint32_t b[N];
            mp.template buckets<N, W>(k, b);
            if constexpr (H) {
#pragma unroll
              for (int j = 0; j < N; ++j) {
                if (b[j] == hb) ++hot_cnt;
                else atomicAdd(&hist[b[j]], 1u);
              }
            } else {
              lds_count_n<N>(hist, b);
            }
          },
          [&](T v) {
            const uint32_t b = mp.template bucket<W>(KeyTraits<T>::enc(v));
            if (H && b == hb) ++hot_cnt;
            else lds_count(hist, b);
          });
      if constexpr (H) {
        const uint32_t wsum = readlane_u32(wave_incl_scan_u32(hot_cnt), 63);
        if (lane_id() == 0 && wsum) atomicAdd(&hist[hb], wsum);
      }
    }
  };
  if (hb != kNone32) {
    if (mp.windowed()) run(std::true_type{}, std::true_type{});
    else run(std::false_type{}, std::true_type{});
  } else {
    if (mp.windowed()) run(std::true_type{}, std::false_type{});
    else run(std::false_type{}, std::false_type{});
  }
  __syncthreads();
  for (uint32_t h = 0; h < tiles_per_wg && w0 + h < G; ++h) {
    uint32_t* row = rows + (size_t)(w0 + h) * B;
    const uint32_t* hist = hist0 + h * B;
    for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) row[i] = hist[i];
  }
}

// ---------------------------------------------------------------------------
// Segmented passes: the input buffer holds keys grouped by range; a range's
// elements occupy [seg_start, seg_end).  Each lane follows the ranges along
// its (increasing) positions, starting from the range that owns the
// workgroup's first element.
// ---------------------------------------------------------------------------
template <class K> struct SegCursor {
  const Range<K>* __restrict__ rg;
  uint32_t r;
  uint32_t seg_end;
  K lo;
  uint32_t base, sh, mult, nbm1;
  __device__ __forceinline__ void load(uint32_t r_) {
    r = r_;
    Range<K> q = rg[r];
    seg_end = q.seg_end; lo = q.lo; base = q.base; sh = q.sh; mult = q.mult; nbm1 = q.nb - 1;
  }
  __device__ __forceinline__ uint32_t bucket(K key, uint32_t pos) {
    while (pos >= seg_end) load(r + 1);
    return base + map_off<K>(key, lo, sh, mult);
  }
};

template <class K>
__global__ void __launch_bounds__(1024) k_countS(const K* __restrict__ buf, const Range<K>* __restrict__ ranges,
                                                 const uint32_t* __restrict__ wg_first, const PassCtl* __restrict__ ctl,
                                                 uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile;
  if (blockIdx.x >= ctl->G) return;
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem);
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  uint32_t t0 = blockIdx.x * tile;
  uint32_t t1 = tile_end(t0, tile, n);
  SegCursor<K> cur;
  cur.rg = ranges;
  cur.load(wg_first[blockIdx.x]);
  stream_tile<K>(buf, t0, t1, [&](K key, uint32_t pos) {
    lds_count(hist, cur.bucket(key, pos));
  });
  __syncthreads();
  uint32_t* row = rows + (size_t)blockIdx.x * B;
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) row[i] = hist[i];
}

// ---------------------------------------------------------------------------
// k_colsum: rows[G][B] -> column totals (the pass histogram).  64 columns per
// workgroup; each wave sums a band of rows with eight loads in flight.  The
// rows themselves are left as raw per-workgroup counts: the compactions claim
// their write offsets from them with global atomics.
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(1024) k_colsum(const uint32_t* __restrict__ rows, const PassCtl* __restrict__ ctl,
                                                 uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[16][64];
  const uint32_t G = ctl->G, B = ctl->B;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const uint32_t col = blockIdx.x * 64 + lane;
  if (blockIdx.x * 64 >= B) return;
  uint32_t s = 0;
  if (col < B) {
    uint32_t r = wv;
    for (; r + 7 * nw < G; r += 8 * nw) {
      uint32_t v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = rows[(size_t)(r + j * nw) * B + col];
#pragma unroll
      for (int j = 0; j < 8; ++j) s += v[j];
    }
    for (; r < G; r += nw) s += rows[(size_t)r * B + col];
  }
  wsum[wv][lane] = s;
  __syncthreads();
  if (wv == 0 && col < B) {
    uint32_t t = 0;
    for (uint32_t j = 0; j < nw; ++j) t += wsum[j][lane];
    totals[col] = t;
  }
}

// ---------------------------------------------------------------------------
// k_plan: device-side findKBuckets + updatekVals + next-pass splitter.
// ---------------------------------------------------------------------------
template <class K> struct PlanArgs {
  const uint32_t* totals;
  const Range<K>* ranges;    // this pass (sorted by base)
  const PassCtl* ctl;        // this pass
  PassCtl* ctl_next;
  Range<K>* next_ranges;
  uint32_t nk;
  uint32_t* k_rank;
  uint32_t* k_state;
  uint32_t* k_aux;
  uint16_t* slot_of_bucket;  // [B]
  uint32_t* slot_dst;        // [slots]
  uint32_t* slot_bucket;     // [slots]
  uint32_t* act_bucket;      // [U]
  uint32_t* act_cls;
  uint32_t* act_cnt;         // elements of each active bucket
  uint32_t* act_dst;
  uint32_t* act_slot;
  K* act_klo;
  K* act_khi;
  uint32_t* small_start;     // [U_small]
  uint32_t* small_len;
  K* small_klo;              // key interval of each small segment
  K* small_khi;
  uint32_t* small_kstart;    // [U_small + 1]
  uint32_t* small_order;     // [U_small] segments in finish order
  uint32_t* klist;           // [nk]
  uint32_t* wg_first_next;   // [G_max]
  void* out;                 // T[nk]
  uint32_t fin_cap;          // largest segment the finish kernel sorts
  uint32_t finish_cap;       // largest k-bucket finished after this pass (>= fin_cap: the rest from global memory)
  uint32_t B_req;            // bucket budget of the next pass
  uint32_t B_cap;            // hard cap (LDS)
  uint32_t G_max;            // workgroups of the streaming kernels
  uint32_t min_tile;         // smallest tile worth a workgroup
  uint32_t force_finish;     // nonzero: stop after this pass (all active buckets must fit fin_cap)
  uint32_t merge_small;      // nonzero: small buckets also become next-pass ranges (one bucket each)
  MapParam<K>* next_params;  // packed map parameters of next_ranges (or null)
  uint64_t* rank_words;      // slot rank words of this pass's buckets (or null)
  PassCtl* mirror;           // host-visible copy of ctl and ctl_next, written last (or null)
  uint32_t* mirror_seq;      // with a mirror: seq is written here after it (host-visible)
  uint32_t seq;
  uint32_t lds_words;        // dynamic LDS of the launch, in words (the tables' bound plus room for caches)
  uint32_t* nan_out;         // the first pass writes its NaN count here (or null)
  // grouped compaction (or null): slot s belongs to group s / kGroupSlots; the
  // slots of a group are contiguous in the output, so a group owns one span
  uint32_t* grp_cursor;      // [groups]: the next free position of each group (advanced by the compaction)
  uint32_t* grp_start;       // [groups + 1]: first output position of each group (total at the end)
};

template <class T>
__global__ void __launch_bounds__(1024) k_plan(PlanArgs<typename KeyTraits<T>::K> a) {
  using K = typename KeyTraits<T>::K;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t scratch[64];
  __shared__ uint32_t sh_tot[8];
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  const PassCtl& ctl = *a.ctl;
  // a pass the previous plan left empty (its control block is still zero):
  // nothing to plan, and the next pass stays empty too
  if (ctl.n == 0) return;
  const uint32_t B = ctl.B;
  uint32_t* P = reinterpret_cast<uint32_t*>(smem);     // inclusive prefix of the histogram
  uint32_t* aidx = P + B;                              // active flag -> index + 1
  uint32_t* rk_bits = aidx + B;                        // rank words under construction (first pass)
  uint32_t* rk_first = rk_bits + rank_word_count(B);
  uint32_t* xtra = rk_first + rank_word_count(B);  // the rest of the launch's LDS
  const uint32_t used = 2 * B + 2 * rank_word_count(B);
  const uint32_t xtra_words = a.lds_words > used ? a.lds_words - used : 0u;
  uint32_t* kfill = P;  // from P5 on (P and its reuse as range bases are dead): ks per small segment

  __shared__ uint32_t ord_cnt[kOrderBins];  // small segments per size class (finish order)
  if (t < kOrderBins) ord_cnt[t] = 0;
  if (t == 0) sh_tot[7] = 0;  // longest small segment
  // P0: histogram prefix, reset tables
  if (a.rank_words)
    for (uint32_t w = t; w < rank_word_count(B); w += nt) {
      rk_bits[w] = 0;
      rk_first[w] = kNone32;
    }
  for (uint32_t i0 = t; i0 < B; i0 += 8 * nt) {  // eight reads in flight per thread
    uint32_t v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) v[u] = i0 + u * nt < B ? a.totals[i0 + u * nt] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t i = i0 + u * nt;
      if (i < B) {
        P[i] = v[u];
        aidx[i] = 0;
        a.slot_of_bucket[i] = (uint16_t)kNone16;
      }
    }
  }
  __syncthreads();
  block_incl_scan_lds(P, B, scratch);

  // P1: findKBuckets for every live k (its rank inside the bucket from the
  // prefix below the bucket: LDS only)
  for (uint32_t k = t; k < a.nk; k += nt) {
    if (a.k_state[k] != kLive) continue;
    uint32_t g = a.k_rank[k];
    uint32_t lo = 0, hi = B - 1;  // first b with P[b] >= g
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (P[mid] >= g) hi = mid; else lo = mid + 1;
    }
    aidx[lo] = 1;
    a.k_aux[k] = lo;
    a.k_rank[k] = g - (lo ? P[lo - 1] : 0u);  // rank inside the bucket (1-based)
  }
  __syncthreads();
  // P is free from here on: reuse it for the ranges' first buckets
  uint32_t* rbase = P;
  for (uint32_t r = t; r < ctl.R; r += nt) rbase[r] = a.ranges[r].base;

  // P2: number the active buckets in key order
  {
    uint32_t per = (B + nt - 1) / nt;
    uint32_t b0 = t * per, b1 = b0 + per < B ? b0 + per : B;
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; ++i) s += aidx[i];
    uint32_t tot;
    uint32_t off = block_excl_scan_u32(s, scratch, &tot);
    for (uint32_t i = b0; i < b1; ++i)
      if (aidx[i]) {
        a.act_bucket[off] = i;
        aidx[i] = ++off;
      }
    if (t == 0) sh_tot[0] = tot;
    __threadfence_block();
    __syncthreads();
  }
  const uint32_t U = sh_tot[0];

  // P3: classify active buckets (one per thread, from the list P2 wrote)
  for (uint32_t q = t; q < U; q += nt) {
    uint32_t b = a.act_bucket[q];
    uint32_t c = a.totals[b];
    // owning range: last r with base <= b
    uint32_t lo = 0, hi = ctl.R - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (rbase[mid] <= b) lo = mid; else hi = mid - 1;
    }
    Range<K> rg = a.ranges[lo];
    K klo = 0, khi = 0;
    bucket_interval<K>(rg, b - rg.base, klo, khi);
    uint32_t cls;
    if (klo == khi) cls = kResolved;
    else if (a.merge_small) cls = kClsBig;
    else if (c <= a.finish_cap) cls = kClsSmall;
    else cls = a.force_finish ? kClsSmall : kClsBig;
    a.act_cls[q] = cls;
    a.act_cnt[q] = c;
    a.act_klo[q] = klo;
    a.act_khi[q] = khi;
  }
  __threadfence_block();
  __syncthreads();

  // P4: layout of the next buffer: big segments first, then small ones
  {
    uint32_t per = (U + nt - 1) / nt;
    uint32_t q0 = t * per, q1 = q0 + per < U ? q0 + per : U;
    uint32_t nbig = 0, cbig = 0, nsml = 0, csml = 0;
    for (uint32_t q = q0; q < q1; ++q) {
      uint32_t cls = a.act_cls[q], c = a.act_cnt[q];
      if (cls == kClsBig) { nbig++; cbig += c; }
      else if (cls == kClsSmall) { nsml++; csml += c; }
    }
    const uint32_t sv[4] = {nbig, cbig, nsml, csml};
    uint32_t se[4], sT[4];
    block_excl_scan_u32x4(sv, se, sT, scratch);
    uint32_t e0 = se[0], e1 = se[1], e2 = se[2], e3 = se[3];
    const uint32_t T0 = sT[0], T1 = sT[1], T2 = sT[2], T3 = sT[3];
    for (uint32_t q = q0; q < q1; ++q) {
      uint32_t cls = a.act_cls[q], c = a.act_cnt[q];
      if (cls == kClsBig) {
        a.act_slot[q] = e0++;
        a.act_dst[q] = e1;
        e1 += c;
      } else if (cls == kClsSmall) {
        a.act_slot[q] = T0 + e2++;
        a.act_dst[q] = T1 + e3;
        e3 += c;
      } else {
        a.act_slot[q] = kNone32;
        a.act_dst[q] = 0;
      }
    }
    if (t == 0) { sh_tot[1] = T0; sh_tot[2] = T1; sh_tot[3] = T2; sh_tot[4] = T3; }
  }
  __threadfence_block();
  __syncthreads();
  const uint32_t U_big = sh_tot[1], N_big = sh_tot[2], U_small = sh_tot[3], N_small = sh_tot[4];

  // next-pass bucket budget: at least 2 buckets per big range when possible
  uint32_t Bn = a.B_req;
  if (Bn < 4 * U_big) Bn = 4 * U_big;
  if (Bn > a.B_cap) Bn = a.B_cap;

  // P5: slot tables, small segment descriptors, next ranges (unplaced).
  // With room in LDS (past the tables, or behind kfill in P; and aidx once P7
  // is done) the small segments' class information and k-list starts are kept
  // there for P8, so it does not read back what this workgroup just wrote to
  // global memory.
  const bool lcache = U_small < B && (U_small <= xtra_words || 2 * U_small <= B);
  uint32_t* sinfo = U_small <= xtra_words ? xtra : P + U_small;  // class information per small segment
  uint32_t* kst = aidx;           // k-list start per small segment (from P8 on)
  uint32_t small_max = 0;  // this thread's longest small segment
  // four active buckets per thread at a time, their loads issued together
  for (uint32_t q0 = t; q0 < U; q0 += 4 * nt) {
  uint32_t sq[4], bq[4], cq[4], dq[4];
  K lq[4], hq[4];
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t q = q0 + u * nt;
    const bool in = q < U;
    sq[u] = in ? a.act_slot[q] : kNone32;
    bq[u] = in ? a.act_bucket[q] : 0u;
    cq[u] = in ? a.act_cnt[q] : 0u;
    dq[u] = in ? a.act_dst[q] : 0u;
    lq[u] = in ? a.act_klo[q] : (K)0;
    hq[u] = in ? a.act_khi[q] : (K)0;
  }
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t s = sq[u];
    if (s == kNone32) continue;
    const uint32_t b = bq[u], c = cq[u], dst = dq[u];
    a.slot_of_bucket[b] = (uint16_t)s;
    a.slot_dst[s] = dst;
    a.slot_bucket[s] = b;
    if (a.rank_words) {
      atomicOr(&rk_bits[b >> 5], 1u << (b & 31));
      atomicMin(&rk_first[b >> 5], s);
    }
    if (s < U_big) {
      Range<K> nr;
      nr.lo = lq[u];
      nr.hi = hq[u];
      nr.seg_start = dst;
      nr.seg_end = dst + c;
      nr.base = 0;
      uint32_t nb;
      // small buckets carried into the next pass (merge_small) are split like
      // the others: their elements pass through the next pass anyway, and
      // finer buckets leave less for the finish
      if (2 * U_big <= Bn) nb = 2 + (uint32_t)(((uint64_t)c * (Bn - 2 * U_big)) / N_big);
      else nb = 1 + (uint32_t)(((uint64_t)c * (Bn - U_big)) / N_big);
      make_map<K>(nr, nb);
      a.next_ranges[s] = nr;
    } else {
      uint32_t j = s - U_big;
      const K klo = lq[u], khi = hq[u];
      a.small_start[j] = dst;
      a.small_len[j] = c;
      small_max = c > small_max ? c : small_max;
      a.small_klo[j] = klo;
      a.small_khi[j] = khi;
      if (lcache) sinfo[j] = order_bin_info<K>(c, klo, khi, a.fin_cap);
      kfill[j] = 0;
    }
  }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {  // one LDS atomic per wave
    const uint32_t o = (uint32_t)__shfl_xor((int)small_max, d, 64);
    small_max = o > small_max ? o : small_max;
  }
  if (lane_id() == 0 && small_max) atomicMax(&sh_tot[7], small_max);
  __threadfence_block();
  __syncthreads();

  // group spans of the grouped compaction: group g starts where its first slot
  // g * kGroupSlots is laid out (slots are contiguous in slot order)
  if (a.grp_cursor) {
    const uint32_t total = N_big + N_small;
    const uint32_t U_slots = U_big + U_small;
    const uint32_t ng = (U_slots + kGroupSlots - 1) / kGroupSlots;
    for (uint32_t g = t; g <= ng; g += nt) {
      const uint32_t pos = g < ng ? a.slot_dst[g * kGroupSlots] : total;
      a.grp_start[g] = pos;
      if (g < ng) a.grp_cursor[g] = pos;
    }
  }

  // slot rank words (32 buckets each): has-a-slot bits, slot of the first
  if (a.rank_words)
    for (uint32_t w = t; w < rank_word_count(B); w += nt) {
      const uint32_t bits = rk_bits[w];
      a.rank_words[w] = ((uint64_t)(bits ? rk_first[w] : 0u) << 32) | bits;
    }

  // P6: bucket bases of the next pass
  uint32_t B_next = 0;
  {
    uint32_t per = (U_big + nt - 1) / nt;
    uint32_t j0 = t * per, j1 = j0 + per < U_big ? j0 + per : U_big;
    uint32_t s = 0;
    for (uint32_t j = j0; j < j1; ++j) s += a.next_ranges[j].nb;
    uint32_t tot;
    uint32_t off = block_excl_scan_u32(s, scratch, &tot);
    for (uint32_t j = j0; j < j1; ++j) {
      a.next_ranges[j].base = off;
      if (a.next_params) {
        Range<K> r = a.next_ranges[j];
        r.base = off;
        a.next_params[j] = to_param<K>(r);
      }
      off += a.next_ranges[j].nb;
    }
    B_next = tot;
  }

  // P7: re-rank live ks, four per thread at a time (their loads issued together)
  T* out = reinterpret_cast<T*>(a.out);
  for (uint32_t k0 = t; k0 < a.nk; k0 += 4 * nt) {
    uint32_t stq[4], bq[4], lq[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = k0 + u * nt;
      stq[u] = k < a.nk ? a.k_state[k] : kDone;
      bq[u] = k < a.nk ? a.k_aux[k] : 0u;
      lq[u] = k < a.nk ? a.k_rank[k] : 0u;
    }
    uint32_t qq[4], cq[4], dq[4], sq[4];
    K kq[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const bool live = stq[u] == kLive;
      qq[u] = live ? aidx[bq[u]] - 1 : 0u;
      cq[u] = live ? a.act_cls[qq[u]] : kResolved;
      dq[u] = live ? a.act_dst[qq[u]] : 0u;
      sq[u] = live ? a.act_slot[qq[u]] : 0u;
      kq[u] = live ? a.act_klo[qq[u]] : (K)0;
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      if (stq[u] != kLive) continue;
      const uint32_t k = k0 + u * nt;
      if (cq[u] == kResolved) {
        out[k] = KeyTraits<T>::dec(kq[u]);
        a.k_state[k] = kDone;
      } else if (cq[u] == kClsBig) {
        a.k_rank[k] = dq[u] + lq[u];
      } else {
        uint32_t j = sq[u] - U_big;
        a.k_rank[k] = lq[u];
        a.k_aux[k] = j;
        a.k_state[k] = kSmallNew;
        atomicAdd(&kfill[j], 1u);
      }
    }
  }
  __threadfence_block();
  __syncthreads();

  // P8: k lists of the small segments (counting sort by segment)
  const bool waves = U_small >= kFinishWaveMinSegs;
  auto binfo = [&](uint32_t j) {
    return lcache ? sinfo[j] : order_bin_info<K>(a.small_len[j], a.small_klo[j], a.small_khi[j], a.fin_cap);
  };
  auto kstart = [&](uint32_t j) { return lcache ? kst[j] : a.small_kstart[j]; };
  {
    uint32_t per = (U_small + nt - 1) / nt;
    uint32_t j0 = t * per, j1 = j0 + per < U_small ? j0 + per : U_small;
    uint32_t s = 0;
    for (uint32_t j = j0; j < j1; ++j) {
      s += kfill[j];
      // few size classes, many segments: one LDS atomic per wave for a class
      // shared by many lanes
      lds_count(ord_cnt, order_bin_of(binfo(j), kfill[j], waves));
    }
    uint32_t tot;
    uint32_t off = block_excl_scan_u32(s, scratch, &tot);
    for (uint32_t j = j0; j < j1; ++j) {
      a.small_kstart[j] = off;
      if (lcache) kst[j] = off;
      off += kfill[j];
      kfill[j] = 0;
    }
    if (t == 0) {
      a.small_kstart[U_small] = tot;
      if (lcache) kst[U_small] = tot;
    }
  }
  __threadfence_block();
  __syncthreads();
  __shared__ uint32_t s_fin_big;
  if (t == 0) {  // largest size class first
    sh_tot[5] = ord_cnt[0];
    sh_tot[6] = ord_cnt[1];
    uint32_t acc = 0;
    for (int c = (int)kOrderBins - 1; c >= 0; --c) {
      const uint32_t v = ord_cnt[c];
      ord_cnt[c] = acc;
      acc += v;
    }
    // the classes from kFinishSplitLen keys up come first: they end where
    // the class below them starts
    s_fin_big = ord_cnt[2 + (kFinishSplitLen >> 10) - 1];
  }
  uint32_t live = 0;
  for (uint32_t k0 = t; k0 < a.nk; k0 += 4 * nt) {  // four ks per thread at a time
    uint32_t stq[4], jq[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = k0 + u * nt;
      stq[u] = k < a.nk ? a.k_state[k] : kDone;
      jq[u] = stq[u] == kSmallNew ? a.k_aux[k] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = k0 + u * nt;
      if (stq[u] == kSmallNew) {
        const uint32_t j = jq[u];
        const uint32_t pos = atomicAdd(&kfill[j], 1u);
        a.klist[kstart(j) + pos] = k;
        a.k_state[k] = kSmall;
      }
      if (stq[u] == kLive) live++;
    }
  }
  uint32_t live_tot;
  block_excl_scan_u32(live, scratch, &live_tot);  // (its barriers also publish ord_cnt)
  // finish order of the small segments: larger ones first, so the longest
  // workgroups start early instead of trailing the launch
  for (uint32_t j = t; j < U_small; j += nt) {
    const uint32_t nks = kstart(j + 1) - kstart(j);
    a.small_order[lds_claim(ord_cnt, order_bin_of(binfo(j), nks, waves), true)] = j;
  }

  // P9: next-pass tiling and first range of every tile
  uint32_t Gn = 0, tile = 0;
  if (N_big > 0) {
    Gn = (N_big + a.min_tile - 1) / a.min_tile;
    if (Gn > a.G_max) Gn = a.G_max;
    if (Gn < 1) Gn = 1;
    tile = (N_big + Gn - 1) / Gn;
    tile = (tile + 63) & ~63u;
    Gn = (N_big + tile - 1) / tile;
  }
  for (uint32_t j = t; j < U_big; j += nt) {
    Range<K> r = a.next_ranges[j];
    uint32_t w0 = (r.seg_start + tile - 1) / tile;
    uint32_t w1 = (r.seg_end + tile - 1) / tile;
    for (uint32_t w = w0; w < w1 && w < Gn; ++w) a.wg_first_next[w] = j;
  }
  if (t == 0) {
    PassCtl* cn = a.ctl_next;
    cn->n = N_big;
    cn->tile = tile;
    cn->G = Gn;
    cn->R = U_big;
    cn->Rp = pow2ceil_u32(U_big);
    cn->B = B_next;
    cn->nan_lo_bucket = kNone32;
    cn->nan_hi_bucket = kNone32;
    PassCtl* cc = const_cast<PassCtl*>(a.ctl);
    cc->U = U;
    cc->U_big = U_big;
    cc->U_small = U_small;
    cc->U_slots = U_big + U_small;
    cc->N_big = N_big;
    cc->N_small = N_small;
    cc->grp_sparse = (uint64_t)(N_big + N_small) * 4 < (uint64_t)ctl.n ? 1u : 0u;
    cc->live_ks = live_tot;
    cc->U_wide = sh_tot[5];
    cc->U_wave = sh_tot[6];
    cc->U_fin_big = s_fin_big;
    cc->small_max = sh_tot[7];
    uint32_t nan = 0;
    if (ctl.nan_lo_bucket != kNone32) nan += a.totals[ctl.nan_lo_bucket];
    if (ctl.nan_hi_bucket != kNone32) nan += a.totals[ctl.nan_hi_bucket];
    cc->nan_count = nan;
    if (a.nan_out) *a.nan_out = nan;
    // the host waits for this kernel, not for the compaction behind it: hand it
    // both control blocks (read back by the thread that wrote them)
    if (a.mirror) {
      a.mirror[0] = *cc;
      a.mirror[1] = *cn;
      __threadfence_system();  // the mirror reaches host memory before the sequence word
      *reinterpret_cast<volatile uint32_t*>(a.mirror_seq) = a.seq;
    }
  }
}

// ---------------------------------------------------------------------------
// Compaction: copy every element of an active (small or big) bucket to its
// segment.  Segment s occupies [seg_cursor[s] at plan time, + count); workgroup
// w claims the run for its own elements of that bucket with one atomicAdd of
// its histogram-row count (segments come out contiguous; the order of the
// workgroups' runs inside a segment is whatever the atomics give, which no
// later step depends on).
// ---------------------------------------------------------------------------
struct CompactTables {
  uint16_t* slot;    // [B]
  uint32_t* cursor;  // [U_slots]
};

__device__ __forceinline__ void compact_init(CompactTables& ct, unsigned char* p, uint32_t B, uint32_t U_slots,
                                             const uint16_t* __restrict__ slot_of_bucket,
                                             uint32_t* __restrict__ seg_cursor,
                                             const uint32_t* __restrict__ slot_bucket,
                                             const uint32_t* __restrict__ row) {
  ct.slot = reinterpret_cast<uint16_t*>(p);
  ct.cursor = reinterpret_cast<uint32_t*>(p + ((size_t)B * 2 + 15) / 16 * 16);
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) ct.slot[i] = slot_of_bucket[i];
  for (uint32_t s = threadIdx.x; s < U_slots; s += blockDim.x) {
    uint32_t c = row[slot_bucket[s]];
    ct.cursor[s] = c ? atomicAdd(&seg_cursor[s], c) : 0u;
  }
}

__host__ __device__ __forceinline__ size_t compact_lds_bytes(uint32_t B, uint32_t U_slots) {
  return ((size_t)B * 2 + 15) / 16 * 16 + ((size_t)U_slots * 4 + 15) / 16 * 16;
}

// First-pass compaction, append layout: workgroup w writes the survivors of
// its tile (elements of any active bucket) contiguously to [R_w, R_w + n_w),
// n_w = sum over active buckets of its histogram row, R_w claimed with one
// global atomicAdd on the pass's append cursor.  Each wave claims a run for its
// survivors with one LDS atomic and writes it with consecutive lanes, so HBM
// sees full lines instead of one partial line per survivor.  The next pass
// finds each element's range from its key, so the order of the runs is free.
template <class T>
__global__ void __launch_bounds__(1024) k_compact0(const T* __restrict__ x, const Range<typename KeyTraits<T>::K>* __restrict__ ranges,
                                                   const typename KeyTraits<T>::K* __restrict__ lut,
                                                   PassCtl* __restrict__ ctl, const uint32_t* __restrict__ rows,
                                                   const uint16_t* __restrict__ slot_of_bucket,
                                                   const uint32_t* __restrict__ slot_bucket,
                                                   typename KeyTraits<T>::K* __restrict__ out) {
  using K = typename KeyTraits<T>::K;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t cursor;
  __shared__ uint32_t part[16];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, U_slots = ctl->U_slots;
  if (blockIdx.x >= ctl->G || U_slots == 0) return;
  __shared__ __align__(16) K s_lut[LutGeom<K>::kEntries];  // static: folded cell addresses
  Pass0Map<K> mp;
  mp.lut = s_lut;
  mp.set_params(smem, ctl->R);
  uint16_t* slot = reinterpret_cast<uint16_t*>(smem + Pass0Map<K>::param_bytes(ctl->R));
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) slot[i] = slot_of_bucket[i];
  mp.load(lut, ranges, ctl->R);
  const uint32_t* rp = rows + (size_t)blockIdx.x * B;
  uint32_t acc = 0;
  for (uint32_t s = threadIdx.x; s < U_slots; s += blockDim.x) acc += rp[slot_bucket[s]];
  {
    uint32_t tot;
    (void)block_excl_scan_u32(acc, part, &tot);
    if (threadIdx.x == 0) cursor = tot ? atomicAdd(&ctl->append, tot) : 0u;
  }
  mp.set_window(ctl);
  __syncthreads();
  uint32_t t0 = blockIdx.x * tile;
  uint32_t t1 = tile_end(t0, tile, n);
  constexpr int N = Vec16<T>::N;
  auto run_tile = [&](auto win) {
  constexpr bool W = decltype(win)::value;
  stream_tile_vec_any<T, true>(
      x, t0, t1,
      [&](const T (&v)[N]) {
        K k[N];
#pragma unroll
        for (int j = 0; j < N; ++j) k[j] = KeyTraits<T>::enc(v[j]);
        uint32_t b[N];
        mp.template buckets<N, W>(k, b);
        bool act[N];
        uint64_t m[N];
        uint32_t tot = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          act[j] = slot[b[j]] != kNone16;
          m[j] = wave_ballot(act[j]);
          tot += (uint32_t)__builtin_popcountll(m[j]);
        }
        if (tot) {  // one run claimed for the wave's survivors of all N loads
          const uint64_t ex = wave_ballot(true);
          const uint32_t first = (uint32_t)__builtin_ctzll(ex);
          // mask_below(ex) is 0 in lane `first`; as a divergent offset it
          // keeps the compiler from wrapping this one-lane atomic in a
          // whole-wave reduction (see lds_claim)
          const uint32_t z = mask_below(ex);
          uint32_t base = 0;
          if (lane_id() == first) base = atomicAdd(&cursor + z, tot);
          // wave-uniform run pointer: each store is scalar base + 32-bit lane offset
          K* run = out + readlane_u32(base, first);
#pragma unroll
          for (int j = 0; j < N; ++j) {
            if (act[j]) run[mask_below(m[j])] = k[j];
            run += (uint32_t)__builtin_popcountll(m[j]);
          }
        }
      },
      [&](T v) {
        K key = KeyTraits<T>::enc(v);
        bool act = slot[mp.template bucket<W>(key)] != kNone16;
        uint64_t m = wave_ballot(act);
        if (m) {
          uint32_t first = (uint32_t)__builtin_ctzll(m);
          uint32_t base = 0;
          if (lane_id() == first) base = atomicAdd(&cursor, (uint32_t)__builtin_popcountll(m));
          base = readlane_u32(base, first);
          if (act) out[base + mask_below(m)] = key;
        }
      });
  };
  if (mp.windowed()) run_tile(std::true_type{});
  else run_tile(std::false_type{});
}

// ---------------------------------------------------------------------------
// Segmented first pass (every k-bucket in a segment of its own, as the
// reference's copyElements lays them out before sorting the remainder,
// DistributedSMOS/bucketMultiselect.cu:779-793), in two levels staged
// through LDS.
//
// Writing every survivor straight into its segment gives a workgroup one write
// stream per segment, and each store instruction of a wave touches up to 64
// cache lines.  Instead slot s belongs to group s / kGroupSlots (the slots of a
// group are contiguous in the output, so a group owns one span of it), and
//   level 1 (k_group_pass0) streams the input and moves every survivor to its
//     group's span, at a run each workgroup claims from its histogram row,
//     writing beside it (same position, one byte) the slot's index inside the
//     group;
//   level 2 (k_group_split) takes one chunk of one group's span per
//     workgroup and moves every key to its slot's segment.
// Both levels stage a chunk of keys in LDS sorted by destination -- a counting
// sort: one LDS atomic per key gives its rank among the chunk's keys bound for
// the same place, a scan of the counts places each destination -- and write it
// out with consecutive lanes on consecutive addresses of one run, so a wave's
// store touches one or two cache lines per run instead of one per key.
// ---------------------------------------------------------------------------
template <class K> struct GroupStage {
  // level 1: 1024 threads, PER keys a thread per chunk
  static constexpr uint32_t PER = sizeof(K) == 4 ? 8u : 4u;
  static constexpr uint32_t NV = PER / Vec16<K>::N;  // 16-byte loads per thread per chunk
  static constexpr uint32_t kThreads = 1024;
  static constexpr uint32_t CH = kThreads * PER;     // keys per chunk
  // level 2: 512 threads, PER2 keys a thread (measured at 2^28 fp32: 0.41 ms
  // with 163 M survivors, against 0.43 with 1024 threads and 0.52 with 256)
  static constexpr uint32_t PER2 = sizeof(K) == 4 ? 16u : 8u;
  static constexpr uint32_t NV2 = PER2 / Vec16<K>::N;
  static constexpr uint32_t kThreads2 = 512;
  static constexpr uint32_t CH2 = kThreads2 * PER2;
  // dynamic LDS of level 1 after the map parameters: slot table, then the
  // staged chunk (keys, slots)
  static __host__ __device__ size_t pass0_bytes(uint32_t B) {
    return ((size_t)B * 2 + 15) / 16 * 16 + (size_t)CH * (sizeof(K) + 2);
  }
  static __host__ __device__ size_t split_bytes() { return (size_t)CH2 * (sizeof(K) + 4); }
};

// r[j] = rank of key j among the keys of the workgroup with the same
// destination d[j] so far (cnt[d] counts them; d == kNone32: no key).  A
// destination shared by at least kPeelMin lanes of the wave (checked on the
// first key) takes one atomic per wave.
template <int N>
__device__ __forceinline__ void stage_rank(uint32_t* cnt, const uint32_t (&d)[N], uint32_t (&r)[N]) {
  const uint64_t a0 = wave_ballot(d[0] != kNone32);
  const uint32_t d0 = a0 ? readlane_u32(d[0], (uint32_t)__builtin_ctzll(a0)) : kNone32;
  const bool peel = a0 && __builtin_popcountll(wave_ballot(d[0] == d0)) >= kPeelMin;
  if (peel) {
#pragma unroll
    for (int j = 0; j < N; ++j) r[j] = lds_claim(cnt, d[j], d[j] != kNone32);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) r[j] = d[j] != kNone32 ? atomicAdd(&cnt[d[j]], 1u) : 0u;
  }
}

// Exclusive scan of cnt[0, m) (m <= 256) by one whole wave: off[i] gets the
// prefix (written by this wave; waves that scan the same counts write the
// same values), the return value is the sum.
__device__ __forceinline__ uint32_t wave_scan_256(const uint32_t* cnt, uint32_t m, uint32_t* off) {
  const uint32_t l = lane_id();
  uint32_t c[4], s = 0;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    c[q] = 4 * l + q < m ? cnt[4 * l + q] : 0u;
    s += c[q];
  }
  const uint32_t inc = wave_incl_scan_u32(s);
  uint32_t e = inc - s;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    if (4 * l + q < m) off[4 * l + q] = e;
    e += c[q];
  }
  return readlane_u32(inc, 63);
}

template <class T>
__global__ void __launch_bounds__(1024) k_group_pass0(const T* __restrict__ x,
                                                      const Range<typename KeyTraits<T>::K>* __restrict__ ranges,
                                                      const typename KeyTraits<T>::K* __restrict__ lut,
                                                      const PassCtl* __restrict__ ctl, const uint32_t* __restrict__ rows,
                                                      const uint16_t* __restrict__ slot_of_bucket,
                                                      const uint32_t* __restrict__ slot_bucket,
                                                      uint32_t* __restrict__ grp_cursor,
                                                      typename KeyTraits<T>::K* __restrict__ out,
                                                      uint8_t* __restrict__ out_id) {
  using K = typename KeyTraits<T>::K;
  using S = GroupStage<K>;
  constexpr int N = Vec16<T>::N;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_cnt[2][kMaxGroups];  // keys of the chunk per group, by chunk parity
  __shared__ uint32_t s_off[kMaxGroups];     // chunk-local start of each group's keys
  __shared__ uint32_t s_pos[kMaxGroups];     // output position of them
  __shared__ uint32_t s_run[kMaxGroups];     // next position of this workgroup's run of each group
  __shared__ uint32_t s_tot;
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, U_slots = ctl->U_slots;
  if (blockIdx.x >= ctl->G || U_slots == 0 || ctl->grp_sparse) return;  // sparse: k_group_scatter's
  const uint32_t ng = (U_slots + kGroupSlots - 1) / kGroupSlots;
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  __shared__ __align__(16) K s_lut[LutGeom<K>::kEntries];
  Pass0Map<K> mp;
  mp.lut = s_lut;
  mp.set_params(smem, ctl->R);
  unsigned char* p = smem + Pass0Map<K>::param_bytes(ctl->R);
  uint16_t* slot = reinterpret_cast<uint16_t*>(p);
  K* stk = reinterpret_cast<K*>(p + ((size_t)B * 2 + 15) / 16 * 16);  // staged keys
  uint16_t* sts = reinterpret_cast<uint16_t*>(stk + S::CH);              // their slots
  for (uint32_t i = t; i < B; i += nt) slot[i] = slot_of_bucket[i];
  for (uint32_t g = t; g < kMaxGroups; g += nt) {
    s_cnt[0][g] = 0;
    s_cnt[1][g] = 0;
    s_run[g] = 0;
  }
  mp.load(lut, ranges, ctl->R);
  __syncthreads();
  // this tile's survivors per group, from its histogram row: one run per group
  const uint32_t* row = rows + (size_t)blockIdx.x * B;
  for (uint32_t sl = t; sl < U_slots; sl += nt) {
    const uint32_t c = row[slot_bucket[sl]];
    if (c) atomicAdd(&s_run[sl / kGroupSlots], c);
  }
  __syncthreads();
  for (uint32_t g = t; g < ng; g += nt) {
    const uint32_t c = s_run[g];
    s_run[g] = c ? atomicAdd(&grp_cursor[g], c) : 0u;
  }
  mp.set_window(ctl);
  __syncthreads();
  const uint32_t t0 = blockIdx.x * tile;
  const uint32_t t1 = tile_end(t0, tile, n);
  const bool vec = (reinterpret_cast<uintptr_t>(x) & 15u) == 0;
  const uint4* xv = reinterpret_cast<const uint4*>(x);
  const uint32_t ch = nt * S::PER;  // keys per chunk (the staging holds S::CH >= ch)
  auto run_tile = [&](auto win) {
    constexpr bool W = decltype(win)::value;
    uint32_t par = 0;
    // one chunk (key q of this thread valid where ok[q]): bucket and slot,
    // rank among the chunk's keys of the same group, stage, write out
    auto chunk = [&](const K (&k)[S::PER], const bool (&ok)[S::PER]) {
      uint32_t d[S::PER], sl[S::PER], r[S::PER];
#pragma unroll
      for (uint32_t u = 0; u < S::NV; ++u) {
        K kk[N];
#pragma unroll
        for (int j = 0; j < N; ++j) kk[j] = k[u * N + j];
        uint32_t b[N];
        mp.template buckets<N, W>(kk, b);
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const uint32_t q = u * N + j;
          sl[q] = ok[q] ? slot[b[j]] : kNone16;
          d[q] = sl[q] != kNone16 ? sl[q] / kGroupSlots : kNone32;
        }
      }
      uint32_t* cnt = s_cnt[par];
      stage_rank<S::PER>(cnt, d, r);
      __syncthreads();  // counts complete; the previous chunk is written out
      if (t < 64) {
        s_tot = wave_scan_256(cnt, ng, s_off);
        for (uint32_t g = t; g < ng; g += 64) {
          s_pos[g] = s_run[g];
          s_run[g] += cnt[g];
          s_cnt[par ^ 1][g] = 0;
        }
      }
      __syncthreads();
#pragma unroll
      for (uint32_t q = 0; q < S::PER; ++q)
        if (d[q] != kNone32) {
          const uint32_t pos = s_off[d[q]] + r[q];
          stk[pos] = k[q];
          sts[pos] = (uint16_t)sl[q];
        }
      __syncthreads();
      const uint32_t tot = s_tot;
      for (uint32_t i = t; i < tot; i += nt) {
        const uint32_t s = sts[i], g = s / kGroupSlots;
        const uint32_t o = s_pos[g] + (i - s_off[g]);
        out[o] = stk[i];
        out_id[o] = (uint8_t)(s % kGroupSlots);
      }
      par ^= 1u;
    };
    // Whole chunks from 16-byte loads issued two chunks ahead: with one
    // workgroup per CU, four 16-byte loads per lane stay in flight through a
    // chunk's barriers.  Every load is unconditional (the two past the last
    // whole chunk reload it) and the three load buffers rotate by unrolling,
    // not by copies, so the compiler waits for exactly the loads the current
    // chunk needs.
    const uint32_t nfull = vec ? (t1 - t0) / ch : 0u;
    if (nfull) {
      auto issue = [&](uint32_t c, uint4 (&q)[S::NV]) {
        const uint32_t v0 = (t0 + (c < nfull ? c : nfull - 1) * ch) / N;
#pragma unroll
        for (uint32_t u = 0; u < S::NV; ++u) q[u] = ld_stream(xv + v0 + u * nt + t);
      };
      // chunk c from qa (loaded), qb holds chunk c + 1 (in flight), qn gets c + 2
      auto step = [&](uint32_t c, const uint4 (&qa)[S::NV], uint4 (&qn)[S::NV]) {
        issue(c + 2, qn);
        K k[S::PER];
        bool ok[S::PER];
#pragma unroll
        for (uint32_t u = 0; u < S::NV; ++u) {
          T v[N];
          Vec16<T>::unpack(qa[u], v);
#pragma unroll
          for (int j = 0; j < N; ++j) {
            k[u * N + j] = KeyTraits<T>::enc(v[j]);
            ok[u * N + j] = true;
          }
        }
        chunk(k, ok);
      };
      uint4 qx[S::NV], qy[S::NV], qz[S::NV];
      issue(0, qx);
      issue(1, qy);
      for (uint32_t c = 0; c < nfull; c += 3) {
        step(c, qx, qz);
        if (c + 1 >= nfull) break;
        step(c + 1, qy, qx);
        if (c + 2 >= nfull) break;
        step(c + 2, qz, qy);
      }
    }
    // the rest (a ragged end, or a base not 16-byte aligned): element loads
    for (uint32_t c0 = t0 + nfull * ch; c0 < t1; c0 += ch) {
      K k[S::PER];
      bool ok[S::PER];
#pragma unroll
      for (uint32_t q = 0; q < S::PER; ++q) {
        const uint32_t e = c0 + q * nt + t;
        ok[q] = e < t1;
        k[q] = KeyTraits<T>::enc(x[ok[q] ? e : t0]);
      }
      chunk(k, ok);
    }
  };
  if (mp.windowed()) run_tile(std::true_type{});
  else run_tile(std::false_type{});
}

// Level 1 when few keys survive (the plan sets ctl->grp_sparse below a
// quarter): the same runs (one per group per workgroup, claimed from its
// histogram row), filled with no staging: each survivor claims its place with
// an LDS atomic on its group's run pointer (wave-aggregated for a shared
// group) and stores key and id there.  A wave's store then touches as many
// lines as its keys have groups, but each workgroup keeps one open line per
// group, so the lines fill in the caches before they are written back; with
// few survivors that costs less than the staging's three barriers per chunk
// (measured: 15 % survivors 0.59 -> 0.53 ms, 0.4 % 0.123 -> 0.086 ms; with
// half the keys surviving it is 20-60 % slower, hence the switch).  No barrier
// after set-up, and the smaller LDS admits two workgroups per CU.
template <class T>
__global__ void __launch_bounds__(1024) k_group_scatter(const T* __restrict__ x,
                                                        const Range<typename KeyTraits<T>::K>* __restrict__ ranges,
                                                        const typename KeyTraits<T>::K* __restrict__ lut,
                                                        const PassCtl* __restrict__ ctl, const uint32_t* __restrict__ rows,
                                                        const uint16_t* __restrict__ slot_of_bucket,
                                                        const uint32_t* __restrict__ slot_bucket,
                                                        uint32_t* __restrict__ grp_cursor,
                                                        typename KeyTraits<T>::K* __restrict__ out,
                                                        uint8_t* __restrict__ out_id) {
  using K = typename KeyTraits<T>::K;
  constexpr int N = Vec16<T>::N;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_run[kMaxGroups];  // next position of this workgroup's run of each group
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, U_slots = ctl->U_slots;
  if (blockIdx.x >= ctl->G || U_slots == 0 || !ctl->grp_sparse) return;  // dense: k_group_pass0's
  const uint32_t ng = (U_slots + kGroupSlots - 1) / kGroupSlots;
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  __shared__ __align__(16) K s_lut[LutGeom<K>::kEntries];
  Pass0Map<K> mp;
  mp.lut = s_lut;
  mp.set_params(smem, ctl->R);
  uint16_t* slot = reinterpret_cast<uint16_t*>(smem + Pass0Map<K>::param_bytes(ctl->R));
  for (uint32_t i = t; i < B; i += nt) slot[i] = slot_of_bucket[i];
  for (uint32_t g = t; g < kMaxGroups; g += nt) s_run[g] = 0;
  mp.load(lut, ranges, ctl->R);
  __syncthreads();
  const uint32_t* row = rows + (size_t)blockIdx.x * B;
  for (uint32_t sl = t; sl < U_slots; sl += nt) {
    const uint32_t c = row[slot_bucket[sl]];
    if (c) atomicAdd(&s_run[sl / kGroupSlots], c);
  }
  __syncthreads();
  for (uint32_t g = t; g < ng; g += nt) {
    const uint32_t c = s_run[g];
    s_run[g] = c ? atomicAdd(&grp_cursor[g], c) : 0u;
  }
  mp.set_window(ctl);
  __syncthreads();
  const uint32_t t0 = blockIdx.x * tile;
  const uint32_t t1 = tile_end(t0, tile, n);
  auto run_tile = [&](auto win) {
    constexpr bool W = decltype(win)::value;
    stream_tile_vec_any<T, true>(
        x, t0, t1,
        [&](const T (&v)[N]) {
          K k[N];
#pragma unroll
          for (int j = 0; j < N; ++j) k[j] = KeyTraits<T>::enc(v[j]);
          uint32_t b[N], sl[N], d[N], r[N];
          mp.template buckets<N, W>(k, b);
          bool any = false;
#pragma unroll
          for (int j = 0; j < N; ++j) {
            sl[j] = slot[b[j]];
            d[j] = sl[j] != kNone16 ? sl[j] / kGroupSlots : kNone32;
            any |= d[j] != kNone32;
          }
          if (wave_ballot(any)) {  // nothing for a wave without survivors
            stage_rank<N>(s_run, d, r);
#pragma unroll
            for (int j = 0; j < N; ++j)
              if (d[j] != kNone32) {
                out[r[j]] = k[j];
                out_id[r[j]] = (uint8_t)(sl[j] % kGroupSlots);
              }
          }
        },
        [&](T v) {
          const K key = KeyTraits<T>::enc(v);
          const uint32_t s = slot[mp.template bucket<W>(key)];
          const bool act = s != kNone16;
          const uint32_t pos = lds_claim(s_run, act ? s / kGroupSlots : 0u, act);
          if (act) {
            out[pos] = key;
            out_id[pos] = (uint8_t)(s % kGroupSlots);
          }
        });
  };
  if (mp.windowed()) run_tile(std::true_type{});
  else run_tile(std::false_type{});
}

// Level 2: workgroup i takes chunk i of the concatenated group chunks (group g
// cut into chunks of at most CH keys) and moves every key to its slot's segment;
// slot_cursor[s] is the next free position of slot s's segment (the plan sets
// it to the segment start).  Workgroups past the last chunk return at once.
// Launched with GroupStage<K>::kThreads2 threads (a chunk is PER2 keys a thread).
template <class K>
__global__ void __launch_bounds__(512) k_group_split(const K* __restrict__ in, const uint8_t* __restrict__ in_id,
                                                     const PassCtl* __restrict__ ctl,
                                                     const uint32_t* __restrict__ grp_start,
                                                     uint32_t* __restrict__ slot_cursor, K* __restrict__ out) {
  using S = GroupStage<K>;
  constexpr uint32_t N = Vec16<K>::N;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_cnt[kGroupSlots], s_off[kGroupSlots], s_pos[kGroupSlots];
  __shared__ uint32_t s_job[3];  // group, first position, end position
  const uint32_t U_slots = ctl->U_slots;
  if (U_slots == 0) return;
  const uint32_t ng = (U_slots + kGroupSlots - 1) / kGroupSlots;
  const uint32_t nt = blockDim.x, t = threadIdx.x;
  K* stk = reinterpret_cast<K*>(smem);                      // staged keys
  uint32_t* sto = reinterpret_cast<uint32_t*>(stk + S::CH2);  // their output positions
  if (t < kGroupSlots) s_cnt[t] = 0;
  // group g's chunks start at its span's start rounded down to a 16-byte
  // boundary plus multiples of CH2, so no chunk covers more than CH2 / N vectors
  if (t < 64) {  // find this workgroup's group: chunk prefix over the groups
    uint32_t c[4], s = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      const uint32_t g = 4 * t + q;
      c[q] = g < ng ? (grp_start[g + 1] - (grp_start[g] & ~(N - 1)) + S::CH2 - 1) / S::CH2 : 0u;
      s += c[q];
    }
    const uint32_t inc = wave_incl_scan_u32(s);
    uint32_t e = inc - s;
    if (t == 0) s_job[0] = kNone32;
    wave_sync();
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      const uint32_t g = 4 * t + q;
      if (blockIdx.x >= e && blockIdx.x < e + c[q]) {
        const uint32_t gs = grp_start[g], ge = grp_start[g + 1];
        const uint32_t cb = (gs & ~(N - 1)) + (blockIdx.x - e) * S::CH2;
        s_job[0] = g;
        s_job[1] = cb > gs ? cb : gs;
        s_job[2] = ge - cb > S::CH2 ? cb + S::CH2 : ge;
      }
      e += c[q];
    }
  }
  __syncthreads();
  const uint32_t g = s_job[0];
  if (g == kNone32) return;
  const uint32_t a = s_job[1], len = s_job[2] - s_job[1];
  // 16-byte loads of the keys (aligned down: the vectors covering [a, a + len),
  // keys outside it dropped) and the matching N-byte loads of their ids; the
  // buffers are allocated in 256-byte multiples, so every vector that holds a
  // key of the chunk lies inside them
  using IdWord = typename std::conditional<N == 4, uint32_t, uint16_t>::type;
  const uint32_t v0 = a / N, v1 = (a + len + N - 1) / N;
  K k[S::PER2];
  uint32_t d[S::PER2], r[S::PER2];
#pragma unroll
  for (uint32_t u = 0; u < S::NV2; ++u) {
    const uint32_t v = v0 + u * nt + t;
    uint4 q = make_uint4(0, 0, 0, 0);
    IdWord w = 0;
    if (v < v1) {
      q = reinterpret_cast<const uint4*>(in)[v];
      w = reinterpret_cast<const IdWord*>(in_id)[v];
    }
    K e[N];
    __builtin_memcpy(e, &q, 16);
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
      const uint32_t pos = v * N + j;
      const bool ok = v < v1 && pos >= a && pos - a < len;
      k[u * N + j] = e[j];
      d[u * N + j] = ok ? (uint32_t)((w >> (8 * j)) & 0xFFu) : kNone32;
    }
  }
  stage_rank<S::PER2>(s_cnt, d, r);
  __syncthreads();
  // one claim per slot of the chunk, then every wave scans the counts itself
  if (t < kGroupSlots) {
    const uint32_t c = s_cnt[t];
    s_pos[t] = c ? atomicAdd(&slot_cursor[g * kGroupSlots + t], c) : 0u;
  }
  (void)wave_scan_256(s_cnt, kGroupSlots, s_off);
  __syncthreads();
#pragma unroll
  for (uint32_t q = 0; q < S::PER2; ++q)
    if (d[q] != kNone32) {
      const uint32_t pos = s_off[d[q]] + r[q];
      stk[pos] = k[q];
      sto[pos] = s_pos[d[q]] + r[q];
    }
  __syncthreads();
  for (uint32_t i = t; i < len; i += nt) out[sto[i]] = stk[i];
}

// Pass 1 reads the append-layout buffer: an element's range is the slot of its
// first-pass bucket (recomputed from the key with the first-pass map), and the
// range's own map gives the pass-1 bucket.  The first pass numbers its slots in
// bucket order, so a bucket's slot is a rank: one 64-bit word per 32 buckets
// holds which of them have a slot (low half) and the slot of the first such
// bucket (high half), slot = high + popcount(low & bits below).  With only the
// level-2 tables the first pass used, the LDS footprint (cell table, first-pass
// parameters, rank words, pass-1 parameters, histogram) leaves room for two
// workgroups per CU.  Every element here comes from a slotted bucket.
template <class K> struct Pass1Map {
  Pass0Map<K> m0;
  const uint64_t* rank0;               // LDS: ceil(B0 / 32) rank words
  const MapParam<K>* p1;               // pass-1 range parameters by slot, global memory
  Pass0Map<K> p1s;                     // ... or as per-field LDS arrays (p1_lds)
  bool p1_lds;
  __device__ __forceinline__ uint32_t slot0(uint32_t b0) const {
    const uint64_t e = rank0[b0 >> 5];
    const uint32_t below = (uint32_t)e & ((1u << (b0 & 31u)) - 1u);
    return (uint32_t)(e >> 32) + (uint32_t)__builtin_popcount(below);
  }
  __device__ __forceinline__ uint32_t range_bucket(uint32_t s0, K key) const {
    return p1_lds ? p1s.range_bucket(s0, key) : param_bucket<K>(p1[s0], key);
  }
  template <bool WIN = false> __device__ __forceinline__ uint32_t bucket(K key) const {
    return range_bucket(slot0(m0.template bucket<WIN>(key)), key);
  }
  template <int N, bool WIN = false> __device__ __forceinline__ void buckets(const K (&key)[N], uint32_t (&b)[N]) const {
    uint32_t b0[N];
    m0.template buckets<N, WIN>(key, b0);
    uint32_t s0[N];
#pragma unroll
    for (int j = 0; j < N; ++j) s0[j] = slot0(b0[j]);
    if (p1_lds) {
#pragma unroll
      for (int j = 0; j < N; ++j) b[j] = p1s.range_bucket(s0[j], key[j]);
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) b[j] = param_bucket<K>(p1[s0[j]], key[j]);
    }
  }
};

// LDS bytes of the pass-1 map; R1 = 0 when the parameters stay in global memory.
template <class K>
__host__ __device__ __forceinline__ size_t pass1_map_bytes(uint32_t B0, uint32_t R0, uint32_t l2, uint32_t R1) {
  return Pass0Map<K>::bytes(R0, l2) + ((size_t)rank_word_count(B0) * 8 + 15) / 16 * 16 + Pass0Map<K>::param_bytes(R1);
}

template <class K>
__device__ __forceinline__ void pass1_map_init(Pass1Map<K>& pm, unsigned char* p, const PassCtl* ctl0,
                                               const K* lut, const Range<K>* rng0, const uint64_t* rank0_g,
                                               const MapParam<K>* p1_g, uint32_t R1, bool p1_in_lds) {
  const uint32_t B0 = ctl0->B, R0 = ctl0->R, l2 = ctl0->l2_used;
  pm.m0.carve(p, R0, l2);  // first-pass cell table first (p is LDS offset 0 in the pass-1 kernels)
  pm.m0.load(lut, rng0, R0, l2);
  pm.m0.set_window(ctl0);
  uint64_t* rk = reinterpret_cast<uint64_t*>(p + Pass0Map<K>::bytes(R0, l2));
  for (uint32_t i = threadIdx.x; i < rank_word_count(B0); i += blockDim.x) rk[i] = rank0_g[i];
  pm.rank0 = rk;
  unsigned char* q = p + Pass0Map<K>::bytes(R0, l2) + ((size_t)rank_word_count(B0) * 8 + 15) / 16 * 16;
  pm.p1 = p1_g;
  pm.p1_lds = p1_in_lds;
  if (p1_in_lds) {
    pm.p1s.set_params(q, R1);
    for (uint32_t i = threadIdx.x; i < R1; i += blockDim.x) pm.p1s.store_param(i, p1_g[i]);
  }
}

// The pass-1 kernels keep the pass-1 range parameters in LDS when the whole
// layout with them fits p1_budget bytes (the host passes the dynamic LDS it
// gave the launch, or 0 to keep them in global memory).
__device__ __forceinline__ bool pass1_params_in_lds(size_t with_params, uint32_t p1_budget) {
  return with_params <= (size_t)p1_budget;
}

template <class K>
__global__ void __launch_bounds__(1024) k_count1(const K* __restrict__ buf, const PassCtl* __restrict__ ctl0,
                                                 const K* __restrict__ lut, const Range<K>* __restrict__ rng0,
                                                 const uint64_t* __restrict__ rank0, const MapParam<K>* __restrict__ p1,
                                                 uint32_t p1_budget, const PassCtl* __restrict__ ctl,
                                                 uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile;
  if (blockIdx.x >= ctl->G) return;
  const size_t hist_bytes = ((size_t)B * 4 + 15) / 16 * 16;
  const bool p1_in_lds = pass1_params_in_lds(
      pass1_map_bytes<K>(ctl0->B, ctl0->R, ctl0->l2_used, ctl->R) + hist_bytes, p1_budget);
  Pass1Map<K> pm;
  pass1_map_init<K>(pm, smem, ctl0, lut, rng0, rank0, p1, ctl->R, p1_in_lds);
  uint32_t* hist = reinterpret_cast<uint32_t*>(
      smem + pass1_map_bytes<K>(ctl0->B, ctl0->R, ctl0->l2_used, p1_in_lds ? ctl->R : 0));
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  uint32_t t0 = blockIdx.x * tile;
  uint32_t t1 = tile_end(t0, tile, n);
  constexpr int N = Vec16<K>::N;
  auto run_tile = [&](auto win) {
    constexpr bool W = decltype(win)::value;
    stream_tile_vec<K>(
        buf, t0, t1,
        [&](const K (&k)[N]) {
          uint32_t b[N];
          pm.template buckets<N, W>(k, b);
          lds_count_n<N>(hist, b);
        },
        [&](K key) { lds_count(hist, pm.template bucket<W>(key)); });
  };
  if (pm.m0.windowed()) run_tile(std::true_type{});
  else run_tile(std::false_type{});
  __syncthreads();
  uint32_t* row = rows + (size_t)blockIdx.x * B;
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) row[i] = hist[i];
}

template <class K>
__global__ void __launch_bounds__(1024) k_compact1(const K* __restrict__ buf, const PassCtl* __restrict__ ctl0,
                                                   const K* __restrict__ lut, const Range<K>* __restrict__ rng0,
                                                   const uint64_t* __restrict__ rank0, const MapParam<K>* __restrict__ p1,
                                                   uint32_t p1_budget, const PassCtl* __restrict__ ctl,
                                                   const uint32_t* __restrict__ rows,
                                                   const uint16_t* __restrict__ slot_of_bucket,
                                                   uint32_t* __restrict__ slot_dst,
                                                   const uint32_t* __restrict__ slot_bucket, K* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, U_slots = ctl->U_slots;
  if (blockIdx.x >= ctl->G || U_slots == 0) return;
  const bool p1_in_lds = pass1_params_in_lds(
      pass1_map_bytes<K>(ctl0->B, ctl0->R, ctl0->l2_used, ctl->R) + compact_lds_bytes(B, U_slots), p1_budget);
  Pass1Map<K> pm;
  pass1_map_init<K>(pm, smem, ctl0, lut, rng0, rank0, p1, ctl->R, p1_in_lds);
  CompactTables ct;
  compact_init(ct, smem + pass1_map_bytes<K>(ctl0->B, ctl0->R, ctl0->l2_used, p1_in_lds ? ctl->R : 0), B, U_slots, slot_of_bucket,
               slot_dst, slot_bucket, rows + (size_t)blockIdx.x * B);
  __syncthreads();
  uint32_t t0 = blockIdx.x * tile;
  uint32_t t1 = tile_end(t0, tile, n);
  constexpr int N = Vec16<K>::N;
  auto run_tile = [&](auto win) {
  constexpr bool W = decltype(win)::value;
  stream_tile_vec<K>(
      buf, t0, t1,
      [&](const K (&k)[N]) {
        uint32_t b[N];
        pm.template buckets<N, W>(k, b);
        uint32_t sl[N];
        bool any = false;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          sl[j] = ct.slot[b[j]];
          any |= sl[j] != kNone16;
        }
        // one LDS atomic per survivor (no duplicate peeling: the survivors of
        // a wave rarely share a segment), nothing for a wave without any
        if (wave_ballot(any)) {
#pragma unroll
          for (int j = 0; j < N; ++j)
            if (sl[j] != kNone16) out[atomicAdd(&ct.cursor[sl[j]], 1u)] = k[j];
        }
      },
      [&](K key) {
        uint32_t s = ct.slot[pm.template bucket<W>(key)];
        bool act = s != kNone16;
        uint32_t p = lds_claim(ct.cursor, s, act);
        if (act) out[p] = key;
      });
  };
  if (pm.m0.windowed()) run_tile(std::true_type{});
  else run_tile(std::false_type{});
}

template <class K>
__global__ void __launch_bounds__(1024) k_compactS(const K* __restrict__ buf, const Range<K>* __restrict__ ranges,
                                                   const uint32_t* __restrict__ wg_first, const PassCtl* __restrict__ ctl,
                                                   const uint32_t* __restrict__ rows,
                                                   const uint16_t* __restrict__ slot_of_bucket,
                                                   uint32_t* __restrict__ slot_dst,
                                                   const uint32_t* __restrict__ slot_bucket, K* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t B = ctl->B, n = ctl->n, tile = ctl->tile, U_slots = ctl->U_slots;
  if (blockIdx.x >= ctl->G || U_slots == 0) return;
  CompactTables ct;
  compact_init(ct, smem, B, U_slots, slot_of_bucket, slot_dst, slot_bucket, rows + (size_t)blockIdx.x * B);
  __syncthreads();
  uint32_t t0 = blockIdx.x * tile;
  uint32_t t1 = tile_end(t0, tile, n);
  SegCursor<K> cur;
  cur.rg = ranges;
  cur.load(wg_first[blockIdx.x]);
  stream_tile<K>(buf, t0, t1, [&](K key, uint32_t pos) {
    uint32_t s = ct.slot[cur.bucket(key, pos)];
    bool act = s != kNone16;
    uint32_t p = lds_claim(ct.cursor, s, act);
    if (act) out[p] = key;
  });
}

// ---------------------------------------------------------------------------
// k_finish: exact answers for the ks of every small segment.
// A launch handles the segments of one size class (len_lo, len_hi].
// The segment (<= fin_cap keys) is loaded into LDS once.  A segment holding
// few ks answers all of them by one radix select: 8-bit digits over only the
// low bits in which the segment's key interval [klo, khi] varies (all keys
// share the bits above).  Each round builds one LDS histogram per distinct
// prefix the ks have reached (ks sorted by rank, so equal prefixes are
// neighbours) in a single sweep over the keys, and each k then takes its digit
// from its prefix's histogram.  A segment holding many ks is sorted instead
// (bitonic network in LDS) and read at every rank.
// ---------------------------------------------------------------------------
constexpr uint32_t kSelectMaxKs = 8;

// Radix select of one rank (the batched kernel's finish): same digits, one
// 256-bin histogram per round.
template <class K>
__device__ K radix_select_lds(const K* __restrict__ s, uint32_t len, K klo, K khi, uint32_t rank, uint32_t* hist,
                              uint32_t* bcast) {
  constexpr uint32_t kb = sizeof(K) * 8;
  uint32_t vary = bitlen64((uint64_t)(klo ^ khi));
  K prefix = vary >= kb ? (K)0 : (K)(klo & (K)(~(K)0 << vary));
  uint32_t top = (vary + 7) & ~7u;
  for (int shift = (int)top - 8; shift >= 0; shift -= 8) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    K hmask = (uint32_t)shift + 8 >= kb ? (K)0 : (K)(~(K)0 << (shift + 8));
    K want = prefix & hmask;
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
      K key = s[i];
      bool m = (key & hmask) == want;
      lds_count(hist, (uint32_t)(key >> shift) & 255u, m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      uint32_t l = threadIdx.x;
      uint32_t c0 = hist[4 * l], c1 = hist[4 * l + 1], c2 = hist[4 * l + 2], c3 = hist[4 * l + 3];
      uint32_t sum = c0 + c1 + c2 + c3;
      uint32_t cum = wave_incl_scan_u32(sum) - sum;
      uint32_t cs[4] = {c0, c1, c2, c3};
      for (uint32_t b = 0; b < 4; ++b) {
        if (cum < rank && cum + cs[b] >= rank) {
          bcast[0] = 4 * l + b;
          bcast[1] = cum;
        }
        cum += cs[b];
      }
    }
    __syncthreads();
    uint32_t d = bcast[0];
    rank -= bcast[1];
    prefix = (K)(prefix | ((K)d << shift));
    __syncthreads();
  }
  return prefix;
}

struct SelectLds {
  alignas(16) uint32_t hist[kSelectMaxKs * 256];
  uint32_t kid[kSelectMaxKs];   // k index, sorted by rank
  uint32_t rank[kSelectMaxKs];  // 1-based rank inside the segment, then inside the prefix
  uint32_t grp[kSelectMaxKs];   // histogram of the k's prefix this round
  uint32_t ng;
  uint64_t pre[kSelectMaxKs];   // digits found so far (as the key type, widened)
  uint64_t gpre[kSelectMaxKs];  // distinct prefixes of this round
};

// Digits of a radix select over the low `vary` bits of a segment's keys (all
// keys share the bits above): the first digit is up to kSelectFirstBits wide
// (every k starts from the same prefix, so one histogram of 2^w bins, the
// whole of SelectLds::hist), the later ones up to 8 bits (one 256-bin
// histogram per distinct prefix the ks have reached).  A first digit of 11
// bits resolves most segments of a bucketing pass in one sweep: their
// intervals span a few thousand key values.
constexpr uint32_t kSelectFirstBits = 11;
static_assert((1u << kSelectFirstBits) <= kSelectMaxKs * 256, "first digit histogram");

__host__ __device__ __forceinline__ void select_first_digit(uint32_t vary, uint32_t& shift, uint32_t& width) {
  width = vary < kSelectFirstBits ? vary : kSelectFirstBits;
  shift = vary - width;
}

// Clears the first digit's histogram: all of SelectLds::hist (the scan reads
// every bin whatever the digit's width; clearing only the bins a narrow digit
// uses and scanning only those measured slower, the extra paths costing the
// finish kernel its two workgroups per CU or more spills).  Whole workgroup;
// a barrier must follow before the counting.
__device__ __forceinline__ void select_zero_first(SelectLds& L) {
  uint4* h = reinterpret_cast<uint4*>(L.hist);
  for (uint32_t i = threadIdx.x; i < kSelectMaxKs * 64; i += blockDim.x) h[i] = make_uint4(0, 0, 0, 0);
}

// Count key (any value of the segment) into the first digit's histogram.
template <class K>
__device__ __forceinline__ void select_count_first(uint32_t* hist, K key, uint32_t shift, uint32_t width, bool active) {
  lds_count(hist, (uint32_t)(key >> shift) & ((1u << width) - 1u), active);
}

// On entry kid/rank hold q (<= kSelectMaxKs) ks sorted by rank; on exit pre[j]
// is the key of rank rank[j] among s[0, len).  Whole workgroup.  s is the
// segment in LDS, or in global memory for a segment too wide and too large
// for the LDS layouts (each digit then reads it from the caches again).  With
// `counted` the caller has already built the first digit's histogram in
// L.hist (zeroed, then select_count_first for every key, then a barrier --
// the one after its writes of kid/rank serves), typically while loading the
// segment; s is then read only for the later digits (not at all when the
// first digit covers every varying bit).
template <class K>
__device__ void radix_select_multi(const K* __restrict__ s, uint32_t len, K klo, K khi, uint32_t q, SelectLds& L,
                                   bool counted = false) {
  constexpr uint32_t kb = sizeof(K) * 8;
  const uint32_t vary = bitlen64((uint64_t)(klo ^ khi));
  const K base = vary >= kb ? (K)0 : (K)(klo & (K)(~(K)0 << vary));
  const uint32_t wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // every key once, four reads in flight per thread (s may be LDS or global)
  auto sweep = [&](auto&& f) {
    const uint32_t nt = blockDim.x;
    uint32_t i = threadIdx.x;
    for (; i + 3 * nt < len; i += 4 * nt) {
      const K k0 = s[i], k1 = s[i + nt], k2 = s[i + 2 * nt], k3 = s[i + 3 * nt];
      f(k0);
      f(k1);
      f(k2);
      f(k3);
    }
    for (; i < len; i += nt) f(s[i]);
  };
  uint32_t sh1, w1;
  select_first_digit(vary, sh1, w1);
  if (w1 == 0) {  // every key equal
    if (threadIdx.x < q) L.pre[threadIdx.x] = base;
  } else {
    if (!counted) {
      select_zero_first(L);
      __syncthreads();
      sweep([&](K key) { select_count_first<K>(L.hist, key, sh1, w1, true); });
      __syncthreads();
    }
    // each k scans the histogram with one wave, 32 bins a lane read as eight
    // 16-byte words (one LDS round trip); the lane where the rank falls walks
    // its bins in registers, half at a time
    const uint32_t l = threadIdx.x & 63;
    const uint4* h4 = reinterpret_cast<const uint4*>(L.hist) + 8 * l;
    for (uint32_t j = wv; j < q; j += nw) {
      const uint32_t want = L.rank[j];
      uint32_t lo = 0, hi = 0;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const uint4 x = h4[v];
        (v < 4 ? lo : hi) += x.x + x.y + x.z + x.w;
      }
      uint32_t cum = wave_incl_scan_u32(lo + hi) - (lo + hi);
      if (cum < want && cum + lo + hi >= want) {  // exactly one lane
        const bool second = cum + lo < want;
        uint32_t c[16];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint4 x = h4[(second ? 4 : 0) + v];
          c[4 * v] = x.x;
          c[4 * v + 1] = x.y;
          c[4 * v + 2] = x.z;
          c[4 * v + 3] = x.w;
        }
        if (second) cum += lo;
        uint32_t b = 32 * l + (second ? 16u : 0u), rr = 0;
        bool done = false;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          if (!done && cum + c[e] >= want) {
            b += (uint32_t)e;
            rr = want - cum;
            done = true;
          }
          cum += c[e];
        }
        L.pre[j] = (K)(base | ((K)b << sh1));
        L.rank[j] = rr;
      }
    }
  }
  __syncthreads();
  for (uint32_t rem = sh1; rem > 0;) {
    const uint32_t w = rem < 8 ? rem : 8u;
    const uint32_t shift = rem - w;
    rem = shift;
    const uint32_t dmask = (1u << w) - 1u;
    const K hmask = shift + w >= kb ? (K)0 : (K)(~(K)0 << (shift + w));
    if (threadIdx.x == 0) {
      uint32_t ng = 0;
      for (uint32_t j = 0; j < q; ++j) {
        const K pj = (K)((K)L.pre[j] & hmask);
        if (ng == 0 || (K)L.gpre[ng - 1] != pj) L.gpre[ng++] = pj;
        L.grp[j] = ng - 1;
      }
      L.ng = ng;
    }
    __syncthreads();
    const uint32_t ng = L.ng;
    K gp[kSelectMaxKs];  // the round's prefixes in registers (entries >= ng unused)
#pragma unroll
    for (uint32_t i = 0; i < kSelectMaxKs; ++i) gp[i] = (K)L.gpre[i];
    for (uint32_t i = threadIdx.x; i < ng * 256; i += blockDim.x) L.hist[i] = 0;
    __syncthreads();
    if (ng == 1) {  // every k shares one prefix: the common case
      sweep([&](K key) { lds_count(L.hist, (uint32_t)(key >> shift) & dmask, (K)(key & hmask) == gp[0]); });
    } else {
      sweep([&](K key) {
        const K km = (K)(key & hmask);
        uint32_t g = ng;  // the prefixes are distinct: at most one matches
#pragma unroll
        for (uint32_t p = 0; p < kSelectMaxKs; ++p)
          if (p < ng && gp[p] == km) g = p;
        lds_count(L.hist, g * 256 + ((uint32_t)(key >> shift) & dmask), g < ng);
      });
    }
    __syncthreads();
    for (uint32_t j = wv; j < q; j += nw) {
      const uint32_t l = threadIdx.x & 63;
      const uint32_t* h = L.hist + L.grp[j] * 256;
      const uint32_t want = L.rank[j];
      uint32_t c[4] = {h[4 * l], h[4 * l + 1], h[4 * l + 2], h[4 * l + 3]};
      const uint32_t sum = c[0] + c[1] + c[2] + c[3];
      uint32_t cum = wave_incl_scan_u32(sum) - sum;
      for (uint32_t b = 0; b < 4; ++b) {
        if (cum < want && cum + c[b] >= want) {
          L.pre[j] = (K)((K)L.pre[j] | ((K)(4 * l + b) << shift));
          L.rank[j] = want - cum;
        }
        cum += c[b];
      }
    }
    __syncthreads();
  }
}

// Reads buf[st, st + len) with 16-byte loads, up to kSegLoadVecs per lane in
// flight (a segment of the finish capacity is one round trip), and calls
// f(i, key, in) for every key of every vector: i its index in the segment, in
// whether it belongs to the segment (the vectors are aligned down/up to 16
// bytes; i wraps above len for the head keys before st).  Every buffer this
// reads is allocated in 256-byte multiples from a 256-byte aligned base, so
// the rounded vectors stay inside it.
constexpr int kSegLoadVecs = 4;

template <class K, class F>
__device__ __forceinline__ void load_segment(const K* __restrict__ buf, uint32_t st, uint32_t len, F&& f) {
  constexpr uint32_t N = 16 / sizeof(K);
  const uint32_t head = st & (N - 1);
  const uint4* src = reinterpret_cast<const uint4*>(buf + (st - head));
  const uint32_t nv = (head + len + N - 1) / N;
  auto put = [&](uint32_t i, const uint4& v) {
    K e[N];
    __builtin_memcpy(e, &v, 16);
#pragma unroll
    for (uint32_t q = 0; q < N; ++q) {
      const uint32_t idx = i * N + q - head;
      f(idx, e[q], idx < len);
    }
  };
  uint4 v[kSegLoadVecs];
#pragma unroll
  for (int j = 0; j < kSegLoadVecs; ++j) {
    uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
    if (i < nv) v[j] = src[i];
  }
#pragma unroll
  for (int j = 0; j < kSegLoadVecs; ++j) {
    uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
    if (i < nv) put(i, v[j]);
  }
  for (uint32_t i = threadIdx.x + kSegLoadVecs * blockDim.x; i < nv; i += blockDim.x) put(i, src[i]);
}

// Answers the ks [q0, q1) of one segment held in LDS as type C, C value = key - base.
// s is writable LDS, except for the global-memory case, which passes at most
// kSelectMaxKs ks and so never reaches the sort.  counted: see
// radix_select_multi (at most kSelectMaxKs ks).
template <class T, class C>
__device__ __forceinline__ void finish_segment(const C* __restrict__ s, uint32_t len, C clo, C chi,
                                               typename KeyTraits<T>::K base, uint32_t q0, uint32_t q1,
                                               uint32_t kk, uint32_t kr, const uint32_t* __restrict__ klist,
                                               uint32_t* __restrict__ k_rank, uint32_t* __restrict__ k_state,
                                               T* __restrict__ out, SelectLds& L, bool counted = false) {
  using K = typename KeyTraits<T>::K;
  const uint32_t q = q1 - q0;
  if (q == 1) {
    if (threadIdx.x == 0) {
      L.kid[0] = kk;
      L.rank[0] = kr;
    }
    __syncthreads();
    radix_select_multi<C>(s, len, clo, chi, 1, L, counted);
    if (threadIdx.x == 0) {
      out[kk] = KeyTraits<T>::dec((K)(base + (K)(C)L.pre[0]));
      k_state[kk] = kDone;
    }
  } else if (q <= kSelectMaxKs) {
    if (threadIdx.x < q) {
      L.kid[threadIdx.x] = kk;
      L.rank[threadIdx.x] = kr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // insertion sort by rank (q <= 8)
      for (uint32_t a = 1; a < q; ++a) {
        const uint32_t ka = L.kid[a], ra = L.rank[a];
        uint32_t b = a;
        for (; b > 0 && L.rank[b - 1] > ra; --b) {
          L.kid[b] = L.kid[b - 1];
          L.rank[b] = L.rank[b - 1];
        }
        L.kid[b] = ka;
        L.rank[b] = ra;
      }
    }
    __syncthreads();
    radix_select_multi<C>(s, len, clo, chi, q, L, counted);
    if (threadIdx.x < q) {
      const uint32_t k = L.kid[threadIdx.x];
      out[k] = KeyTraits<T>::dec((K)(base + (K)(C)L.pre[threadIdx.x]));
      k_state[k] = kDone;
    }
  } else {
    C* w = const_cast<C*>(s);
    const uint32_t M = pow2ceil_u32(len);
    for (uint32_t i = len + threadIdx.x; i < M; i += blockDim.x) w[i] = (C)~(C)0;
    __syncthreads();
    bitonic_sort_lds<C, false, false>(w, M);  // within the 64-VGPR budget: plain LDS stages
    for (uint32_t qq = q0 + threadIdx.x; qq < q1; qq += blockDim.x) {
      const uint32_t k = klist[qq];
      out[k] = KeyTraits<T>::dec((K)(base + (K)s[k_rank[k] - 1]));
      k_state[k] = kDone;
    }
  }
}

// Eight waves per SIMD (two workgroups per CU) caps the registers at 64 VGPRs
// (left alone the compiler takes 91 SGPRs, which admits only one 1024-thread
// workgroup per CU).  The wide segments' global-memory path lives in its own
// kernel: inlined here it made this one spill.  Dynamic LDS: fin_cap / 2 keys.
template <class T>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8))) k_finish(const typename KeyTraits<T>::K* __restrict__ buf, const PassCtl* __restrict__ ctl,
                                                 const uint32_t* __restrict__ small_start,
                                                 const uint32_t* __restrict__ small_len,
                                                 const typename KeyTraits<T>::K* __restrict__ small_klo,
                                                 const typename KeyTraits<T>::K* __restrict__ small_khi,
                                                 const uint32_t* __restrict__ small_kstart,
                                                 const uint32_t* __restrict__ small_order,
                                                 const uint32_t* __restrict__ klist, uint32_t* __restrict__ k_rank,
                                                 uint32_t* __restrict__ k_state, T* __restrict__ out,
                                                 uint32_t fin_cap, uint32_t* __restrict__ err,
                                                 uint32_t part = 0) {
  using K = typename KeyTraits<T>::K;
  using H = typename HalfKey<K>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ SelectLds L;
  // finish order: [whole-workgroup segments | U_wave one-wave segments | U_wide (k_finish_wide)];
  // part 0 takes all of it, part 1 only the whole-workgroup segments of at
  // least kFinishSplitLen keys, part 2 the rest (see kFinishSplitMinSegs)
  const uint32_t U_small = ctl->U_small, U_wave = ctl->U_wave;
  const uint32_t U_grp = U_small - U_wave - ctl->U_wide;
  const uint32_t split = ctl->U_fin_big < U_grp ? ctl->U_fin_big : U_grp;
  const uint32_t jj_lo = part == 2 ? split : 0u, jj_hi = part == 1 ? split : U_grp;
  // the first digit's histogram starts cleared; each segment that counts
  // into it clears it again once done
  select_zero_first(L);
  __syncthreads();
  // The next segment's descriptor (and the order entry of the one after it)
  // is read while this one is finished, so a segment starts with its own
  // loads instead of two dependent round trips for where it lies (4-byte
  // keys: with 8-byte keys the held descriptor spills and costs more than
  // it saves).
  constexpr bool kPre = sizeof(K) == 4;
  struct Seg {
    uint32_t st, len, q0, q1;
    K klo, khi;
  };
  auto order_of = [&](uint32_t x) { return small_order ? small_order[x] : x; };
  auto seg_of = [&](uint32_t j) {
    Seg g;
    g.st = small_start[j];
    g.len = small_len[j];
    g.klo = small_klo[j];
    g.khi = small_khi[j];
    g.q0 = small_kstart[j];
    g.q1 = small_kstart[j + 1];
    return g;
  };
  const uint32_t gs = gridDim.x;
  Seg sn{};
  uint32_t jn2 = 0;
  if (kPre) {
    if (jj_lo + blockIdx.x < jj_hi) sn = seg_of(order_of(jj_lo + blockIdx.x));
    if (jj_lo + blockIdx.x + gs < jj_hi) jn2 = order_of(jj_lo + blockIdx.x + gs);
  }
  for (uint32_t jj = jj_lo + blockIdx.x; jj < jj_hi; jj += gs) {
    const Seg g = kPre ? sn : seg_of(order_of(jj));
    if (kPre) {
      if (jj + gs < jj_hi) sn = seg_of(jn2);
      if (jj + 2 * gs < jj_hi) jn2 = order_of(jj + 2 * gs);
    }
    const uint32_t st = g.st, len = g.len;
    if (len > fin_cap) {  // cannot happen: such segments are ordered into k_finish_wide's range; reported
      if (threadIdx.x == 0) *err = 1u;  // a flag word in host-visible memory
      continue;
    }
    const K klo = g.klo, khi = g.khi;
    const uint32_t q0 = g.q0, q1 = g.q1;
    if (!finish_fits_half<K>(len, klo, khi, fin_cap)) continue;  // k_finish_wide's
    // the ks and their ranks are fetched while the segment loads
    uint32_t kk = 0, kr = 0;
    if (threadIdx.x < q1 - q0 && q1 - q0 <= kSelectMaxKs) {
      kk = klist[q0 + threadIdx.x];
      kr = k_rank[kk];
    }
    // The segment is loaded into LDS as C values (key - base).  With at most
    // kSelectMaxKs ks the load also counts the radix select's first digit,
    // and stores the keys only when later digits need them.
    auto run = [&](auto* sv, auto clo, auto chi, K base) {
      using C = typename std::remove_pointer<decltype(sv)>::type;
      const bool sel = q1 - q0 <= kSelectMaxKs;
      uint32_t sh1, w1;
      select_first_digit(bitlen64((uint64_t)(clo ^ chi)), sh1, w1);
      const bool keep = !sel || sh1 > 0;
      load_segment<K>(buf, st, len, [&](uint32_t i, K key, bool in) {
        const C c = (C)(key - base);
        if (in && keep) sv[i] = c;
        if (sel) select_count_first<C>(L.hist, c, sh1, w1, in);
      });
      finish_segment<T, C>(sv, len, clo, chi, base, q0, q1, kk, kr, klist, k_rank, k_state, out, L, sel);
      if (sel) select_zero_first(L);  // for the next segment (the barrier below orders it)
    };
    if (len > fin_cap / 2) run(reinterpret_cast<H*>(smem), (H)0, (H)(khi - klo), klo);  // narrow interval
    else run(reinterpret_cast<K*>(smem), klo, khi, (K)0);
    __syncthreads();
  }
  // segments of at most kFinishWaveLen keys and kFinishWaveKs ks: one wave
  // each, radix-selected from global memory (L2) with a 256-bin histogram of
  // the wave's own in LDS, no workgroup barrier
  if (U_wave == 0 || part == 1) return;
  const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t* wh = reinterpret_cast<uint32_t*>(smem) + w * 256;
  uint32_t* bc = reinterpret_cast<uint32_t*>(smem) + nw * 256 + w * 2;
  for (uint32_t jj = U_grp + blockIdx.x * nw + w; jj < U_grp + U_wave; jj += gridDim.x * nw) {
    const uint32_t j = small_order[jj];
    const uint32_t st = small_start[j], len = small_len[j];
    const K klo = small_klo[j], khi = small_khi[j];
    const uint32_t q0 = small_kstart[j], q1 = small_kstart[j + 1];
    for (uint32_t qq = q0; qq < q1; ++qq) {
      const uint32_t k = klist[qq];
      const K key = wave_radix_select<K>(buf + st, len, klo, khi, k_rank[k], wh, bc);
      if (lane_id() == 0) {
        out[k] = KeyTraits<T>::dec(key);
        k_state[k] = kDone;
      }
    }
  }
}

// Stable LSD radix sort of a[0, len) by ONE workgroup in global memory (8-bit
// digits over bits [0, vary) -- the bits where the segment's keys differ),
// ping-ponging between a and tmp.  Returns the buffer holding the sorted keys.
// Each digit pass: an LDS histogram of the digit, its exclusive scan, then the
// keys in index order, one per thread per tile: a key's place among equal
// digits of its wave comes from eight ballots (lanes with the same digit), the
// waves' counts per digit are ranked in wave order in LDS, so equal digits
// keep their order (stable) and every key lands at
// base[digit] + (keys of that digit in earlier tiles and waves) + rank in wave.
struct RadixSortLds {
  uint32_t base[256];
  uint32_t wcnt[16][256];  // per wave: keys of each digit in the current tile, then their first position
};

template <class K>
__device__ const K* wg_radix_sort_global(K* __restrict__ a, K* __restrict__ tmp, uint32_t len, uint32_t vary,
                                         RadixSortLds& L) {
  const uint32_t nt = blockDim.x, t = threadIdx.x, w = t >> 6, nw = nt >> 6;
  K* src = a;
  K* dst = tmp;
  for (uint32_t shift = 0; shift < vary; shift += 8) {
    for (uint32_t i = t; i < 256; i += nt) L.base[i] = 0;
    __syncthreads();
    for (uint32_t i = t; i < len; i += nt) atomicAdd(&L.base[(uint32_t)(src[i] >> shift) & 255u], 1u);
    __syncthreads();
    if (t < 64) {  // exclusive scan of the 256 digit totals, four per lane
      const uint32_t c0 = L.base[4 * t], c1 = L.base[4 * t + 1], c2 = L.base[4 * t + 2], c3 = L.base[4 * t + 3];
      const uint32_t sum = c0 + c1 + c2 + c3;
      const uint32_t ex = wave_incl_scan_u32(sum) - sum;
      L.base[4 * t] = ex;
      L.base[4 * t + 1] = ex + c0;
      L.base[4 * t + 2] = ex + c0 + c1;
      L.base[4 * t + 3] = ex + c0 + c1 + c2;
    }
    __syncthreads();
    for (uint32_t t0 = 0; t0 < len; t0 += nt) {
      for (uint32_t i = t; i < nw * 256; i += nt) L.wcnt[i >> 8][i & 255] = 0;
      __syncthreads();
      const uint32_t i = t0 + t;
      const bool valid = i < len;
      const K key = valid ? src[i] : (K)0;
      const uint32_t d = (uint32_t)(key >> shift) & 255u;
      uint64_t peers = wave_ballot(valid);
#pragma unroll
      for (uint32_t b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = wave_ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t rank = mask_below(peers);
      if (valid && rank == 0) L.wcnt[w][d] = (uint32_t)__builtin_popcountll(peers);
      __syncthreads();
      for (uint32_t dd = t; dd < 256; dd += nt) {  // positions in wave order, then the next tile's base
        uint32_t off = L.base[dd];
        for (uint32_t v = 0; v < nw; ++v) {
          const uint32_t c = L.wcnt[v][dd];
          L.wcnt[v][dd] = off;
          off += c;
        }
        L.base[dd] = off;
      }
      __syncthreads();
      if (valid) dst[L.wcnt[w][d] + rank] = key;
      __syncthreads();  // every position read before the next tile clears them
    }
    K* sw = src;
    src = dst;
    dst = sw;
  }
  return src;
}

// Small segments too wide for half-width offsets and too large for full keys
// in the finish kernel's LDS (rare: a few per call at most), and segments
// larger than the finish capacity (the one-pass bucketMultiselect, or a pass
// cap that forced finishing), from global memory: up to kSelectMaxKs ks by
// radix select (one sweep per digit for all of them); more ks (a k-dense
// segment, e.g. the top segment of a top-k select) by a stable radix sort of
// the segment through the same positions of the spare buffer `tmp`, read at
// every rank.  They are the last U_wide entries of the finish order; one
// workgroup per segment.
template <class T>
__global__ void __launch_bounds__(1024) k_finish_wide(typename KeyTraits<T>::K* __restrict__ buf,
                                                      typename KeyTraits<T>::K* __restrict__ tmp,
                                                      const PassCtl* __restrict__ ctl,
                                                      const uint32_t* __restrict__ small_start,
                                                      const uint32_t* __restrict__ small_len,
                                                      const typename KeyTraits<T>::K* __restrict__ small_klo,
                                                      const typename KeyTraits<T>::K* __restrict__ small_khi,
                                                      const uint32_t* __restrict__ small_kstart,
                                                      const uint32_t* __restrict__ small_order,
                                                      const uint32_t* __restrict__ klist, uint32_t* __restrict__ k_rank,
                                                      uint32_t* __restrict__ k_state, T* __restrict__ out,
                                                      uint32_t fin_cap) {
  using K = typename KeyTraits<T>::K;
  __shared__ SelectLds L;
  __shared__ RadixSortLds RL;
  const uint32_t U_small = ctl->U_small, U_wide = ctl->U_wide;
  for (uint32_t jj = U_small - U_wide + blockIdx.x; jj < U_small; jj += gridDim.x) {
    const uint32_t j = small_order ? small_order[jj] : jj;
    const uint32_t st = small_start[j], len = small_len[j];
    const K klo = small_klo[j], khi = small_khi[j];
    if (!finish_wide_segment<K>(len, klo, khi, fin_cap)) continue;  // k_finish's
    const uint32_t q0 = small_kstart[j], q1 = small_kstart[j + 1];
    if (q1 - q0 > kSelectMaxKs && tmp) {
      const K* sorted = wg_radix_sort_global<K>(buf + st, tmp + st, len, bitlen64((uint64_t)(klo ^ khi)), RL);
      for (uint32_t qq = q0 + threadIdx.x; qq < q1; qq += blockDim.x) {
        const uint32_t k = klist[qq];
        out[k] = KeyTraits<T>::dec(sorted[k_rank[k] - 1]);
        k_state[k] = kDone;
      }
      __syncthreads();
      continue;
    }
    for (uint32_t c0 = q0; c0 < q1; c0 += kSelectMaxKs) {
      const uint32_t c1 = q1 - c0 < kSelectMaxKs ? q1 : c0 + kSelectMaxKs;
      uint32_t kk = 0, kr = 0;
      if (threadIdx.x < c1 - c0) {
        kk = klist[c0 + threadIdx.x];
        kr = k_rank[kk];
      }
      finish_segment<T, K>(buf + st, len, klo, khi, (K)0, c0, c1, kk, kr, klist, k_rank, k_state, out, L);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Tiny inputs (n <= finish capacity): one segment holding everything.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_tiny_setup(const T* __restrict__ x, uint32_t n, uint32_t nk, typename KeyTraits<T>::K* __restrict__ buf,
                             uint32_t* __restrict__ k_rank, uint32_t* __restrict__ k_state,
                             uint32_t* __restrict__ small_start, uint32_t* __restrict__ small_len,
                             typename KeyTraits<T>::K* __restrict__ small_klo, typename KeyTraits<T>::K* __restrict__ small_khi,
                             uint32_t* __restrict__ small_kstart, uint32_t* __restrict__ klist, PassCtl* ctl,
                             uint32_t wide) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) buf[i] = KeyTraits<T>::enc(x[i]);
  __shared__ uint32_t scratch[32];
  uint32_t cnt = 0;
  for (uint32_t k = threadIdx.x; k < nk; k += blockDim.x) cnt += (k_state[k] == kLive);
  uint32_t tot;
  uint32_t off = block_excl_scan_u32(cnt, scratch, &tot);
  for (uint32_t k = threadIdx.x; k < nk; k += blockDim.x)
    if (k_state[k] == kLive) {
      klist[off++] = k;
      k_state[k] = kSmall;
    }
  if (threadIdx.x == 0) {
    small_start[0] = 0;
    small_len[0] = n;
    small_klo[0] = 0;
    small_khi[0] = KeyMax<typename KeyTraits<T>::K>::v;
    small_kstart[0] = 0;
    small_kstart[1] = tot;
    ctl->U_small = 1;
    ctl->U_big = 0;
    ctl->N_big = 0;
    ctl->U = 1;
    ctl->U_wide = wide;
  }
}

}  // namespace ggms
