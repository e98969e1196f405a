// Device-side building blocks shared by the selection kernels:
//   * order-preserving integer keys for every value type,
//   * the fixed-point linear-slope bucket map used by every splitter range,
//   * wave64 helpers (ballot / lane masks) for aggregated LDS atomics.
//
// Selection is ordering work, so every comparison and every bucket decision is
// made on integer keys: the map  value -> key  is strictly monotone, which makes
// the whole pipeline exact by construction (no float slope arithmetic on values).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GGMS_HD __host__ __device__ __forceinline__

namespace ggms {

// ---------------------------------------------------------------------------
// Keys.  fp32 -> u32, fp64 -> u64: flip the sign bit of non-negative values,
// flip all bits of negative values.  -0.0 orders just below +0.0; NaNs land
// outside [key(-inf), key(+inf)] and are reported, never selected.
// ---------------------------------------------------------------------------
template <class T> struct KeyTraits;

template <> struct KeyTraits<float> {
  using K = uint32_t;
  static constexpr bool kFloat = true;
  static GGMS_HD K enc(float v) {  // negative: ~u; otherwise u | sign bit
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
  }
  static GGMS_HD float dec(K k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
  }
  static constexpr K kNegInf = 0x007FFFFFu;  // enc(-inf)
  static constexpr K kPosInf = 0xFF800000u;  // enc(+inf)
};

template <> struct KeyTraits<double> {
  using K = uint64_t;
  static constexpr bool kFloat = true;
  static GGMS_HD K enc(double v) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    return u ^ ((uint64_t)((int64_t)u >> 63) | 0x8000000000000000ull);
  }
  static GGMS_HD double dec(K k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __builtin_bit_cast(double, u);
  }
  static constexpr K kNegInf = 0x000FFFFFFFFFFFFFull;
  static constexpr K kPosInf = 0xFFF0000000000000ull;
};

template <> struct KeyTraits<uint32_t> {
  using K = uint32_t;
  static constexpr bool kFloat = false;
  static GGMS_HD K enc(uint32_t v) { return v; }
  static GGMS_HD uint32_t dec(K k) { return k; }
  static constexpr K kNegInf = 0u;
  static constexpr K kPosInf = 0xFFFFFFFFu;
};

template <> struct KeyTraits<int32_t> {
  using K = uint32_t;
  static constexpr bool kFloat = false;
  static GGMS_HD K enc(int32_t v) { return __builtin_bit_cast(uint32_t, v) ^ 0x80000000u; }
  static GGMS_HD int32_t dec(K k) { return __builtin_bit_cast(int32_t, k ^ 0x80000000u); }
  static constexpr K kNegInf = 0u;
  static constexpr K kPosInf = 0xFFFFFFFFu;
};

template <class K> struct KeyMax;
template <> struct KeyMax<uint32_t> { static constexpr uint32_t v = 0xFFFFFFFFu; };
template <> struct KeyMax<uint64_t> { static constexpr uint64_t v = 0xFFFFFFFFFFFFFFFFull; };

// ---------------------------------------------------------------------------
// One range of a splitter: keys [lo, hi] spread over nb buckets starting at
// global bucket `base`.  The slope is fixed point:
//     off(key) = mult ? mulhi32((key - lo) >> sh, mult) : (key - lo)
// with sh chosen so (hi - lo) >> sh fits 32 bits and
//     mult = floor(nb * 2^32 / (((hi - lo) >> sh) + 1))  (< 2^32),
// which keeps off in [0, nb) and makes off non-decreasing in key.
// mult == 0 marks the identity map used when the range holds <= nb keys.
// For segmented passes seg_start/seg_end give the range's slice of the pass
// buffer (its elements are stored contiguously there).
// ---------------------------------------------------------------------------
template <class K> struct Range {
  K lo;
  K hi;
  uint32_t base;
  uint32_t nb;
  uint32_t sh;
  uint32_t mult;
  uint32_t seg_start;
  uint32_t seg_end;
};

GGMS_HD uint32_t bitlen64(uint64_t w) { return w ? 64u - (uint32_t)__builtin_clzll(w) : 0u; }

// End of the tile starting at t0 (t0 < n): min(t0 + tile, n) without forming
// t0 + tile, which could pass 2^32 on the last tile.
GGMS_HD uint32_t tile_end(uint32_t t0, uint32_t tile, uint32_t n) { return n - t0 > tile ? t0 + tile : n; }

// Fill nb/sh/mult of r for the requested bucket count (>= 1).
template <class K> GGMS_HD void make_map(Range<K>& r, uint32_t nb_req) {
  uint64_t w = (uint64_t)(r.hi - r.lo);
  if (nb_req < 1) nb_req = 1;
  if (w < (uint64_t)nb_req) {  // at most nb_req keys: one bucket per key
    r.nb = (uint32_t)w + 1u;
    r.sh = 0;
    r.mult = 0;
    return;
  }
  uint32_t bl = bitlen64(w);
  uint32_t sh = bl > 32 ? bl - 32 : 0;
  uint64_t wt = w >> sh;  // <= 2^32 - 1, and > nb_req - 1
  r.nb = nb_req;
  r.sh = sh;
  r.mult = (uint32_t)(((uint64_t)nb_req << 32) / (wt + 1));
}

template <class K> GGMS_HD uint32_t map_off(K key, K lo, uint32_t sh, uint32_t mult) {
  uint32_t t = (uint32_t)((K)(key - lo) >> sh);
  return mult ? (uint32_t)(((uint64_t)t * (uint64_t)mult) >> 32) : t;
}

// First key of bucket `off` of range r (off < r.nb); may exceed r.hi for an
// empty trailing bucket.
template <class K> GGMS_HD uint64_t bucket_first_key_off(const Range<K>& r, uint32_t off) {
  if (r.mult == 0) return (uint64_t)off;
  uint64_t t = (((uint64_t)off << 32) + (uint64_t)r.mult - 1ull) / (uint64_t)r.mult;
  return t << r.sh;  // offset from lo, < 2^64 since t <= 2^32
}

// Key interval [klo, khi] of bucket `off` of range r.  Returns false if empty.
template <class K> GGMS_HD bool bucket_interval(const Range<K>& r, uint32_t off, K& klo, K& khi) {
  uint64_t w = (uint64_t)(r.hi - r.lo);
  uint64_t a = bucket_first_key_off(r, off);
  if (a > w) return false;
  uint64_t b = (off + 1 < r.nb) ? bucket_first_key_off(r, off + 1) : w + 1;  // exclusive, may be 2^64 for u64
  uint64_t last = (off + 1 < r.nb) ? (b - 1) : w;
  if (last > w) last = w;
  if (off + 1 < r.nb && b == 0) last = w;  // wrapped: only possible when w == 2^64-1
  klo = (K)(r.lo + (K)a);
  khi = (K)(r.lo + (K)last);
  return klo <= khi;
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mask_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }
// Orders the LDS accesses of the lanes of one wave around this point (the
// wave's LDS instructions execute in order; this keeps the compiler from
// moving them across it and waits for the ones issued).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Duplicate peeling: when at least kPeelMin active lanes of a wave share the
// first active lane's bucket, that bucket takes one atomic for all of them, so
// a dominant value costs one LDS atomic instead of 64 serialised ones.  The
// test is a single compare + ballot and the branch is wave-uniform, so random
// data pays almost nothing for it.
constexpr uint32_t kPeelMin = 16;

// hist[b] += 1 for every active lane with `active` set.
__device__ __forceinline__ void lds_count(uint32_t* hist, uint32_t b, bool active = true) {
  uint64_t am = wave_ballot(active);
  if (am == 0) return;
  uint32_t first = (uint32_t)__builtin_ctzll(am);
  uint32_t b0 = readlane_u32(b, first);
  bool same = active && (b == b0);
  uint64_t m0 = wave_ballot(same);
  uint32_t c0 = (uint32_t)__builtin_popcountll(m0);
  if (c0 >= kPeelMin) {
    // b == b0 in lane `first`; the lane's own b keeps the address divergent
    // (see lds_claim)
    if (lane_id() == first) atomicAdd(&hist[b], c0);
    if (active && !same) atomicAdd(&hist[b], 1u);
  } else if (active) {
    atomicAdd(&hist[b], 1u);
  }
}

// pos = cursor[s]++ for every active lane (LDS), with the same peeling.
__device__ __forceinline__ uint32_t lds_claim(uint32_t* cursor, uint32_t s, bool active) {
  uint64_t am = wave_ballot(active);
  if (am == 0) return 0;
  uint32_t first = (uint32_t)__builtin_ctzll(am);
  uint32_t s0 = readlane_u32(s, first);
  bool same = active && (s == s0);
  uint64_t m0 = wave_ballot(same);
  uint32_t c0 = (uint32_t)__builtin_popcountll(m0);
  uint32_t pos = 0;
  if (c0 >= kPeelMin) {
    uint32_t base0 = 0;
    // s == s0 in lane `first`.  Through the lane's own s the address is
    // divergent, so the compiler does not wrap this one-lane atomic in the
    // whole-wave reduction it builds for uniform addresses (count the active
    // lanes, elect one, broadcast), which costs ~5 VALU and ~5 SALU each time.
    if (lane_id() == first) base0 = atomicAdd(&cursor[s], c0);
    base0 = readlane_u32(base0, first);
    if (same) pos = base0 + mask_below(m0);
    else if (active) pos = atomicAdd(&cursor[s], 1u);
  } else if (active) {
    pos = atomicAdd(&cursor[s], 1u);
  }
  return pos;
}

}  // namespace ggms
