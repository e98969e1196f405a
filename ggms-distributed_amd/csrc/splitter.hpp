// First-pass splitter: sample the input, sort the sample in LDS, place pivots
// and turn them into a partition of the whole key space into linear-slope
// ranges.
//
// Reference: generatePivots (DistributedSMOS/iterativeSMOS.cu:1011-1054) draws
// 1024 random elements, sorts them and takes p0 = min, p1 = s[21],
// p_i = s[70(i-1)+21], p15 = s[1001], p16 = max, each of the 16 ranges getting
// 512 of 8192 buckets and slope 512 / (p_{i+1} - p_i) (iterativeSMOS.cu:1258-1270).
// The one-GPU share of distributedSMOS takes a strided sample instead
// (DistributedSMOS/distributedSMOS_Kernel.cu:369-378) and re-samples it with
// generatePivots (DistributedSMOS/distributedSMOS.cpp:450-451).
//
// Differences, all of them exactness-neutral (any monotone partition gives
// exact results; only the amount of work changes):
//   * pivots and slopes live in key space and the slope is fixed point;
//   * the sample extremes replace the global min/max (no extra full pass):
//     keys below the sample minimum / above the sample maximum fall into two
//     edge ranges, and keys outside [key(-inf), key(+inf)] (NaNs) into two
//     one-bucket ranges that the plan kernel reports;
//   * a pivot value that repeats (a value holding >= 1/16 of the sample)
//     becomes its own one-key range, so a dominant duplicate is resolved
//     without being copied.
#pragma once
#include "select_kernels.hpp"

namespace ggms {

// MLP layout for the learned splitter (all fp32, row-major):
//   W1[32][65], b1[32], W2[32], b2[1]  -> 2145 floats
constexpr int kMlpIn = 65, kMlpHidden = 32, kMlpWin = 64;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden + 1;

// LDS of k_plan0: sorted sample, pivots (<= 2048), range starts, requested and
// final bucket counts (at most 2 * 1025 + 6 ranges).
constexpr uint32_t kPlan0MaxRanges = 2064;
template <class K> __host__ __device__ constexpr size_t plan0_lds_bytes(uint32_t S_pad) {
  return (size_t)S_pad * sizeof(K) + 2048 * sizeof(K) + kPlan0MaxRanges * sizeof(K) +
         (kPlan0MaxRanges * 8 > kMlpParams * 4 ? kPlan0MaxRanges * 8 : kMlpParams * 4);
}

template <class K>
__device__ __forceinline__ double key_to_double(K k);
template <>
__device__ __forceinline__ double key_to_double<uint32_t>(uint32_t k) { return (double)k; }
template <>
__device__ __forceinline__ double key_to_double<uint64_t>(uint64_t k) { return (double)k; }

// Learned pivot: the MLP reads a window of 64 consecutive sorted sample keys
// centred on the target quantile, normalised to [0, 1] by the window's own
// extremes, plus the target's fractional position inside the window, and
// predicts where between the two order statistics that bracket the target
// the true quantile lies (sigmoid output).  The prediction therefore always
// falls between those two sample keys.
template <class K>
__device__ K learned_pivot(const K* s, uint32_t S, double rank_frac, const float* w) {
  double pos = rank_frac * (double)(S - 1);
  int ip = (int)pos;
  if (ip >= (int)S - 1) ip = (int)S - 2;
  if (ip < 0) ip = 0;
  double fr = pos - (double)ip;
  int w0 = ip - kMlpWin / 2 + 1;
  if (w0 < 0) w0 = 0;
  if (w0 + kMlpWin > (int)S) w0 = (int)S - kMlpWin;
  K klo = s[w0], khi = s[w0 + kMlpWin - 1];
  double span = key_to_double<K>(khi - klo);
  float in[kMlpIn];
  for (int i = 0; i < kMlpWin; ++i)
    in[i] = span > 0 ? (float)(key_to_double<K>(s[w0 + i] - klo) / span) : 0.0f;
  in[kMlpWin] = (float)(((double)(ip - w0) + fr) / (double)(kMlpWin - 1));
  const float* W1 = w;
  const float* b1 = w + kMlpHidden * kMlpIn;
  const float* W2 = b1 + kMlpHidden;
  const float b2 = W2[kMlpHidden];
  float o = b2;
  for (int h = 0; h < kMlpHidden; ++h) {
    float z = b1[h];
    for (int i = 0; i < kMlpIn; ++i) z += W1[h * kMlpIn + i] * in[i];
    o += W2[h] * (z > 0.0f ? z : 0.0f);
  }
  float f = 1.0f / (1.0f + __expf(-o));
  K a = s[ip], b = s[ip + 1];
  double d = key_to_double<K>(b - a) * (double)f;
  K off = (K)d;
  if (off > (K)(b - a)) off = (K)(b - a);
  return (K)(a + off);
}

template <class T>
__global__ void __launch_bounds__(1024) k_plan0(const T* __restrict__ x, Plan0Args a,
                                                Range<typename KeyTraits<T>::K>* __restrict__ ranges,
                                                typename KeyTraits<T>::K* __restrict__ lut,
                                                PassCtl* __restrict__ ctl) {
  using K = typename KeyTraits<T>::K;
  using KT = KeyTraits<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t bcast_R;
  const uint32_t n = ctl->n;
  const uint32_t S = a.S;
  const uint32_t Spad = pow2ceil_u32(S);
  K* s = reinterpret_cast<K*>(smem);
  K* q = s + Spad;  // pivot keys, up to 2048, then the range list
  const K lowest = KT::kFloat ? KT::kNegInf : (K)0;
  const K highest = KT::kFloat ? KT::kPosInf : KeyMax<K>::v;

  // 1. sample
  for (uint32_t i = threadIdx.x; i < Spad; i += blockDim.x) {
    K key = KeyMax<K>::v;
    if (i < S) {
      uint32_t pos;
      if (a.mode == kModeStrided) {
        uint32_t S1 = a.strideS;
        uint32_t stride = n / S1 > 0 ? n / S1 : 1;
        uint32_t j = (uint32_t)(((uint64_t)philox_draw(a.seed, kStreamSubSample, i).v[0] * S1) >> 32);
        uint64_t p = (uint64_t)j * stride;
        pos = p < n ? (uint32_t)p : n - 1;
      } else {
        pos = (uint32_t)(((uint64_t)philox_draw(a.seed, kStreamSample, i).v[0] * n) >> 32);
      }
      key = KT::enc(x[pos]);
      if (key < lowest) key = lowest;
      if (key > highest) key = highest;
    }
    s[i] = key;
  }
  __syncthreads();
  bitonic_sort_lds<K>(s, Spad);

  // 2. pivots
  uint32_t Q = 0;
  if (a.mode == kModeLinear || a.mode == kModeStrided) {
    const uint32_t P = a.P;
    Q = P;
    if (threadIdx.x == 0) {
      if (P <= 3) {
        q[0] = s[0];
        if (P == 3) q[1] = s[S / 2];
        q[P - 1] = s[S - 1];
      } else {
        uint32_t eo = (22u * S + 512u) / 1024u;
        if (eo < 1) eo = 1;
        if (2 * eo >= S) eo = S / 4 > 0 ? S / 4 : 1;
        uint32_t po = (S - 2 * eo) / (P - 3);
        q[0] = s[0];
        q[1] = s[eo - 1];
        for (uint32_t i = 2; i < P - 2; ++i) {
          uint32_t idx = po * (i - 1) + eo - 1;
          q[i] = s[idx < S ? idx : S - 1];
        }
        q[P - 2] = s[S - eo - 1];
        q[P - 1] = s[S - 1];
      }
    }
  } else {
    const uint32_t Pl = a.P;  // pieces
    Q = Pl + 1;
    // network weights staged in LDS (every thread reads the same weight at the
    // same time: broadcast reads)
    float* wl = reinterpret_cast<float*>(q + 2048 + kPlan0MaxRanges);
    if (a.mode == kModeLearned) {
      for (uint32_t i = threadIdx.x; i < (uint32_t)kMlpParams; i += blockDim.x) wl[i] = a.mlp[i];
      __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j <= Pl; j += blockDim.x) {
      K v;
      if (j == 0) v = s[0];
      else if (j == Pl) v = s[S - 1];
      else if (a.mode == kModeLearned) v = learned_pivot<K>(s, S, (double)j / (double)Pl, wl);
      else {
        double pos = (double)j / (double)Pl * (double)(S - 1);
        v = s[(uint32_t)(pos + 0.5)];
      }
      q[j] = v;
    }
    __syncthreads();
    // predictions may be out of order by a key or two; sort them
    uint32_t Qp = pow2ceil_u32(Q);
    for (uint32_t j = Q + threadIdx.x; j < Qp; j += blockDim.x) q[j] = KeyMax<K>::v;
    __syncthreads();
    bitonic_sort_lds<K>(q, Qp);
  }
  __syncthreads();

  // 3. range list: starts and requested bucket counts, emitted in order by one
  //    thread into LDS (cheap: no map arithmetic), then finished in parallel.
  K* rlo = q + 2048;                                                  // [kPlan0MaxRanges]
  // the network weights (only used before this point) share this space
  uint32_t* rreq = reinterpret_cast<uint32_t*>(rlo + kPlan0MaxRanges);  // [kPlan0MaxRanges]
  uint32_t* rnb = rreq + kPlan0MaxRanges;                             // [kPlan0MaxRanges]
  if (threadIdx.x == 0) {
    enum { kSingle = 0, kEdge = 1, kInterior = 2, kNan = 3 };
    uint32_t m_int = 0;
    uint32_t R = 0;
    uint32_t share = 1;
    for (int emit = 0; emit < 2; ++emit) {
      auto cb = [&](K lo, int kind) {
        if (!emit) {
          if (kind == kInterior) m_int++;
          return;
        }
        rlo[R] = lo;
        rreq[R] = kind == kInterior ? share : kind == kEdge ? a.E : 1u;
        R++;
      };
      if (KT::kFloat) cb((K)0, kNan);
      if (q[0] > lowest) cb(lowest, kEdge);
      uint32_t i = 0;
      while (i < Q) {
        K v = q[i];
        uint32_t j = i;
        while (j + 1 < Q && q[j + 1] == v) ++j;
        if (j + 1 >= Q) {
          cb(v, kSingle);
          break;
        }
        K nxt = q[j + 1];
        if (j > i) {
          cb(v, kSingle);
          if ((K)(nxt - v) >= 2) cb((K)(v + 1), kInterior);
        } else {
          cb(v, kInterior);
        }
        i = j + 1;
      }
      if (q[Q - 1] < highest) cb((K)(q[Q - 1] + 1), kEdge);
      if (KT::kFloat) cb((K)(highest + 1), kNan);
      if (!emit) share = m_int ? (a.B / m_int > 0 ? a.B / m_int : 1) : 1;
    }
    bcast_R = R;
  }
  __syncthreads();
  const uint32_t R = bcast_R;
  // 4. map parameters (parallel), bucket bases (block scan), global copy
  __shared__ uint32_t scan_tmp[32];
  uint32_t per = (R + blockDim.x - 1) / blockDim.x;
  uint32_t r0 = threadIdx.x * per, r1 = r0 + per < R ? r0 + per : R;
  uint32_t mysum = 0;
  for (uint32_t r = r0; r < r1; ++r) {
    Range<K> rr;
    rr.lo = rlo[r];
    rr.hi = r + 1 < R ? (K)(rlo[r + 1] - 1) : KeyMax<K>::v;
    make_map<K>(rr, rreq[r]);
    rnb[r] = rr.nb;
    mysum += rr.nb;
  }
  uint32_t Btot;
  uint32_t off = block_excl_scan_u32(mysum, scan_tmp, &Btot);
  for (uint32_t r = r0; r < r1; ++r) {
    Range<K> rr;
    rr.lo = rlo[r];
    rr.hi = r + 1 < R ? (K)(rlo[r + 1] - 1) : KeyMax<K>::v;
    make_map<K>(rr, rreq[r]);
    rr.base = off;
    rr.seg_start = 0;
    rr.seg_end = 0;
    off += rr.nb;
    ranges[r] = rr;
  }
  if (threadIdx.x == 0) {
    ctl->R = R;
    ctl->Rp = pow2ceil_u32(R);
    ctl->B = Btot;
    ctl->nan_lo_bucket = KT::kFloat ? 0u : kNone32;
    ctl->nan_hi_bucket = KT::kFloat ? Btot - 1 : kNone32;
  }
  // 5. key-prefix cell table, from the LDS copy of the range starts
  build_lut<K>(rlo, R, lut);
}

}  // namespace ggms
