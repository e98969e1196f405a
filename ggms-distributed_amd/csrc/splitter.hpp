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
// learned pivots are evaluated in chunks of this many (features of a chunk in LDS)
constexpr uint32_t kMlpChunk = 128;
constexpr uint32_t kMlpFeatStride = kMlpIn + 1;
template <class K> __host__ __device__ constexpr size_t plan0_lds_core(uint32_t S_pad) {
  return (size_t)S_pad * sizeof(K) + 2048 * sizeof(K) + kPlan0MaxRanges * sizeof(K) +
         (kPlan0MaxRanges * 8 > (kMlpParams + kMlpChunk * kMlpFeatStride) * 4
              ? kPlan0MaxRanges * 8
              : (kMlpParams + kMlpChunk * kMlpFeatStride) * 4);
}
// plus the cell-table counters
template <class K> __host__ __device__ constexpr size_t plan0_lds_bytes(uint32_t S_pad) {
  return plan0_lds_core<K>(S_pad) + (size_t)lut_count_words<K>() * 4;
}

template <class K>
__device__ __forceinline__ double key_to_double(K k);
template <>
__device__ __forceinline__ double key_to_double<uint32_t>(uint32_t k) { return (double)k; }
template <>
__device__ __forceinline__ double key_to_double<uint64_t>(uint64_t k) { return (double)k; }

// Learned pivots: the MLP reads a window of 64 consecutive sorted sample keys
// centred on the target quantile, normalised to [0, 1] by the window's own
// extremes, plus the target's fractional position inside the window, and
// predicts where between the two order statistics that bracket the target
// the true quantile lies (sigmoid output).  The prediction therefore always
// falls between those two sample keys.
// All learned pivots: q[0] = s[0], q[Pl] = s[S-1] and, for j = 1 .. Pl-1,
// the network's pivot for rank fraction j / Pl.  Pivots go in chunks: the
// chunk's features are built in LDS (one thread per feature), then one thread
// per (pivot, hidden unit) evaluates a hidden unit and the 32 units of a pivot
// are summed across lanes.  feat: kMlpChunk * kMlpFeatStride floats of LDS.
template <class K>
__device__ __forceinline__ void learned_window(uint32_t j, uint32_t Pl, uint32_t S, double& pos, int& ip, int& w0) {
  pos = (double)j / (double)Pl * (double)(S - 1);
  ip = (int)pos;
  if (ip >= (int)S - 1) ip = (int)S - 2;
  if (ip < 0) ip = 0;
  w0 = ip - kMlpWin / 2 + 1;
  if (w0 < 0) w0 = 0;
  if (w0 + kMlpWin > (int)S) w0 = (int)S - kMlpWin;
}

template <class K>
__device__ void learned_pivots(const K* s, uint32_t S, uint32_t Pl, const float* w, K* q, float* feat) {
  const float* W1 = w;
  const float* b1 = w + kMlpHidden * kMlpIn;
  const float* W2 = b1 + kMlpHidden;
  const float b2 = W2[kMlpHidden];
  if (threadIdx.x == 0) {
    q[0] = s[0];
    q[Pl] = s[S - 1];
  }
  for (uint32_t c0 = 1; c0 < Pl; c0 += kMlpChunk) {
    const uint32_t nc = Pl - c0 < kMlpChunk ? Pl - c0 : kMlpChunk;
    for (uint32_t t = threadIdx.x; t < nc * kMlpIn; t += blockDim.x) {
      const uint32_t pj = t / kMlpIn, i = t % kMlpIn;
      double pos;
      int ip, w0;
      learned_window<K>(c0 + pj, Pl, S, pos, ip, w0);
      float f;
      if (i < (uint32_t)kMlpWin) {
        K klo = s[w0], khi = s[w0 + kMlpWin - 1];
        double span = key_to_double<K>(khi - klo);
        f = span > 0 ? (float)(key_to_double<K>(s[w0 + i] - klo) / span) : 0.0f;
      } else {
        f = (float)(((double)(ip - w0) + (pos - (double)ip)) / (double)(kMlpWin - 1));
      }
      feat[pj * kMlpFeatStride + i] = f;
    }
    __syncthreads();
    const uint32_t work = (nc * kMlpHidden + blockDim.x - 1) / blockDim.x * blockDim.x;
    for (uint32_t t = threadIdx.x; t < work; t += blockDim.x) {
      const uint32_t pj = t / kMlpHidden, h = t % kMlpHidden;
      float contrib = 0.0f;
      if (pj < nc) {
        const float* in = feat + pj * kMlpFeatStride;
        float z = b1[h];
        for (int i = 0; i < kMlpIn; ++i) z += W1[h * kMlpIn + i] * in[i];
        contrib = W2[h] * (z > 0.0f ? z : 0.0f);
      }
#pragma unroll
      for (int d = kMlpHidden / 2; d > 0; d >>= 1) contrib += __shfl_xor(contrib, d, kMlpHidden);
      if (pj < nc && h == 0) {
        const uint32_t j = c0 + pj;
        float f = 1.0f / (1.0f + __expf(-(b2 + contrib)));
        double pos;
        int ip, w0;
        learned_window<K>(j, Pl, S, pos, ip, w0);
        K lo = s[ip], hi = s[ip + 1];
        double d = key_to_double<K>(hi - lo) * (double)f;
        K off = (K)d;
        if (off > (K)(hi - lo)) off = (K)(hi - lo);
        q[j] = (K)(lo + off);
      }
    }
    __syncthreads();
  }
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
  // k_setup's work folded in (the host does it only for nk <= kSetupKsPerThread
  // * blockDim): control blocks now, the ks' reads issued now and used after
  // the sample gathers, so their round trip (host memory) overlaps those
  constexpr uint32_t kCtlWords = sizeof(PassCtl) / 4;
  uint32_t kv[kSetupKsPerThread];
  if (a.setup) {
    for (uint32_t i = threadIdx.x; i < a.ctl_words; i += blockDim.x) {
      if (i == 0) ctl[0] = a.c0;
      else if (i >= kCtlWords) reinterpret_cast<uint32_t*>(ctl)[i] = 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kSetupKsPerThread; ++u) {
      const uint32_t i = threadIdx.x + u * blockDim.x;
      kv[u] = i < a.nk ? a.ks[i] : 0u;
    }
  }
  const uint32_t n = a.setup ? a.c0.n : ctl->n;
  const uint32_t S = a.S;
  const uint32_t Spad = pow2ceil_u32(S);
  K* s = reinterpret_cast<K*>(smem);
  K* q = s + Spad;  // pivot keys, up to 2048, then the range list
  uint32_t* cells = reinterpret_cast<uint32_t*>(smem + plan0_lds_core<K>(Spad));  // cell-table counters
  const K lowest = KT::kFloat ? KT::kNegInf : (K)0;
  const K highest = KT::kFloat ? KT::kPosInf : KeyMax<K>::v;

  // the learned form's network weights go to LDS now, their reads overlapping
  // the sample gathers (the area is not used before the pivots)
  if (a.mode == kModeLearned) {
    float* wl0 = reinterpret_cast<float*>(q + 2048 + kPlan0MaxRanges);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kMlpParams; i += blockDim.x) wl0[i] = a.mlp[i];
  }

  // 1. sample: up to 8 positions per thread, all loads issued before the
  //    first use (random reads over the whole vector: each one is a full
  //    memory round trip, so they must overlap)
  constexpr uint32_t kSampleUnroll = 8;
  constexpr uint32_t kSampleRun = 16;
  for (uint32_t i0 = threadIdx.x; i0 < Spad; i0 += kSampleUnroll * blockDim.x) {
    uint32_t pos[kSampleUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kSampleUnroll; ++u) {
      uint32_t i = i0 + u * blockDim.x;
      uint32_t p = 0;
      if (i < S) {
        if (a.mode == kModeStrided) {
          uint32_t S1 = a.strideS;
          uint32_t stride = n / S1 > 0 ? n / S1 : 1;
          uint32_t j = (uint32_t)(((uint64_t)philox_draw(a.seed, kStreamSubSample, i).v[0] * S1) >> 32);
          uint64_t pp = (uint64_t)j * stride;
          p = pp < n ? (uint32_t)pp : n - 1;
        } else if (a.mode == kModeLearned || a.mode == kModeSketch) {
          // runs of kSampleRun consecutive elements from random starts: one
          // 64-byte read (and one address translation) per run instead of per
          // element; the linear form keeps independent positions
          const uint32_t c = i / kSampleRun;
          const uint32_t start =
              (uint32_t)(((uint64_t)philox_draw(a.seed, kStreamSample, c).v[0] * (n - kSampleRun + 1)) >> 32);
          p = start + i % kSampleRun;
        } else {
          p = (uint32_t)(((uint64_t)philox_draw(a.seed, kStreamSample, i).v[0] * n) >> 32);
        }
      }
      pos[u] = p;
    }
    T v[kSampleUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kSampleUnroll; ++u) v[u] = x[pos[u]];
#pragma unroll
    for (uint32_t u = 0; u < kSampleUnroll; ++u) {
      uint32_t i = i0 + u * blockDim.x;
      if (i >= Spad) break;
      K key = KeyMax<K>::v;
      if (i < S) {
        key = KT::enc(v[u]);
        if (key < lowest) key = lowest;
        if (key > highest) key = highest;
      }
      s[i] = key;
    }
  }
  if (a.setup) {
#pragma unroll
    for (uint32_t u = 0; u < kSetupKsPerThread; ++u) {
      const uint32_t i = threadIdx.x + u * blockDim.x;
      if (i >= a.nk) break;
      const uint32_t k = kv[u];
      if (k < 1 || k > n) {
        *a.bad = 1u;
        a.k_rank[i] = 1;
        a.k_state[i] = kDone;
      } else {
        a.k_rank[i] = n - k + 1;
        a.k_state[i] = kLive;
      }
    }
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
    // network weights, staged in LDS at the start (every thread reads the same
    // weight at the same time: broadcast reads)
    float* wl = reinterpret_cast<float*>(q + 2048 + kPlan0MaxRanges);
    if (a.mode == kModeLearned) {
      learned_pivots<K>(s, S, Pl, wl, q, reinterpret_cast<float*>(wl + kMlpParams));
    } else {
      for (uint32_t j = threadIdx.x; j <= Pl; j += blockDim.x) {
        K v;
        if (j == 0) v = s[0];
        else if (j == Pl) v = s[S - 1];
        else {
          double pos = (double)j / (double)Pl * (double)(S - 1);
          v = s[(uint32_t)(pos + 0.5)];
        }
        q[j] = v;
      }
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
  // Every run of equal pivots [i, j] emits, in pivot order: a one-key range
  // when it is the last run or a repeated value (then also the interior range
  // above it when the gap to the next pivot holds two or more keys), else one
  // interior range.  Runs are found in parallel (a run starts where the pivot
  // changes) and placed by a block scan; NaN and edge ranges bracket them.
  __shared__ uint32_t scan_tmp[32];
  // longest run of equal pivots (length - 1) << 16 | its first pivot: the
  // most frequent sample value, which the cell-table window puts on a cell edge
  __shared__ uint32_t hot_run;
  __shared__ uint32_t hot_bucket;  // bucket of a dominant value's one-key range (k_count0 counts it in registers)
  if (threadIdx.x == 0) {
    hot_run = 0;
    hot_bucket = kNone32;
  }
  __syncthreads();
  {
    enum : uint32_t { kSingle = 0, kEdge = 1, kInterior = 2, kNan = 3 };
    const uint32_t pre = (KT::kFloat ? 1u : 0u) + (q[0] > lowest ? 1u : 0u);
    const uint32_t per_q = (Q + blockDim.x - 1) / blockDim.x;
    const uint32_t i0 = threadIdx.x * per_q, i1 = i0 + per_q < Q ? i0 + per_q : Q;
    uint32_t cnt = 0, nint = 0;
    for (uint32_t i = i0; i < i1; ++i) {
      if (i > 0 && q[i] == q[i - 1]) continue;
      uint32_t j = i;
      while (j + 1 < Q && q[j + 1] == q[i]) ++j;
      if (j > i) atomicMax(&hot_run, ((j - i) << 16) | i);
      if (j + 1 >= Q) cnt += 1;
      else if (j > i) {
        cnt += 1;
        if ((K)(q[j + 1] - q[i]) >= 2) { cnt += 1; nint += 1; }
      } else { cnt += 1; nint += 1; }
    }
    uint32_t total, ints;
    uint32_t at = pre + block_excl_scan_u32(cnt, scan_tmp, &total);
    (void)block_excl_scan_u32(nint, scan_tmp, &ints);
    const uint32_t share = ints ? (a.B / ints > 0 ? a.B / ints : 1u) : 1u;
    for (uint32_t i = i0; i < i1; ++i) {
      if (i > 0 && q[i] == q[i - 1]) continue;
      const K v = q[i];
      uint32_t j = i;
      while (j + 1 < Q && q[j + 1] == v) ++j;
      if (j + 1 >= Q) {
        rlo[at] = v; rreq[at++] = 1;
      } else if (j > i) {
        rlo[at] = v; rreq[at++] = 1;
        if ((K)(q[j + 1] - v) >= 2) { rlo[at] = (K)(v + 1); rreq[at++] = share; }
      } else {
        rlo[at] = v; rreq[at++] = share;
      }
    }
    if (threadIdx.x == 0) {
      uint32_t r = 0;
      if (KT::kFloat) { rlo[r] = (K)0; rreq[r++] = 1; }
      if (q[0] > lowest) { rlo[r] = lowest; rreq[r++] = a.E; }
      r = pre + total;
      if (q[Q - 1] < highest) { rlo[r] = (K)(q[Q - 1] + 1); rreq[r++] = a.E; }
      if (KT::kFloat) { rlo[r] = (K)(highest + 1); rreq[r++] = 1; }
      bcast_R = r;
    }
    (void)kSingle; (void)kEdge; (void)kInterior; (void)kNan;
  }
  __syncthreads();
  const uint32_t R = bcast_R;
  // a value holding at least a quarter of the pivots is dominant
  const bool hot = hot_run && 4 * ((hot_run >> 16) + 1) >= Q;
  const K hv = hot ? q[hot_run & 0xFFFFu] : (K)0;
  // 4. map parameters (parallel), bucket bases (block scan), global copy
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
    if (hot && rr.lo == hv && rr.hi == hv) hot_bucket = rr.base;
  }
  if (threadIdx.x == 0) {
    ctl->R = R;
    ctl->Rp = pow2ceil_u32(R);
    ctl->B = Btot;
    ctl->nan_lo_bucket = KT::kFloat ? 0u : kNone32;
    ctl->nan_hi_bucket = KT::kFloat ? Btot - 1 : kNone32;
  }
  // 5. key-prefix cell table, from the LDS copy of the range starts, over the
  //    window the interior ranges span (sample minimum through the start of
  //    the top edge range): narrow inputs (small integers, one sign of floats)
  //    get cells as fine as their span allows instead of sharing a handful of
  //    top-key-bit cells; a span of the whole key space keeps the top bits
  using G = LutGeom<K>;
  const K wlo = q[0], whi = q[Q - 1];
  const uint32_t need = bitlen64((uint64_t)(K)(whi - wlo) + 1u);
  uint32_t wsh = need > G::L1BITS ? need - G::L1BITS : 0u;
  K wbase = wlo;
  if (wsh < G::S1 && hot_run) {
    // a repeated pivot value v is a one-key range [v, v] followed by a range
    // starting at v + 1: with v on a cell's low edge that cell is a one-start
    // leaf instead of a level-2 table, so the (many) keys equal to v take the
    // one-compare path.  Move the window's base down to the cell edge below v
    // (one cell more when the span then overflows the window).
    const K v = q[hot_run & 0xFFFFu];
    for (uint32_t w = wsh; w <= wsh + 1 && w < G::S1; ++w) {
      const K cw = (K)((K)1 << w);
      const K below = (K)(((K)(v - wlo) + cw - 1) / cw * cw);  // v - base, a multiple of the cell width
      if (below > v) break;                                      // no room below v
      const K base = (K)(v - below);
      if ((K)(whi - base) <= (K)((((K)G::L1CELLS) << w) - 1)) {
        wbase = base;
        wsh = w;
        break;
      }
    }
  }
  if (wsh >= G::S1) {
    wsh = G::S1;
    wbase = 0;
  } else {
    const K width = (K)(((K)G::L1CELLS) << wsh);
    if ((K)(KeyMax<K>::v - wbase) < (K)(width - 1)) wbase = (K)(KeyMax<K>::v - (width - 1));
  }
  const uint32_t l2 = build_lut_counted<K>(rlo, R, lut, cells, wbase, wsh);
  if (threadIdx.x == 0) {
    ctl->hot_bucket = hot_bucket;
    ctl->l2_used = l2;
    ctl->lut_base_lo = (uint32_t)(uint64_t)wbase;
    ctl->lut_base_hi = (uint32_t)((uint64_t)wbase >> 32);
    ctl->lut_shift = wsh;
    // the ranges holding keys outside the window: below it, the ranges up to
    // the one holding wbase - 1; above it, from the one holding the first key
    // past the window (rlo[0] == 0: every key has a range)
    const K wend = wsh >= G::S1 ? KeyMax<K>::v : (K)(wbase + ((((K)G::L1CELLS) << wsh) - 1));
    uint32_t rb = 0, ra = R - 1;
    if (wbase > 0) rb = last_start_le<K>(rlo, 0, R - 1, (K)(wbase - 1));
    if (wend < KeyMax<K>::v) ra = last_start_le<K>(rlo, 0, R - 1, (K)(wend + 1));
    ctl->lut_r_below = rb;
    ctl->lut_r_above = ra;
  }
}

}  // namespace ggms
