// Emulated workgroup runs of the finish kernel's multi-rank radix select
// (radix_select_multi) against std::sort: segments of 1..16384 keys inside a
// random key interval, 1..8 ranks each (duplicates allowed), 32- and 64-bit
// keys and the finish's 16-bit offsets, narrow intervals full of duplicates
// and the full key range.
#include "hip/hip_runtime.h"
#include <algorithm>
#include <random>
#include "select_kernels.hpp"
// counted: the caller builds the first digit's histogram (as the finish does
// while loading a segment) and, when that digit covers every varying bit,
// passes no keys at all (radix_select_multi must not read them).
template <class K> int run(uint64_t seed, uint32_t nt, bool counted) {
  std::mt19937_64 rng(seed);
  const uint32_t len = 1 + (uint32_t)(rng() % 16384);
  const int kind = (int)(rng() % 5);
  K klo, khi;
  if (kind == 0) { klo = 0; khi = (K)~(K)0; }                                   // full range
  else {
    const uint32_t bits = kind == 1 ? 3 : kind == 2 ? 12 : kind == 4 ? 11 : 1 + (uint32_t)(rng() % (sizeof(K) * 8 - 1));
    const K w = (K)(((K)1 << bits) - 1);
    klo = (K)rng();
    if ((K)(klo + w) < klo) klo = (K)(klo - w);                                 // keep [klo, klo + w] unwrapped
    khi = (K)(klo + (K)(rng() % ((uint64_t)w + 1)));
  }
  std::vector<K> a(len);
  const uint64_t span = (uint64_t)(khi - klo);
  for (auto& v : a) v = (K)(klo + (K)(span == ~0ull ? rng() : rng() % (span + 1)));
  if (len > 1) a[rng() % len] = klo, a[rng() % len] = khi;  // the interval's ends may or may not occur
  std::vector<K> sorted = a;
  std::sort(sorted.begin(), sorted.end());
  const uint32_t q = 1 + (uint32_t)(rng() % ggms::kSelectMaxKs);
  static ggms::SelectLds L;
  std::vector<uint32_t> ranks(q);
  for (auto& r : ranks) r = 1 + (uint32_t)(rng() % len);
  if (q > 1 && (rng() & 1)) ranks[1] = ranks[0];  // a repeated rank
  std::sort(ranks.begin(), ranks.end());
  for (uint32_t j = 0; j < q; ++j) { L.kid[j] = j; L.rank[j] = ranks[j]; }
  blockDim = dim3(nt); g_bar.reset(nt);
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      threadIdx = dim3(t);
      const K* s = a.data();
      if (counted) {
        uint32_t sh1, w1;
        ggms::select_first_digit(ggms::bitlen64((uint64_t)(klo ^ khi)), sh1, w1);
        ggms::select_zero_first(L);
        __syncthreads();
        for (uint32_t i = t; i < len; i += nt) ggms::select_count_first<K>(L.hist, a[i], sh1, w1, true);
        __syncthreads();
        if (sh1 == 0) s = nullptr;
      }
      ggms::radix_select_multi<K>(s, len, klo, khi, q, L, counted);
      g_bar.exit_thread();
    });
  for (auto& x : th) x.join();
  int bad = 0;
  for (uint32_t j = 0; j < q; ++j)
    if ((K)L.pre[j] != sorted[ranks[j] - 1]) {
      if (bad < 3) printf("FAIL K=%zu len=%u q=%u kind=%d counted=%d rank=%u got %llx want %llx\n", sizeof(K), len, q, kind, (int)counted, ranks[j],
                          (unsigned long long)(K)L.pre[j], (unsigned long long)sorted[ranks[j] - 1]);
      ++bad;
    }
  return bad ? 1 : 0;
}
// wave_radix_select: every wave of the workgroup selects its own rank in its
// own segment (no workgroup barrier in between), against std::sort
template <class K> int run_wave(uint64_t seed, uint32_t nt) {
  std::mt19937_64 rng(seed);
  const uint32_t nw = nt / 64;
  std::vector<std::vector<K>> seg(nw);
  std::vector<K> lo(nw), hi(nw), want(nw), got(nw);
  std::vector<uint32_t> rank(nw);
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t len = 1 + (uint32_t)(rng() % 700);
    const uint32_t bits = 1 + (uint32_t)(rng() % (sizeof(K) * 8));
    const K wd = bits >= sizeof(K) * 8 ? (K)~(K)0 : (K)(((K)1 << bits) - 1);
    K klo = wd == (K)~(K)0 ? (K)0 : (K)rng();
    if ((K)(klo + wd) < klo) klo = (K)(klo - wd);  // keep [klo, klo + wd] unwrapped
    const K khi = (K)(klo + wd);
    seg[w].resize(len);
    for (auto& v : seg[w]) v = (K)(klo + (K)(wd == (K)~(K)0 ? rng() : rng() % ((uint64_t)wd + 1)));
    if (rng() & 1) for (auto& v : seg[w]) if (rng() % 3 == 0) v = seg[w][0];  // duplicates
    std::vector<K> srt = seg[w];
    std::sort(srt.begin(), srt.end());
    rank[w] = 1 + (uint32_t)(rng() % len);
    want[w] = srt[rank[w] - 1];
    lo[w] = klo;
    hi[w] = khi;
  }
  static uint32_t hist[16 * 256];
  static uint32_t bc[16][2];
  blockDim = dim3(nt); g_bar.reset(nt);
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      threadIdx = dim3(t);
      const uint32_t w = t >> 6;
      const K k = ggms::wave_radix_select<K>(seg[w].data(), (uint32_t)seg[w].size(), lo[w], hi[w], rank[w], hist + 256 * w, bc[w]);
      if ((t & 63) == 0) got[w] = k;
      g_bar.exit_thread();
    });
  for (auto& x : th) x.join();
  int bad = 0;
  for (uint32_t w = 0; w < nw; ++w)
    if (got[w] != want[w]) {
      if (bad < 3) printf("WAVE FAIL K=%zu len=%zu rank=%u got %llx want %llx\n", sizeof(K), seg[w].size(), rank[w],
                          (unsigned long long)got[w], (unsigned long long)want[w]);
      ++bad;
    }
  return bad ? 1 : 0;
}
int main(int argc, char** argv) {
  const int trials = argc > 1 ? atoi(argv[1]) : 60;
  int bad = 0;
  for (int i = 0; i < trials; ++i) {
    bad += run<uint32_t>(1000 + i, (i & 1) ? 256 : 128, (i & 2) != 0);
    bad += run<uint64_t>(5000 + i, (i & 1) ? 128 : 256, (i & 2) != 0);
    bad += run<uint16_t>(9000 + i, (i & 1) ? 256 : 128, (i & 2) != 0);
    bad += run_wave<uint32_t>(13000 + i, 256);
    bad += run_wave<uint64_t>(17000 + i, 128);
  }
  printf(bad ? "select FAIL\n" : "select ok\n");
  return bad;
}
