#include "hip/hip_runtime.h"
#include <algorithm>
#include <random>
#include <cmath>
#include "splitter.hpp"
using namespace ggms;
int run(uint32_t mode, uint32_t S, uint32_t P, uint32_t B, int kind, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd;
  uint32_t n = 1 << 20;
  std::vector<float> x(n);
  for (auto& v : x) {
    float f = nd(rng);
    if (kind == 1) f = (rng() % 10 < 9) ? 2.5f : f;
    if (kind == 2) f = (float)(rng() % 5);
    if (kind == 3) f = (float)(rng() % 1024);  // narrow span: fine cells
    if (kind == 4) f = std::fabs(f);            // one sign: a window of part of the key space
    if (kind == 6) f = (rng() % 100 < 99) ? 1.0f : f;  // 99 % one value, 1 % N(0, 1)
    if (kind == 5) {                             // narrow peaks: many starts per cell (level-2 tables)
      const float mu[3] = {37.5f, -12.25f, 80.125f};
      f = mu[rng() % 3] + 0.01f * f;
    }
    x[&v - x.data()] = f;
  }
  std::vector<float> w(kMlpParams);
  for (auto& v : w) v = (float)((int)(rng() % 2001) - 1000) / 1000.0f;
  Plan0Args a{}; a.mode = mode; a.S = S; a.P = P; a.B = B; a.E = 4; a.strideS = 8192; a.seed = seed; a.mlp = w.data();
  std::vector<Range<uint32_t>> ranges(4096);
  std::vector<uint32_t> lut(LutGeom<uint32_t>::kEntries);
  PassCtl ctl{}; ctl.n = n;
  uint32_t Sp = 1; while (Sp < S) Sp <<= 1;
  std::vector<unsigned char> smem(plan0_lds_bytes<uint32_t>(Sp) + 64);
  g_dyn_smem = smem.data();
  blockDim = dim3(1024); g_bar.reset(1024);
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < 1024; ++t) th.emplace_back([&, t] { threadIdx = dim3(t); k_plan0<float>(x.data(), a, ranges.data(), lut.data(), &ctl); g_bar.exit_thread(); });
  for (auto& t : th) t.join();
  // check ranges partition key space and LUT maps each key to the right range
  uint32_t R = ctl.R, bad = 0, base = 0;
  if (ranges[0].lo != 0) bad++;
  for (uint32_t r = 0; r < R; ++r) {
    if (r + 1 < R && !(ranges[r].lo < ranges[r + 1].lo)) bad++;
    if (r + 1 < R && ranges[r].hi != ranges[r + 1].lo - 1) bad++;
    if (ranges[r].base != base) bad++;
    base += ranges[r].nb;
  }
  if (base != ctl.B) bad++;
  // reference map by binary search vs LUT-based map
  std::vector<uint32_t> starts(R); for (uint32_t r = 0; r < R; ++r) starts[r] = ranges[r].lo;
  std::vector<unsigned char> prm(Pass0Map<uint32_t>::param_bytes(R) + 16);
  Pass0Map<uint32_t> mp; mp.lut = lut.data(); mp.set_params(prm.data(), R); mp.set_window(&ctl);
  const bool narrow = kind >= 2 && kind <= 4;  // kinds 2-4 hold one sign only
  if (kind <= 5 && narrow != mp.windowed()) { printf("kind %d: window shift %u unexpected\n", kind, ctl.lut_shift); bad++; }
  if (kind == 6 && mp.windowed()) {  // the dominant value sits on a cell edge: a leaf, not a table
    const uint32_t d = KeyTraits<float>::enc(1.0f) - ctl.lut_base_lo;
    const uint32_t e = lut[d >> ctl.lut_shift];
    if ((d & ((1u << ctl.lut_shift) - 1)) != 0 || (e & LutGeom<uint32_t>::FLAG)) {
      printf("kind 6: dominant value at cell offset %u, entry %08x\n", d & ((1u << ctl.lut_shift) - 1), e);
      bad++;
    }
  }
  // the dominant-value bucket: kinds 1 (90 % one value) and 6 (99 %) name the
  // one-key range of that value; inputs without a dominant value name none
  {
    const float dom = kind == 1 ? 2.5f : 1.0f;
    if (kind == 1 || kind == 6) {
      const uint32_t kd = KeyTraits<float>::enc(dom);
      const uint32_t r = std::upper_bound(starts.begin(), starts.end(), kd) - starts.begin() - 1;
      if (ctl.hot_bucket == kNone32 || ranges[r].lo != kd || ranges[r].hi != kd || ranges[r].base != ctl.hot_bucket) {
        printf("kind %d: dominant bucket %u, range [%08x, %08x] base %u\n", kind, ctl.hot_bucket, ranges[r].lo,
               ranges[r].hi, ranges[r].base);
        bad++;
      }
    } else if ((kind == 0 || kind == 4) && ctl.hot_bucket != kNone32) {
      printf("kind %d: dominant bucket %u without a dominant value\n", kind, ctl.hot_bucket);
      bad++;
    }
  }
  for (uint32_t r = 0; r < R; ++r) mp.store_param(r, ranges[r]);
  emu_solo = true;  // the map checked from this thread: a wave of one lane
  for (int t = 0; t < 200000; ++t) {
    uint32_t key = (uint32_t)rng();
    if (t % 2) key = KeyTraits<float>::enc(x[rng() % n]);
    uint32_t r = std::upper_bound(starts.begin(), starts.end(), key) - starts.begin() - 1;
    uint32_t want = ranges[r].base + map_off<uint32_t>(key, ranges[r].lo, ranges[r].sh, ranges[r].mult);
    threadIdx = dim3(0);
    if (t % 4 == 3) key = starts[rng() % R] + (uint32_t)(rng() % 3) - 1;  // at and around range starts
    if (t % 4 == 3) {
      r = std::upper_bound(starts.begin(), starts.end(), key) - starts.begin() - 1;
      want = ranges[r].base + map_off<uint32_t>(key, ranges[r].lo, ranges[r].sh, ranges[r].mult);
    }
    uint32_t got = mp.windowed() ? mp.bucket<true>(key) : mp.bucket<false>(key);
    if (got != want) { if (bad < 5) printf("key %08x got %u want %u (r=%u)\n", key, got, want, r); bad++; }
  }
  // the four-key form the streaming kernels use (its level-2 and search
  // blocks are shared by the four keys)
  for (int t = 0; t < 50000; ++t) {
    uint32_t k4[4], b4[4], w4[4];
    for (int j = 0; j < 4; ++j) {
      const int kind4 = (int)(rng() % 3);
      k4[j] = kind4 == 0 ? (uint32_t)rng() : kind4 == 1 ? KeyTraits<float>::enc(x[rng() % n])
                                                        : starts[rng() % R] + (uint32_t)(rng() % 3) - 1;
      const uint32_t r = std::upper_bound(starts.begin(), starts.end(), k4[j]) - starts.begin() - 1;
      w4[j] = ranges[r].base + map_off<uint32_t>(k4[j], ranges[r].lo, ranges[r].sh, ranges[r].mult);
    }
    threadIdx = dim3(0);
    if (mp.windowed()) mp.buckets<4, true>(k4, b4); else mp.buckets<4, false>(k4, b4);
    for (int j = 0; j < 4; ++j)
      if (b4[j] != w4[j]) { if (bad < 5) printf("4-key form: key %08x got %u want %u\n", k4[j], b4[j], w4[j]); bad++; }
  }
  emu_solo = false;
  printf("mode %u S %u P %u kind %d: R=%u B=%u %s\n", mode, S, P, kind, R, ctl.B, bad ? "FAIL" : "ok");
  return bad ? 1 : 0;
}
int main() {
  int bad = 0;

  for (int kind = 0; kind < 7; ++kind) {
    bad += run(kModeLinear, 1024, 17, 8192, kind, 1 + kind);
    bad += run(kModeSketch, 8192, 256, 16384, kind, 11 + kind);
    bad += run(kModeLearned, 8192, 256, 16384, kind, 21 + kind);
    bad += run(kModeLearned, 8192, 512, 16384, kind, 31 + kind);
  }
  return bad;
}
