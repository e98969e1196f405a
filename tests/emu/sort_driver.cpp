#include "hip/hip_runtime.h"
#include <algorithm>
#include <random>
#include "select_kernels.hpp"
template <class K, bool SW> int run1(uint32_t M, uint32_t nt, uint64_t seed) {
  std::mt19937_64 rng(seed);
  // SW (swizzled layout between the register phases): 64 canaries behind the M keys must survive
  const uint32_t room = M;
  std::vector<K> a(room + 64, (K)0x5A5A5A5A5A5A5A5Aull);
  for (uint32_t i = 0; i < M; ++i) a[i] = (K)(rng() % (seed % 3 == 0 ? 7 : ~0ull));
  std::vector<K> want(a.begin(), a.begin() + M); std::sort(want.begin(), want.end());
  blockDim = dim3(nt); g_bar.reset(nt);
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < nt; ++t) th.emplace_back([&, t] { threadIdx = dim3(t); ggms::bitonic_sort_lds<K, SW>(a.data(), M); g_bar.exit_thread(); });
  for (auto& x : th) x.join();
  bool ok = std::equal(want.begin(), want.end(), a.begin());
  for (uint32_t i = room; i < room + 64; ++i) ok &= a[i] == (K)0x5A5A5A5A5A5A5A5Aull;
  if (!ok) { printf("FAIL M=%u nt=%u K=%zu sw=%d\n", M, nt, sizeof(K), (int)SW); return 1; }
  return 0;
}
template <class K> int run(uint32_t M, uint32_t nt, uint64_t seed) { return run1<K, false>(M, nt, seed) + run1<K, true>(M, nt, seed + 1); }
int main() {
  int bad = 0;
  for (uint32_t nt : {256u, 1024u})
    for (uint32_t M = 1; M <= 16384; M <<= 1) {
      bad += run<uint32_t>(M, nt, M + nt);
      bad += run<uint32_t>(M, nt, 3 * (M + nt));
      if (M <= 8192) bad += run<uint64_t>(M, nt, M * 7 + nt);
      bad += run<uint16_t>(M, nt, M * 5 + nt);
    }
  bad += run<uint16_t>(32768, 1024, 77);  // 32 keys per thread in registers
  bad += run<uint16_t>(32768, 1024, 78);
  printf(bad ? "sort FAIL\n" : "sort ok\n");
  return bad;
}
