#include "hip/hip_runtime.h"
#include <algorithm>
#include <random>
#include <functional>
#include <memory>
#include "ggms.h"
#include "batch_kernel.hpp"



using namespace ggms;

// All of the grid's workgroups run at once, each with its own LDS (dynamic
// and __shared__ slots) filled with garbage first, as on the GPU; the
// persistent queue is then contended like on the device.
template <class T>
void run_kernel(BatchArgs a, uint32_t grid, uint32_t threads, size_t lds) {
  if (grid > EMU_MAX_WG) { fprintf(stderr, "emulated grid too large\n"); std::abort(); }
  gridDim = dim3(grid);
  blockDim = dim3(threads);
  std::unique_ptr<EmuWG[]> wgs(new EmuWG[grid]);
  // exactly the launch's dynamic LDS: an access past it is caught when the
  // driver is built with AddressSanitizer
  std::vector<std::vector<unsigned char>> smem(grid, std::vector<unsigned char>(lds, 0xA5));
  for (uint32_t b = 0; b < grid; ++b) {
    wgs[b].id = b;
    wgs[b].bar.reset(threads);
    wgs[b].dyn_smem = smem[b].data();
    emu_poison_wg(b, 0xA5);
  }
  std::vector<std::thread> th;
  th.reserve((size_t)grid * threads);
  for (uint32_t b = 0; b < grid; ++b)
    for (uint32_t t = 0; t < threads; ++t)
      th.emplace_back([&, b, t] {
        emu_wg = &wgs[b];
        blockIdx = dim3(b);
        threadIdx = dim3(t);
        k_batch<T>(a);
        g_bar.exit_thread();
      });
  for (auto& x : th) x.join();
}

static ggms_batch_class default_class(uint32_t c) {
  ggms_batch_class k{};
  k.num_buckets = c >= 20 ? 8192u : (c >= 17 ? 4096u : 2048u);
  k.sample_size = c >= 20 ? 2048u : 1024u;
  k.pieces = c >= 20 ? 64u : 32u;
  k.rebucket = kBatchNB;
  return k;
}

struct Problem { std::vector<float> v; std::vector<uint32_t> ks; };

static ggms_batch_class g_override{};
static int run_batch(const std::vector<Problem>& ps, uint32_t wgs, uint32_t threads, const char* label,
                     uint32_t extra = 3) {
  uint32_t P = ps.size();
  std::vector<uint64_t> off(P + 1, 0);
  std::vector<uint32_t> koff(P + 1, 0);
  for (uint32_t p = 0; p < P; ++p) { off[p + 1] = off[p] + ps[p].v.size(); koff[p + 1] = koff[p] + ps[p].ks.size(); }
  std::vector<float> vals(off[P] + 4);
  std::vector<uint32_t> ks(koff[P] + 1);
  for (uint32_t p = 0; p < P; ++p) {
    std::copy(ps[p].v.begin(), ps[p].v.end(), vals.begin() + off[p]);
    std::copy(ps[p].ks.begin(), ps[p].ks.end(), ks.begin() + koff[p]);
  }
  std::vector<float> out(koff[P] + 1, -12345.f);
  std::vector<uint32_t> order(P);
  for (uint32_t p = 0; p < P; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
  // workgroups beyond the problems find the queue empty at once and take the
  // kernel's exit protocol beside the busy ones, as most of a GPU grid does
  uint32_t grid = std::min(std::min(wgs, P + extra), EMU_MAX_WG);
  uint32_t queue[64] = {0};
  std::vector<uint32_t> status(P, 99);
  std::vector<uint32_t> sA(off[P] + 4), sB(off[P] + 4);
  uint32_t stack_cap = kBatchMaxKs + 8;
  std::vector<BatchItem<uint32_t>> stacks((size_t)grid * stack_cap);
  ggms_batch_class cfg[GGMS_BATCH_CLASSES];
  for (uint32_t c = 0; c < GGMS_BATCH_CLASSES; ++c) {
    ggms_batch_class k = g_override.num_buckets ? g_override : default_class(c);
    k.num_buckets = std::max<uint32_t>(16, std::min<uint32_t>(k.num_buckets, kBatchNB - 2 * kBatchMaxPieces - 16));
    cfg[c] = k;
  }
  BatchArgs a{};
  a.values = vals.data(); a.off = off.data(); a.ks = ks.data(); a.koff = koff.data(); a.out = out.data();
  a.order = order.data(); a.queue = queue; a.status = status.data(); a.scratchA = sA.data(); a.scratchB = sB.data();
  a.stacks = stacks.data(); a.nprob = P; a.stack_cap = stack_cap; a.seed = 0x5EED; a.cfg = cfg;
  // twice: the first launch registers every __shared__ declaration it reaches,
  // so the second starts with all of its LDS poisoned; both must answer
  int bad = 0;
  for (int launch = 0; launch < 2; ++launch) {
  std::fill(out.begin(), out.end(), -12345.f);
  std::fill(status.begin(), status.end(), 99u);
  run_kernel<float>(a, grid, threads, BatchLds<uint32_t>::bytes());
  if (queue[0] != 0 || queue[1] != 0) { printf("%s: queue not reset (%u %u)\n", label, queue[0], queue[1]); bad++; }
  for (uint32_t p = 0; p < P; ++p) {
    bool nan = false;
    for (float f : ps[p].v) nan |= std::isnan(f);
    bool kbad = false;
    for (uint32_t k : ps[p].ks) kbad |= (k < 1 || k > ps[p].v.size());
    uint32_t want_st = kbad ? 1 : (nan ? 2 : 0);
    if (ps[p].ks.empty()) want_st = 0;
    if (status[p] != want_st) { printf("%s: problem %u status %u want %u\n", label, p, status[p], want_st); bad++; continue; }
    if (want_st) continue;
    std::vector<float> s = ps[p].v;
    std::sort(s.begin(), s.end(), [](float x, float y) { return KeyTraits<float>::enc(x) < KeyTraits<float>::enc(y); });
    for (uint32_t j = 0; j < ps[p].ks.size(); ++j) {
      float w = s[s.size() - ps[p].ks[j]];
      float g = out[koff[p] + j];
      if (memcmp(&w, &g, 4)) { if (bad < 10) printf("%s: problem %u k=%u got %g want %g\n", label, p, ps[p].ks[j], g, w); bad++; }
    }
  }
  }
  printf("%s: %u problems, %u concurrent workgroups, two poisoned launches, %s\n", label, P, grid, bad ? "FAIL" : "ok");
  return bad;
}

// A batch written by the test (u32 problems, then per problem u32 n, u32 nk,
// n floats, nk ks): the GPU tests' own batches, value for value.
static std::vector<Problem> read_batch(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); std::abort(); }
  auto rd = [&](void* p, size_t n) { if (fread(p, 1, n, f) != n) { fprintf(stderr, "short batch file\n"); std::abort(); } };
  uint32_t P = 0;
  rd(&P, 4);
  std::vector<Problem> ps(P);
  for (auto& pr : ps) {
    uint32_t n = 0, nk = 0;
    rd(&n, 4);
    rd(&nk, 4);
    pr.v.resize(n);
    pr.ks.resize(nk);
    rd(pr.v.data(), (size_t)n * 4);
    rd(pr.ks.data(), (size_t)nk * 4);
  }
  fclose(f);
  return ps;
}

int main(int argc, char** argv) {
  if (argc > 2) {  // replay batch files: 1024 threads, the library's grid (one workgroup per problem), then three more
    int bad = 0;
    for (int i = 2; i < argc; ++i) {
      const std::vector<Problem> ps = read_batch(argv[i]);
      char lab[64];
      snprintf(lab, sizeof lab, "file%d", i - 2);
      bad += run_batch(ps, 256, 1024, lab, 0);
      snprintf(lab, sizeof lab, "file%d+3", i - 2);
      bad += run_batch(ps, 256, 1024, lab, 3);
    }
    emu_print_wave_stats();
    return bad ? 1 : 0;
  }
  std::mt19937_64 rng(10);
  std::normal_distribution<float> nd;
  int bad = 0;
  {
    std::vector<float> good(50000);
    for (auto& f : good) f = nd(rng);
    std::vector<float> nan_big = good; nan_big[1234] = NAN;
    std::vector<float> nan_small(good.begin(), good.begin() + 500); nan_small[7] = NAN;
    std::vector<Problem> ps;
    ps.push_back({good, {1, 2, 50000}});
    ps.push_back({nan_big, {5}});
    ps.push_back({std::vector<float>(good.begin(), good.begin() + 1000), {0}});
    ps.push_back({std::vector<float>(good.begin(), good.begin() + 1000), {1001}});
    ps.push_back({nan_small, {3}});
    ps.push_back({std::vector<float>(good.begin(), good.begin() + 20000), {17, 20000}});
    uint32_t threads = argc > 1 ? atoi(argv[1]) : 1024;
    bad += run_batch(ps, 256, threads, "errors");
    // the GPU test's batch: NaN problems beside valid ones only
    std::vector<Problem> pn;
    pn.push_back({good, {1, 2, 50000}});
    pn.push_back({nan_big, {5}});
    pn.push_back({std::vector<float>(good.begin(), good.begin() + 1000), {3}});
    pn.push_back({nan_small, {3}});
    pn.push_back({std::vector<float>(good.begin(), good.begin() + 20000), {17, 20000}});
    bad += run_batch(pn, 256, 1024, "nan");
  }
  // random batches: mixed sizes, many ks, duplicates
  for (int trial = 0; trial < 3; ++trial) {
    std::vector<Problem> ps;
    std::uniform_real_distribution<double> ud(0, 1);
    for (int p = 0; p < 12; ++p) {
      uint32_t n = (uint32_t)std::pow(2.0, 6 + ud(rng) * 12);
      Problem pr;
      pr.v.resize(n);
      int kind = p % 4;
      for (uint32_t i = 0; i < n; ++i) {
        float f = nd(rng);
        if (kind == 1) f = (ud(rng) < 0.9) ? 2.5f : f;
        if (kind == 2) f = (float)(i % 5);
        if (kind == 3) f = std::exp(f);
        pr.v[i] = f;
      }
      uint32_t nk = 1 + (uint32_t)(ud(rng) * (p % 3 == 0 ? 2000 : 300));
      for (uint32_t j = 0; j < nk; ++j) pr.ks.push_back(1 + (uint32_t)(ud(rng) * n) % n);
      ps.push_back(pr);
    }
    if (trial == 1) g_override = ggms_batch_class{16, 3, 2, 16};
    if (trial == 2) g_override = ggms_batch_class{1024, 512, 16, 2};
    char lab[32];
    snprintf(lab, sizeof lab, "random%d", trial);
    bad += run_batch(ps, 256, 256, lab);
    g_override = ggms_batch_class{};
  }
  emu_print_wave_stats();
  return bad ? 1 : 0;
}
