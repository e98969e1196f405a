// Reference-shaped caller of include/ggms.hpp: the call sites of the reference's
// timing harness (DistributedSMOS/iterativeSMOSTimingFunctions.cu:147,220 and
// DistributedSMOS/distributedSMOSTimingFunctions.cpp:1516) written against the
// shim.  Input v[i] = i + 1, so the k-th largest of n values is n - k + 1
// (the known-answer harness of DistributedSMOS/iterativeSMOS_copy.cu:2320-2352).
// Prints "shim ok" and exits 0 when every call returns the reference's (T)1
// with the right values and a bad k returns (T)0 with GGMS_EINVAL.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <vector>

#include "ggms.hpp"

template <class T>
static bool run(const char* name) {
  const int n = 100003;
  std::vector<T> h(n);
  for (int i = 0; i < n; ++i) h[i] = (T)(i + 1);
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * n) != hipSuccess) return false;
  if (hipMemcpy(d, h.data(), sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return false;
  uint32_t ks[5] = {1, 2, 500, 77777, (uint32_t)n};
  bool ok = true;
  for (int alg = 0; alg < 4; ++alg) {
    T out[5] = {};
    T r = alg == 0   ? IterativeSMOS::iterativeSMOSWrapper<T>(d, n, ks, 5, out, 256, 1024)
          : alg == 1 ? BucketMultiselect::bucketMultiselectWrapper<T>(d, n, ks, 5, out, 256, 1024)
          : alg == 2 ? DistributedSMOS::distributedSMOSWrapper<T>(d, n, ks, 5, out, 256, 1024, 0)
                     : DistributedBucketMultiselect::distributedBucketMultiselectWrapper<T>(d, n, ks, 5, out, 256,
                                                                                            1024, 0);
    ok &= r == (T)1 && ggms_last_wrapper_status() == GGMS_OK;
    for (int j = 0; j < 5; ++j) ok &= out[j] == (T)(n - (int)ks[j] + 1);
    if (!ok) std::printf("%s alg %d: return %g, status %d\n", name, alg, (double)r, (int)ggms_last_wrapper_status());
  }
  uint32_t bad[2] = {3, 0};
  T out2[2] = {};
  T r = IterativeSMOS::iterativeSMOSWrapper<T>(d, n, bad, 2, out2, 256, 1024);
  ok &= r == (T)0 && ggms_last_wrapper_status() == GGMS_EINVAL;
  (void)hipFree(d);
  return ok;
}

int main() {
  bool ok = run<float>("float") && run<double>("double") && run<unsigned int>("uint") && run<int>("int");
  std::printf(ok ? "shim ok\n" : "shim FAILED\n");
  return ok ? 0 : 1;
}
