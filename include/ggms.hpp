// ggms.hpp -- the reference's C++ call shape over the C ABI (header only).
//
//   IterativeSMOS::iterativeSMOSWrapper<T>(d_vector, length, kVals_ori, numKs, outputs, blocks, threads)
//     DistributedSMOS/iterativeSMOS.cuh:218-220 (instantiated for int, uint, float, double at
//     DistributedSMOS/iterativeSMOS.cu:1604-1612)
//   BucketMultiselect::bucketMultiselectWrapper<T>(...)
//     DistributedSMOS/bucketMultiselect.cuh:105-107
//   DistributedSMOS::distributedSMOSWrapper<T>(..., rank)
//     DistributedSMOS/distributedSMOS.hpp:22-25 (one GPU holds the whole vector)
//   DistributedBucketMultiselect::distributedBucketMultiselectWrapper<T>(..., rank)
//     DistributedSMOS/distributedBucketMultiselect.hpp:12-14 (one GPU holds the whole vector)
//
// Semantics kept: d_vector is a device pointer of `length` values, kVals_ori a
// host array of numKs ranks with k = 1 the largest, outputs a host array of
// numKs values in kVals_ori order.  Return value: the reference's (T)1 on
// success (DistributedSMOS/iterativeSMOS.cu:1600, distributedSMOS.cpp:1265);
// (T)0 on failure, with the reason in ggms_last_wrapper_status().  The input
// is not modified (the reference's iterativeSMOS clobbers it).
#pragma once
#include <stdint.h>

#include "ggms.h"

namespace ggms_detail {
template <class T> struct Suffix;
template <> struct Suffix<float> {
  static int it(float* v, int l, const uint32_t* k, int nk, float* o, int b, int t) { return ggms_iterativeSMOSWrapper_f32(v, l, k, nk, o, b, t); }
  static int bm(float* v, int l, const uint32_t* k, int nk, float* o, int b, int t) { return ggms_bucketMultiselectWrapper_f32(v, l, k, nk, o, b, t); }
  static int ds(float* v, int l, const uint32_t* k, int nk, float* o, int b, int t, int r) { return ggms_distributedSMOSWrapper_f32(v, l, k, nk, o, b, t, r); }
  static int db(float* v, int l, const uint32_t* k, int nk, float* o, int b, int t, int r) { return ggms_distributedBucketMultiselectWrapper_f32(v, l, k, nk, o, b, t, r); }
};
template <> struct Suffix<double> {
  static int it(double* v, int l, const uint32_t* k, int nk, double* o, int b, int t) { return ggms_iterativeSMOSWrapper_f64(v, l, k, nk, o, b, t); }
  static int bm(double* v, int l, const uint32_t* k, int nk, double* o, int b, int t) { return ggms_bucketMultiselectWrapper_f64(v, l, k, nk, o, b, t); }
  static int ds(double* v, int l, const uint32_t* k, int nk, double* o, int b, int t, int r) { return ggms_distributedSMOSWrapper_f64(v, l, k, nk, o, b, t, r); }
  static int db(double* v, int l, const uint32_t* k, int nk, double* o, int b, int t, int r) { return ggms_distributedBucketMultiselectWrapper_f64(v, l, k, nk, o, b, t, r); }
};
template <> struct Suffix<unsigned int> {
  static int it(unsigned int* v, int l, const uint32_t* k, int nk, unsigned int* o, int b, int t) { return ggms_iterativeSMOSWrapper_u32(v, l, k, nk, o, b, t); }
  static int bm(unsigned int* v, int l, const uint32_t* k, int nk, unsigned int* o, int b, int t) { return ggms_bucketMultiselectWrapper_u32(v, l, k, nk, o, b, t); }
  static int ds(unsigned int* v, int l, const uint32_t* k, int nk, unsigned int* o, int b, int t, int r) { return ggms_distributedSMOSWrapper_u32(v, l, k, nk, o, b, t, r); }
  static int db(unsigned int* v, int l, const uint32_t* k, int nk, unsigned int* o, int b, int t, int r) { return ggms_distributedBucketMultiselectWrapper_u32(v, l, k, nk, o, b, t, r); }
};
template <> struct Suffix<int> {
  static int it(int* v, int l, const uint32_t* k, int nk, int* o, int b, int t) { return ggms_iterativeSMOSWrapper_i32(v, l, k, nk, o, b, t); }
  static int bm(int* v, int l, const uint32_t* k, int nk, int* o, int b, int t) { return ggms_bucketMultiselectWrapper_i32(v, l, k, nk, o, b, t); }
  static int ds(int* v, int l, const uint32_t* k, int nk, int* o, int b, int t, int r) { return ggms_distributedSMOSWrapper_i32(v, l, k, nk, o, b, t, r); }
  static int db(int* v, int l, const uint32_t* k, int nk, int* o, int b, int t, int r) { return ggms_distributedBucketMultiselectWrapper_i32(v, l, k, nk, o, b, t, r); }
};
}  // namespace ggms_detail

namespace ggms_detail {
template <class T> inline T ref_result(int status) { return status == GGMS_OK ? (T)1 : (T)0; }
}  // namespace ggms_detail

namespace IterativeSMOS {
template <typename T>
inline T iterativeSMOSWrapper(T* d_vector, int length, uint32_t* kVals_ori, int numKs, T* outputs, int blocks,
                              int threads) {
  return ggms_detail::ref_result<T>(ggms_detail::Suffix<T>::it(d_vector, length, kVals_ori, numKs, outputs, blocks, threads));
}
}  // namespace IterativeSMOS

namespace BucketMultiselect {
template <typename T>
inline T bucketMultiselectWrapper(T* d_vector, int length, uint32_t* kVals_ori, int numKs, T* outputs, int blocks,
                                  int threads) {
  return ggms_detail::ref_result<T>(ggms_detail::Suffix<T>::bm(d_vector, length, kVals_ori, numKs, outputs, blocks, threads));
}
}  // namespace BucketMultiselect

namespace DistributedSMOS {
template <typename T>
inline T distributedSMOSWrapper(T* d_vector, int length, uint32_t* kVals_ori, int numKs, T* outputs, int blocks,
                                int threads, int rank) {
  return ggms_detail::ref_result<T>(ggms_detail::Suffix<T>::ds(d_vector, length, kVals_ori, numKs, outputs, blocks, threads, rank));
}
}  // namespace DistributedSMOS

namespace DistributedBucketMultiselect {
template <typename T>
inline T distributedBucketMultiselectWrapper(T* d_vector, int length, uint32_t* kVals_ori, int numKs, T* outputs,
                                             int blocks, int threads, int rank) {
  return ggms_detail::ref_result<T>(ggms_detail::Suffix<T>::db(d_vector, length, kVals_ori, numKs, outputs, blocks, threads, rank));
}
}  // namespace DistributedBucketMultiselect
