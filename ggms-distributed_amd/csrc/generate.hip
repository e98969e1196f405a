// Seeded synthetic inputs (replaces DistributedSMOS/generateProblems.cu, whose
// values come from cuRAND XORWOW seeded by wall-clock time).
//
// Every element is a pure function of (seed, index): Philox4x32-10 draws,
// turned into values with explicit IEEE double arithmetic (this file is built
// with -ffp-contract=off and calls no device math library), so a CPU
// implementation of the same definitions reproduces the bits exactly.
//
// Distributions (dtype-independent definitions; fp32 outputs are the fp64
// value rounded to nearest, except UNIFORM fp32 which takes 24 bits):
//   UNIFORM        [0,1): element i uses draw(i/2), words 2(i%2), 2(i%2)+1
//                  (u32: draw(i/4) word i%4, optionally scaled into [0, p0))
//   NORMAL         Box-Muller: pair i/2 from draw(i/2); cos for even i, sin odd
//   HALF_NORMAL    |NORMAL|                   (generateHalfNormalFloats :381-409)
//   LOGNORMAL      exp(NORMAL)
//   MIXTURE        1-8 Gaussians (means U[-100,100], std U[0.01,10]) chosen per
//                  element, plus 5% lognormal tails +-exp(2 + 1.5 z)
//   NEAR_CONSTANT  p0 except a p1 fraction of elements that are NORMAL
//   CONSTANT       p0                          (generateAllOnesFloats)
//   CAUCHY         tan of a uniform angle      (generateCauchyFloats)
//   MOD_INDEX      i mod p0                    (generateUniformZeroToFourUints)
//   ASCENDING      i                           (generateSortedArrayUints shape)
#include <hip/hip_runtime.h>

#include <cstring>

#include "ggms.h"
#include "philox.hpp"

namespace ggms {
namespace gen {

__host__ __device__ __forceinline__ double dbits(uint64_t b) { return __builtin_bit_cast(double, b); }
__host__ __device__ __forceinline__ uint64_t ubits(double d) { return __builtin_bit_cast(uint64_t, d); }

__host__ __device__ double det_log(double x) {
  uint64_t b = ubits(x);
  int e = (int)((b >> 52) & 0x7FF) - 1022;
  double m = dbits((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  const double ln2_hi = dbits(0x3FE62E42FEE00000ull), ln2_lo = dbits(0x3DEA39EF35793C76ull);
  double de = (double)e;
  return de * ln2_hi + (de * ln2_lo + 2.0 * s * p);
}

__host__ __device__ double det_exp(double x) {
  const double inv_ln2 = 1.44269504088896338700;
  const double ln2_hi = dbits(0x3FE62E42FEE00000ull), ln2_lo = dbits(0x3DEA39EF35793C76ull);
  double kd = floor(x * inv_ln2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  const double c[20] = {1.0 / 121645100408832000.0, 1.0 / 6402373705728000.0, 1.0 / 355687428096000.0,
                        1.0 / 20922789888000.0, 1.0 / 1307674368000.0, 1.0 / 87178291200.0,
                        1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                        1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0,
                        1.0 / 24.0, 1.0 / 6.0, 1.0 / 2.0, 1.0, 1.0};
  double p = 1.0 / 2432902008176640000.0;
#pragma unroll
  for (int i = 0; i < 20; ++i) p = p * r + c[i];
  return p * dbits((uint64_t)(k + 1023) << 52);
}

__host__ __device__ double det_sqrt(double x) {
  if (x <= 0.0) return 0.0;
  uint64_t b = ubits(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double y = dbits((uint64_t)((e / 2) + 1023) << 52);
#pragma unroll
  for (int i = 0; i < 7; ++i) y = 0.5 * (y + x / y);
  return y;
}

__host__ __device__ void det_sincos2pi(double u, double* sn, double* cs) {
  double t = u * 8.0;
  int q = (int)t;
  if (q > 7) q = 7;
  double f = t - (double)q;
  if (q & 1) f = 1.0 - f;
  const double pi_4 = 0.78539816339744830962;
  double a = f * pi_4;
  double a2 = a * a;
  const double sc[9] = {1.0 / 355687428096000.0, -1.0 / 1307674368000.0, 1.0 / 6227020800.0,
                        -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0, 1.0};
  double ps = -1.0 / 121645100408832000.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) ps = ps * a2 + sc[i];
  double s = a * ps;
  const double cc[10] = {-1.0 / 6402373705728000.0, 1.0 / 20922789888000.0, -1.0 / 87178291200.0,
                         1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0, -1.0 / 720.0,
                         1.0 / 24.0, -1.0 / 2.0, 1.0};
  double pc = 1.0 / 2432902008176640000.0;
#pragma unroll
  for (int i = 0; i < 10; ++i) pc = pc * a2 + cc[i];
  double c = pc;
  double rs, rc;
  switch (q) {
    case 0: rs = s; rc = c; break;
    case 1: rs = c; rc = s; break;
    case 2: rs = c; rc = -s; break;
    case 3: rs = s; rc = -c; break;
    case 4: rs = -s; rc = -c; break;
    case 5: rs = -c; rc = -s; break;
    case 6: rs = -c; rc = s; break;
    default: rs = -s; rc = c; break;
  }
  *sn = rs;
  *cs = rc;
}

__host__ __device__ __forceinline__ double u53w(uint32_t hi, uint32_t lo) {
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)m * (1.0 / 9007199254740992.0);
}

__host__ __device__ double normal_at(uint64_t seed, uint32_t stream, uint64_t i) {
  U32x4 w = philox_draw(seed, stream, i >> 1);
  double u1 = u53w(w.v[0], w.v[1]) + (1.0 / 9007199254740992.0);
  double u2 = u53w(w.v[2], w.v[3]);
  double r = det_sqrt(-2.0 * det_log(u1));
  double sn, cs;
  det_sincos2pi(u2, &sn, &cs);
  return (i & 1) ? r * sn : r * cs;
}

struct MixParams {
  int m;
  double mu[8];
  double sigma[8];
};

struct GenArgs {
  int dist;
  uint64_t seed;
  double p0, p1;
  MixParams mp;
};

// fp64 value of element i; *raw receives the 32-bit word UNIFORM fp32 uses
__device__ double value_at(const GenArgs& a, uint64_t i, uint32_t* raw) {
  switch (a.dist) {
    case GGMS_DIST_UNIFORM: {
      U32x4 w = philox_draw(a.seed, kStreamValues, i >> 1);
      uint32_t hi = w.v[(i & 1) * 2], lo = w.v[(i & 1) * 2 + 1];
      *raw = hi;
      return u53w(hi, lo);
    }
    case GGMS_DIST_NORMAL:
      return normal_at(a.seed, kStreamValues, i);
    case GGMS_DIST_HALF_NORMAL:
      return fabs(normal_at(a.seed, kStreamValues, i));
    case GGMS_DIST_LOGNORMAL:
      return det_exp(normal_at(a.seed, kStreamValues, i));
    case GGMS_DIST_MIXTURE: {
      U32x4 w = philox_draw(a.seed, kStreamValues2, i);
      double z = normal_at(a.seed, kStreamValues, i);
      if ((double)w.v[0] < 0.05 * 4294967296.0) {
        double v = det_exp(2.0 + 1.5 * z);
        return (w.v[1] & 1u) ? -v : v;
      }
      int c = (int)(((uint64_t)w.v[1] * (uint64_t)a.mp.m) >> 32);
      return a.mp.mu[c] + a.mp.sigma[c] * z;
    }
    case GGMS_DIST_NEAR_CONSTANT: {
      U32x4 w = philox_draw(a.seed, kStreamValues2, i);
      if ((double)w.v[0] < a.p1 * 4294967296.0) return normal_at(a.seed, kStreamValues, i);
      return a.p0;
    }
    case GGMS_DIST_CONSTANT:
      return a.p0;
    case GGMS_DIST_CAUCHY: {
      U32x4 w = philox_draw(a.seed, kStreamValues, i >> 1);
      double u = u53w(w.v[(i & 1) * 2], w.v[(i & 1) * 2 + 1]);
      double sn, cs;
      det_sincos2pi(0.5 * u + 0.75, &sn, &cs);
      return sn / cs;
    }
    case GGMS_DIST_MOD_INDEX: {
      uint64_t m = a.p0 > 0 ? (uint64_t)a.p0 : 1;
      return (double)(i % m);
    }
    case GGMS_DIST_ASCENDING:
      return (double)i;
  }
  return 0.0;
}

template <class T> struct Emit;
template <> struct Emit<float> {
  __device__ static float f(const GenArgs& a, uint64_t i) {
    uint32_t raw = 0;
    double d = value_at(a, i, &raw);
    if (a.dist == GGMS_DIST_UNIFORM) return (float)(raw >> 8) * (1.0f / 16777216.0f);
    return (float)d;
  }
};
template <> struct Emit<double> {
  __device__ static double f(const GenArgs& a, uint64_t i) {
    uint32_t raw = 0;
    return value_at(a, i, &raw);
  }
};
template <> struct Emit<uint32_t> {
  __device__ static uint32_t f(const GenArgs& a, uint64_t i) {
    if (a.dist == GGMS_DIST_MOD_INDEX) return (uint32_t)(i % (uint64_t)(a.p0 > 0 ? a.p0 : 1));
    if (a.dist == GGMS_DIST_ASCENDING) return (uint32_t)i;
    if (a.dist == GGMS_DIST_CONSTANT) return (uint32_t)a.p0;
    U32x4 w = philox_draw(a.seed, kStreamValues, i >> 2);
    uint32_t v = w.v[i & 3];
    if (a.p0 > 0) v = (uint32_t)(((uint64_t)v * (uint64_t)a.p0) >> 32);
    return v;
  }
};
template <> struct Emit<int32_t> {
  __device__ static int32_t f(const GenArgs& a, uint64_t i) { return (int32_t)Emit<uint32_t>::f(a, i); }
};

template <class T>
__global__ void __launch_bounds__(256) k_generate(T* __restrict__ out, uint64_t n, GenArgs a) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = Emit<T>::f(a, i);
}

// Host side: mixture parameters (same Philox draws, host double arithmetic).
static void mixture_params(uint64_t seed, MixParams* mp) {
  auto u53h = [](uint32_t hi, uint32_t lo) {
    uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
    return (double)m * (1.0 / 9007199254740992.0);
  };
  U32x4 w = philox_draw(seed, kStreamParams, 0);
  mp->m = 1 + (int)(((uint64_t)w.v[0] * 8u) >> 32);
  for (int j = 0; j < 8; ++j) {
    w = philox_draw(seed, kStreamParams, 1 + (uint64_t)j);
    volatile double u0 = u53h(w.v[0], w.v[1]);
    volatile double u1 = u53h(w.v[2], w.v[3]);
    volatile double t0 = 200.0 * u0;
    volatile double t1 = 9.99 * u1;
    mp->mu[j] = -100.0 + t0;
    mp->sigma[j] = 0.01 + t1;
  }
}

}  // namespace gen
}  // namespace ggms

using namespace ggms;
using namespace ggms::gen;

extern "C" ggms_status ggms_generate(ggms_dtype t, ggms_dist d, void* d_out, uint64_t n, uint64_t seed, double p0,
                                     double p1, void* stream) {
  if (!d_out && n) return GGMS_EINVAL;
  if ((unsigned)d > (unsigned)GGMS_DIST_ASCENDING) return GGMS_EINVAL;
  if (n == 0) return GGMS_OK;
  GenArgs a{};
  a.dist = (int)d;
  a.seed = seed;
  a.p0 = p0;
  a.p1 = p1;
  if (d == GGMS_DIST_MIXTURE) mixture_params(seed, &a.mp);
  hipStream_t st = (hipStream_t)stream;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case GGMS_F32: k_generate<float><<<(unsigned)blocks, 256, 0, st>>>((float*)d_out, n, a); break;
    case GGMS_F64: k_generate<double><<<(unsigned)blocks, 256, 0, st>>>((double*)d_out, n, a); break;
    case GGMS_U32: k_generate<uint32_t><<<(unsigned)blocks, 256, 0, st>>>((uint32_t*)d_out, n, a); break;
    case GGMS_I32: k_generate<int32_t><<<(unsigned)blocks, 256, 0, st>>>((int32_t*)d_out, n, a); break;
    default: return GGMS_EINVAL;
  }
  if (hipGetLastError() != hipSuccess) return GGMS_EDEVICE;
  return GGMS_OK;
}

extern "C" ggms_status ggms_mixture_params(uint64_t seed, int* m, double* mu, double* sigma) {
  if (!m || !mu || !sigma) return GGMS_EINVAL;
  MixParams mp{};
  mixture_params(seed, &mp);
  *m = mp.m;
  memcpy(mu, mp.mu, sizeof(mp.mu));
  memcpy(sigma, mp.sigma, sizeof(mp.sigma));
  return GGMS_OK;
}

// ---------------------------------------------------------------------------
// k-rank generators (host).  Formulas of DistributedSMOS/generateProblems.cu:
// generateKUniformRandom :584-604, generateKUniform :608-620, generateKNormal
// :622-651, generateKCluster :654-696, generateKSectioned :698-720; uniform
// floats come from Philox stream kStreamKs instead of cuRAND.  Three more
// shapes serve the adversarial and small configurations: distinct sorted
// uniform ranks, evenly spaced ranks ceil(i*n/nk), and ranks drawn uniformly
// inside a window of 1% of the ranks at a random offset.
// ---------------------------------------------------------------------------
namespace {
inline float kuni(uint64_t seed, uint64_t i) {
  U32x4 w = philox_draw(seed, kStreamKs, i >> 2);
  return (float)(w.v[i & 3] >> 8) * (1.0f / 16777216.0f);
}
inline uint32_t kclamp(int64_t v, uint64_t n) {
  if (v <= 0) return 1;
  if ((uint64_t)v > n) return (uint32_t)n;
  return (uint32_t)v;
}
}  // namespace

#include <algorithm>
#include <vector>

extern "C" ggms_status ggms_generate_ks(ggms_kdist kd, uint32_t nk, uint64_t n, uint64_t seed, uint32_t* k) {
  if (nk == 0) return GGMS_OK;
  if (!k || n == 0 || n > 0xFFFFFFFFull) return GGMS_EINVAL;
  switch (kd) {
    case GGMS_K_UNIFORM_RANDOM:
      for (uint32_t i = 0; i < nk; ++i) k[i] = kclamp((int64_t)(uint32_t)(kuni(seed, i) * (float)n), n);
      return GGMS_OK;
    case GGMS_K_UNIFORM:
      k[0] = 1;
      for (uint32_t i = 1; i + 1 < nk; ++i) k[i] = kclamp((int64_t)(uint32_t)((i / (float)(nk - 1)) * n), n);
      if (nk > 1) k[nk - 1] = (uint32_t)n;
      return GGMS_OK;
    case GGMS_K_NORMAL:
      for (uint32_t i = 0; i < nk; ++i) {
        float z = (float)normal_at(seed, kStreamKs, i);
        k[i] = kclamp((int64_t)(uint32_t)((std::fabs(z) + 7.0f) * (n / 14.0)), n);
      }
      return GGMS_OK;
    case GGMS_K_CLUSTER: {
      if (n < 11) return GGMS_EINVAL;
      uint32_t fi = 0;
      int64_t maxassigned = 0;
      for (uint32_t i = 0; i < nk; ++i) k[i] = 0;
      for (uint32_t i = 4; i < nk; i += 9) {
        k[i] = (uint32_t)(kuni(seed, fi++) * (float)(n - 10) + 5);
        for (uint32_t j = 1; j < 5; ++j) {
          k[i - j] = k[i] - j;
          if (i + j < nk) k[i + j] = k[i] + j;
          maxassigned = std::max<int64_t>(maxassigned, (int64_t)(i + j));
        }
      }
      for (uint32_t i = 0; i < nk; ++i) k[i] = kclamp((int64_t)k[i], n);
      if (maxassigned < (int64_t)nk - 1) {
        k[maxassigned + 1] = (uint32_t)(kuni(seed, fi++) * (float)(n - 10) + 5);
        for (int64_t j = maxassigned + 2; j < (int64_t)nk; ++j)
          k[j] = std::min<uint32_t>(k[j - 1] + 1, (uint32_t)n - 1);
        for (uint32_t i = 0; i < nk; ++i) k[i] = kclamp((int64_t)k[i], n);
      }
      return GGMS_OK;
    }
    case GGMS_K_SECTIONED:
      k[0] = kclamp((int64_t)(uint32_t)(kuni(seed, 0) * (float)(n > nk ? n - nk : 0)), n);
      for (uint32_t i = 1; i < nk; ++i) k[i] = kclamp((int64_t)k[i - 1] + 1, n);
      return GGMS_OK;
    case GGMS_K_DISTINCT_SORTED: {
      if (nk > n) return GGMS_EINVAL;
      std::vector<uint32_t> v;
      uint64_t j = 0;
      while (v.size() < nk) {
        U32x4 w = philox_draw(seed, kStreamKs, j++);
        v.push_back((uint32_t)(((uint64_t)w.v[0] * n) >> 32) + 1);
        if (v.size() == nk) {
          std::sort(v.begin(), v.end());
          v.erase(std::unique(v.begin(), v.end()), v.end());
        }
      }
      for (uint32_t i = 0; i < nk; ++i) k[i] = v[i];
      return GGMS_OK;
    }
    case GGMS_K_EVEN:
      for (uint32_t i = 0; i < nk; ++i) k[i] = (uint32_t)(((uint64_t)(i + 1) * n + nk - 1) / nk);
      return GGMS_OK;
    case GGMS_K_WINDOW: {
      uint64_t w = std::max<uint64_t>(nk, n / 100);
      if (w > n) w = n;
      U32x4 d = philox_draw(seed, kStreamKs, 0);
      uint64_t off = ((uint64_t)d.v[0] * (n - w + 1)) >> 32;
      for (uint32_t i = 0; i < nk; ++i) {
        d = philox_draw(seed, kStreamKs, 1 + (uint64_t)i);
        k[i] = (uint32_t)(off + 1 + (((uint64_t)d.v[0] * w) >> 32));
      }
      return GGMS_OK;
    }
  }
  return GGMS_EINVAL;
}
