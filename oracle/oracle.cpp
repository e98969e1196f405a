// CPU oracle for the multiselect path.  TEST INFRASTRUCTURE ONLY: it is used
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as
// the checker and the CPU baseline.  The product library never calls it.
//
// Independent restatement (it shares no code with the HIP library) of:
//   * sort-and-choose: ascending sort, then out_i = sorted[n - k_i]
//       DistributedSMOS/iterativeSMOSTimingFunctions.cu:75-120 (sort :85-86,
//       copyInChunk :66-73).  Values are sorted through order-preserving
//       integer keys so the result is a total order (-0.0 < +0.0).
//   * k conversion: k-th largest -> k-th smallest, length - k + 1
//       DistributedSMOS/iterativeSMOS.cu:1587-1588.
//   * an nth_element multiselect (sorted ks, recursive partition) as a second
//     CPU baseline.
//   * Philox4x32-10 (Salmon et al., SC'11) and the library's synthetic value
//     and k-rank distributions.  The k distributions follow
//     DistributedSMOS/generateProblems.cu:584-720; value distributions follow
//     the names of generateProblems.cu:86-111, 381-409, 456-481, 43-49 (the
//     reference draws them from cuRAND, whose stream is not reproduced).
//   * the fixed-point linear-slope bucket map, to recount a first-pass
//     histogram from the range table the library reports.
//   * the learned splitter's MLP forward / backward (float32, for gradient
//     checks of the HIP training kernels).
//
// Parity status of each function is stated at its definition.
#include <parallel/algorithm>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

// ---------------------------------------------------------------------------
// keys (pinned by tests/test_oracle.py: monotonicity over all special values,
// round trip, and agreement with numpy's sort order)
// ---------------------------------------------------------------------------
static inline uint32_t key_f32(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  return (u >> 31) ? ~u : (u ^ 0x80000000u);
}
static inline float unkey_f32(uint32_t k) {
  uint32_t u = (k >> 31) ? (k ^ 0x80000000u) : ~k;
  float v;
  std::memcpy(&v, &u, 4);
  return v;
}
static inline uint64_t key_f64(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return (u >> 63) ? ~u : (u ^ 0x8000000000000000ull);
}
static inline double unkey_f64(uint64_t k) {
  uint64_t u = (k >> 63) ? (k ^ 0x8000000000000000ull) : ~k;
  double v;
  std::memcpy(&v, &u, 8);
  return v;
}

enum { D_F32 = 0, D_F64 = 1, D_U32 = 2, D_I32 = 3 };

extern "C" void oracle_keys(int dtype, const void* x, uint64_t n, void* keys_out) {
  if (dtype == D_F32) {
    const float* a = (const float*)x;
    uint32_t* o = (uint32_t*)keys_out;
    for (uint64_t i = 0; i < n; ++i) o[i] = key_f32(a[i]);
  } else if (dtype == D_F64) {
    const double* a = (const double*)x;
    uint64_t* o = (uint64_t*)keys_out;
    for (uint64_t i = 0; i < n; ++i) o[i] = key_f64(a[i]);
  } else if (dtype == D_U32) {
    std::memcpy(keys_out, x, n * 4);
  } else {
    const uint32_t* a = (const uint32_t*)x;
    uint32_t* o = (uint32_t*)keys_out;
    for (uint64_t i = 0; i < n; ++i) o[i] = a[i] ^ 0x80000000u;
  }
}

static int is_nan_key32(uint32_t k) { return k < 0x007FFFFFu || k > 0xFF800000u; }
static int is_nan_key64(uint64_t k) { return k < 0x000FFFFFFFFFFFFFull || k > 0xFFF0000000000000ull; }

template <class K>
static void psort(std::vector<K>& v, int threads) {
  if (threads > 1) {
    omp_set_num_threads(threads);
    __gnu_parallel::sort(v.begin(), v.end());
  } else {
    std::sort(v.begin(), v.end());
  }
}

// Sort-and-choose.  Returns 0, 1 (bad k) or 2 (NaN in input).
// Parity: pinned by the closed-form known-answer tests (v[i] = i+1, constant
// vectors, k = 1 / k = n) and by numpy.sort on random inputs.
extern "C" int oracle_sort_and_choose(int dtype, const void* x, uint64_t n, const uint32_t* ks, uint32_t nk, void* out,
                           int threads) {
  for (uint32_t i = 0; i < nk; ++i)
    if (ks[i] < 1 || ks[i] > n) return 1;
  if (dtype == D_F64) {
    std::vector<uint64_t> k(n);
    oracle_keys(dtype, x, n, k.data());
    for (uint64_t i = 0; i < n; ++i)
      if (is_nan_key64(k[i])) return 2;
    psort(k, threads);
    double* o = (double*)out;
    for (uint32_t i = 0; i < nk; ++i) o[i] = unkey_f64(k[n - ks[i]]);
    return 0;
  }
  std::vector<uint32_t> k(n);
  oracle_keys(dtype, x, n, k.data());
  if (dtype == D_F32)
    for (uint64_t i = 0; i < n; ++i)
      if (is_nan_key32(k[i])) return 2;
  psort(k, threads);
  for (uint32_t i = 0; i < nk; ++i) {
    uint32_t kk = k[n - ks[i]];
    if (dtype == D_F32) ((float*)out)[i] = unkey_f32(kk);
    else if (dtype == D_U32) ((uint32_t*)out)[i] = kk;
    else ((int32_t*)out)[i] = (int32_t)(kk ^ 0x80000000u);
  }
  return 0;
}

// nth_element multiselect: ascending ranks sorted, each segment between two
// consecutive ranks partitioned independently (parallel over segments after
// a sequential split at the median rank levels).
template <class K>
static void nth_rec(K* a, uint64_t lo, uint64_t hi, const uint64_t* r, uint32_t r0, uint32_t r1, int depth) {
  // a[lo..hi) contains ranks r[r0..r1) (absolute positions)
  if (r0 >= r1 || hi - lo <= 1) return;
  uint32_t mid = (r0 + r1) / 2;
  uint64_t p = r[mid];
  std::nth_element(a + lo, a + p, a + hi);
  if (depth < 6) {
#pragma omp task
    nth_rec(a, lo, p, r, r0, mid, depth + 1);
#pragma omp task
    nth_rec(a, p + 1, hi, r, mid + 1, r1, depth + 1);
#pragma omp taskwait
  } else {
    nth_rec(a, lo, p, r, r0, mid, depth + 1);
    nth_rec(a, p + 1, hi, r, mid + 1, r1, depth + 1);
  }
}

template <class K>
static void nth_multi(std::vector<K>& k, const uint32_t* ks, uint32_t nk, uint64_t n, std::vector<K>& res,
                      int threads) {
  std::vector<uint64_t> r(nk);
  for (uint32_t i = 0; i < nk; ++i) r[i] = n - ks[i];
  std::vector<uint64_t> u(r);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  omp_set_num_threads(threads > 0 ? threads : 1);
#pragma omp parallel
#pragma omp single
  nth_rec(k.data(), 0, n, u.data(), 0, (uint32_t)u.size(), 0);
  res.resize(nk);
  for (uint32_t i = 0; i < nk; ++i) res[i] = k[r[i]];
}

// Parity: pinned against oracle_sort_and_choose (tests/test_oracle.py).
extern "C" int oracle_nth_multiselect(int dtype, const void* x, uint64_t n, const uint32_t* ks, uint32_t nk, void* out,
                           int threads) {
  for (uint32_t i = 0; i < nk; ++i)
    if (ks[i] < 1 || ks[i] > n) return 1;
  if (dtype == D_F64) {
    std::vector<uint64_t> k(n), res;
    oracle_keys(dtype, x, n, k.data());
    nth_multi(k, ks, nk, n, res, threads);
    for (uint32_t i = 0; i < nk; ++i) ((double*)out)[i] = unkey_f64(res[i]);
    return 0;
  }
  std::vector<uint32_t> k(n), res;
  oracle_keys(dtype, x, n, k.data());
  nth_multi(k, ks, nk, n, res, threads);
  for (uint32_t i = 0; i < nk; ++i) {
    uint32_t kk = res[i];
    if (dtype == D_F32) ((float*)out)[i] = unkey_f32(kk);
    else if (dtype == D_U32) ((uint32_t*)out)[i] = kk;
    else ((int32_t*)out)[i] = (int32_t)(kk ^ 0x80000000u);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Philox4x32-10.  Parity: pinned by the published Random123 known-answer
// vectors (tests/golden/philox_kat.json).
// ---------------------------------------------------------------------------
extern "C" void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void draw(uint64_t seed, uint32_t stream, uint64_t idx, uint32_t* w) {
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u};
  uint32_t k[2] = {(uint32_t)seed ^ 0x67676D73u, (uint32_t)(seed >> 32)};
  oracle_philox4x32_10(c, k, w);
}

// ---------------------------------------------------------------------------
// Deterministic transcendental functions in plain IEEE double arithmetic
// (no FMA, no library calls) so the HIP generators and this oracle agree bit
// for bit.  Accuracy is pinned against libm in tests/test_oracle.py.
// ---------------------------------------------------------------------------
static inline double dm_bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
static inline uint64_t dm_u64(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }

// natural log for positive normal x
static double dm_log(double x) {
  uint64_t b = dm_u64(x);
  int e = (int)((b >> 52) & 0x7FF) - 1022;
  double m = dm_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);  // [0.5, 1)
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
  const double ln2_hi = dm_bits(0x3FE62E42FEE00000ull), ln2_lo = dm_bits(0x3DEA39EF35793C76ull);
  double de = (double)e;
  return de * ln2_hi + (de * ln2_lo + 2.0 * s * p);
}

// exp for |x| < 700
static double dm_exp(double x) {
  const double inv_ln2 = 1.44269504088896338700;
  const double ln2_hi = dm_bits(0x3FE62E42FEE00000ull), ln2_lo = dm_bits(0x3DEA39EF35793C76ull);
  double kd = std::floor(x * inv_ln2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 2432902008176640000.0;  // 1/20!
  const double inv_fact[20] = {1.0 / 121645100408832000.0, 1.0 / 6402373705728000.0, 1.0 / 355687428096000.0,
                               1.0 / 20922789888000.0, 1.0 / 1307674368000.0, 1.0 / 87178291200.0,
                               1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                               1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0,
                               1.0 / 24.0, 1.0 / 6.0, 1.0 / 2.0, 1.0, 1.0};
  for (int i = 0; i < 20; ++i) p = p * r + inv_fact[i];
  return p * dm_bits((uint64_t)(k + 1023) << 52);
}

// sqrt by Heron iterations from an exponent-halved initial guess
static double dm_sqrt(double x) {
  if (x <= 0.0) return 0.0;
  uint64_t b = dm_u64(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double y = dm_bits((uint64_t)((e / 2) + 1023) << 52);
  for (int i = 0; i < 7; ++i) y = 0.5 * (y + x / y);
  return y;
}

// sin and cos of 2*pi*u for u in [0, 1): octant reduction, Taylor on [0, pi/4]
static void dm_sincos2pi(double u, double* sn, double* cs) {
  double t = u * 8.0;
  int q = (int)t;
  if (q > 7) q = 7;
  double f = t - (double)q;
  if (q & 1) f = 1.0 - f;
  const double pi_4 = 0.78539816339744830962;
  double a = f * pi_4;
  double a2 = a * a;
  // sin(a) = a * (1 - a^2/3! + a^4/5! - ... - a^18/19!), Horner in a^2
  const double sc[9] = {1.0 / 355687428096000.0, -1.0 / 1307674368000.0, 1.0 / 6227020800.0,
                        -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0, 1.0};
  double ps = -1.0 / 121645100408832000.0;  // -1/19!
  for (int i = 0; i < 9; ++i) ps = ps * a2 + sc[i];
  double s = a * ps;
  double pc = 1.0 / 2432902008176640000.0;  // 1/20!
  const double cc[10] = {-1.0 / 6402373705728000.0, 1.0 / 20922789888000.0, -1.0 / 87178291200.0,
                         1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0, -1.0 / 720.0,
                         1.0 / 24.0, -1.0 / 2.0, 1.0};
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

extern "C" void oracle_detmath(int fn, const double* x, uint64_t n, double* out, double* out2) {
  for (uint64_t i = 0; i < n; ++i) {
    if (fn == 0) out[i] = dm_log(x[i]);
    else if (fn == 1) out[i] = dm_exp(x[i]);
    else if (fn == 2) out[i] = dm_sqrt(x[i]);
    else dm_sincos2pi(x[i], &out[i], &out2[i]);
  }
}

// ---------------------------------------------------------------------------
// value generators (same definitions as the library's HIP generators;
// parity: CPU/GPU bit identity plus distribution tests)
// ---------------------------------------------------------------------------
enum {
  G_UNIFORM = 0, G_NORMAL = 1, G_HALF_NORMAL = 2, G_LOGNORMAL = 3, G_MIXTURE = 4, G_NEAR_CONSTANT = 5,
  G_CONSTANT = 6, G_CAUCHY = 7, G_MOD_INDEX = 8, G_ASCENDING = 9
};

static inline double u53w(uint32_t hi, uint32_t lo) {
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)m * (1.0 / 9007199254740992.0);
}

// standard normal of element i (Box-Muller, pairs share one Philox draw)
static double normal_at(uint64_t seed, uint32_t stream, uint64_t i) {
  uint32_t w[4];
  draw(seed, stream, i >> 1, w);
  double u1 = u53w(w[0], w[1]) + (1.0 / 9007199254740992.0);  // (0, 1]
  double u2 = u53w(w[2], w[3]);
  double r = dm_sqrt(-2.0 * dm_log(u1));
  double sn, cs;
  dm_sincos2pi(u2, &sn, &cs);
  return (i & 1) ? r * sn : r * cs;
}

struct MixParams {
  int m;
  double mu[8];
  double sigma[8];
};

extern "C" void oracle_mixture_params(uint64_t seed, int* m, double* mu, double* sigma) {
  uint32_t w[4];
  draw(seed, 5u, 0, w);
  *m = 1 + (int)(((uint64_t)w[0] * 8u) >> 32);
  for (int j = 0; j < 8; ++j) {
    draw(seed, 5u, 1 + (uint64_t)j, w);
    mu[j] = -100.0 + 200.0 * u53w(w[0], w[1]);
    sigma[j] = 0.01 + 9.99 * u53w(w[2], w[3]);
  }
}

// element i of distribution `dist` as a double (integer distributions return
// the integer value exactly)
static double gen_value(int dist, uint64_t seed, uint64_t i, double p0, double p1, const MixParams* mp,
                        uint32_t* raw) {
  uint32_t w[4];
  switch (dist) {
    case G_UNIFORM:
      draw(seed, 0u, i >> 1, w);
      *raw = w[(i & 1) * 2];
      return u53w(w[(i & 1) * 2], w[(i & 1) * 2 + 1]);
    case G_NORMAL:
      return normal_at(seed, 0u, i);
    case G_HALF_NORMAL:
      return std::fabs(normal_at(seed, 0u, i));
    case G_LOGNORMAL:
      return dm_exp(normal_at(seed, 0u, i));
    case G_MIXTURE: {
      draw(seed, 1u, i, w);
      double z = normal_at(seed, 0u, i);
      if ((double)w[0] < 0.05 * 4294967296.0) {
        double v = dm_exp(2.0 + 1.5 * z);
        return (w[1] & 1u) ? -v : v;
      }
      int c = (int)(((uint64_t)w[1] * (uint64_t)mp->m) >> 32);
      return mp->mu[c] + mp->sigma[c] * z;
    }
    case G_NEAR_CONSTANT: {
      draw(seed, 1u, i, w);
      if ((double)w[0] < p1 * 4294967296.0) return normal_at(seed, 0u, i);
      return p0;
    }
    case G_CONSTANT:
      return p0;
    case G_CAUCHY: {
      draw(seed, 0u, i >> 1, w);
      double u = u53w(w[(i & 1) * 2], w[(i & 1) * 2 + 1]);
      double sn, cs;
      dm_sincos2pi(0.5 * u + 0.75, &sn, &cs);  // angle in [1.5 pi, 2.5 pi): tan covers R
      return sn / cs;
    }
    default:
      return 0.0;
  }
}

// Fills out[0..n) for dtype.  p0/p1: distribution parameters (constant value
// / outlier fraction; modulus or range for integer distributions).
extern "C" int oracle_generate(int dtype, int dist, uint64_t n, uint64_t seed, double p0, double p1, void* out) {
  MixParams mp{};
  if (dist == G_MIXTURE) oracle_mixture_params(seed, &mp.m, mp.mu, mp.sigma);
  for (uint64_t i = 0; i < n; ++i) {
    if (dtype == D_U32 || dtype == D_I32) {
      uint32_t v;
      uint32_t w[4];
      if (dist == G_MOD_INDEX) v = (uint32_t)(i % (uint64_t)(p0 > 0 ? p0 : 1));
      else if (dist == G_ASCENDING) v = (uint32_t)i;
      else if (dist == G_CONSTANT) v = (uint32_t)p0;
      else {
        draw(seed, 0u, i >> 2, w);
        v = w[i & 3];
        if (p0 > 0) v = (uint32_t)(((uint64_t)v * (uint64_t)p0) >> 32);
      }
      ((uint32_t*)out)[i] = dtype == D_I32 ? (uint32_t)((int32_t)v) : v;
      continue;
    }
    uint32_t raw = 0;
    double d;
    if (dist == G_MOD_INDEX) d = (double)(i % (uint64_t)(p0 > 0 ? p0 : 1));
    else if (dist == G_ASCENDING) d = (double)i;
    else d = gen_value(dist, seed, i, p0, p1, &mp, &raw);
    if (dtype == D_F32) {
      float f;
      if (dist == G_UNIFORM) f = (float)(raw >> 8) * (1.0f / 16777216.0f);
      else f = (float)d;
      ((float*)out)[i] = f;
    } else {
      ((double*)out)[i] = d;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// k generators (DistributedSMOS/generateProblems.cu:584-720 formulas; the
// uniform floats come from Philox stream 4 instead of cuRAND)
// ---------------------------------------------------------------------------
enum { K_UNIFORM_RANDOM = 0, K_UNIFORM = 1, K_NORMAL = 2, K_CLUSTER = 3, K_SECTIONED = 4, K_DISTINCT_SORTED = 5,
       K_EVEN = 6, K_WINDOW = 7 };

static inline float kuni(uint64_t seed, uint64_t i) {
  uint32_t w[4];
  draw(seed, 4u, i >> 2, w);
  return (float)(w[i & 3] >> 8) * (1.0f / 16777216.0f);
}

static inline uint32_t kclamp(int64_t v, uint64_t n) {
  if (v <= 0) return 1;
  if ((uint64_t)v > n) return (uint32_t)n;
  return (uint32_t)v;
}

extern "C" int oracle_generate_ks(int kdist, uint32_t nk, uint64_t n, uint64_t seed, uint32_t* k) {
  if (nk == 0) return 0;
  switch (kdist) {
    case K_UNIFORM_RANDOM:
      for (uint32_t i = 0; i < nk; ++i) k[i] = kclamp((int64_t)(uint32_t)(kuni(seed, i) * (float)n), n);
      return 0;
    case K_UNIFORM:
      k[0] = 1;
      for (uint32_t i = 1; i + 1 < nk; ++i) k[i] = kclamp((int64_t)(uint32_t)((i / (float)(nk - 1)) * n), n);
      if (nk > 1) k[nk - 1] = (uint32_t)n;
      return 0;
    case K_NORMAL:
      for (uint32_t i = 0; i < nk; ++i) {
        float z = (float)normal_at(seed, 4u, i);
        k[i] = kclamp((int64_t)(uint32_t)((std::fabs(z) + 7.0f) * (n / 14.0)), n);
      }
      return 0;
    case K_CLUSTER: {
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
        for (int64_t j = maxassigned + 2; j < (int64_t)nk; ++j) k[j] = std::min<uint32_t>(k[j - 1] + 1, (uint32_t)n - 1);
        for (uint32_t i = 0; i < nk; ++i) k[i] = kclamp((int64_t)k[i], n);
      }
      return 0;
    }
    case K_SECTIONED: {
      k[0] = kclamp((int64_t)(uint32_t)(kuni(seed, 0) * (float)(n > nk ? n - nk : 0)), n);
      for (uint32_t i = 1; i < nk; ++i) k[i] = kclamp((int64_t)k[i - 1] + 1, n);
      return 0;
    }
    case K_DISTINCT_SORTED: {
      if (nk > n) return 1;
      std::vector<uint32_t> v;
      uint64_t j = 0;
      while (v.size() < nk) {
        uint32_t w[4];
        draw(seed, 4u, j++, w);
        v.push_back((uint32_t)(((uint64_t)w[0] * n) >> 32) + 1);
        if (v.size() == nk) {
          std::sort(v.begin(), v.end());
          v.erase(std::unique(v.begin(), v.end()), v.end());
        }
      }
      for (uint32_t i = 0; i < nk; ++i) k[i] = v[i];
      return 0;
    }
    case K_EVEN:
      for (uint32_t i = 0; i < nk; ++i) k[i] = (uint32_t)(((uint64_t)(i + 1) * n + nk - 1) / nk);
      return 0;
    case K_WINDOW: {
      uint64_t w = std::max<uint64_t>(nk, n / 100);
      if (w > n) w = n;
      uint32_t d[4];
      draw(seed, 4u, 0, d);
      uint64_t off = ((uint64_t)d[0] * (n - w + 1)) >> 32;
      for (uint32_t i = 0; i < nk; ++i) {
        draw(seed, 4u, 1 + (uint64_t)i, d);
        k[i] = (uint32_t)(off + 1 + (((uint64_t)d[0] * w) >> 32));
      }
      return 0;
    }
  }
  return 1;
}

// ---------------------------------------------------------------------------
// bucket map and first-pass histogram from a reported range table.
// ranges: R records {lo, hi (u64), base, nb, sh, mult (u32)} = 32 bytes each.
// Parity: the map is pinned by brute-force properties (monotone, in range,
// covering) in tests/test_oracle.py.
// ---------------------------------------------------------------------------
struct RangeRec {
  uint64_t lo, hi;
  uint32_t base, nb, sh, mult;
};

extern "C" uint32_t oracle_map_bucket(const void* ranges, uint32_t R, uint64_t key) {
  const RangeRec* rg = (const RangeRec*)ranges;
  // last range with lo <= key
  uint32_t a = 0, b = R - 1;
  while (a < b) {
    uint32_t m = (a + b + 1) / 2;
    if (rg[m].lo <= key) a = m; else b = m - 1;
  }
  const RangeRec& r = rg[a];
  uint64_t t64 = (key - r.lo) >> r.sh;
  uint32_t t = (uint32_t)t64;
  uint32_t off = r.mult ? (uint32_t)(((uint64_t)t * r.mult) >> 32) : t;
  if (off > r.nb - 1) off = r.nb - 1;
  return r.base + off;
}

extern "C" int oracle_first_pass_hist(int dtype, const void* x, uint64_t n, const void* ranges, uint32_t R, uint32_t B,
                           uint32_t* hist) {
  std::memset(hist, 0, (size_t)B * 4);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t key;
    if (dtype == D_F32) key = key_f32(((const float*)x)[i]);
    else if (dtype == D_F64) key = key_f64(((const double*)x)[i]);
    else if (dtype == D_U32) key = ((const uint32_t*)x)[i];
    else key = ((const uint32_t*)x)[i] ^ 0x80000000u;
    uint32_t b = oracle_map_bucket(ranges, R, key);
    if (b >= B) return 1;
    hist[b]++;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Learned splitter MLP (65 -> 32 ReLU -> 1, sigmoid output, squared error on
// the sigmoid).  Parameter layout: W1[32][65], b1[32], W2[32], b2.
// forward: out[b] = sigmoid(b2 + sum_h W2[h] relu(b1[h] + W1[h] . x_b))
// backward: mean squared error 0.5 * mean_b (out_b - y_b)^2, gradient into g.
// Parity: pinned by central finite differences in tests/test_oracle.py.
// ---------------------------------------------------------------------------
enum { MI = 65, MH = 32, MP = MH * MI + MH + MH + 1 };

extern "C" void oracle_mlp_forward(const float* w, const float* x, uint32_t nb, float* out) {
  const float* W1 = w;
  const float* b1 = w + MH * MI;
  const float* W2 = b1 + MH;
  float b2 = W2[MH];
  for (uint32_t b = 0; b < nb; ++b) {
    float o = b2;
    for (int h = 0; h < MH; ++h) {
      float z = b1[h];
      for (int i = 0; i < MI; ++i) z += W1[h * MI + i] * x[(size_t)b * MI + i];
      o += W2[h] * (z > 0.0f ? z : 0.0f);
    }
    out[b] = 1.0f / (1.0f + std::exp(-o));
  }
}

extern "C" double oracle_mlp_loss_grad(const float* w, const float* x, const float* y, uint32_t nb, float* g) {
  const float* W1 = w;
  const float* b1 = w + MH * MI;
  const float* W2 = b1 + MH;
  float b2 = W2[MH];
  std::vector<double> gd(MP, 0.0);
  double loss = 0.0;
  for (uint32_t b = 0; b < nb; ++b) {
    float z[MH], a[MH];
    float o = b2;
    for (int h = 0; h < MH; ++h) {
      float s = b1[h];
      for (int i = 0; i < MI; ++i) s += W1[h * MI + i] * x[(size_t)b * MI + i];
      z[h] = s;
      a[h] = s > 0.0f ? s : 0.0f;
      o += W2[h] * a[h];
    }
    double p = 1.0 / (1.0 + std::exp(-(double)o));
    double e = p - (double)y[b];
    loss += 0.5 * e * e;
    double d_o = e * p * (1.0 - p) / nb;
    gd[MH * MI + MH + MH] += d_o;
    for (int h = 0; h < MH; ++h) {
      gd[MH * MI + MH + h] += d_o * a[h];
      if (z[h] > 0.0f) {
        double dz = d_o * W2[h];
        gd[MH * MI + h] += dz;
        for (int i = 0; i < MI; ++i) gd[h * MI + i] += dz * x[(size_t)b * MI + i];
      }
    }
  }
  if (g)
    for (int i = 0; i < MP; ++i) g[i] = (float)gd[i];
  return loss / nb;
}

