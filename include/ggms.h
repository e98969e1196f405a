/*
 * ggms.h -- C ABI of the MI355X multiple-order-statistics selection library.
 *
 * Given a device vector of n values (1 <= n <= 2^31 - 1, the reference's int
 * length) and a list of k ranks, return the k-th value for every rank.  Ranks
 * follow the reference wrapper convention: k = 1 is the LARGEST value, k = n
 * the smallest
 * (reference: DistributedSMOS/iterativeSMOS.cu:1587-1588 converts k-th largest
 * to k-th smallest with length - k + 1; the sort-and-choose oracle picks
 * sorted[n - k], DistributedSMOS/iterativeSMOSTimingFunctions.cu:67-73).
 *
 * Results are bit-exact: values are compared through an order-preserving
 * integer key (fp32 -> u32, fp64 -> u64), so -0.0 orders below +0.0.  An input
 * holding a NaN is rejected with GGMS_ENAN.
 *
 * All pointers named d_* are device pointers (hipMalloc'd); *_host pointers are
 * host memory.  No function keeps a pointer after it returns.  The input vector
 * is never modified (the reference's iterativeSMOS clobbers it; this library
 * keeps it intact).
 */
#ifndef GGMS_H
#define GGMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GGMS_OK = 0,
  GGMS_EINVAL = 1,   /* bad argument: null pointer, n == 0, k outside [1, n], unknown enum */
  GGMS_ENAN = 2,     /* the input vector holds a NaN */
  GGMS_ENOMEM = 3,   /* device or host allocation failed */
  GGMS_EDEVICE = 4,  /* HIP runtime error (launch, copy, sync) */
  GGMS_ENOTSUP = 5   /* configuration not supported (e.g. learned splitter without weights) */
} ggms_status;

typedef enum { GGMS_F32 = 0, GGMS_F64 = 1, GGMS_U32 = 2, GGMS_I32 = 3 } ggms_dtype;

typedef enum {
  GGMS_ALG_BUCKET = 0,           /* bucketMultiselect: ONE bucketing pass; every k-bucket is compacted into its
                                    own segment and finished there (LDS radix select / sort, or a radix select
                                    from global memory for a segment above the LDS capacity) */
  GGMS_ALG_ITERATIVE = 1,        /* iterativeSMOS: re-bucket the k-buckets until they are small */
  GGMS_ALG_DISTRIBUTED_1GPU = 2, /* distributedSMOS collapsed to one GPU: strided sample, no exchange */
  GGMS_ALG_DISTRIBUTED_BUCKET_1GPU = 3 /* distributedBucketMultiselect collapsed to one GPU: strided sample,
                                    one bucketing pass, k-buckets finished as in GGMS_ALG_BUCKET */
} ggms_alg;

typedef enum {
  GGMS_SPLIT_LINEAR = 0,  /* reference splitter: 1024 random samples, 17 pivots, linear slope per range */
  GGMS_SPLIT_LEARNED = 1, /* MLP-predicted pivots from a sorted-sample sketch; linear slope inside ranges */
  GGMS_SPLIT_SKETCH = 2   /* same sample and range count as LEARNED, pivots read straight off the sample */
} ggms_splitter;

/* Tunables.  Zero means "library default" for every numeric field. */
typedef struct {
  uint32_t num_buckets;   /* buckets per pass (0 = auto: 14336, or 16384 when the first pass is segmented and n >= 2^27; max 16384; the reference uses 8192) */
  uint32_t num_pivots;    /* linear splitter pivots (default 17 => 16 ranges) */
  uint32_t sample_size;   /* splitter sample (default 1024, 2048 for the linear splitter on 4-byte keys with n >= 2^27 and few ks; 8192 for the one-GPU distributed form) */
  uint32_t block_threads; /* workgroup size of the streaming kernels (default 1024) */
  uint32_t blocks_per_cu; /* streaming workgroups per CU (0 = auto: 2 in the first pass, 1 after; 1 throughout when the first pass is segmented and n >= 2^27) */
  uint32_t splitter;      /* ggms_splitter (default LINEAR) */
  uint32_t learned_pieces;/* ranges produced by the learned / sketch splitter (0 = auto: 64, or 128 when the first pass is segmented and n >= 2^27) */
  uint32_t max_passes;    /* safety cap on bucketing passes (default 64) */
  uint64_t seed;          /* sampling seed (default 0x5EED) */
  void* stream;           /* hipStream_t to run on; NULL = the null stream */
  uint32_t profile;       /* nonzero: time every kernel with HIP events on the launch stream (see ggms_stats) */
  uint32_t schedule;      /* ggms_schedule (default HOST) */
  uint32_t device_passes; /* DEVICE schedule: bucketing passes launched (0 = 3, max 8); the last one finishes
                             every k-bucket left whatever its size */
  uint32_t first_pass;    /* ggms_first_pass: layout of the first pass's output (iterative forms; default AUTO) */
  uint32_t reserved[4];
} ggms_opts;

/* Layout of the first pass's output in the iterative forms (ITERATIVE and
 * DISTRIBUTED_1GPU; the one-pass forms are always SEGMENTED).
 *   MERGED:    every k-bucket goes on to one second pass that re-buckets them
 *              all together (the reference's iterativeSMOS loop); few ks.
 *   SEGMENTED: every k-bucket is compacted into a segment of its own; those
 *              within the finish capacity are finished right away, the others
 *              re-bucketed segment by segment; many ks.
 *   AUTO:      SEGMENTED when there is at least one k per four first-pass
 *              buckets, MERGED otherwise.  Results are identical either way. */
typedef enum { GGMS_FIRST_PASS_AUTO = 0, GGMS_FIRST_PASS_MERGED = 1, GGMS_FIRST_PASS_SEGMENTED = 2 } ggms_first_pass;

/* How a call's passes are driven.
 *   HOST:   the host reads one control block per pass (from pinned memory, while
 *           the pass's compaction runs) and launches the next pass with exact
 *           shapes; stops as soon as every k is answered.
 *   DEVICE: every launch of every pass is enqueued at once with shapes sized
 *           for the largest pass; the kernels read their control blocks and a
 *           pass that turned out empty returns at once.  No host read between
 *           kernels: ggms_select_device with this schedule is asynchronous and
 *           can be captured in a hipGraph (reserve the workspace first with
 *           ggms_ctx_reserve; read the status with ggms_ctx_device_status once
 *           the stream has completed).  Replaces the reference's host loop
 *           (DistributedSMOS/iterativeSMOS.cu:1348-1512). */
typedef enum { GGMS_SCHEDULE_HOST = 0, GGMS_SCHEDULE_DEVICE = 1 } ggms_schedule;

/* Per-call statistics of the last selection run on a context. */
#define GGMS_MAX_PASS_STATS 64
/* kernel kinds timed when ggms_opts.profile is set */
enum {
  GGMS_KERNEL_PLAN0 = 0, GGMS_KERNEL_COUNT0 = 1, GGMS_KERNEL_COUNTS = 2, GGMS_KERNEL_COLSUM = 3, GGMS_KERNEL_PLAN = 4,
  GGMS_KERNEL_COMPACT0 = 5, GGMS_KERNEL_COMPACTS = 6, GGMS_KERNEL_FINISH = 7, GGMS_KERNEL_OTHER = 8,
  GGMS_KERNEL_KINDS = 9
};
typedef struct {
  uint32_t passes;                          /* bucketing passes run */
  uint32_t finished_segments;               /* k-buckets resolved by the LDS sort */
  uint64_t n_in[GGMS_MAX_PASS_STATS];       /* elements streamed by the count kernel of each pass */
  uint32_t active[GGMS_MAX_PASS_STATS];     /* active (k-holding) buckets found by each pass */
  uint64_t n_kept[GGMS_MAX_PASS_STATS];     /* elements copied forward by each pass's compaction */
  uint32_t buckets[GGMS_MAX_PASS_STATS];    /* buckets used by each pass */
  float count_ms[GGMS_MAX_PASS_STATS];      /* profile: count kernel time of each pass */
  float compact_ms[GGMS_MAX_PASS_STATS];    /* profile: compaction kernel time of each pass */
  float kernel_ms[GGMS_KERNEL_KINDS];       /* profile: total time per kernel kind */
  uint32_t kernel_launches[GGMS_KERNEL_KINDS];
} ggms_stats;

typedef struct ggms_ctx ggms_ctx;

/* Fill *opts with library defaults. */
void ggms_opts_default(ggms_opts* opts);

/* Human-readable status name. */
const char* ggms_status_str(ggms_status s);

/* A context owns the device workspace (grown on demand, reused across calls)
 * for one device.  A context is not re-entrant: use one per host thread. */
ggms_status ggms_ctx_create(int device, ggms_ctx** out);
void ggms_ctx_destroy(ggms_ctx* ctx);
ggms_status ggms_ctx_stats(const ggms_ctx* ctx, ggms_stats* out);

/* First-pass tiling of an n-element vector over at most `workgroups` streaming
 * workgroups, and no more workgroups than 16384-element pieces: every
 * workgroup but the last owns `tile` consecutive elements (a multiple of 64),
 * the last owns the rest; *groups workgroups in all.  Host-only (no device is
 * touched).
 * GGMS_EINVAL for n == 0 or n > 2^31 - 1 (the reference's int length). */
ggms_status ggms_first_pass_tiling(uint64_t n, uint32_t workgroups, uint32_t* tile, uint32_t* groups);

/* Device workspace bytes a call with these arguments needs. */
uint64_t ggms_workspace_bytes(ggms_dtype t, uint64_t n, uint32_t nk, const ggms_opts* opts);

/* Sizes the context's device workspace and pinned host block for calls of up
 * to these arguments, so that later calls (in particular DEVICE-schedule
 * calls under hipGraph capture, which never allocate) find them ready. */
ggms_status ggms_ctx_reserve(ggms_ctx* ctx, ggms_dtype t, uint64_t n, uint32_t nk, const ggms_opts* opts);

/* Status of the context's last DEVICE-schedule ggms_select_device call, from
 * the words its kernels wrote; valid once its stream has completed. */
ggms_status ggms_ctx_device_status(const ggms_ctx* ctx);

/*
 * Multiselect.  d_vec: n values of type t on the device.  ks_host: nk ranks in
 * [1, n] (k = 1 is the largest), any order, duplicates allowed.  out_host: nk
 * values written in the order of ks_host.  opts may be NULL.
 * Replaces IterativeSMOS::iterativeSMOSWrapper (DistributedSMOS/iterativeSMOS.cuh:218-220),
 * BucketMultiselect::bucketMultiselectWrapper (DistributedSMOS/bucketMultiselect.cuh:105-107)
 * and DistributedSMOS::distributedSMOSWrapper (DistributedSMOS/distributedSMOS.hpp:22-25)
 * through the alg argument.
 */
ggms_status ggms_select(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n,
                        const uint32_t* ks_host, uint32_t nk, void* out_host, const ggms_opts* opts);

/* Same, with ks and outputs in device memory; asynchronous with respect to the
 * host except for one small status read per bucketing pass (HOST schedule), or
 * fully asynchronous and capturable (DEVICE schedule, see ggms_schedule). */
ggms_status ggms_select_device(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n,
                               const uint32_t* d_ks, uint32_t nk, void* d_out, const ggms_opts* opts);

/*
 * Allocation-free multiselect, for callers that cannot allocate device or
 * pinned memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same meaning as
 * ggms_select_device (ks and outputs in device memory, k = 1 the largest),
 * run with the DEVICE schedule on the current device, on a fixed scratch
 * array of GGMS_NOALLOC_SCRATCH_BYTES that the library holds on every device
 * it runs on until the process ends.  Calls no allocating routine; waits for
 * opts->stream and returns the call's status.  Limit: the call's workspace
 * (ggms_workspace_bytes with opts->schedule = GGMS_SCHEDULE_DEVICE) plus 256
 * bytes must fit the scratch -- about 4.2 million values of any type
 * (2^22 fp32 / u32 / i32 / fp64 with up to tens of thousands of ks); larger
 * calls return GGMS_ENOMEM without running.  Every call on a device shares
 * that device's scratch: calls that overlap on one device race (use one host
 * thread and one stream per device, or serialise).
 */
#define GGMS_NOALLOC_SCRATCH_BYTES (96u << 20)
ggms_status ggms_select_noalloc(ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n, const uint32_t* d_ks,
                                uint32_t nk, void* d_out, const ggms_opts* opts);

/*
 * Top-k select: out_host[i] = the (i + 1)-th largest value, i < K (the K
 * largest values, largest first), by the multiselect engine `alg` over the
 * ranks 1..K.  Replaces the reference's top-k timing path, which feeds
 * ks = 1..K to bucketMultiselectWrapper
 * (timeBucketTopkselect, DistributedSMOS/iterativeSMOSTimingFunctions.cu:306-333).
 * K == 0 does nothing; K > n is GGMS_EINVAL.
 */
ggms_status ggms_select_topk(ggms_ctx* ctx, ggms_alg alg, ggms_dtype t, const void* d_vec, uint64_t n, uint32_t K,
                             void* out_host, const ggms_opts* opts);

/*
 * Reference call shapes, one per instantiated type, using the library's
 * default context of the current device.  blocks / threads are accepted for
 * signature compatibility (the reference passes SM count and max threads) and
 * only threads is honoured, as the workgroup size when it is a multiple of 64.
 * Return value: GGMS_OK or an error status (the reference returns T and
 * ignores errors); the calling thread's last status is also kept for
 * ggms_last_wrapper_status (the C++ shim in ggms.hpp returns the reference's
 * (T)1 on success and (T)0 on failure, and callers read the reason there).
 *   IterativeSMOS::iterativeSMOSWrapper   DistributedSMOS/iterativeSMOS.cu:1579-1612
 *   BucketMultiselect::bucketMultiselectWrapper DistributedSMOS/bucketMultiselect.cu:852-899
 *   DistributedSMOS::distributedSMOSWrapper DistributedSMOS/distributedSMOS.cpp:1208-1280
 */
int ggms_iterativeSMOSWrapper_f32(float* d_vector, int length, const uint32_t* kVals_ori, int numKs, float* outputs, int blocks, int threads);
int ggms_iterativeSMOSWrapper_f64(double* d_vector, int length, const uint32_t* kVals_ori, int numKs, double* outputs, int blocks, int threads);
int ggms_iterativeSMOSWrapper_u32(uint32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, uint32_t* outputs, int blocks, int threads);
int ggms_iterativeSMOSWrapper_i32(int32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, int32_t* outputs, int blocks, int threads);
int ggms_bucketMultiselectWrapper_f32(float* d_vector, int length, const uint32_t* kVals_ori, int numKs, float* outputs, int blocks, int threads);
int ggms_bucketMultiselectWrapper_f64(double* d_vector, int length, const uint32_t* kVals_ori, int numKs, double* outputs, int blocks, int threads);
int ggms_bucketMultiselectWrapper_u32(uint32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, uint32_t* outputs, int blocks, int threads);
int ggms_bucketMultiselectWrapper_i32(int32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, int32_t* outputs, int blocks, int threads);
int ggms_distributedSMOSWrapper_f32(float* d_vector, int length, const uint32_t* kVals_ori, int numKs, float* outputs, int blocks, int threads, int rank);
int ggms_distributedSMOSWrapper_f64(double* d_vector, int length, const uint32_t* kVals_ori, int numKs, double* outputs, int blocks, int threads, int rank);
int ggms_distributedSMOSWrapper_u32(uint32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, uint32_t* outputs, int blocks, int threads, int rank);
int ggms_distributedSMOSWrapper_i32(int32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, int32_t* outputs, int blocks, int threads, int rank);
/*   DistributedBucketMultiselect::distributedBucketMultiselectWrapper
 *     DistributedSMOS/distributedBucketMultiselect.hpp:12-14, .cpp:929-1018 (one GPU holds the whole vector) */
int ggms_distributedBucketMultiselectWrapper_f32(float* d_vector, int length, const uint32_t* kVals_ori, int numKs, float* outputs, int blocks, int threads, int rank);
int ggms_distributedBucketMultiselectWrapper_f64(double* d_vector, int length, const uint32_t* kVals_ori, int numKs, double* outputs, int blocks, int threads, int rank);
int ggms_distributedBucketMultiselectWrapper_u32(uint32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, uint32_t* outputs, int blocks, int threads, int rank);
int ggms_distributedBucketMultiselectWrapper_i32(int32_t* d_vector, int length, const uint32_t* kVals_ori, int numKs, int32_t* outputs, int blocks, int threads, int rank);
/* Status of the calling thread's last ggms_*Wrapper_* call (GGMS_OK before any). */
ggms_status ggms_last_wrapper_status(void);

/*
 * Introspection of the first bucketing pass (used by the parity tests to check
 * the splitter and the fused count kernel against the oracle): runs sampling,
 * pivot/range construction and the count kernel, and copies the range table and
 * the bucket histogram to the host.  ranges_host receives *num_ranges records of
 * ggms_range_u64 (keys widened to 64 bits); hist_host receives *num_buckets
 * counts.  Capacities are given in *num_ranges / *num_buckets on entry.
 */
typedef struct {
  uint64_t lo, hi;  /* key interval, inclusive */
  uint32_t base;    /* first bucket */
  uint32_t nb;      /* bucket count */
  uint32_t sh;      /* shift of (key - lo) */
  uint32_t mult;    /* fixed-point slope; 0 = identity */
} ggms_range_u64;

ggms_status ggms_debug_first_pass(ggms_ctx* ctx, ggms_dtype t, const void* d_vec, uint64_t n, const ggms_opts* opts,
                                  ggms_range_u64* ranges_host, uint32_t* num_ranges, uint32_t* hist_host,
                                  uint32_t* num_buckets);

/* ------------------------------------------------------------------------
 * Batched selection: many independent problems of uneven sizes solved by ONE
 * persistent launch (no reference counterpart: the reference answers one
 * problem per call, IterativeSMOS::iterativeSMOSWrapper,
 * DistributedSMOS/iterativeSMOS.cu:1069-1577; each problem here has the same
 * meaning as one such call).
 *
 * Problem p owns values[offsets[p] .. offsets[p+1]) (device array, one dtype
 * for the batch), k ranks ks[koffsets[p] .. koffsets[p+1]) (device array,
 * k = 1 is the largest, 1 <= k <= size, at most GGMS_BATCH_MAX_KS per
 * problem) and writes out[koffsets[p] + j] = the ks[koffsets[p] + j]-th
 * largest of its values.  offsets / koffsets are host arrays of
 * num_problems + 1 entries.  Problem sizes are at most 2^31 - 1.
 *
 * A plan fixes the buffers and shapes; its launch can be repeated (new
 * values in the same buffers) and, with use_graph, is a hipGraph replay.
 * Per-size-class parameters (class c = problems with ceil(log2(size)) == c,
 * the last class also takes every larger size) normally come from the
 * autotuner's table.  The k ranks are read and checked when the plan is made
 * (GGMS_EINVAL for a k outside [1, size]); a NaN in a problem's values is
 * meant to be a per-problem GGMS_ENAN, reported by ggms_batch_plan_status
 * after the launch, but a launch over a batch holding NaN values has not
 * completed on the GPU in testing (cause not found): keep NaN out of
 * batched inputs, or use ggms_select, which rejects it with GGMS_ENAN.
 * ------------------------------------------------------------------------ */
#define GGMS_BATCH_CLASSES 24
#define GGMS_BATCH_MAX_KS 2048

typedef struct {
  uint32_t num_buckets;  /* first-pass buckets (16 .. 8048) */
  uint32_t sample_size;  /* sample keys (pieces + 1 .. 2048) */
  uint32_t pieces;       /* equal-quantile pieces of the sample (2 .. 64) */
  uint32_t rebucket;     /* buckets of each later pass over an oversized k-bucket (2 .. 8192; 0 = 8192) */
} ggms_batch_class;

typedef struct {
  uint32_t workgroups;     /* persistent workgroups; 0 = one per CU */
  uint32_t block_threads;  /* threads per workgroup: 256, 512 or 1024 (0 = 1024) */
  uint32_t use_graph;      /* capture the launch in a hipGraph when the plan is made */
  uint32_t reserved;
  uint64_t seed;           /* sampling seed */
  ggms_batch_class classes[GGMS_BATCH_CLASSES];  /* num_buckets == 0: built-in default */
} ggms_batch_opts;

typedef struct ggms_batch_plan ggms_batch_plan;

void ggms_batch_opts_default(ggms_batch_opts* opts);
void ggms_batch_default_classes(ggms_batch_class* classes);
ggms_status ggms_batch_plan_create(int device, ggms_dtype t, const void* d_values, uint32_t num_problems,
                                   const uint64_t* h_offsets, const uint32_t* d_ks, const uint32_t* h_koffsets,
                                   void* d_out, const ggms_batch_opts* opts, ggms_batch_plan** plan);
/* Asynchronous on stream (NULL: default stream). */
ggms_status ggms_batch_plan_launch(ggms_batch_plan* plan, void* stream);
/* Copies per-problem status (num_problems entries) to h_status; returns the first non-OK one. */
ggms_status ggms_batch_plan_status(ggms_batch_plan* plan, uint32_t* h_status);
void ggms_batch_plan_destroy(ggms_batch_plan* plan);
/* One-shot: plan, launch, wait, status, destroy. */
ggms_status ggms_select_batched(int device, ggms_dtype t, const void* d_values, uint32_t num_problems,
                                const uint64_t* h_offsets, const uint32_t* d_ks, const uint32_t* h_koffsets,
                                void* d_out, const ggms_batch_opts* opts);


/* ------------------------------------------------------------------------
 * Seeded synthetic inputs (replace DistributedSMOS/generateProblems.cu:43-513,
 * which draw from cuRAND seeded by wall-clock time).  Element i is a pure
 * function of (seed, i) through Philox4x32-10; see generate.hip for the exact
 * definitions.  p0 / p1 parametrise some distributions:
 *   UNIFORM       u32/i32: p0 > 0 scales into [0, p0)
 *   NEAR_CONSTANT value p0, except a fraction p1 of elements drawn NORMAL
 *   CONSTANT      value p0
 *   MOD_INDEX     i mod p0
 * ------------------------------------------------------------------------ */
typedef enum {
  GGMS_DIST_UNIFORM = 0, GGMS_DIST_NORMAL = 1, GGMS_DIST_HALF_NORMAL = 2, GGMS_DIST_LOGNORMAL = 3,
  GGMS_DIST_MIXTURE = 4, GGMS_DIST_NEAR_CONSTANT = 5, GGMS_DIST_CONSTANT = 6, GGMS_DIST_CAUCHY = 7,
  GGMS_DIST_MOD_INDEX = 8, GGMS_DIST_ASCENDING = 9
} ggms_dist;

/* Fill d_out[0..n) on the device (asynchronous on stream). */
ggms_status ggms_generate(ggms_dtype t, ggms_dist d, void* d_out, uint64_t n, uint64_t seed, double p0, double p1,
                          void* stream);

/* Parameters of the GGMS_DIST_MIXTURE components for a seed: m in [1, 8], mu[8], sigma[8]. */
ggms_status ggms_mixture_params(uint64_t seed, int* m, double* mu, double* sigma);

/* k-rank lists (host).  Reference k distributions of
 * DistributedSMOS/generateProblems.cu:584-720 plus DISTINCT_SORTED (distinct
 * uniform ranks, ascending), EVEN (ceil(i*n/nk)) and WINDOW (uniform inside a
 * window of max(nk, n/100) ranks at a random offset). */
typedef enum {
  GGMS_K_UNIFORM_RANDOM = 0, GGMS_K_UNIFORM = 1, GGMS_K_NORMAL = 2, GGMS_K_CLUSTER = 3, GGMS_K_SECTIONED = 4,
  GGMS_K_DISTINCT_SORTED = 5, GGMS_K_EVEN = 6, GGMS_K_WINDOW = 7
} ggms_kdist;

ggms_status ggms_generate_ks(ggms_kdist kd, uint32_t nk, uint64_t n, uint64_t seed, uint32_t* ks_host);


/* ------------------------------------------------------------------------
 * Learned splitter (GGMS_SPLIT_LEARNED).  A 65 -> 32 (ReLU) -> 1 (sigmoid)
 * fp32 MLP; parameters W1[32][65], b1[32], W2[32], b2 (2145 floats).  The
 * first pass evaluates it on windows of the sorted sample to place its pivots
 * (no reference counterpart: this is the project's extension of the
 * reference splitter, generatePivots, DistributedSMOS/iterativeSMOS.cu:1011-1054).
 * Training kernels operate on device arrays: x[nb][65], y[nb].
 * ------------------------------------------------------------------------ */
uint32_t ggms_mlp_param_count(void);
/* Install the weights the learned splitter of `device` uses (host array). */
ggms_status ggms_set_learned_weights(int device, const float* w_host, uint32_t n);
ggms_status ggms_clear_learned_weights(int device);
/* out[b] = sigmoid(MLP(x[b])) */
ggms_status ggms_mlp_forward(const float* d_w, const float* d_x, uint32_t nb, float* d_out, void* stream);
/* Workspace (floats) ggms_mlp_loss_grad needs for a batch of nb. */
uint64_t ggms_mlp_workspace_floats(uint32_t nb);
/* grad = d/dw of mean_b 0.5 (sigmoid(MLP(x_b)) - y_b)^2, loss = that mean. */
ggms_status ggms_mlp_loss_grad(const float* d_w, const float* d_x, const float* d_y, uint32_t nb, float* d_grad,
                               float* d_loss, float* d_work, void* stream);
/* One Adam step (step counts from 1). */
ggms_status ggms_mlp_adam_step(float* d_w, const float* d_grad, float* d_m, float* d_v, uint32_t step, float lr,
                               float beta1, float beta2, float eps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GGMS_H */
