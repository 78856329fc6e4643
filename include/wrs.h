/*
 * wrs.h -- C ABI of the MI355X (gfx950) weighted-random-sampling library.
 *
 * Construct-then-sample, as the reference kistenklaus/merian-wrs exposes it:
 * weights in -> PSA alias table (or ITS CDF) out -> samples out.  Every entry
 * point is asynchronous on the given HIP stream, takes plain device pointers
 * owned by the caller, and returns a wrs_status; nothing throws across the ABI.
 * A plan owns the device workspace of the builds (look-back records, partition
 * indices and prefixes, per-build scalars).  One plan must not be used by two
 * streams at the same time; distinct plans are independent.
 *
 * Each entry names the reference interface it replaces (file:line in the
 * reference repository).
 */
#ifndef WRS_H
#define WRS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* wrs_stream_t; /* == hipStream_t; NULL = default stream */

typedef enum {
    WRS_OK = 0,
    WRS_EINVAL = 1,  /* bad argument (null pointer, n == 0, n > plan max, misaligned) */
    WRS_ENOMEM = 2,  /* device allocation failed */
    WRS_EHIP = 3,    /* a HIP runtime call or launch failed; see wrs_last_error() */
    WRS_ENOTSUP = 4, /* no gfx950 device / unsupported configuration */
    WRS_EWAIT = 5,   /* a bounded device-side wait of a build gave up: its outputs are
                        invalid (reported by wrs_plan_status / wrs_batch_plan_status) */
} wrs_status;

/* Arithmetic of a build.  The reference GPU kernels are plain FP32 throughout
 * (src/device/prefix_partition/decoupled/shader.comp:96-193). */
typedef enum {
    WRS_PREC_FP32 = 0,        /* FP32 mean, scan, carry, split and pack (reference arithmetic) */
    WRS_PREC_FP64_CARRY = 1,  /* FP64 total/mean, exact look-back carry, FP64 split/pack, FP32 tile scan */
    WRS_PREC_COMPENSATED = 2, /* as FP64_CARRY with compensated (two-sum hi+lo) FP32 tile sums: default */
    WRS_PREC_FP64 = 3,        /* as FP64_CARRY with FP64 tile sums */
} wrs_precision;

/* Alias table row, 8 bytes: keep bucket i with probability p, else take a.
 * Reference: AliasTableEntry{float p; uint a} src/host/types/alias_table.hpp:12-44,
 * std430 layout src/device/wrs/alias/psa/layout/alias_table.hpp:13-19. */
typedef struct {
    float p;
    uint32_t a;
} wrs_alias_entry;

/* Split point: first i lights and j heavies, spill = residual of heavy j.
 * Reference Split{uint i; uint j; float spill} src/host/types/split.hpp:9-40;
 * the spill is kept in FP64 here. */
typedef struct {
    uint32_t i;
    uint32_t j;
    double spill;
} wrs_split;

/* Launch geometry of a build; zero fields take the defaults.  Chosen by the
 * autotuner below or by hand.  The table of a build can depend on merge_tile
 * only through rounding of the block-boundary splits (both results are valid
 * tables at the same error); builds with equal tunings are bit-identical.
 * Reference: the PSAConfig knobs splitSize and the pack/split variants
 * (src/device/wrs/alias/psa/PSA.hpp:19-41), and the split sizes its
 * benchmarks sweep (src/bench/psa_split2.cpp:34-270). */
typedef struct {
    uint32_t merge_tile;    /* split granularity: buckets per merge-pack block,
                               1024 (default), 2048 or 4096 */
    uint32_t merge_per;     /* buckets per merge-pack thread: 2, 4 (default) or 8
                               (merge_tile 4096: 4 or 8) */
    uint32_t partition;     /* prefix-partition form: 0 reduce-then-scan (default), 1 single
                               pass with decoupled look-back (per segment in a batch) */
    uint32_t direct;        /* pack form, WRS_PACK_*: DIRECT (the default) the partition
                               writes the heavies and every 64th light's position only and
                               the merge pack reads the lights from the weights; CLASSIC it
                               writes every index and weight in partition order (12N more
                               bytes; same table).  SERIAL, SUBGROUP and INLINE are the reference's
                               own strategies, kept as baselines: K = ceil(N / split_size)
                               split points and a serial sweep of each split interval --
                               SERIAL writes the K split records first and sweeps each
                               interval with one thread (ScalarSplit + ScalarPack,
                               split/scalar/shader.comp:111-133, pack/scalar/float.comp:
                               60-180), SUBGROUP sweeps the same records with 16 lanes per
                               interval that load its heavies and lights a chunk at a time
                               (ScalarSplit + SubgroupPack, pack/subgroup/shader.comp:
                               68-246), INLINE finds the split points in the sweeping
                               kernel (InlineSplitPack, splitpack/inline/shader.comp:
                               225-270).  SERIAL and INLINE give identical tables;
                               SUBGROUP's residual steps as r + (w - mean), as
                               SubgroupPack's does, where the others step (r + w) - mean;
                               all differ from the merge forms' tables by the rounding of
                               the sweep's running residual.  The first build of such a
                               plan allocates the full prefix arrays (8*max_n bytes,
                               4*max_n in WRS_PREC_FP32 and WRS_PREC_FP64_CARRY) and, for
                               SERIAL and SUBGROUP, 16*K split records.  Debug builds use
                               CLASSIC; batch plans build DIRECT with partition 0 only
                               (the sweep forms and STREAM are rejected there).
                               STREAM evaluates the same sweep with exact integer keys
                               (the mean and every weight times 2^(53 - ilogb(mean)))
                               and writes the table tile by tile in element order: an
                               aggregate pass, the heavies' keys, one pass over the
                               weights for the heavies' rows and one for the table; no
                               split search and no merge blocks.  Its alias indices are
                               those of the exact sweep (the merge forms' agree except
                               where FP64 rounding of their block offsets decides a
                               tie); its heavy probabilities round the exact residual
                               once.  The first build of such a plan allocates 30 bytes
                               per weight (keys, indices, rows and crossing offsets of
                               up to max_n - 1 heavies) and 1.3 KB per tile of 4096
                               weights; partition is ignored. */
} wrs_tuning;
#define WRS_PACK_DEFAULT 0u /* the direct form */
#define WRS_PACK_DIRECT 1u
#define WRS_PACK_CLASSIC 2u
#define WRS_PACK_SERIAL 3u
#define WRS_PACK_INLINE 4u
#define WRS_PACK_SUBGROUP 5u
#define WRS_PACK_STREAM 6u

/* Build configuration.  Reference: PSAConfig (src/device/wrs/alias/psa/PSA.hpp:19-41),
 * whose splitSize sets the split granularity. */
typedef struct {
    uint32_t split_size;     /* split granularity (K = ceil(N/split_size)) of the debug split
                                output and of the SERIAL, SUBGROUP and INLINE pack forms;
                                0 -> 32 */
    wrs_precision precision; /* default WRS_PREC_COMPENSATED */
    wrs_tuning tuning;       /* all zero: default geometry */
} wrs_config;

typedef struct wrs_plan wrs_plan;

/* Creates a plan able to build distributions of up to max_n weights
 * (max_n < 2^32).  Allocates about 8.4*max_n bytes of device memory: the
 * indices and weights in partition order (8*max_n), the prefix samples, light
 * positions, boundary splits and look-back records (max_n/3 together).  The
 * full prefix arrays (4 or 8 bytes per weight) follow on the first debug build,
 * sample-statistics call or SERIAL/INLINE build that needs them.
 * Replaces the construction of device::PSA / device::WRS and their
 * Buffers::allocate (src/device/wrs/alias/psa/PSA.hpp:43-147,
 * src/device/wrs/WRS.hpp:60-88). */
wrs_status wrs_plan_create(const wrs_config* cfg, uint64_t max_n, wrs_plan** out);
wrs_status wrs_plan_destroy(wrs_plan* plan);

/* PSA alias-table build: weights[n] (finite, >= 0, not all zero) -> table[n].
 * Passes: FP64 total, heavy/light prefix-partition, split + pack as a parallel
 * merge (exact splits at every merge_tile buckets, then an LDS merge per block).
 * Replaces device::PSA::run (src/device/wrs/alias/psa/PSA.hpp:170-237) and
 * device::WRS::build for an AliasTable config (src/device/wrs/WRS.hpp:121-143). */
wrs_status wrs_psa_build(wrs_plan* plan, const float* weights, uint64_t n,
                         wrs_alias_entry* table, wrs_stream_t stream);

/* Changes the launch geometry of an existing plan (e.g. to the autotuner's
 * choice) without reallocating: the workspace does not depend on it.  No
 * reference counterpart (the reference fixes its configuration when it
 * creates a pipeline). */
wrs_status wrs_plan_set_tuning(wrs_plan* plan, const wrs_tuning* tuning);

/* Device pointers for the intermediate results of a build (any may be NULL). */
typedef struct {
    double* mean;          /* [1]  mean (pivot and bucket size) */
    uint32_t* heavy_count; /* [1]  H */
    uint32_t* indices;     /* [n]  partition indices, heavies at [0,H), light r at n-1-r */
    float* prefix_hi;      /* [n]  partition prefixes, same layout */
    float* prefix_lo;      /* [n]  low words, prefix = hi + lo (WRS_PREC_COMPENSATED and WRS_PREC_FP64) */
    wrs_split* splits;     /* [K]  splits[k-1] = S(k), k = 1..K, splits[K-1] = (L, H-1, 0) */
} wrs_psa_debug;

/* wrs_psa_build that also copies out the intermediates (for parity tests). */
wrs_status wrs_psa_build_debug(wrs_plan* plan, const float* weights, uint64_t n,
                               wrs_alias_entry* table, const wrs_psa_debug* dbg,
                               wrs_stream_t stream);

/* FP64 total and mean of weights[n] (deterministic fixed-order reduction).
 * Replaces Mean<float>::run / AtomicMean (src/device/mean/Mean.hpp:118-140,
 * src/device/mean/atomic/shader.comp:69-82). out2 = {total, mean}. */
wrs_status wrs_mean(wrs_plan* plan, const float* weights, uint64_t n, double* out2,
                    wrs_stream_t stream);

/* Heavy/light prefix-partition around `pivot` (heavy <=> w > pivot; pass NAN
 * to use the mean).  indices/prefix_hi[n] in the reference layout (heavies
 * forward, lights reversed from the tail), prefix_lo[n] optional, heavy_count[1].
 * Replaces PrefixPartition<float>::run (src/device/prefix_partition/
 * PrefixPartition.hpp:38-206; kernel decoupled/shader.comp:313-362). */
wrs_status wrs_prefix_partition(wrs_plan* plan, const float* weights, uint64_t n, double pivot,
                                uint32_t* indices, float* prefix_hi, float* prefix_lo,
                                uint32_t* heavy_count, wrs_stream_t stream);

/* Inclusive prefix sum out[i] = in[0] + ... + in[i] (hi word, optional lo word).
 * Replaces PrefixSum<float>::run / DecoupledPrefixSum (src/device/prefix_sum/
 * decoupled/DecoupledPrefixSum.hpp:174-197, shader.comp:321-333). */
wrs_status wrs_scan_inclusive(wrs_plan* plan, const float* in, uint64_t n, float* out_hi,
                              float* out_lo, wrs_stream_t stream);

/* ITS build: normalised CDF cdf[i] = (w[0]+..+w[i]) / sum(w) (hi, optional lo).
 * Replaces ITS::build (src/device/wrs/its/ITS.hpp:90-108), which scans into an
 * un-normalised CMF. */
wrs_status wrs_its_build(wrs_plan* plan, const float* weights, uint64_t n, float* cdf_hi,
                         float* cdf_lo, wrs_stream_t stream);

/* Alias sampling: out[t] for t < s, Philox4x32-10 with key = seed and
 * counter = counter_base + t.  Replaces SampleAliasTable::run
 * (src/device/wrs/alias/sampling/AliasTableSampling.hpp:75-85) and
 * WRS::sample (src/device/wrs/WRS.hpp:145-168). */
wrs_status wrs_alias_sample(const wrs_alias_entry* table, uint64_t n, uint64_t s, uint64_t seed,
                            uint64_t counter_base, uint32_t* out, wrs_stream_t stream);

/* ITS sampling: out[t] = the smallest i with cdf(i) >= u_t (at most n - 1),
 * cdf(i) = cdf_hi[i] + cdf_lo[i] in FP64 (cdf_lo optional), u_t in (0, 1]
 * from Philox4x32-10 (key = seed, counter = counter_base + t).  cdf_lo holds
 * the low words of a split double CDF (hi = (float)c, lo = (float)(c - hi),
 * so |lo| <= ulp(hi) / 2), as wrs_its_build writes them.  Above 2^16 entries
 * (and a 16-byte aligned cdf_hi) the search runs through line indexes of the
 * CDF that the call builds in a buffer of its own per (device, stream),
 * n / 31 floats, kept for the next call on that stream and grown in stream
 * order.  Replaces InverseTransformSampling::run (src/device/wrs/its/sampling/
 * InverseTransformSampling.hpp:123-138). */
wrs_status wrs_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n, uint64_t s,
                          uint64_t seed, uint64_t counter_base, uint32_t* out,
                          wrs_stream_t stream);

/* wrs_its_sample without allocation: the same samples (bit-identical), with
 * the line indexes in a static device array of the library (4.2 MB per GPU,
 * held until the process ends) instead of a buffer allocated per stream, so
 * the call makes no device allocation of any kind.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Limit:
 * 1 <= n <= WRS_ITS_NOALLOC_MAX_N (WRS_EINVAL beyond).  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (two streams, two
 * threads) race; calls in stream order do not. */
#define WRS_ITS_NOALLOC_MAX_N (1ull << 25)
wrs_status wrs_its_sample_noalloc(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                  uint64_t s, uint64_t seed, uint64_t counter_base,
                                  uint32_t* out, wrs_stream_t stream);

/* Windowed ("cooperative") sampling, for throughput where the samples need
 * not be independent of each other.  Samples come in groups of 64 consecutive
 * counters (group g = counter / 64, with its own Philox stream); all samples
 * of a group draw from one window chosen per group, so every sample's marginal
 * distribution is exact but the samples of a group share the window.
 *  - alias: the window is `window` consecutive buckets (cyclic, min(window, n))
 *    from a uniform start; each sample takes a uniform bucket in it and the
 *    coin of wrs_alias_sample.
 *  - ITS: the group's uniform is narrowed to at most `window` CDF entries by
 *    64-part searches (parts != 0) or by halving; each sample draws uniformly
 *    within the window's CDF interval and searches only the window.
 * Replace the cooperative paths of SampleAliasTable / InverseTransformSampling
 * (src/device/wrs/alias/sampling/shader.comp:68-90,103-110, configured by
 * SampleAliasTableConfig::cooperativeSampleSize, AliasTableSampling.hpp:26-33;
 * src/device/wrs/its/sampling/shader.comp:100-129, configured by
 * InverseTransformSamplingConfig, InverseTransformSampling.hpp:69-81), whose
 * alias window is clamped to [0, N-1] (edge buckets over-sampled) and whose ITS
 * parts can miss the top of the range.  window >= 1. */
wrs_status wrs_alias_sample_window(const wrs_alias_entry* table, uint64_t n, uint64_t s,
                                   uint64_t seed, uint64_t counter_base, uint32_t window,
                                   uint32_t* out, wrs_stream_t stream);
wrs_status wrs_its_sample_window(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                 uint64_t s, uint64_t seed, uint64_t counter_base,
                                 uint32_t window, int parts, uint32_t* out,
                                 wrs_stream_t stream);

/* Cutpoint method (Chen & Asau's guide table) over a normalised ITS CDF
 * (wrs_its_build output; cdf_lo optional, as for wrs_its_sample).  guide has
 * m + 1 entries (1 <= m < 2^32): guide[g] = smallest i with cdf(i) >= g/m
 * (n - 1 if none), guide[m] = n - 1.  Replaces the guiding-table pass of
 * Cutpoint::build (src/device/wrs/cutpoint/Cutpoint.hpp:93-127,
 * guiding_table/shader.comp:37-54; its guidingTableSize is m). */
wrs_status wrs_cutpoint_build(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                              uint32_t* guide, uint64_t m, wrs_stream_t stream);
/* Guided ITS sampling: out[t] is bit-identical to wrs_its_sample's with the
 * same seed and counters (the guide only narrows the lower-bound search to
 * [guide[g], guide[g+1]], g = floor(u m)).  Replaces Cutpoint::sample
 * (Cutpoint.hpp:129-146, sampling/shader.comp:118-177), whose second uniform
 * per guide cell weights every cell by 1/m instead of by its mass. */
wrs_status wrs_cutpoint_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                               const uint32_t* guide, uint64_t m, uint64_t s, uint64_t seed,
                               uint64_t counter_base, uint32_t* out, wrs_stream_t stream);

/* Philox4x32-10 blocks: out[4t..4t+3] = philox(ctr = (lo(base+t), hi(base+t),
 * stream_word, 0), key = seed).  Replaces the PRNG module (src/device/prng/
 * philox/Philox.hpp:47-104), which is a non-standard 2x32 7-round variant. */
wrs_status wrs_philox(uint64_t seed, uint32_t stream_word, uint64_t base, uint64_t n,
                      uint32_t* out, wrs_stream_t stream);

/* Seeded synthetic weights on the device (the reference generates benchmark
 * weights with its PRNG kernel, src/bench/psa_split2.cpp:148-228). */
typedef enum {
    WRS_GEN_UNIFORM01 = 0, /* U(0,1), 24-bit */
    WRS_GEN_LOGNORMAL = 1, /* exp(a + b z), z ~ N(0,1) by Box-Muller */
    WRS_GEN_ZIPF = 2,      /* rank^-a, ranks a seeded permutation of 1..n */
    WRS_GEN_SPIKE = 3,     /* 1% of positions U(1e5,1e6), the rest U(1e-3,1) */
} wrs_generator;
wrs_status wrs_generate(wrs_generator kind, double a, double b, uint64_t seed, uint64_t n,
                        float* out, wrs_stream_t stream);

/* ---- batched (segmented) builds of many independent distributions ------
 * Segment b is weights[offsets[b] .. offsets[b+1]) (offsets[0] == 0, every
 * segment non-empty and at most 2^23 weights); its table occupies the same
 * range of `tables` with segment-local indices.  The plan copies the offsets
 * and derives its work lists from them.  Every segment gets exactly the
 * arithmetic of wrs_psa_build, so the tables are bit-identical to building
 * the segments one at a time with a plan of the same merge_tile.  No
 * reference counterpart: the reference builds one distribution per dispatch
 * sequence. */
typedef struct wrs_batch_plan wrs_batch_plan;
wrs_status wrs_batch_plan_create(const wrs_config* cfg, const uint64_t* offsets_host,
                                 uint32_t nseg, wrs_batch_plan** out);
wrs_status wrs_batch_plan_destroy(wrs_batch_plan* plan);
wrs_status wrs_psa_build_batched(wrs_batch_plan* plan, const float* weights,
                                 wrs_alias_entry* tables, wrs_stream_t stream);
/* per_segment samples of each segment: out[b*per_segment + t] is a segment-local index,
 * Philox counter = counter_base + b*per_segment + t (the counters a per-segment
 * wrs_alias_sample with that counter_base would use).  offsets_dev: nseg+1
 * offsets in device memory. */
wrs_status wrs_alias_sample_batched(const wrs_alias_entry* tables, const uint64_t* offsets_dev,
                                    uint32_t nseg, uint32_t per_segment, uint64_t seed,
                                    uint64_t counter_base, uint32_t* out, wrs_stream_t stream);

/* ---- sample statistics (on the device) ----
 * Histogram of s samples over n bins: hist has n + 1 uint64 counters, hist[n]
 * counting samples >= n; clear != 0 zeroes them first, else the counts
 * accumulate (several sampling passes).  Replaces the atomic histogram of the
 * reference's statistics (src/device/statistics/histogram/atomic/shader.comp:1-38,
 * Histogram.hpp). */
wrs_status wrs_histogram(const uint32_t* samples, uint64_t s, uint64_t n, uint64_t* hist,
                         int clear, wrs_stream_t stream);
/* Statistics of a histogram of s samples against weights[n] (q = w / sum w),
 * FP64 in a fixed order, written to out6 (device memory, 6 doubles):
 *   [0] chi-square over pools of consecutive bins that each expect at least
 *       min_expected samples (the last partial pool merged into the one before;
 *       min_expected <= 0: one pool per bin, the reference's statistic)
 *   [1] degrees of freedom (pools - 1)   [2] sum (hist[i]/s - q_i)^2
 *   [3] its expectation for an exact sampler, sum q_i (1 - q_i) / s
 *   [4] samples in bins of zero weight   [5] pools
 * RMSE = sqrt([2] / n), its expectation sqrt([3] / n).  Uses the plan's
 * normalised scan (n <= max_n).  Replaces ChiSquare / MeanSquaredError
 * (src/device/statistics/chi_square/reduce/shader.comp:40-73, rmse/mse/shader.comp:52-77),
 * whose float atomics make the sums order-dependent. */
wrs_status wrs_sample_stats(wrs_plan* plan, const float* weights, uint64_t n, const uint64_t* hist,
                            uint64_t s, double min_expected, double* out6, wrs_stream_t stream);

/* ---- autotuner ----
 * Times candidate geometries (wrs_tuning) on the caller's own data with HIP
 * events and returns the fastest, by coordinate descent: the prefix-partition
 * form at the default merge shape, then the pack form (classic, direct, and
 * for a single build the reference's SERIAL, SUBGROUP and INLINE sweeps at the
 * config's split size), then every supported merge shape with the better forms (none
 * if a sweep form won), then the two fastest candidates again, each after 20
 * builds of its own, keeping the faster.  Each candidate is the median of 9
 * builds after one warm-up; 20 untimed builds of the first candidate run
 * before the search.
 * Results are cached per process under a key of device, precision and problem
 * shape (n and split size, or the segment count and a histogram of log2
 * segment sizes); with the environment variable WRS_TUNE_CACHE naming a file they are
 * also read from and appended to that file.  Synchronises `stream`;
 * allocates its own scratch.  log (optional, log_cap entries) receives every
 * candidate timed, *log_len their number (0 for a cache hit).  No reference
 * counterpart: the reference fixes its launch shapes at pipeline creation. */
typedef struct {
    wrs_tuning tuning;
    float ms; /* median build time */
} wrs_tune_entry;
#define WRS_TUNE_NO_CACHE 1u /* measure even if the cache holds the key (and update it) */
wrs_status wrs_autotune_psa(const wrs_config* cfg, const float* weights, uint64_t n,
                            uint32_t flags, wrs_tuning* best, wrs_tune_entry* log,
                            uint32_t log_cap, uint32_t* log_len, wrs_stream_t stream);
wrs_status wrs_autotune_batched(const wrs_config* cfg, const uint64_t* offsets_host,
                                uint32_t nseg, const float* weights, uint32_t flags,
                                wrs_tuning* best, wrs_tune_entry* log, uint32_t log_cap,
                                uint32_t* log_len, wrs_stream_t stream);

/* ---- status of the device-side waits ----
 * A build waits on the device in two places: the last block of the total
 * reads the other blocks' totals (stores already issued), and the look-back
 * scans read their predecessors' records.  Each wait is bounded (about a
 * second of polling); when one gives up, the build records the failure in the
 * plan, its remaining kernels return at once (no table or CDF is written) and
 * the next build starts clean.  The reference bounds its mean look-back the
 * same way (MAX_IT, src/device/mean/decoupled/float.comp:143,194-197) but
 * reports nothing.
 * wrs_plan_status synchronises `stream` and returns WRS_EWAIT if the plan's
 * most recent build (or scan) failed so, WRS_OK otherwise. */
wrs_status wrs_plan_status(wrs_plan* plan, wrs_stream_t stream);
wrs_status wrs_batch_plan_status(wrs_batch_plan* plan, wrs_stream_t stream);
/* Failure injection for tests of the error path: wait_polls bounds every wait
 * of the plan's builds (0: the default), hooks (bits) withhold a publication
 * so that a wait must give up: 1 the first block's total, 2 the first record
 * of the tile-aggregate scan, 4 the first block's total published under the
 * previous build's tag (as a late store of that build would land), which the
 * reader must not take.  Not for production use. */
wrs_status wrs_plan_set_test_hooks(wrs_plan* plan, uint32_t wait_polls, uint32_t hooks);

/* ---- introspection ---- */
const char* wrs_last_error(void);         /* thread-local message of the last failure */
const char* wrs_version(void);
/* Per-phase device times (ms) of the most recent PSA build on this plan when
 * timing is enabled (HIP events on the build's stream): [0] reduce,
 * [1] finalize, [2] tile aggregates, [3] tile scan, [4] partition emit,
 * [5] block-boundary splits, [6] merge pack.  Writes min(count,
 * WRS_PHASE_COUNT) values.  Replaces the merian profiler scopes around each
 * dispatch (src/device/wrs/alias/psa/splitpack/inline/InlineSplitPack.hpp:109-150). */
#define WRS_PHASE_COUNT 7
wrs_status wrs_plan_set_timing(wrs_plan* plan, int enable);
wrs_status wrs_plan_phase_times(wrs_plan* plan, float* ms, uint32_t count);

#ifdef __cplusplus
}
#endif
#endif /* WRS_H */
