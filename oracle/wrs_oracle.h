/*
 * wrs_oracle.h -- CPU oracle for the weighted-random-sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it.  The product (merian-wrs_amd/) never links or calls it.
 *
 * It restates, function by function, the sequential CPU reference path of
 * kistenklaus/merian-wrs (the headers under src/host/reference/), plus the Philox4x32-10
 * counter-based generator (Salmon et al., SC'11, Random123) and the seeded
 * synthetic weight generators the tests and benchmarks use.
 *
 * Each definition names the reference file:line it follows.  Two scalar
 * instantiations exist for the templated reference code: *_f32 (T = P = float,
 * the reference's own test instantiation) and *_f64 (T = P = double, the
 * instantiation the reference's tests-of-tests use, src/host/assert/test.cpp:317-346).
 *
 * Documented deviations from the reference text:
 *   - ceilMulDiv uses 64-bit intermediates (src/host/reference/split.hpp:59-65
 *     overflows in uint32 once K > 2^16); value-identical where it does not.
 *   - two split rules: the reference CPU's (split.hpp:14-57) and the reference
 *     GPU kernels' lower bound (split/scalar/shader.comp:72-109); only the
 *     latter always reproduces the sequential sweep (see oro_split_lb_*).
 *   - split(): the reference reads out of bounds / asserts on paths that are
 *     unreachable in exact arithmetic (j == H, j == 0 with sigma > target).
 *     Those paths return a clamped in-bounds split (see oro_split_*).
 *   - pack2(): interval ends are clamped to be >= the interval start, as the
 *     reference's InlineSplitPack shader does (splitpack/inline/shader.comp:266-267).
 *   - psa_alias_table(): H == 0 or L == 0 (all weights equal up to rounding of
 *     the mean) yields the identity table, as sweeping_alias_table.hpp:66-73
 *     does; the reference PSA path asserts there.
 */
#ifndef WRS_ORACLE_H
#define WRS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (Random123) ---------------------------------------- */
void oro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* out[4*t .. 4*t+3] = philox(ctr = (lo(base+t), hi(base+t), stream, 0), key = seed) */
void oro_philox_block(uint64_t seed, uint32_t stream, uint64_t base, uint64_t n, uint32_t* out);

/* ---- seeded synthetic generators (definitions shared with the product) -- */
void oro_gen_uniform01(uint64_t seed, uint64_t n, float* out);
void oro_gen_lognormal(uint64_t seed, double mu, double sigma, uint64_t n, float* out);
uint64_t oro_feistel_perm(uint64_t seed, uint64_t n, uint64_t i);
void oro_gen_zipf(uint64_t seed, double s, uint64_t n, float* out);
void oro_gen_spike(uint64_t seed, uint64_t n, float* out);
void oro_segment_sizes(uint64_t seed, uint32_t b, uint64_t lo, uint64_t hi, uint64_t* sizes);

/* ---- reductions (src/host/reference/reduce.hpp:81-92) ------------------- */
float oro_kahan_reduce_f32(const float* x, uint64_t n);
double oro_kahan_reduce_f64(const double* x, uint64_t n);

/* ---- stable heavy/light partition (src/host/reference/partition.hpp:35-53)
 * idx_cpu: CPU layout (heavies ascending at [0,H), lights ascending at [H,N)).
 * Returns H.  The comparison is `w > pivot` in the instantiation's type. */
uint32_t oro_stable_partition_indices_f32(const float* w, uint32_t n, float pivot, uint32_t* idx_cpu);
uint32_t oro_stable_partition_indices_f64(const double* w, uint32_t n, double pivot, uint32_t* idx_cpu);

/* ---- Kahan inclusive prefix sum (src/host/reference/prefix_sum.hpp:13-34) */
void oro_prefix_sum_f32(const float* in, uint64_t n, float* out);
void oro_prefix_sum_f64(const double* in, uint64_t n, double* out);

/* ---- split search (src/host/reference/split.hpp:14-57, 59-86) ----------- */
typedef struct { uint32_t i; uint32_t j; float spill; } oro_split_f32_t;
typedef struct { uint32_t i; uint32_t j; double spill; } oro_split_f64_t;
uint64_t oro_ceil_mul_div(uint64_t a, uint64_t b, uint64_t c);
oro_split_f32_t oro_split_f32(const float* hp, uint32_t H, const float* lp, uint32_t L, float mean, uint32_t n);
oro_split_f64_t oro_split_f64(const double* hp, uint32_t H, const double* lp, uint32_t L, double mean, uint32_t n);
/* Lower-bound split of the reference GPU kernels (split/scalar/shader.comp:72-109):
 * the largest j with sigma(min(n-j,L), j) <= mean*n.  The reference CPU split
 * accepts any j with sigma(i,j) <= mean*n < sigma(i,j+1); that set can hold a
 * second j (one heavy fewer) that is not the state of the sequential sweep, and
 * the CPU binary search sometimes returns it (non-monotone splits, wrong table).
 * The lower-bound j is always the sweep state. */
oro_split_f32_t oro_split_lb_f32(const float* hp, uint32_t H, const float* lp, uint32_t L, float mean, uint32_t n);
oro_split_f64_t oro_split_lb_f64(const double* hp, uint32_t H, const double* lp, uint32_t L, double mean, uint32_t n);
#define ORO_SPLIT_REFERENCE_CPU 0 /* split.hpp:14-57 */
#define ORO_SPLIT_LOWER_BOUND 1   /* split/scalar/shader.comp:72-109 */
void oro_splitK_f32(const float* hp, uint32_t H, const float* lp, uint32_t L, float mean,
                    uint32_t N, uint32_t K, int rule, oro_split_f32_t* splits);
void oro_splitK_f64(const double* hp, uint32_t H, const double* lp, uint32_t L, double mean,
                    uint32_t N, uint32_t K, int rule, oro_split_f64_t* splits);
/* (i, j) after n buckets of the split-free sequential sweep, n = 0..N-1 */
void oro_sweep_states_f64(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                          const double* w, double mean, uint32_t* st_i, uint32_t* st_j);

/* ---- pack (src/host/reference/pack.hpp:14-83, 179-202) ----------------- */
void oro_pack_splits_f32(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                         const float* w, uint32_t N, float mean,
                         const oro_split_f32_t* splits, uint32_t K, float* p, uint32_t* a);
void oro_pack_splits_f64(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                         const double* w, uint32_t N, double mean,
                         const oro_split_f64_t* splits, uint32_t K, double* p, uint32_t* a);
/* SubgroupPack (src/device/wrs/alias/psa/pack/subgroup/shader.comp:68-246) as a
 * sequential sweep: pack2 with r = r + (weight - mean) and the spill taken as
 * it is after the first interval (see the definition for the deviation). */
void oro_pack_grouped_f32(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                          const float* w, uint32_t N, float mean,
                          const oro_split_f32_t* splits, uint32_t K, float* p, uint32_t* a);
void oro_pack_grouped_f64(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                          const double* w, uint32_t N, double mean,
                          const oro_split_f64_t* splits, uint32_t K, double* p, uint32_t* a);

/* ---- full PSA build (src/host/reference/psa_alias_table.hpp:16-67) -----
 * reference_scale != 0 multiplies the weights by 1e-4 first, as the reference
 * does at psa_alias_table.hpp:25-28.  Optional debug outputs may be NULL. */
typedef struct {
    double mean;          /* mean used as pivot and bucket size */
    uint32_t heavy_count; /* H */
    uint32_t* idx_cpu;    /* [N] partition indices, CPU layout (may be NULL) */
    double* heavy_prefix; /* [N] heavy prefix in [0,H) (may be NULL) */
    double* light_prefix; /* [N] light prefix in [0,L) (may be NULL) */
    oro_split_f64_t* splits; /* [K] (may be NULL) */
} oro_psa_debug;
void oro_psa_alias_table_f32(const float* w, uint32_t n, uint32_t K, int reference_scale,
                             int split_rule, float* p, uint32_t* a, oro_psa_debug* dbg);
/* mean_override: if not NaN, used instead of the Kahan mean (to compare a GPU
 * build "given the pivot"). */
void oro_psa_alias_table_f64(const float* w, uint32_t n, uint32_t K, int reference_scale,
                             int split_rule, double mean_override, double* p, uint32_t* a,
                             oro_psa_debug* dbg);

/* The same sweep with exact integer keys (the product's streamed pack form,
 * WRS_PACK_STREAM): the mean and every weight times 2^(53 - ilogb(mean)),
 * rounded to integers; heavy <=> w > mean.  f32 != 0: p divides in FP32 (the
 * FP32 mode), else multiplies by the FP64 reciprocal. */
void oro_psa_stream_table(const float* w, uint32_t n, double mean, int f32, float* p,
                          uint32_t* a);

/* ---- sampling ----------------------------------------------------------- */
/* alias: src/host/reference/sample_alias_table.hpp:15-41 acceptance rule
 * (u2 > p -> alias) with the Philox stream of the product. */
void oro_sample_alias(const float* p, const uint32_t* a, uint64_t n, uint64_t s, uint64_t seed,
                      uint64_t counter_base, uint32_t* out);
/* ITS: lower bound (smallest idx with cdf(idx) >= u), as the reference shader's
 * binarySearch (src/device/wrs/its/sampling/shader.comp:131-140). cdf_lo may be NULL. */
void oro_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n, uint64_t s,
                    uint64_t seed, uint64_t counter_base, uint32_t* out);

/* ---- implied distribution (src/host/reference/inverse_alias_table.hpp:11-27),
 * FP64, normalised so that sum(q) == 1 in exact arithmetic. */
void oro_implied_distribution_f32(const float* p, const uint32_t* a, uint64_t n, double* q);
void oro_implied_distribution_f64(const double* p, const uint32_t* a, uint64_t n, double* q);

#ifdef __cplusplus
}
#endif
#endif
