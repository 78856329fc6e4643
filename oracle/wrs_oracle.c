/*
 * wrs_oracle.c -- CPU oracle (test infrastructure only; see wrs_oracle.h).
 *
 * Build: oracle/build_oracle.sh (gcc -O2 -ffp-contract=off, so every float
 * operation below is a single IEEE operation in the written order, matching the
 * sequential reference code it restates).
 */
#include "wrs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORO_MIN(a, b) ((a) < (b) ? (a) : (b))
#define ORO_MAX(a, b) ((a) > (b) ? (a) : (b))

/* ======================================================================== */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11.  Multipliers and Weyl constants as published.    */
/* ======================================================================== */
void oro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t m0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t m1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)m1;
        const uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)m0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static void philox_at(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                      uint32_t out[4]) {
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oro_philox4x32_10(ctr, key, out);
}

void oro_philox_block(uint64_t seed, uint32_t stream, uint64_t base, uint64_t n, uint32_t* out) {
    for (uint64_t t = 0; t < n; ++t) {
        const uint64_t c = base + t;
        philox_at(seed, (uint32_t)c, (uint32_t)(c >> 32), stream, 0u, out + 4 * t);
    }
}

/* Counter-word tags that keep the generator streams apart from the sampler
 * streams (which use tag 0). */
#define TAG_GEN 0x47454E31u /* "GEN1" */
#define TAG_FST 0x46535431u /* "FST1" */
#define TAG_SEG 0x53454731u /* "SEG1" */

static float u24(uint32_t x) { return (float)(x >> 8) * 0x1.0p-24f; }
static double u53(uint32_t hi, uint32_t lo) {
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * 0x1.0p-53;
}
static double u53_open0(uint32_t hi, uint32_t lo) {
    return (double)(((((uint64_t)hi << 32) | lo) >> 11) + 1u) * 0x1.0p-53;
}

void oro_gen_uniform01(uint64_t seed, uint64_t n, float* out) {
    uint32_t r[4];
    for (uint64_t i = 0; i < n; ++i) {
        philox_at(seed, (uint32_t)i, (uint32_t)(i >> 32), 0u, TAG_GEN, r);
        out[i] = u24(r[0]);
    }
}

void oro_gen_lognormal(uint64_t seed, double mu, double sigma, uint64_t n, float* out) {
    uint32_t r[4];
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < n; ++i) {
        philox_at(seed, (uint32_t)i, (uint32_t)(i >> 32), 0u, TAG_GEN, r);
        const double u1 = u53_open0(r[0], r[1]);
        const double u2 = u53(r[2], r[3]);
        const double z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
        out[i] = (float)exp(mu + sigma * z);
    }
}

/* Seeded pseudo-random bijection of [0,n): a 4-round balanced Feistel network
 * on 2*half bits whose round function is a Philox output, restricted to [0,n)
 * by cycle walking (every orbit of a permutation returns to [0,n)). */
uint64_t oro_feistel_perm(uint64_t seed, uint64_t n, uint64_t i) {
    unsigned bits = 1;
    while (bits < 64 && (1ull << bits) < n) ++bits;
    const unsigned half = (bits + 1) / 2;
    const uint64_t mask = (1ull << half) - 1ull;
    uint64_t x = i;
    uint32_t r[4];
    do {
        uint64_t L = x >> half, R = x & mask;
        for (uint32_t round = 0; round < 4; ++round) {
            philox_at(seed, (uint32_t)R, round, 0u, TAG_FST, r);
            const uint64_t F = (((uint64_t)r[1] << 32) | r[0]) & mask;
            const uint64_t nl = R;
            R = L ^ F;
            L = nl;
        }
        x = (L << half) | R;
    } while (x >= n);
    return x;
}

void oro_gen_zipf(uint64_t seed, double s, uint64_t n, float* out) {
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t rank = 1u + oro_feistel_perm(seed, n, i);
        out[i] = (float)pow((double)rank, -s);
    }
}

void oro_gen_spike(uint64_t seed, uint64_t n, float* out) {
    uint32_t r[4];
    const uint64_t spikes = n / 100u;
    for (uint64_t i = 0; i < n; ++i) {
        philox_at(seed, (uint32_t)i, (uint32_t)(i >> 32), 0u, TAG_GEN, r);
        const double u = u53(r[0], r[1]);
        const int spike = oro_feistel_perm(seed, n, i) < spikes;
        out[i] = spike ? (float)(1e5 + 9e5 * u) : (float)(1e-3 + (1.0 - 1e-3) * u);
    }
}

void oro_segment_sizes(uint64_t seed, uint32_t b, uint64_t lo, uint64_t hi, uint64_t* sizes) {
    uint32_t r[4];
    const double ratio = (double)hi / (double)lo;
    for (uint32_t s = 0; s < b; ++s) {
        philox_at(seed, s, 0u, TAG_SEG, 0u, r);
        const double u = u53(r[0], r[1]);
        uint64_t v = (uint64_t)floor((double)lo * pow(ratio, u));
        sizes[s] = ORO_MIN(ORO_MAX(v, lo), hi);
    }
}

/* ======================================================================== */
/* reduce.hpp:81-92 kahan_reduction                                          */
/* ======================================================================== */
#define DEFINE_KAHAN(T, NAME)                                                  \
    T NAME(const T* x, uint64_t n) {                                           \
        T sum = 0, c = 0;                                                      \
        for (uint64_t k = 0; k < n; ++k) {                                     \
            const T y = x[k] - c;                                              \
            const T t = sum + y;                                               \
            c = (t - sum) - y;                                                 \
            sum = t;                                                           \
        }                                                                      \
        return sum;                                                            \
    }
DEFINE_KAHAN(float, oro_kahan_reduce_f32)
DEFINE_KAHAN(double, oro_kahan_reduce_f64)

/* ======================================================================== */
/* partition.hpp:35-53 stable_partition_indicies: heavies forward from 0,    */
/* lights backward from N-1, then the light region is reversed.             */
/* ======================================================================== */
#define DEFINE_PARTITION(T, NAME)                                              \
    uint32_t NAME(const T* w, uint32_t n, T pivot, uint32_t* idx) {            \
        uint32_t h = 0;                                                        \
        int64_t l = (int64_t)n - 1;                                            \
        for (uint32_t i = 0; i < n; ++i) {                                     \
            if (w[i] > pivot) {                                                \
                idx[h++] = i;                                                  \
            } else {                                                           \
                idx[l--] = i;                                                  \
            }                                                                  \
        }                                                                      \
        for (uint32_t lo = h, hi = n; lo + 1 < hi; ++lo, --hi) {               \
            const uint32_t t = idx[lo];                                        \
            idx[lo] = idx[hi - 1];                                             \
            idx[hi - 1] = t;                                                   \
        }                                                                      \
        return h;                                                              \
    }
DEFINE_PARTITION(float, oro_stable_partition_indices_f32)
DEFINE_PARTITION(double, oro_stable_partition_indices_f64)

/* ======================================================================== */
/* prefix_sum.hpp:13-34 Kahan inclusive prefix                               */
/* ======================================================================== */
#define DEFINE_PREFIX(T, NAME)                                                 \
    void NAME(const T* in, uint64_t n, T* out) {                               \
        T sum = 0, c = 0;                                                      \
        for (uint64_t k = 0; k < n; ++k) {                                     \
            const T y = in[k] - c;                                             \
            const T t = sum + y;                                               \
            c = (t - sum) - y;                                                 \
            sum = t;                                                           \
            out[k] = sum;                                                      \
        }                                                                      \
    }
DEFINE_PREFIX(float, oro_prefix_sum_f32)
DEFINE_PREFIX(double, oro_prefix_sum_f64)

/* ======================================================================== */
/* split.hpp:59-65 ceilMulDiv, in 64 bit (F13: the uint32 form overflows).   */
/* ======================================================================== */
uint64_t oro_ceil_mul_div(uint64_t a, uint64_t b, uint64_t c) {
    const uint64_t div = a / c;
    const uint64_t rem = a % c;
    return div * b + (rem * b + c - 1u) / c;
}

/* ======================================================================== */
/* split.hpp:14-57 split(): binary search over the number of heavies j with  */
/* i = min(n - j, L) lights, for sigma(i,j) <= mean*n < sigma(i,j+1).        */
/* The probe sequence (a, b, j = (a+b)/2) is the reference's, so that when   */
/* several j satisfy the condition the same one is returned.                 */
/* Paths the reference cannot take in exact arithmetic (it asserts or reads  */
/* out of bounds there) return the clamped split of guard_split().           */
/* ======================================================================== */
#define DEFINE_SPLIT(T, ST, NAME)                                              \
    static ST NAME##_guard(const T* hp, uint32_t H, const T* lp, uint32_t L,   \
                           T mean, uint32_t n, int64_t j, T target) {          \
        ST s;                                                                  \
        int64_t jj = j < 0 ? 0 : j;                                            \
        if (jj > (int64_t)H - 1) jj = (int64_t)H - 1;                          \
        int64_t ii = (int64_t)n - jj;                                          \
        if (ii > (int64_t)L) ii = L;                                           \
        const T light = ii == 0 ? (T)0 : lp[ii - 1];                           \
        T spill = (light + hp[jj]) - target;                                   \
        if (!(spill > (T)0)) spill = mean;                                     \
        s.i = (uint32_t)ii;                                                    \
        s.j = (uint32_t)jj;                                                    \
        s.spill = spill;                                                       \
        return s;                                                              \
    }                                                                          \
    ST NAME(const T* hp, uint32_t H, const T* lp, uint32_t L, T mean,          \
            uint32_t n) {                                                      \
        int64_t a = 0;                                                         \
        int64_t b = ORO_MIN((int64_t)n, (int64_t)H);                           \
        const T target = mean * (T)n;                                          \
        for (;;) {                                                             \
            const int64_t j = (a + b) / 2;                                     \
            const int64_t i = ORO_MIN((int64_t)n - j, (int64_t)L);             \
            if (a > b) {                                                       \
                if (j < 0 || i >= (int64_t)L || j + 1 >= (int64_t)H)           \
                    return NAME##_guard(hp, H, lp, L, mean, n, j, target);     \
                ST s;                                                          \
                s.i = (uint32_t)i;                                             \
                s.j = (uint32_t)j;                                             \
                s.spill = (lp[i] + hp[j + 1]) - target;                        \
                return s;                                                      \
            }                                                                  \
            const T heavy = j == 0 ? (T)0 : hp[j - 1];                         \
            const T light = i == 0 ? (T)0 : lp[i - 1];                         \
            const T sigma = light + heavy;                                     \
            if (j == (int64_t)H)                                               \
                return NAME##_guard(hp, H, lp, L, mean, n, j, target);         \
            const T sigma2 = light + hp[j];                                    \
            if (sigma <= target) {                                             \
                if (sigma2 > target) {                                         \
                    ST s;                                                      \
                    s.i = (uint32_t)i;                                         \
                    s.j = (uint32_t)j;                                         \
                    s.spill = sigma2 - target;                                 \
                    return s;                                                  \
                }                                                              \
                a = j + 1;                                                     \
            } else {                                                           \
                if (j == 0)                                                    \
                    return NAME##_guard(hp, H, lp, L, mean, n, j, target);     \
                b = j - 1;                                                     \
            }                                                                  \
        }                                                                      \
    }
DEFINE_SPLIT(float, oro_split_f32_t, oro_split_f32)
DEFINE_SPLIT(double, oro_split_f64_t, oro_split_f64)

/* ======================================================================== */
/* Lower-bound split of the reference GPU kernels                            */
/* (src/device/wrs/alias/psa/split/scalar/shader.comp:72-109, same search in */
/* splitpack/inline/shader.comp:81-114): the largest j with                  */
/* sigma(min(n-j, L), j) <= mean*n, searched over j in [0, min(n, H-1)].     */
/* sigma(n-j, j) grows with j, so this j is unique: it is the state of the    */
/* sequential sweep after n buckets (oro_sweep_states).  Deviation from the   */
/* shader text: i is clamped to L, not L-1 (the shader's L-1 gives           */
/* i + j = n - 1 when every light is used).                                  */
/* ======================================================================== */
#define DEFINE_SPLIT_LB(T, ST, NAME)                                           \
    ST NAME(const T* hp, uint32_t H, const T* lp, uint32_t L, T mean,          \
            uint32_t n) {                                                      \
        const T target = mean * (T)n;                                          \
        int64_t a = 0;                                                         \
        int64_t b = ORO_MIN((int64_t)n, (int64_t)H - 1);                       \
        while (a <= b) {                                                       \
            const int64_t j = (a + b) / 2;                                     \
            const int64_t i = ORO_MIN((int64_t)n - j, (int64_t)L);             \
            const T heavy = j == 0 ? (T)0 : hp[j - 1];                         \
            const T light = i == 0 ? (T)0 : lp[i - 1];                         \
            if (light + heavy <= target) a = j + 1;                            \
            else b = j - 1;                                                    \
        }                                                                      \
        const int64_t j = b < 0 ? 0 : b;                                       \
        const int64_t i = ORO_MIN((int64_t)n - j, (int64_t)L);                 \
        const T light = i == 0 ? (T)0 : lp[i - 1];                             \
        ST s;                                                                  \
        s.i = (uint32_t)i;                                                     \
        s.j = (uint32_t)j;                                                     \
        s.spill = (light + hp[j]) - target;                                    \
        if (!(s.spill > (T)0)) s.spill = mean;                                 \
        return s;                                                              \
    }
DEFINE_SPLIT_LB(float, oro_split_f32_t, oro_split_lb_f32)
DEFINE_SPLIT_LB(double, oro_split_f64_t, oro_split_lb_f64)

/* split.hpp:67-86 splitK: K splits, the k-th at n = ceil(N*k/K) for
 * k = 1..K-1, the last one fixed to (L, H-1, 0). */
#define DEFINE_SPLITK(T, ST, NAME, SPLIT, SPLIT_LB)                            \
    void NAME(const T* hp, uint32_t H, const T* lp, uint32_t L, T mean,        \
              uint32_t N, uint32_t K, int rule, ST* splits) {                  \
        for (uint32_t k = 1; k + 1 <= K; ++k) {                                \
            const uint32_t n = (uint32_t)oro_ceil_mul_div(N, k, K);            \
            splits[k - 1] = rule == ORO_SPLIT_LOWER_BOUND                      \
                                ? SPLIT_LB(hp, H, lp, L, mean, n)              \
                                : SPLIT(hp, H, lp, L, mean, n);                \
        }                                                                      \
        splits[K - 1].i = L;                                                   \
        splits[K - 1].j = H - 1;                                               \
        splits[K - 1].spill = (T)0;                                            \
    }
DEFINE_SPLITK(float, oro_split_f32_t, oro_splitK_f32, oro_split_f32, oro_split_lb_f32)
DEFINE_SPLITK(double, oro_split_f64_t, oro_splitK_f64, oro_split_f64, oro_split_lb_f64)

/* ======================================================================== */
/* pack.hpp:14-83 pack2 and :179-202 packSplits.                            */
/* ======================================================================== */
#define DEFINE_PACK(T, ST, NAME)                                               \
    static void NAME##_pack2(const uint32_t* heavy, uint32_t H,                \
                             const uint32_t* light, const T* w, T mean,        \
                             uint32_t i0, uint32_t i1, uint32_t j0,            \
                             uint32_t j1, T spill, T* p, uint32_t* a) {        \
        if (i1 < i0) i1 = i0;                                                  \
        if (j1 < j0) j1 = j0;                                                  \
        uint32_t i = i0, j = j0;                                               \
        T r = spill;                                                           \
        if (r == (T)0) r = w[heavy[j]];                                        \
        while (i != i1 || j != j1) {                                           \
            int packHeavy;                                                     \
            if (j == j1) packHeavy = 0;                                        \
            else if (i == i1) packHeavy = 1;                                   \
            else packHeavy = r <= mean;                                        \
            const uint32_t h = heavy[j];                                       \
            const uint32_t widx = packHeavy ? heavy[j + 1] : light[i];         \
            const T weight = w[widx];                                          \
            if (packHeavy) {                                                   \
                p[h] = r / mean;                                               \
                a[h] = widx;                                                   \
                j += 1;                                                        \
            } else {                                                           \
                p[widx] = weight / mean;                                       \
                a[widx] = h;                                                   \
                i += 1;                                                        \
            }                                                                  \
            r = (r + weight) - mean;                                           \
        }                                                                      \
        if (j1 == H - 1) {                                                     \
            const uint32_t h = heavy[j];                                       \
            p[h] = (T)1;                                                       \
            a[h] = h;                                                          \
        }                                                                      \
    }                                                                          \
    void NAME(const uint32_t* heavy, uint32_t H, const uint32_t* light,        \
              uint32_t L, const T* w, uint32_t N, T mean, const ST* splits,    \
              uint32_t K, T* p, uint32_t* a) {                                 \
        (void)L;                                                               \
        (void)N;                                                               \
        ST prev;                                                               \
        prev.i = 0;                                                            \
        prev.j = 0;                                                            \
        prev.spill = (T)0;                                                     \
        for (uint32_t k = 0; k < K; ++k) {                                     \
            NAME##_pack2(heavy, H, light, w, mean, prev.i, splits[k].i,        \
                         prev.j, splits[k].j, prev.spill, p, a);               \
            prev = splits[k];                                                  \
        }                                                                      \
    }
DEFINE_PACK(float, oro_split_f32_t, oro_pack_splits_f32)
DEFINE_PACK(double, oro_split_f64_t, oro_pack_splits_f64)

/* ======================================================================== */
/* SubgroupPack (src/device/wrs/alias/psa/pack/subgroup/shader.comp:68-246)  */
/* over the same K split records.  The shader steps one interval with a      */
/* group of lanes that hold the next heavies and lights in registers; read   */
/* as a sequential sweep it is pack2 with two differences:                   */
/*   - the residual steps as r = r + (weight - mean) (shader.comp:240), not  */
/*     (r + weight) - mean (pack.hpp:80);                                    */
/*   - interval 0 starts from heavy 0's weight and every other interval from */
/*     its split's spill as it is (shader.comp:86-94, 118-125), where pack2  */
/*     replaces a zero spill by the current heavy's weight (pack.hpp:32-37). */
/* Deviation: with the interval's lights used up and r > mean, the shader's  */
/* `hasHeavy && w <= mean` (shader.comp:147) takes one more light, from      */
/* registers that hold no light of the interval; here, as in pack2, the      */
/* heavy is taken.  Interval ends are clamped as in pack2.                   */
/* ======================================================================== */
#define DEFINE_PACK_GROUPED(T, ST, NAME)                                       \
    void NAME(const uint32_t* heavy, uint32_t H, const uint32_t* light,        \
              uint32_t L, const T* w, uint32_t N, T mean, const ST* splits,    \
              uint32_t K, T* p, uint32_t* a) {                                 \
        (void)L;                                                               \
        (void)N;                                                               \
        for (uint32_t k = 0; k < K; ++k) {                                     \
            const uint32_t i0 = k ? splits[k - 1].i : 0u;                      \
            const uint32_t j0 = k ? splits[k - 1].j : 0u;                      \
            uint32_t i1 = splits[k].i, j1 = splits[k].j;                       \
            if (i1 < i0) i1 = i0;                                              \
            if (j1 < j0) j1 = j0;                                              \
            uint32_t i = i0, j = j0;                                           \
            T r = k ? splits[k - 1].spill : w[heavy[j0]];                      \
            while (i != i1 || j != j1) {                                       \
                int packHeavy;                                                 \
                if (j == j1) packHeavy = 0;                                    \
                else if (i == i1) packHeavy = 1;                               \
                else packHeavy = r <= mean;                                    \
                const uint32_t h = heavy[j];                                   \
                const uint32_t widx = packHeavy ? heavy[j + 1] : light[i];     \
                const T weight = w[widx];                                      \
                if (packHeavy) {                                               \
                    p[h] = r / mean;                                           \
                    a[h] = widx;                                               \
                    j += 1;                                                    \
                } else {                                                       \
                    p[widx] = weight / mean;                                   \
                    a[widx] = h;                                               \
                    i += 1;                                                    \
                }                                                              \
                r = r + (weight - mean);                                       \
            }                                                                  \
            if (j1 == H - 1) {                                                 \
                const uint32_t h = heavy[j];                                   \
                p[h] = (T)1;                                                   \
                a[h] = h;                                                      \
            }                                                                  \
        }                                                                      \
    }
DEFINE_PACK_GROUPED(float, oro_split_f32_t, oro_pack_grouped_f32)
DEFINE_PACK_GROUPED(double, oro_split_f64_t, oro_pack_grouped_f64)

/* The sequential sweep of pack2 over one interval (0,0) -> (L, H-1), i.e.
 * the PSA pack with no splits: st_i[n], st_j[n] = lights and heavies packed
 * after n buckets (n = 0..N-1).  This is the ground truth a split must hit. */
void oro_sweep_states_f64(const uint32_t* heavy, uint32_t H, const uint32_t* light, uint32_t L,
                          const double* w, double mean, uint32_t* st_i, uint32_t* st_j) {
    uint32_t i = 0, j = 0, n = 0;
    double r = w[heavy[0]];
    const uint32_t i1 = L, j1 = H - 1;
    st_i[0] = 0;
    st_j[0] = 0;
    while (i != i1 || j != j1) {
        int packHeavy;
        if (j == j1) packHeavy = 0;
        else if (i == i1) packHeavy = 1;
        else packHeavy = r <= mean;
        const double weight = packHeavy ? w[heavy[j + 1]] : w[light[i]];
        if (packHeavy) ++j;
        else ++i;
        r = (r + weight) - mean;
        ++n;
        st_i[n] = i;
        st_j[n] = j;
    }
}

/* ======================================================================== */
/* psa_alias_table.hpp:16-67                                                 */
/* ======================================================================== */
void oro_psa_alias_table_f32(const float* w_in, uint32_t n, uint32_t K, int reference_scale,
                             int split_rule, float* p, uint32_t* a, oro_psa_debug* dbg) {
    float* w = (float*)malloc(sizeof(float) * n);
    for (uint32_t i = 0; i < n; ++i) {
        /* `weights[i] *= 0.0001` with T = float: float * double literal,
         * evaluated in double and rounded back to float. */
        w[i] = reference_scale ? (float)((double)w_in[i] * 0.0001) : w_in[i];
    }
    const float total = oro_kahan_reduce_f32(w, n);
    const float mean = total / (float)n;
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    const uint32_t H = oro_stable_partition_indices_f32(w, n, mean, idx);
    const uint32_t L = n - H;
    if (dbg) {
        dbg->mean = mean;
        dbg->heavy_count = H;
        if (dbg->idx_cpu) memcpy(dbg->idx_cpu, idx, sizeof(uint32_t) * n);
    }
    if (H == 0 || L == 0) {
        for (uint32_t i = 0; i < n; ++i) {
            p[i] = 1.0f;
            a[i] = i;
        }
        free(idx);
        free(w);
        return;
    }
    float* vals = (float*)malloc(sizeof(float) * n);
    float* pre = (float*)malloc(sizeof(float) * n);
    for (uint32_t r = 0; r < n; ++r) vals[r] = w[idx[r]];
    oro_prefix_sum_f32(vals, H, pre);
    oro_prefix_sum_f32(vals + H, L, pre + H);
    oro_split_f32_t* splits = (oro_split_f32_t*)malloc(sizeof(oro_split_f32_t) * K);
    oro_splitK_f32(pre, H, pre + H, L, mean, n, K, split_rule, splits);
    if (dbg) {
        if (dbg->heavy_prefix)
            for (uint32_t r = 0; r < H; ++r) dbg->heavy_prefix[r] = pre[r];
        if (dbg->light_prefix)
            for (uint32_t r = 0; r < L; ++r) dbg->light_prefix[r] = pre[H + r];
        if (dbg->splits)
            for (uint32_t k = 0; k < K; ++k) {
                dbg->splits[k].i = splits[k].i;
                dbg->splits[k].j = splits[k].j;
                dbg->splits[k].spill = splits[k].spill;
            }
    }
    oro_pack_splits_f32(idx, H, idx + H, L, w, n, mean, splits, K, p, a);
    free(splits);
    free(pre);
    free(vals);
    free(idx);
    free(w);
}

void oro_psa_alias_table_f64(const float* w_in, uint32_t n, uint32_t K, int reference_scale,
                             int split_rule, double mean_override, double* p, uint32_t* a,
                             oro_psa_debug* dbg) {
    double* w = (double*)malloc(sizeof(double) * n);
    for (uint32_t i = 0; i < n; ++i) {
        w[i] = reference_scale ? (double)w_in[i] * 0.0001 : (double)w_in[i];
    }
    double mean;
    if (isnan(mean_override)) {
        const double total = oro_kahan_reduce_f64(w, n);
        mean = total / (double)n;
    } else {
        mean = mean_override;
    }
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    const uint32_t H = oro_stable_partition_indices_f64(w, n, mean, idx);
    const uint32_t L = n - H;
    if (dbg) {
        dbg->mean = mean;
        dbg->heavy_count = H;
        if (dbg->idx_cpu) memcpy(dbg->idx_cpu, idx, sizeof(uint32_t) * n);
    }
    if (H == 0 || L == 0) {
        for (uint32_t i = 0; i < n; ++i) {
            p[i] = 1.0;
            a[i] = i;
        }
        free(idx);
        free(w);
        return;
    }
    double* vals = (double*)malloc(sizeof(double) * n);
    double* pre = (double*)malloc(sizeof(double) * n);
    for (uint32_t r = 0; r < n; ++r) vals[r] = w[idx[r]];
    oro_prefix_sum_f64(vals, H, pre);
    oro_prefix_sum_f64(vals + H, L, pre + H);
    oro_split_f64_t* splits = (oro_split_f64_t*)malloc(sizeof(oro_split_f64_t) * K);
    oro_splitK_f64(pre, H, pre + H, L, mean, n, K, split_rule, splits);
    if (dbg) {
        if (dbg->heavy_prefix) memcpy(dbg->heavy_prefix, pre, sizeof(double) * H);
        if (dbg->light_prefix) memcpy(dbg->light_prefix, pre + H, sizeof(double) * L);
        if (dbg->splits) memcpy(dbg->splits, splits, sizeof(oro_split_f64_t) * K);
    }
    oro_pack_splits_f64(idx, H, idx + H, L, w, n, mean, splits, K, p, a);
    free(splits);
    free(pre);
    free(vals);
    free(idx);
    free(w);
}

/* The sequential sweep of pack.hpp:14-83 (one interval (0,0) -> (L, H-1),
 * heavy first on ties, last heavy (1, itself) as pack.hpp:78-81) evaluated
 * with exact integer keys: K = 53 - ilogb(mean), mean_k = mean 2^K (an
 * integer), every light weight rounded to the nearest integer of w 2^K (exact
 * at and above mean 2^-30), every heavy one to the nearest multiple of 2^30
 * (exact for normal floats), X(i) = i mean_k - sum of the first i lights, Y(j) = sum of
 * the first j heavies - j mean_k.  Heavy q is packed before light i iff
 * Y(q+1) <= X(i) (the sweep's "residual <= mean"), with residual
 * Y(q+1) - X(i) + mean_k rounded once to FP64.  Light p = w / mean (FP32
 * division when f32, else w * (1/mean) in FP64), heavy p = residual / mean the
 * same way, both rounded to float and clamped to [0, 1]. */
typedef unsigned __int128 oro_u128;

static oro_u128 qkey(float w, int K) {
    const float y = rintf(ldexpf(w, K));
    return (oro_u128)y;
}
/* a heavy weight (w > mean) times 2^K is a multiple of 2^30: rounded as w 2^(K-30) */
static oro_u128 qkey_heavy(float w, int K) { return qkey(w, K - 30) << 30; }

static float stream_ratio(double x, double mean, int f32) {
    if (f32) return (float)((float)x / (float)mean);
    return (float)(x * (1.0 / mean));
}

void oro_psa_stream_table(const float* w, uint32_t n, double mean, int f32, float* p,
                          uint32_t* a) {
    uint32_t* hid = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* lid = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t H = 0, L = 0;
    for (uint32_t e = 0; e < n; ++e) {
        if ((double)w[e] > mean) hid[H++] = e;
        else lid[L++] = e;
    }
    if (H == 0 || L == 0) {
        for (uint32_t e = 0; e < n; ++e) {
            p[e] = 1.0f;
            a[e] = e;
        }
        free(hid);
        free(lid);
        return;
    }
    const int K = 53 - ilogb(mean);
    const oro_u128 mk = (oro_u128)ldexp(mean, K);
    oro_u128 X = 0;                      /* X(i) */
    oro_u128 Yn = qkey_heavy(w[hid[0]], K) - mk; /* Y(q+1) */
    uint32_t i = 0, q = 0;
    for (;;) {
        if (q + 1u < H && (i == L || Yn <= X)) {
            const oro_u128 R = Yn + mk - X;
            const float r = stream_ratio(ldexp((double)R, -K), mean, f32);
            p[hid[q]] = fminf(fmaxf(r, 0.0f), 1.0f);
            a[hid[q]] = hid[q + 1u];
            ++q;
            Yn += qkey_heavy(w[hid[q]], K) - mk;
        } else if (i < L) {
            const float x = w[lid[i]];
            p[lid[i]] = fminf(stream_ratio((double)x, mean, f32), 1.0f);
            a[lid[i]] = hid[q];
            X += mk - qkey(x, K);
            ++i;
        } else {
            break;
        }
    }
    p[hid[H - 1u]] = 1.0f;
    a[hid[H - 1u]] = hid[H - 1u];
    free(hid);
    free(lid);
}

/* ======================================================================== */
/* Sampling                                                                   */
/* ======================================================================== */
void oro_sample_alias(const float* p, const uint32_t* a, uint64_t n, uint64_t s, uint64_t seed,
                      uint64_t counter_base, uint32_t* out) {
    uint32_t r[4];
    for (uint64_t t = 0; t < s; ++t) {
        const uint64_t c = counter_base + t;
        philox_at(seed, (uint32_t)c, (uint32_t)(c >> 32), 0u, 0u, r);
        const uint64_t x = ((uint64_t)r[1] << 32) | r[0];
        const uint64_t ix = (uint64_t)(((unsigned __int128)x * n) >> 64);
        const float u2 = u24(r[2]);
        /* sample_alias_table.hpp:15-24: u2 > p -> alias, else the index. */
        out[t] = (u2 > p[ix]) ? a[ix] : (uint32_t)ix;
    }
}

void oro_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n, uint64_t s,
                    uint64_t seed, uint64_t counter_base, uint32_t* out) {
    uint32_t r[4];
    for (uint64_t t = 0; t < s; ++t) {
        const uint64_t c = counter_base + t;
        philox_at(seed, (uint32_t)c, (uint32_t)(c >> 32), 0u, 0u, r);
        const double u = u53_open0(r[0], r[1]);
        uint64_t lo = 0, hi = n - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            const double v = (double)cdf_hi[mid] + (cdf_lo ? (double)cdf_lo[mid] : 0.0);
            if (u > v) lo = mid + 1;
            else hi = mid;
        }
        out[t] = (uint32_t)lo;
    }
}

/* ======================================================================== */
/* inverse_alias_table.hpp:11-27 alias_table_to_normalized_weights, in FP64, */
/* divided by N so that it is a distribution.                                */
/* ======================================================================== */
#define DEFINE_IMPLIED(T, NAME)                                                \
    void NAME(const T* p, const uint32_t* a, uint64_t n, double* q) {          \
        for (uint64_t i = 0; i < n; ++i) q[i] = 0.0;                           \
        for (uint64_t i = 0; i < n; ++i) {                                     \
            double pi = (double)p[i];                                          \
            if (pi < 0.0) pi = 0.0;                                            \
            if (pi > 1.0) pi = 1.0;                                            \
            q[i] += pi;                                                        \
            if (a[i] < n) q[a[i]] += 1.0 - pi;                                 \
        }                                                                      \
        for (uint64_t i = 0; i < n; ++i) q[i] /= (double)n;                    \
    }
DEFINE_IMPLIED(float, oro_implied_distribution_f32)
DEFINE_IMPLIED(double, oro_implied_distribution_f64)
