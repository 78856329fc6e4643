// stats.hip -- on-device sample statistics: histogram, pooled chi-square and
// the squared error of the empirical distribution, so that S = 2^30 samples
// can be validated without copying them to the host.
//
// Replaces src/device/statistics/histogram/atomic/shader.comp:1-38 (uint64
// atomic histogram), chi_square/reduce/shader.comp:40-73 (per-bin chi-square,
// float atomics) and rmse/mse/shader.comp:52-77 (squared error, float
// atomics).  Here every sum is FP64 in a fixed order (deterministic), bins of
// zero expectation are counted instead of divided by, and the chi-square pools
// consecutive bins until each pool expects at least m samples (pool of bin i =
// floor(S cdf(i) / m), the last partial pool merged into the one before; m <= 0
// gives one pool per bin, the reference's statistic).
//
// The pooled statistic is a segmented reduction: every tile of 4096 bins, and
// then every run of tiles, is summarised by its first and last pool (both
// possibly open towards the neighbours), the sum of the chi-square terms of the
// pools closed inside it, and the number of pool boundaries inside it; the
// combination of two summaries is associative, so the tiles are combined in a
// fixed tree.
#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

struct Pool {
    uint64_t id;
    uint64_t o;  // observed samples
    double e;    // expected samples
};

__device__ __forceinline__ double pool_term(const Pool& p) {
    if (!(p.e > 0.0)) return 0.0;
    const double d = (double)p.o - p.e;
    return d * d / p.e;
}

struct Summary {
    Pool head, tail;     // first and last pool (the same pool when single)
    double inner;        // chi-square terms of the pools closed inside
    uint64_t bounds;     // pool boundaries inside
    uint32_t single;     // one pool only
    uint32_t empty;      // no bins
    double sse, esse;    // sum (o/S - q)^2, sum q (1 - q) / S
    uint64_t zero_hits;  // samples in bins of zero weight
};

__device__ __forceinline__ Summary sum_empty() {
    Summary r;
    r.head = r.tail = Pool{0, 0, 0.0};
    r.inner = 0.0;
    r.bounds = 0;
    r.single = 1;
    r.empty = 1;
    r.sse = r.esse = 0.0;
    r.zero_hits = 0;
    return r;
}

__device__ __forceinline__ Summary combine(const Summary& a, const Summary& b) {
    if (a.empty) return b;
    if (b.empty) return a;
    Summary r;
    r.empty = 0;
    r.sse = a.sse + b.sse;
    r.esse = a.esse + b.esse;
    r.zero_hits = a.zero_hits + b.zero_hits;
    if (a.tail.id == b.head.id) {
        const Pool mid{a.tail.id, a.tail.o + b.head.o, a.tail.e + b.head.e};
        r.bounds = a.bounds + b.bounds;
        if (a.single && b.single) {
            r.head = r.tail = mid;
            r.inner = 0.0;
            r.single = 1;
        } else if (a.single) {
            r.head = mid;
            r.tail = b.tail;
            r.inner = b.inner;
            r.single = 0;
        } else if (b.single) {
            r.head = a.head;
            r.tail = mid;
            r.inner = a.inner;
            r.single = 0;
        } else {
            r.head = a.head;
            r.tail = b.tail;
            r.inner = a.inner + pool_term(mid) + b.inner;
            r.single = 0;
        }
    } else {
        r.bounds = a.bounds + b.bounds + 1;
        r.head = a.head;
        r.tail = b.tail;
        r.inner = a.inner + (a.single ? 0.0 : pool_term(a.tail)) +
                  (b.single ? 0.0 : pool_term(b.head)) + b.inner;
        r.single = 0;
    }
    return r;
}

constexpr int kStatBlock = 256;
constexpr int kStatPer = 16;
constexpr uint32_t kStatTile = kStatBlock * kStatPer;
constexpr int kSumWords = 16;  // one Summary per tile in the state buffer (<= 128 bytes)
static_assert(sizeof(Summary) <= kSumWords * 8, "summary record");

// Combines the block's kStatBlock summaries (s[t] covers the range after
// s[t-1]) into s[0] by a fixed pairwise tree: level d joins s[t] and s[t + d]
// for t a multiple of 2d, so the order of the ranges is kept.
__device__ __forceinline__ void block_combine(Summary* s) {
    const int t = threadIdx.x;
#pragma unroll
    for (int d = 1; d < kStatBlock; d <<= 1) {
        if ((t & (2 * d - 1)) == 0) s[t] = combine(s[t], s[t + d]);
        __syncthreads();
    }
}

}  // namespace

// Histogram with a per-block LDS cache of bins: a sample's bin is counted in
// the block's hash slot for that bin when the slot is free or already holds
// it, else by a global atomic; the slots are added to the histogram when the
// block ends.  A skewed distribution sends most samples to a few bins, and
// without the cache every sample of a hot bin is an atomic on one address
// (serialised in one L2 channel); with it a hot bin costs one global atomic
// per block.  A block counts at most s / kHistMaxBlocks + 256 kHistPerThread
// samples, so its 32-bit slot counts cannot overflow below 2^43 samples (a
// device holds fewer than 2^37 4-byte samples).
constexpr int kHistSlots = 2048;
constexpr uint32_t kHistEmpty = 0xffffffffu;
constexpr uint64_t kHistPerThread = 64;   // samples per thread before the grid grows
constexpr uint64_t kHistMaxBlocks = 4096;

__global__ __launch_bounds__(256) void k_histogram(const uint32_t* __restrict__ samples,
                                                   uint64_t s, uint64_t n,
                                                   unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_key[kHistSlots];
    __shared__ uint32_t s_cnt[kHistSlots];
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x) {
        s_key[i] = kHistEmpty;
        s_cnt[i] = 0;
    }
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < s; t += stride) {
        const uint64_t x = samples[t];
        const uint32_t key = (uint32_t)(x < n ? x : n);  // n: the out-of-range counter
        bool cached = false;
        if (key != kHistEmpty) {
            const uint32_t h = (key * 2654435761u) >> (32 - 11);  // Fibonacci hash, 11 bits
            uint32_t cur = s_key[h];
            if (cur == kHistEmpty) {
                const uint32_t old = atomicCAS(&s_key[h], kHistEmpty, key);
                cur = old == kHistEmpty ? key : old;
            }
            if (cur == key) {
                atomicAdd(&s_cnt[h], 1u);
                cached = true;
            }
        }
        if (!cached) atomicAdd(hist + key, 1ull);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x)
        if (s_cnt[i] != 0) atomicAdd(hist + s_key[i], (unsigned long long)s_cnt[i]);
}
static_assert(kHistSlots == 1 << 11, "hash width");

// Per tile of kStatTile bins: thread t takes bins 16t..16t+15 in order, the
// block combines the thread summaries in range order (block_combine).
// q(i) = w[i] * inv_total, cum(i) = S * cdf(i) with cdf from the plan's scan.
template <bool LO>
__global__ __launch_bounds__(kStatBlock) void k_stats_tile(
    const unsigned long long* __restrict__ hist, const float* __restrict__ w,
    const float* __restrict__ cdf_hi, const float* __restrict__ cdf_lo, uint64_t n, double S,
    double min_expected, const Params* __restrict__ prm, uint64_t* __restrict__ recs) {
    __shared__ Summary s_sum[kStatBlock];
    const uint64_t tile = blockIdx.x;
    const uint64_t b0 = tile * kStatTile + (uint64_t)threadIdx.x * kStatPer;
    const double inv_total = prm->inv_total;
    const bool pooled = min_expected > 0.0;
    // the last pool is merged into the one before it
    double c_last = (double)cdf_hi[n - 1];
    if (LO) c_last += (double)cdf_lo[n - 1];
    const double p_last = floor(S * c_last / min_expected);
    const uint64_t cap = pooled ? (p_last >= 1.0 ? (uint64_t)p_last - 1 : 0) : ~0ull;

    // every load of the thread's bins first (indices clamped to the last bin),
    // so that the lines of a lane's 16 consecutive bins are fetched once
    uint64_t ov[kStatPer];
    float wv[kStatPer], chv[kStatPer], clv[kStatPer];
#pragma unroll
    for (int k = 0; k < kStatPer; ++k) {
        const uint64_t i = b0 + k < n ? b0 + k : n - 1;
        ov[k] = hist[i];
        wv[k] = w[i];
        chv[k] = pooled ? cdf_hi[i] : 0.0f;
        clv[k] = LO && pooled ? cdf_lo[i] : 0.0f;
    }
    Summary me = sum_empty();
#pragma unroll
    for (int k = 0; k < kStatPer; ++k) {
        const uint64_t i = b0 + k;
        if (i >= n) continue;
        const uint64_t o = ov[k];
        const float wi = wv[k];
        const double q = (double)wi * inv_total;
        uint64_t id = i;
        if (pooled) {
            double c = (double)chv[k];
            if (LO) c += (double)clv[k];
            const double p = floor(S * c / min_expected);
            id = p >= (double)cap ? cap : (uint64_t)(p > 0.0 ? p : 0.0);
        }
        Summary one;
        one.head = one.tail = Pool{id, o, S * q};
        one.inner = 0.0;
        one.bounds = 0;
        one.single = 1;
        one.empty = 0;
        const double d = (double)o / S - q;
        one.sse = d * d;
        one.esse = q * (1.0 - q) / S;
        one.zero_hits = wi == 0.0f ? o : 0;
        me = combine(me, one);
    }
    s_sum[threadIdx.x] = me;
    __syncthreads();
    block_combine(s_sum);
    if (threadIdx.x == 0) *reinterpret_cast<Summary*>(recs + tile * kSumWords) = s_sum[0];
}

// One block: thread t combines tiles [t*per, (t+1)*per) in order, then the
// thread summaries are combined in range order (block_combine); out = {chi2, dof, sse,
// expected sse, zero-mass hits, pools}.
__global__ __launch_bounds__(kStatBlock) void k_stats_final(const uint64_t* __restrict__ recs,
                                                            uint64_t tiles,
                                                            double* __restrict__ out) {
    __shared__ Summary s_sum[kStatBlock];
    const uint64_t per = (tiles + kStatBlock - 1) / kStatBlock;
    Summary me = sum_empty();
    for (uint64_t k = threadIdx.x * per; k < (threadIdx.x + 1) * per && k < tiles; ++k)
        me = combine(me, *reinterpret_cast<const Summary*>(recs + k * kSumWords));
    s_sum[threadIdx.x] = me;
    __syncthreads();
    block_combine(s_sum);
    if (threadIdx.x == 0) {
        const Summary acc = s_sum[0];
        const double chi2 = acc.inner + pool_term(acc.head) + (acc.single ? 0.0 : pool_term(acc.tail));
        const uint64_t pools = acc.empty ? 0 : acc.bounds + 1;
        out[0] = chi2;
        out[1] = pools ? (double)(pools - 1) : 0.0;
        out[2] = acc.sse;
        out[3] = acc.esse;
        out[4] = (double)acc.zero_hits;
        out[5] = (double)pools;
    }
}

hipError_t launch_histogram(const uint32_t* samples, uint64_t s_count, uint64_t n,
                            uint64_t* hist, int clear, hipStream_t s) {
    if (clear) {
        const hipError_t e = launch_fill_u64(hist, n + 1, 0, s);
        if (e != hipSuccess) return e;
    }
    if (s_count == 0) return hipSuccess;
    uint64_t g = (s_count + 256 * kHistPerThread - 1) / (256 * kHistPerThread);
    if (g > kHistMaxBlocks) g = kHistMaxBlocks;
    k_histogram<<<(uint32_t)g, 256, 0, s>>>(samples, s_count, n,
                                            reinterpret_cast<unsigned long long*>(hist));
    return hipGetLastError();
}

uint64_t stats_tiles_for(uint64_t n) { return (n + kStatTile - 1) / kStatTile; }

hipError_t launch_stats(const uint64_t* hist, const float* w, const float* cdf_hi,
                        const float* cdf_lo, uint64_t n, uint64_t s_count, double min_expected,
                        const Params* prm, uint64_t* recs, double* out, hipStream_t s) {
    const uint64_t tiles = stats_tiles_for(n);
    const auto* h = reinterpret_cast<const unsigned long long*>(hist);
    if (cdf_lo)
        k_stats_tile<true><<<(uint32_t)tiles, kStatBlock, 0, s>>>(h, w, cdf_hi, cdf_lo, n,
                                                                  (double)s_count, min_expected,
                                                                  prm, recs);
    else
        k_stats_tile<false><<<(uint32_t)tiles, kStatBlock, 0, s>>>(h, w, cdf_hi, cdf_lo, n,
                                                                   (double)s_count, min_expected,
                                                                   prm, recs);
    k_stats_final<<<1, kStatBlock, 0, s>>>(recs, tiles, out);
    return hipGetLastError();
}

}  // namespace wrsk
