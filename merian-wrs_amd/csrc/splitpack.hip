// splitpack.hip -- the reference splitK layout of PSA split points (debug output).
//
// Split points follow the CPU reference's splitK convention
// (src/host/reference/split.hpp:67-86): K = ceil(N / split_size) intervals,
// S(0) = (0, 0, 0), S(k) = split(ceil(N*k/K)) for 0 < k < K, S(K) = (L, H-1, 0).
// split() is the lower-bound search of the reference GPU kernels (see
// split_search); the prefixes are read from the prefix-partition layout (heavy
// prefix at j-1, light prefix of i lights at N-i) as the GPU split does
// (split/scalar/shader.comp:43-59,86-88).
//
// The build itself packs by a parallel merge (mergepack.hip), which needs no
// split array; this kernel writes the K split points of a build for the
// parity tests, reading the full prefix arrays the debug build requests.  In
// the FP64 modes the target mean*n, the sigmas and the spill are FP64 (the
// FP64-mode prefix is read as hi + lo); in FP32 mode they are FP32.
#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

template <int PREC>
struct Real {
    using T = double;
};
template <>
struct Real<kFP32> {
    using T = float;
};

template <int PREC>
__device__ __forceinline__ typename Real<PREC>::T pre_at(const float* __restrict__ hi,
                                                         const float* __restrict__ lo,
                                                         uint32_t q) {
    if constexpr (PREC == kFP32) return hi[q];
    else if constexpr (prec_has_lo(PREC)) return (double)hi[q] + (double)lo[q];
    else return (double)hi[q];
}

__device__ __forceinline__ uint64_t ceil_mul_div(uint64_t a, uint64_t b, uint64_t c) {
    const uint64_t div = a / c;
    const uint64_t rem = a % c;
    return div * b + (rem * b + c - 1u) / c;
}

template <int PREC>
struct SplitV {
    uint32_t i, j;
    typename Real<PREC>::T spill;
};

// Split at n buckets: the largest j in [0, min(n, H-1)] with
// sigma(min(n-j, L), j) <= mean*n -- the lower-bound search of the reference
// GPU kernels (split/scalar/shader.comp:72-109, splitpack/inline/shader.comp:
// 81-114), with i clamped to L (the shaders clamp to L-1, which breaks
// i + j = n when every light is used).  sigma(n-j, j) grows with j, so this j
// is unique and is exactly the state of the sequential sweep after n buckets.
// (The reference CPU split, split.hpp:14-57, accepts any j with
// sigma(i,j) <= mean*n < sigma(i,j+1); that set can contain j-1 as well, which
// the CPU binary search sometimes returns, giving non-monotone splits and a
// wrong table.)  spill = sigma(i, j+1) - mean*n, the residual of heavy j.
template <int PREC>
__device__ SplitV<PREC> split_search(const float* __restrict__ hi, const float* __restrict__ lo,
                                     uint32_t N, uint32_t H, uint32_t L,
                                     typename Real<PREC>::T mean, uint32_t n) {
    using T = typename Real<PREC>::T;
    const T target = mean * (T)n;
    int64_t a = 0;
    int64_t b = min((int64_t)n, (int64_t)H - 1);
    while (a <= b) {
        const int64_t j = (a + b) / 2;
        const int64_t i = min((int64_t)n - j, (int64_t)L);
        const T heavy = j == 0 ? (T)0 : pre_at<PREC>(hi, lo, (uint32_t)j - 1u);
        const T light = i == 0 ? (T)0 : pre_at<PREC>(hi, lo, N - (uint32_t)i);
        if (light + heavy <= target) a = j + 1;
        else b = j - 1;
    }
    const int64_t j = b < 0 ? 0 : b;
    const int64_t i = min((int64_t)n - j, (int64_t)L);
    const T light = i == 0 ? (T)0 : pre_at<PREC>(hi, lo, N - (uint32_t)i);
    T spill = (light + pre_at<PREC>(hi, lo, (uint32_t)j)) - target;
    if (!(spill > (T)0)) spill = mean;
    return SplitV<PREC>{(uint32_t)i, (uint32_t)j, spill};
}

}  // namespace

// S(k) of a distribution with H heavies and K intervals.
template <int PREC>
__device__ __forceinline__ SplitV<PREC> split_point(const float* __restrict__ hi,
                                                    const float* __restrict__ lo, uint32_t N,
                                                    uint32_t K, uint32_t H,
                                                    typename Real<PREC>::T mean, uint64_t k) {
    using T = typename Real<PREC>::T;
    const uint32_t L = N - H;
    if (k == 0 || H == 0 || L == 0) return SplitV<PREC>{0u, 0u, (T)0};
    if (k == K) return SplitV<PREC>{L, H - 1u, (T)0};
    return split_search<PREC>(hi, lo, N, H, L, mean, (uint32_t)ceil_mul_div(N, k, K));
}

// splits[k-1] = S(k) for k = 1..K (S(K) = (L, H-1, 0)), one thread per split.
template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_splits(uint32_t N, uint32_t K,
                                                         const Params* __restrict__ prm,
                                                         const float* __restrict__ hi,
                                                         const float* __restrict__ lo,
                                                         SplitRec* __restrict__ splits) {
    using T = typename Real<PREC>::T;
    const uint64_t k = (uint64_t)blockIdx.x * kSplitBlock + threadIdx.x + 1u;
    if (k > K || prm->error) return;
    const SplitV<PREC> sp = split_point<PREC>(hi, lo, N, K, prm->heavy_count, (T)prm->mean, k);
    splits[k - 1] = SplitRec{sp.i, sp.j, (double)sp.spill};
}

hipError_t launch_splits(uint64_t n, uint32_t K, int prec, const Params* prm, const float* hi,
                         const float* lo, SplitRec* splits, hipStream_t s) {
    const uint32_t blocks = (K + kSplitBlock - 1) / kSplitBlock;
    const uint32_t N = (uint32_t)n;
    // modes with a hi + lo prefix read both words (kFP64 shares kComp's code)
    if (prec_has_lo(prec))
        k_splits<kComp><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, splits);
    else if (prec == kFP64Carry)
        k_splits<kFP64Carry><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, splits);
    else
        k_splits<kFP32><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, splits);
    return hipGetLastError();
}

}  // namespace wrsk
