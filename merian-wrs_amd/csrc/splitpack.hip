// splitpack.hip -- fused PSA split + pack (pass 3 of a PSA build).
//
// Split points follow the CPU reference's splitK convention
// (src/host/reference/split.hpp:67-86): K = ceil(N / split_size) intervals,
// S(0) = (0, 0, 0), S(k) = split(ceil(N*k/K)) for 0 < k < K, S(K) = (L, H-1, 0).
// split() is the reference's binary search (split.hpp:14-57) with its exact
// probe sequence, so that the same (i, j) is returned when several satisfy
// sigma(i,j) <= mean*n < sigma(i,j+1); the prefixes are read from the
// prefix-partition layout (heavy prefix at j-1, light prefix of i lights at
// N-i) as the GPU split does (split/scalar/shader.comp:43-59,86-88).
//
// pack() is the reference sweep pack2 (src/host/reference/pack.hpp:14-83): from
// S(k) to S(k+1), a light bucket (w_l/mean, current heavy) or a heavy bucket
// (residual/mean, next heavy), residual = (residual + weight) - mean, and the
// last heavy of all gets (1, itself) (pack.hpp:78-81).
//
// One thread per split point, 256 per block: each thread searches its own
// S(k), takes S(k+1) from its neighbour through LDS (as the reference's
// InlineSplitPack does with subgroupShuffle + shared memory,
// splitpack/inline/shader.comp:115-128), and packs interval k; the 256th
// point of a block is the first point of the next block, so a block packs
// 255 intervals.  In the FP64 modes the target mean*n, the sigmas, the spill
// and the running residual are FP64 (the compensated prefix is read as
// hi + lo); in FP32 mode all of them are FP32, as in the reference.
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
    else if constexpr (PREC == kComp) return (double)hi[q] + (double)lo[q];
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

// Clamped split for the paths the reference cannot take in exact arithmetic
// (it asserts or reads out of bounds there).
template <int PREC>
__device__ SplitV<PREC> split_guard(const float* hi, const float* lo, uint32_t N, uint32_t H,
                                    uint32_t L, typename Real<PREC>::T mean, uint32_t n,
                                    int64_t j, typename Real<PREC>::T target) {
    using T = typename Real<PREC>::T;
    int64_t jj = j < 0 ? 0 : j;
    if (jj > (int64_t)H - 1) jj = (int64_t)H - 1;
    int64_t ii = (int64_t)n - jj;
    if (ii > (int64_t)L) ii = L;
    const T light = ii == 0 ? (T)0 : pre_at<PREC>(hi, lo, N - (uint32_t)ii);
    T spill = (light + pre_at<PREC>(hi, lo, (uint32_t)jj)) - target;
    if (!(spill > (T)0)) spill = mean;
    return SplitV<PREC>{(uint32_t)ii, (uint32_t)jj, spill};
}

template <int PREC>
__device__ SplitV<PREC> split_search(const float* __restrict__ hi, const float* __restrict__ lo,
                                     uint32_t N, uint32_t H, uint32_t L,
                                     typename Real<PREC>::T mean, uint32_t n) {
    using T = typename Real<PREC>::T;
    int64_t a = 0;
    int64_t b = min((int64_t)n, (int64_t)H);
    const T target = mean * (T)n;
    for (;;) {
        const int64_t j = (a + b) / 2;
        const int64_t i = min((int64_t)n - j, (int64_t)L);
        if (a > b) {
            if (j < 0 || i >= (int64_t)L || j + 1 >= (int64_t)H)
                return split_guard<PREC>(hi, lo, N, H, L, mean, n, j, target);
            // reference: lightPrefix[i] + heavyPrefix[j+1]
            const T s2 = pre_at<PREC>(hi, lo, N - 1u - (uint32_t)i) +
                         pre_at<PREC>(hi, lo, (uint32_t)j + 1u);
            return SplitV<PREC>{(uint32_t)i, (uint32_t)j, s2 - target};
        }
        const T heavy = j == 0 ? (T)0 : pre_at<PREC>(hi, lo, (uint32_t)j - 1u);
        const T light = i == 0 ? (T)0 : pre_at<PREC>(hi, lo, N - (uint32_t)i);
        const T sigma = light + heavy;
        if (j == (int64_t)H) return split_guard<PREC>(hi, lo, N, H, L, mean, n, j, target);
        const T sigma2 = light + pre_at<PREC>(hi, lo, (uint32_t)j);
        if (sigma <= target) {
            if (sigma2 > target) return SplitV<PREC>{(uint32_t)i, (uint32_t)j, sigma2 - target};
            a = j + 1;
        } else {
            if (j == 0) return split_guard<PREC>(hi, lo, N, H, L, mean, n, j, target);
            b = j - 1;
        }
    }
}

}  // namespace

template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_splitpack(
    const float* __restrict__ w, uint32_t N, uint32_t K, const Params* __restrict__ prm,
    const uint32_t* __restrict__ idx, const float* __restrict__ hi, const float* __restrict__ lo,
    AliasEntry* __restrict__ table, SplitRec* __restrict__ splits) {
    using T = typename Real<PREC>::T;
    __shared__ uint32_t s_i[kSplitBlock], s_j[kSplitBlock];
    __shared__ T s_sp[kSplitBlock];

    const int t = threadIdx.x;
    const uint64_t p = (uint64_t)blockIdx.x * (kSplitBlock - 1) + (uint64_t)t;
    const uint32_t H = prm->heavy_count;
    const uint32_t L = N - H;
    const T mean = (T)prm->mean;
    const bool ident = (H == 0 || L == 0);

    SplitV<PREC> sp{0u, 0u, (T)0};
    if (!ident && p <= K) {
        if (p == K) sp = SplitV<PREC>{L, H - 1u, (T)0};
        else if (p > 0) sp = split_search<PREC>(hi, lo, N, H, L, mean,
                                                (uint32_t)ceil_mul_div(N, p, K));
    }
    s_i[t] = sp.i;
    s_j[t] = sp.j;
    s_sp[t] = sp.spill;
    __syncthreads();
    if (t == kSplitBlock - 1 || p >= K) return;

    if (ident) {  // all weights equal the mean up to its rounding: identity table
        const uint32_t b0 = p == 0 ? 0u : (uint32_t)ceil_mul_div(N, p, K);
        const uint32_t b1 = (uint32_t)ceil_mul_div(N, p + 1, K);
        for (uint32_t x = b0; x < b1; ++x) table[x] = AliasEntry{1.0f, x};
        return;
    }

    const uint32_t i0 = sp.i, j0 = sp.j;
    uint32_t i1 = s_i[t + 1], j1 = s_j[t + 1];
    if (splits) splits[p] = SplitRec{i1, j1, (double)s_sp[t + 1]};
    if (i1 < i0) i1 = i0;
    if (j1 < j0) j1 = j0;

    uint32_t i = i0, j = j0;
    T r = sp.spill;
    if (r == (T)0) r = (T)w[idx[j]];
    while (i != i1 || j != j1) {
        bool packHeavy;
        if (j == j1) packHeavy = false;
        else if (i == i1) packHeavy = true;
        else packHeavy = r <= mean;
        const uint32_t h = idx[j];
        const uint32_t widx = packHeavy ? idx[j + 1] : idx[N - 1u - i];
        const T weight = (T)w[widx];
        if (packHeavy) {
            table[h] = AliasEntry{(float)(r / mean), widx};
            ++j;
        } else {
            table[widx] = AliasEntry{(float)(weight / mean), h};
            ++i;
        }
        r = (r + weight) - mean;
    }
    if (j1 == H - 1u) {
        const uint32_t h = idx[j];
        table[h] = AliasEntry{1.0f, h};
    }
}

hipError_t launch_splitpack(const float* w, uint64_t n, uint32_t K, int prec, const Params* prm,
                            const uint32_t* idx, const float* hi, const float* lo,
                            AliasEntry* table, SplitRec* splits, hipStream_t s) {
    const uint32_t blocks = (K + (kSplitBlock - 2)) / (kSplitBlock - 1);
    const uint32_t nn = (uint32_t)n;
    if (prec == kComp)
        k_splitpack<kComp><<<blocks, kSplitBlock, 0, s>>>(w, nn, K, prm, idx, hi, lo, table, splits);
    else if (prec == kFP64Carry)
        k_splitpack<kFP64Carry><<<blocks, kSplitBlock, 0, s>>>(w, nn, K, prm, idx, hi, lo, table,
                                                              splits);
    else
        k_splitpack<kFP32><<<blocks, kSplitBlock, 0, s>>>(w, nn, K, prm, idx, hi, lo, table, splits);
    return hipGetLastError();
}

}  // namespace wrsk
