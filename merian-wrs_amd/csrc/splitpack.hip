// splitpack.hip -- fused PSA split + pack (pass 3 of a PSA build).
//
// Split points follow the CPU reference's splitK convention
// (src/host/reference/split.hpp:67-86): K = ceil(N / split_size) intervals,
// S(0) = (0, 0, 0), S(k) = split(ceil(N*k/K)) for 0 < k < K, S(K) = (L, H-1, 0).
// split() is the lower-bound search of the reference GPU kernels (see
// split_search); the prefixes are read from the prefix-partition layout (heavy
// prefix at j-1, light prefix of i lights at N-i) as the GPU split does
// (split/scalar/shader.comp:43-59,86-88).
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

// Packs interval k, from S(k) = s0 to S(k+1) = (i1, j1) (pack2, pack.hpp:14-83).
template <int PREC>
__device__ __forceinline__ void pack_interval(const float* __restrict__ w, uint32_t N, uint32_t K,
                                              uint32_t H, typename Real<PREC>::T mean,
                                              const uint32_t* __restrict__ idx,
                                              AliasEntry* __restrict__ table, uint64_t k,
                                              SplitV<PREC> s0, uint32_t i1, uint32_t j1) {
    using T = typename Real<PREC>::T;
    if (H == 0 || H == N) {  // all weights equal the mean up to its rounding: identity
        const uint32_t b0 = k == 0 ? 0u : (uint32_t)ceil_mul_div(N, k, K);
        const uint32_t b1 = (uint32_t)ceil_mul_div(N, k + 1, K);
        for (uint32_t x = b0; x < b1; ++x) table[x] = AliasEntry{1.0f, x};
        return;
    }
    const uint32_t i0 = s0.i, j0 = s0.j;
    if (i1 < i0) i1 = i0;
    if (j1 < j0) j1 = j0;
    uint32_t i = i0, j = j0;
    T r = s0.spill;
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
    const T mean = (T)prm->mean;
    SplitV<PREC> sp{0u, 0u, (T)0};
    if (p <= K) sp = split_point<PREC>(hi, lo, N, K, H, mean, p);
    s_i[t] = sp.i;
    s_j[t] = sp.j;
    s_sp[t] = sp.spill;
    __syncthreads();
    if (t == kSplitBlock - 1 || p >= K) return;
    if (splits) splits[p] = SplitRec{s_i[t + 1], s_j[t + 1], (double)s_sp[t + 1]};
    pack_interval<PREC>(w, N, K, H, mean, idx, table, p, sp, s_i[t + 1], s_j[t + 1]);
}

// Segmented form over the flattened split points of all segments: segment g
// owns points seg_pt0[g] .. seg_pt0[g] + K_g (K_g + 1 points, K_g intervals).
template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_seg_splitpack(
    const float* __restrict__ w, const uint64_t* __restrict__ seg_off,
    const uint64_t* __restrict__ seg_pt0, uint32_t nseg, uint64_t total_points,
    uint32_t split_size, const Params* __restrict__ seg_prm, const uint32_t* __restrict__ idx,
    const float* __restrict__ hi, const float* __restrict__ lo, AliasEntry* __restrict__ table) {
    using T = typename Real<PREC>::T;
    __shared__ uint32_t s_i[kSplitBlock], s_j[kSplitBlock];
    __shared__ T s_sp[kSplitBlock];
    const int t = threadIdx.x;
    const uint64_t p = (uint64_t)blockIdx.x * (kSplitBlock - 1) + (uint64_t)t;
    SplitV<PREC> sp{0u, 0u, (T)0};
    uint32_t g = 0;
    uint64_t k = 0, K = 0, off = 0;
    uint32_t N = 0, H = 0;
    T mean = (T)0;
    if (p < total_points) {
        uint32_t a = 0, b = nseg;  // last g with seg_pt0[g] <= p
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (seg_pt0[mid] <= p) a = mid;
            else b = mid;
        }
        g = a;
        k = p - seg_pt0[g];
        off = seg_off[g];
        N = (uint32_t)(seg_off[g + 1] - off);
        K = (N + split_size - 1) / split_size;
        H = seg_prm[g].heavy_count;
        mean = (T)seg_prm[g].mean;
        sp = split_point<PREC>(hi + off, lo ? lo + off : nullptr, N, (uint32_t)K, H, mean, k);
    }
    s_i[t] = sp.i;
    s_j[t] = sp.j;
    s_sp[t] = sp.spill;
    __syncthreads();
    if (t == kSplitBlock - 1 || p >= total_points || k >= K) return;
    pack_interval<PREC>(w + off, N, (uint32_t)K, H, mean, idx + off, table + off, k, sp,
                        s_i[t + 1], s_j[t + 1]);
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


hipError_t launch_seg_splitpack(const float* w, const uint64_t* seg_off, const uint64_t* seg_pt0,
                                uint32_t nseg, uint64_t total_points, uint32_t split_size,
                                int prec, const Params* seg_prm, const uint32_t* idx,
                                const float* hi, const float* lo, AliasEntry* table,
                                hipStream_t s) {
    const uint32_t blocks = (uint32_t)((total_points + (kSplitBlock - 2)) / (kSplitBlock - 1));
#define WRS_LAUNCH_SSP(P)                                                                     \
    k_seg_splitpack<P><<<blocks, kSplitBlock, 0, s>>>(w, seg_off, seg_pt0, nseg, total_points, \
                                                      split_size, seg_prm, idx, hi, lo, table)
    if (prec == kComp) WRS_LAUNCH_SSP(kComp);
    else if (prec == kFP64Carry) WRS_LAUNCH_SSP(kFP64Carry);
    else WRS_LAUNCH_SSP(kFP32);
#undef WRS_LAUNCH_SSP
    return hipGetLastError();
}

}  // namespace wrsk
