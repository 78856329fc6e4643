// splitpack.hip -- PSA split points in the reference splitK layout, and the
// reference's own split + pack strategies as alternative pack forms.
//
// Split points follow the CPU reference's splitK convention
// (src/host/reference/split.hpp:67-86): K = ceil(N / split_size) intervals,
// S(0) = (0, 0, 0), S(k) = split(ceil(N*k/K)) for 0 < k < K, S(K) = (L, H-1, 0).
// split() is the lower-bound search of the reference GPU kernels (see
// split_search); the prefixes are read from the prefix-partition layout (heavy
// prefix at j-1, light prefix of i lights at N-i) as the GPU split does
// (split/scalar/shader.comp:43-59,86-88).  In the FP64 modes the target
// mean*n, the sigmas and the spill are FP64 (the FP64-mode prefix is read as
// hi + lo); in FP32 mode they are FP32.
//
// The default builds pack by a parallel merge (mergepack.hip), which needs no
// split array.  This file serves
//  - the split points of a debug build (k_splits, for the parity tests);
//  - the serial pack form (WRS_PACK_SERIAL): the reference's ScalarSplit +
//    ScalarPack (split/scalar/shader.comp:111-133, pack/scalar/float.comp:
//    60-180): k_splits writes the K split records, then one thread per split
//    interval sweeps it (k_pack_serial);
//  - the subgroup form (WRS_PACK_SUBGROUP): ScalarSplit + the reference's
//    SubgroupPack (pack/subgroup/shader.comp:68-246): the same split records,
//    each interval swept by a group of lanes that load its heavies and
//    lights a chunk at a time (k_pack_subgroup);
//  - the inline form (WRS_PACK_INLINE): the reference's InlineSplitPack
//    (splitpack/inline/shader.comp:225-270), splits and sweeps in one kernel
//    with no split array: a block finds 257 consecutive split points, keeps
//    them in LDS, and its 256 threads sweep the 256 intervals between them.
// The serial and inline forms run the same sweep over the same split points,
// so their tables are identical; the subgroup form's residual steps in
// SubgroupPack's rounding order.  All differ from the merge pack's tables only
// by the rounding of the sweep's running residual (the merge computes every
// residual from exact fixed-point prefix sums).  They are baselines: each
// interval is walked split_size buckets one dependent step at a time.
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

// ---- the sweep forms -----------------------------------------------------
namespace {

template <int PREC>
__device__ __forceinline__ float sweep_ratio(typename Real<PREC>::T x,
                                             typename Real<PREC>::T mean) {
    return fminf(fmaxf((float)(x / mean), 0.0f), 1.0f);
}

// Buckets [S(k).n, S(k+1).n) with every weight equal to the mean up to its
// rounding (H == 0 or L == 0): the identity rows.
__device__ __forceinline__ void identity_rows(uint32_t N, uint32_t K, uint64_t k,
                                              AliasEntry* __restrict__ table) {
    const uint32_t n0 = k == 0 ? 0u : (uint32_t)ceil_mul_div(N, k, K);
    const uint32_t n1 = (uint32_t)ceil_mul_div(N, k + 1u, K);
    for (uint32_t x = n0; x < n1; ++x) table[x] = AliasEntry{1.0f, x};
}

// The sweep of one split interval (i0, j0, spill) -> (i1, j1), as the
// reference packs it (pack/scalar/float.comp:60-158; CPU pack.hpp:14-83): with
// heavy j current and residual r, the next bucket is heavy j itself when
// r <= mean (p = r/mean, alias = heavy j+1), else the next light i
// (p = w/mean, alias = heavy j); then r = (r + w) - mean with w the weight
// just taken.  The interval's first residual is the split's spill (heavy j's
// residual after the split's buckets), or heavy j's weight at S(0).  Indices
// and weights are read in partition order (idx, pw: heavy j at j, light i at
// N-1-i).  The interval that ends at heavy H-1 gives it (1, itself)
// (float.comp:160-171, pack.hpp:78-81).  Ends are clamped to the start, as
// splitpack/inline/shader.comp:266-267 does (the lower-bound splits are
// monotone, so the clamp never acts).
template <int PREC>
__device__ void sweep_interval(uint32_t i0, uint32_t i1, uint32_t j0, uint32_t j1,
                               typename Real<PREC>::T spill, typename Real<PREC>::T mean,
                               uint32_t N, uint32_t H, const uint32_t* __restrict__ idx,
                               const float* __restrict__ pw, AliasEntry* __restrict__ table) {
    using T = typename Real<PREC>::T;
    i1 = max(i1, i0);
    j1 = max(j1, j0);
    uint32_t i = i0, j = j0;
    T r = spill;
    if (r == (T)0) r = (T)pw[j];
    uint32_t h = idx[j];
    while (i != i1 || j != j1) {
        const bool heavy = j == j1 ? false : (i == i1 ? true : r <= mean);
        const uint32_t slot = heavy ? j + 1u : N - 1u - i;
        const uint32_t x = idx[slot];
        const T w = (T)pw[slot];
        if (heavy) {
            table[h] = AliasEntry{sweep_ratio<PREC>(r, mean), x};
            h = x;
            ++j;
        } else {
            table[x] = AliasEntry{sweep_ratio<PREC>(w, mean), h};
            ++i;
        }
        r = (r + w) - mean;
    }
    if (j1 == H - 1u) table[h] = AliasEntry{1.0f, h};
}

}  // namespace

// Serial form, pack pass: interval k (S(k) -> S(k+1), S(0) = (0, 0, 0)) by
// thread k, from the K records k_splits wrote (splits[k-1] = S(k)).
template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_pack_serial(
    uint32_t N, uint32_t K, const Params* __restrict__ prm, const SplitRec* __restrict__ splits,
    const uint32_t* __restrict__ idx, const float* __restrict__ pw,
    AliasEntry* __restrict__ table) {
    using T = typename Real<PREC>::T;
    const uint64_t k = (uint64_t)blockIdx.x * kSplitBlock + threadIdx.x;
    if (k >= K || prm->error) return;
    const uint32_t H = prm->heavy_count;
    if (H == 0 || H == N) {
        identity_rows(N, K, k, table);
        return;
    }
    const SplitRec a = k == 0 ? SplitRec{0u, 0u, 0.0} : splits[k - 1];
    const SplitRec b = splits[k];
    sweep_interval<PREC>(a.i, b.i, a.j, b.j, (T)a.spill, (T)prm->mean, N, H, idx, pw, table);
}

// Subgroup form, pack pass: the reference's SubgroupPack (pack/subgroup/
// shader.comp:68-246) over the K records k_splits wrote.  kPackLanes lanes
// share interval k: each holds one of the next kPackLanes heavies and one of
// the next kPackLanes lights (index and weight, loaded a chunk at a time with
// consecutive lanes on consecutive partition slots), all of them step the
// interval's sweep in lockstep and read the element a step takes from the
// lane that holds it, and the lane holding an element keeps its row until the
// chunk is used up, when the chunk's rows are stored together.  A heavy
// completed by the first heavy of a new chunk is stored at once by lane 0.
// The sweep is sweep_interval's with SubgroupPack's two differences: the
// residual steps as r = r + (w - mean) (shader.comp:240), and an interval
// after the first starts from its spill as it is (shader.comp:118-125; the
// lower-bound spill is never zero, so this matches the scalar pack's start).
// The shader's `hasHeavy && w <= mean` (shader.comp:147) would take a light
// after the interval's last one; as in the scalar pack a heavy is taken.
template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_pack_subgroup(
    uint32_t N, uint32_t K, const Params* __restrict__ prm, const SplitRec* __restrict__ splits,
    const uint32_t* __restrict__ idx, const float* __restrict__ pw,
    AliasEntry* __restrict__ table) {
    using T = typename Real<PREC>::T;
    constexpr int G = kPackLanes;
    static_assert(kSplitBlock % G == 0 && kWave % G == 0, "lane groups inside a wave");
    if (prm->error) return;
    const uint32_t rl = threadIdx.x % G;
    const uint64_t k = ((uint64_t)blockIdx.x * kSplitBlock + threadIdx.x) / G;
    if (k >= K) return;  // whole groups leave together
    const uint32_t H = prm->heavy_count;
    if (H == 0 || H == N) {
        if (rl == 0) identity_rows(N, K, k, table);
        return;
    }
    const uint32_t L = N - H;
    const T mean = (T)prm->mean;
    const SplitRec a = k == 0 ? SplitRec{0u, 0u, 0.0} : splits[k - 1];
    const SplitRec b = splits[k];
    const uint32_t i0 = a.i, i1 = max(b.i, a.i);
    const uint32_t j0 = a.j, j1 = max(b.j, a.j);
    // chunks: heavies at slots jb + rl, lights ib + rl (slot N-1-(ib+rl))
    uint32_t jb = j0, ib = i0;
    uint32_t hx = 0, lx = 0;
    float hw = 0.0f, lw = 0.0f;
    if (jb + rl < H) {
        hx = idx[jb + rl];
        hw = pw[jb + rl];
    }
    if (ib + rl < L) {
        lx = idx[N - 1u - (ib + rl)];
        lw = pw[N - 1u - (ib + rl)];
    }
    uint32_t h = __shfl(hx, 0, G);
    T r = k == 0 ? (T)__shfl(hw, 0, G) : (T)a.spill;
    AliasEntry he{0.0f, 0u}, le{0.0f, 0u};
    bool hv = false, lv = false;  // this lane's rows not yet stored
    uint32_t i = i0, j = j0;
    while (i != i1 || j != j1) {
        const bool heavy = j == j1 ? false : (i == i1 ? true : r <= mean);
        float w;
        if (heavy) {
            const uint32_t nx = j + 1u;  // < H: j1 <= H-1
            if (nx - jb == (uint32_t)G) {
                if (hv) table[hx] = he;
                hv = false;
                jb = nx;
                if (jb + rl < H) {
                    hx = idx[jb + rl];
                    hw = pw[jb + rl];
                }
                const uint32_t x = __shfl(hx, 0, G);
                w = __shfl(hw, 0, G);
                if (rl == 0) table[h] = AliasEntry{sweep_ratio<PREC>(r, mean), x};
                h = x;
            } else {
                const uint32_t x = __shfl(hx, (int)(nx - jb), G);
                w = __shfl(hw, (int)(nx - jb), G);
                if (rl == j - jb) {
                    he = AliasEntry{sweep_ratio<PREC>(r, mean), x};
                    hv = true;
                }
                h = x;
            }
            j = nx;
        } else {
            if (i - ib == (uint32_t)G) {
                if (lv) table[lx] = le;
                lv = false;
                ib = i;  // < i1 <= L
                if (ib + rl < L) {
                    lx = idx[N - 1u - (ib + rl)];
                    lw = pw[N - 1u - (ib + rl)];
                }
            }
            w = __shfl(lw, (int)(i - ib), G);
            if (rl == i - ib) {
                le = AliasEntry{sweep_ratio<PREC>((T)lw, mean), h};
                lv = true;
            }
            ++i;
        }
        r = r + ((T)w - mean);
    }
    if (hv) table[hx] = he;
    if (lv) table[lx] = le;
    if (j1 == H - 1u && rl == 0) table[h] = AliasEntry{1.0f, h};
}

// Inline form: block b finds S(k) for k = 256 b .. 256 b + 256 (thread t the
// t-th, thread 0 also the last) into LDS, then thread t sweeps interval
// 256 b + t.  The reference's InlineSplitPack ends the last interval at
// (L, H) and lets the last heavy alias the slot after it
// (splitpack/inline/shader.comp:243-246, 133-150); here it ends at
// (L, H-1) as splitK does, so that the last heavy gets (1, itself) in every
// form.
template <int PREC>
__global__ __launch_bounds__(kSplitBlock) void k_splitpack_inline(
    uint32_t N, uint32_t K, const Params* __restrict__ prm, const float* __restrict__ hi,
    const float* __restrict__ lo, const uint32_t* __restrict__ idx, const float* __restrict__ pw,
    AliasEntry* __restrict__ table) {
    using T = typename Real<PREC>::T;
    __shared__ uint32_t s_i[kSplitBlock + 1], s_j[kSplitBlock + 1];
    __shared__ T s_sp[kSplitBlock + 1];
    if (prm->error) return;
    const uint32_t H = prm->heavy_count;
    const uint64_t k0 = (uint64_t)blockIdx.x * kSplitBlock;
    const uint32_t t = threadIdx.x;
    if (H == 0 || H == N) {
        if (k0 + t < K) identity_rows(N, K, k0 + t, table);
        return;
    }
    const T mean = (T)prm->mean;
    for (uint32_t c = t; c <= (uint32_t)kSplitBlock; c += kSplitBlock) {
        const uint64_t k = k0 + c;
        if (k <= K) {
            const SplitV<PREC> sp = split_point<PREC>(hi, lo, N, K, H, mean, k);
            s_i[c] = sp.i;
            s_j[c] = sp.j;
            s_sp[c] = sp.spill;
        }
    }
    __syncthreads();
    if (k0 + t >= K) return;
    sweep_interval<PREC>(s_i[t], s_i[t + 1], s_j[t], s_j[t + 1], s_sp[t], mean, N, H, idx, pw,
                         table);
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

hipError_t launch_pack_serial(uint64_t n, uint32_t K, int prec, const Params* prm,
                              const SplitRec* splits, const uint32_t* idx, const float* pw,
                              AliasEntry* table, hipStream_t s) {
    const uint32_t blocks = (K + kSplitBlock - 1) / kSplitBlock;
    const uint32_t N = (uint32_t)n;
    if (prec == kFP32)
        k_pack_serial<kFP32><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, splits, idx, pw, table);
    else
        k_pack_serial<kComp><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, splits, idx, pw, table);
    return hipGetLastError();
}

hipError_t launch_pack_subgroup(uint64_t n, uint32_t K, int prec, const Params* prm,
                                const SplitRec* splits, const uint32_t* idx, const float* pw,
                                AliasEntry* table, hipStream_t s) {
    const uint64_t lanes = (uint64_t)K * kPackLanes;
    const uint32_t blocks = (uint32_t)((lanes + kSplitBlock - 1) / kSplitBlock);
    const uint32_t N = (uint32_t)n;
    if (prec == kFP32)
        k_pack_subgroup<kFP32><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, splits, idx, pw, table);
    else
        k_pack_subgroup<kComp><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, splits, idx, pw, table);
    return hipGetLastError();
}

hipError_t launch_splitpack_inline(uint64_t n, uint32_t K, int prec, const Params* prm,
                                   const float* hi, const float* lo, const uint32_t* idx,
                                   const float* pw, AliasEntry* table, hipStream_t s) {
    const uint32_t blocks = (K + kSplitBlock - 1) / kSplitBlock;
    const uint32_t N = (uint32_t)n;
    if (prec_has_lo(prec))
        k_splitpack_inline<kComp><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, idx, pw,
                                                                  table);
    else if (prec == kFP64Carry)
        k_splitpack_inline<kFP64Carry><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, idx,
                                                                       pw, table);
    else
        k_splitpack_inline<kFP32><<<blocks, kSplitBlock, 0, s>>>(N, K, prm, hi, lo, idx, pw,
                                                                  table);
    return hipGetLastError();
}

}  // namespace wrsk
