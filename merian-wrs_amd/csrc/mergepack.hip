// mergepack.hip -- PSA split + pack as a parallel merge (pass 3 of a build).
//
// The reference packs each split interval with a serial sweep
// (src/host/reference/pack.hpp:14-83, pack/scalar/float.comp:60-180): with
// heavy j current and residual r, the next bucket is heavy j itself if
// r <= mean (p = r/mean, alias = next heavy), else the next light i
// (p = w_i/mean, alias = heavy j).  Writing LP(i) for the sum of the first i
// lights and HP(j) for the sum of the first j heavies, the residual in state
// (i, j) is r = LP(i) + HP(j+1) - (i+j)*mean, and "r <= mean" is
//
//     Y(j+1) <= X(i),   X(i) = i*mean - LP(i),   Y(j) = HP(j) - j*mean,
//
// with X non-decreasing (lights <= mean) and Y increasing (heavies > mean).
// So the sweep is the merge of the lights keyed X(i) with the heavies keyed
// Y(j+1), heavy first on ties -- and the split of the reference GPU kernels
// (largest j with sigma(n-j, j) <= mean*n, split/scalar/shader.comp:72-109) is
// exactly the merge-path diagonal of that merge.  This file therefore builds
// the table as a two-level merge:
//
//  k_coarse   one half-wave per block boundary n = b*tile (tile = buckets
//             per merge block, a multiple of kSample): the exact split
//             (i, j) with LP(i), HP(j).  A 64-ary search over the sparse
//             FP64 prefix samples the partition wrote (every kSample-th rank;
//             at j = m*kSample the light rank n - j is also a multiple of
//             kSample, so both prefixes are samples) brackets j to a window of
//             kSample heavies; the window is resolved exactly by a scan of
//             the 64 heavy and light weights around it.  Weights are read from
//             pw, the weights in partition order written by the partition
//             next to the indices, so every access here is contiguous (the
//             direct form reads the lights from the weights themselves).
//  k_repair_bounds_lb  makes the splits monotone where rounding broke it.
//  k_mergepack one block per `tile` buckets (tile / per threads, `per`
//             buckets each; the shape is a plan parameter): the block's light
//             and heavy index ranges come from its two boundary splits; their
//             indices and weights (both contiguous runs of idx and pw, loaded
//             in one round trip) are staged in LDS, the keys X and Y
//             relative to the block start are rebuilt by an exact integer
//             block scan (a DPP wave scan), and every item's entry follows
//             from how many items of the other kind precede it in the merge:
//             a binary search over the other sorted key list in LDS (a light
//             continues from its predecessor's count, so it usually needs one
//             comparison), with no walk of the merged order.  The entries are
//             written back in rank order (coalesced: consecutive light ranks
//             are consecutive ascending indices).  The last heavy of all gets
//             (1, itself), as in pack.hpp:78-81.
//
// The split search runs in FP64 in the FP64-carry modes and in FP32 in FP32
// mode, as the reference kernels do; the merge keys inside a block are exact
// 64-bit fixed-point sums in every mode (so a block's merge is always a
// permutation of its slots), and a boundary repair pass keeps the block ranges
// valid when rounding has made the splits non-monotone.
#include <algorithm>
#include <type_traits>

#include "device_common.hpp"
#include "internal.hpp"
#include "lookback.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

template <int PREC>
struct RealT {
    using T = double;
};
template <>
struct RealT<kFP32> {
    using T = float;
};

// prefix of the first j heavies / first i lights where j, i are multiples of
// kSample (samples) or i == L (the light total)
template <class T>
__device__ __forceinline__ T hp_sample(const double* __restrict__ samp_h, uint32_t j) {
    return j == 0 ? (T)0 : (T)samp_h[j / kSample - 1u];
}
template <class T>
__device__ __forceinline__ T lp_sample(const double* __restrict__ samp_l, uint32_t i, uint32_t L,
                                       T light_sum) {
    if (i == 0) return (T)0;
    if (i == L) return light_sum;
    return (T)samp_l[i / kSample - 1u];
}


// Direct form: element index of the light of rank r < L.  It is r plus the
// heavies before it; heavy h (element idx[h], heavies in element order) has
// f(h) = idx[h] - h lights before it, f is non-decreasing, and heavy h comes
// before light r iff f(h) <= r.  lpos[k] = x(k kSample) brackets the count to
// [lpos[k] - k kSample, lpos[k+1] - (k+1) kSample] (H after the last
// sampled light), searched by bisection (no steps where no heavy lies between
// the two sampled lights).
__device__ __forceinline__ uint32_t light_pos(const uint32_t* __restrict__ idx,
                                              const uint32_t* __restrict__ lpos, uint32_t r,
                                              uint32_t L, uint32_t H) {
    const uint32_t k = r / kSample;
    const uint32_t r1 = (k + 1u) * kSample;
    // heavies before light k kSample and before light (k+1) kSample (at most H)
    uint32_t a = min(lpos[k] - k * kSample, H);
    uint32_t b = r1 < L ? min(lpos[k + 1u] - r1, H) : H;
    while (a < b) {                               // first h in [a, b) with f(h) > r
        const uint32_t mid = a + (b - a) / 2u;
        if (idx[mid] - mid <= r) a = mid + 1u;
        else b = mid;
    }
    return r + a;
}

}  // namespace

// ---- the exact split at a block boundary, two boundaries per wave ------------
// The split search is bound by its scattered sample loads and a launch has
// more boundaries than the device holds waves, so a wave runs two boundaries:
// each half-wave of 32 lanes works as 64 virtual lanes with two registers per
// lane (virtual lane v = sl in register 0, v = 32 + sl in register 1).  The
// scans below perform the additions of a 64-lane Hillis-Steele scan in the
// same order, so the result does not depend on this mapping.
struct VLane {
    int sl;  // lane in the half-wave
    int h;   // half: 0 or 1
};

// ballot over the 64 virtual lanes of this half
__device__ __forceinline__ uint64_t vballot(bool p0, bool p1, const VLane& v) {
    const uint64_t b0 = (__ballot(p0) >> (32 * v.h)) & 0xffffffffull;
    const uint64_t b1 = (__ballot(p1) >> (32 * v.h)) & 0xffffffffull;
    return b0 | (b1 << 32);
}

// the value virtual lane src holds (src in [0, 64), any per lane)
template <class T>
__device__ __forceinline__ T vshfl(T x0, T x1, int src, const VLane& v) {
    const int l = 32 * v.h + (src & 31);
    const T a = __shfl(x0, l, kWave);
    const T b = __shfl(x1, l, kWave);
    return src < 32 ? a : b;
}

// exclusive scan over the 64 virtual lanes, with the additions of a 64-lane
// Hillis-Steele scan: step d adds virtual lane p - d to lane p (left operand)
template <class T>
__device__ __forceinline__ void vexcl_scan(T& x0, T& x1, const VLane& v) {
    T i0 = x0, i1 = x1;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const int l = 32 * v.h + ((v.sl - d) & 31);
        const T a0 = __shfl(i0, l, kWave);  // register 0 of lane (sl - d) mod 32
        const T a1 = __shfl(i1, l, kWave);
        const T n0 = v.sl >= d ? a0 + i0 : i0;          // p = sl
        const T n1 = (v.sl >= d ? a1 : a0) + i1;        // p = 32 + sl (>= d)
        i0 = n0;
        i1 = n1;
    }
    i1 = i0 + i1;  // d = 32
    const int l = 32 * v.h + ((v.sl - 1) & 31);
    const T b0 = __shfl(i0, l, kWave);
    const T b1 = __shfl(i1, l, kWave);
    x0 = v.sl >= 1 ? b0 : (T)0;
    x1 = v.sl >= 1 ? b1 : b0;  // p = 32: incl(31), register 0 of lane 31
}

// The exact split at n buckets of the distribution (pw, samples) of size N
// with H heavies; executed by one half-wave, result valid in each of its lanes.
// Phase 1: a 32-ary search over the FP64 prefix samples (every kSample-th
// rank) for the largest multiple j of kSample with LP(n - j) + HP(j) <= n mean
// (monotone in j, so any search order finds the same j; 32 candidates a round
// cost fewer scattered sample loads than 64 for one more round at most);
// phase 2: the 64 heavy and light weights around it resolve the exact j.  The
// direct form reads the lights from ls (the heavy indices idx locate them) and
// records the element index of the split's light.
//
// Direct form, phase 2 of a split: the 64 lights of ranks [anchor - kSample,
// anchor), anchor a multiple of kSample with kSample <= anchor < L, are the
// lights of the element range [lpos[anchor/kSample - 1], lpos[anchor/kSample])
// -- exactly kSample lights, with the heavies that lie between them.  When the
// range holds at most kGatherSpan elements the half-wave loads it whole, one
// round of contiguous loads where light_pos would take a dependent bisection
// per light, finds its lights by the partition's test (!(w > pivot)) and hands
// light anchor-1-e to virtual lane e through LDS.  Returns false (nothing
// written) when the case does not apply; the caller then locates each light
// by light_pos, which gives the same lights.  Uniform per half-wave.
constexpr uint32_t kGatherSpan = 8u * 32u;  // up to eight loads per lane of a half-wave

template <class T>
__device__ __forceinline__ bool gather_lights(const LightSrc& ls, float pivot, uint32_t anchor,
                                              uint32_t L, const VLane& v, uint32_t (&lx)[2],
                                              T (&lw)[2]) {
    __shared__ uint32_t s_x[256 / 32][kSample];
    __shared__ float s_w[256 / 32][kSample];
    static_assert(kSample == kWave, "a virtual lane per light of a sample chunk");
    if (anchor < kSample || anchor >= L) return false;
    const uint32_t k = anchor / kSample;
    const uint32_t x0 = ls.lpos[k - 1u], x1 = ls.lpos[k];
    const uint32_t span = x1 - x0;  // kSample lights and the heavies between them
    if (span > kGatherSpan) return false;
    const uint32_t hv = threadIdx.x / 32u;  // this half-wave's LDS rows (blocks of 256)
    uint32_t before = 0;  // lights at smaller offsets
    const uint32_t below = (1u << v.sl) - 1u;
    // the previous split of this half-wave has read its rows
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // passes of kChunk consecutive elements (one for most ranges), each as
    // rounds of 32 loaded together
    constexpr int kChunk = 4 * 32;
    for (uint32_t o0 = 0; o0 < span; o0 += kChunk) {  // uniform per half-wave
        float x[kChunk / 32];
#pragma unroll
        for (int c = 0; c < kChunk / 32; ++c) {
            const uint32_t o = o0 + (uint32_t)v.sl + 32u * (uint32_t)c;
            x[c] = o < span ? ls.w[x0 + o] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < kChunk / 32; ++c) {
            const uint32_t o = o0 + (uint32_t)v.sl + 32u * (uint32_t)c;
            const bool light = o < span && !(x[c] > pivot);
            const uint32_t m = (uint32_t)(__ballot(light) >> (32 * v.h));
            const uint32_t r = before + (uint32_t)__popc(m & below);
            if (light && r < kSample) {
                s_x[hv][r] = x0 + o;
                s_w[hv][r] = x[c];
            }
            before += (uint32_t)__popc(m);
        }
    }
    // the rows of this half-wave are written and read by its own lanes only
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const uint32_t e = (uint32_t)v.sl + 32u * (uint32_t)c;  // light anchor - 1 - e
        lx[c] = s_x[hv][kSample - 1u - e];
        lw[c] = (T)s_w[hv][kSample - 1u - e];
    }
    return true;
}

// A bracket [jb_lo, jb_hi] expected to hold the split's j (the split of the
// previous boundary, and that plus the buckets in between: splits are
// monotone) narrows phase 1 to the sample multiples in it, one round of
// nearby loads instead of rounds of scattered ones.  The narrowed answer is
// used only where it is the full search's (the predicate true at the
// bracket's low end unless that is the range's, false past its answer inside
// the bracket unless the bracket ends with the range); otherwise the full
// search runs, so the split never depends on the bracket.
template <int PREC, bool DIRECT>
__device__ __forceinline__ Coarse coarse_split_half(const float* __restrict__ pw,
                                                    const double* __restrict__ samp_h,
                                                    const double* __restrict__ samp_l, uint32_t N,
                                                    uint32_t H, typename RealT<PREC>::T mean,
                                                    typename RealT<PREC>::T light_sum, uint32_t n,
                                                    const uint32_t* __restrict__ idx,
                                                    const LightSrc& ls, float pivot,
                                                    uint32_t jb_lo = 0u,
                                                    uint32_t jb_hi = 0xffffffffu) {
    using T = typename RealT<PREC>::T;
    const uint32_t L = N - H;
    const VLane v{lane_id() & 31, lane_id() >> 5};
    if (n == 0 || H == 0 || L == 0) {
        const uint32_t x0 = DIRECT && H != 0 && L != 0 ? ls.lpos[0] : 0u;  // light 0
        return Coarse{0u, 0u, 0.0, 0.0, x0, 0u};
    }
    if (n >= N) return Coarse{L, H - 1u, (double)light_sum, 0.0, N, 0u};
    const T target = mean * (T)n;
    const uint32_t lo = n > L ? n - L : 0u;
    const uint32_t hi = min(n, H - 1u);
    // phase 1: largest multiple m*kSample in [lo, hi] with sigma <= target
    const uint32_t m_lo = (uint32_t)(((uint64_t)lo + kSample - 1u) / kSample);
    const uint32_t m_hi = hi / kSample;
    // largest m in [a, b] with the predicate true (-1 if none): 32 candidates
    // a round, one per lane of the half-wave
    auto search = [&](uint32_t a, uint32_t b) -> int64_t {
        int64_t best = -1;
        for (;;) {
            const uint32_t R = b - a + 1u;
            const uint32_t step = (R + 31u) / 32u;
            const uint32_t m = a + (uint32_t)v.sl * step;
            bool pr = false;
            if (m <= b) {
                const uint32_t j = m * kSample;
                const uint32_t i = min(n - j, L);
                pr = lp_sample<T>(samp_l, i, L, light_sum) + hp_sample<T>(samp_h, j) <= target;
            }
            const uint64_t bal = vballot(pr, false, v);
            if (bal == 0) break;
            const int t = 63 - __builtin_clzll(bal);
            best = (int64_t)a + (int64_t)t * step;
            if (step == 1u) break;
            a = (uint32_t)best + 1u;
            b = min(b, (uint32_t)best + step - 1u);
            if (a > b) break;
        }
        return best;
    };
    int64_t best = -1;
    if (m_lo <= m_hi) {
        const uint32_t ba = max(m_lo, jb_lo / kSample);
        const uint32_t bb = jb_hi == 0xffffffffu ? m_hi : min(m_hi, jb_hi / kSample + 1u);
        bool done = false;
        if (ba <= bb && (ba > m_lo || bb < m_hi)) {
            best = search(ba, bb);
            done = (best >= 0 || ba == m_lo) && (best < (int64_t)bb || bb == m_hi);
        }
        if (!done) best = search(m_lo, m_hi);
    }
    const uint32_t base = best >= 0 ? (uint32_t)best * kSample : (lo / kSample) * kSample;
    // phase 2: resolve j in [base, base + 63] with the gathered weights
    const uint32_t i_top = n - base;  // multiple of kSample
    const uint32_t anchor = min(i_top, L);
    const T lp_anchor = lp_sample<T>(samp_l, anchor, L, light_sum);
    const T hp_base = hp_sample<T>(samp_h, base);
    T hw[2], lw[2];
    uint32_t lx[2] = {0u, 0u};
    bool gathered = false;
    if constexpr (DIRECT) gathered = gather_lights<T>(ls, pivot, anchor, L, v, lx, lw);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int vl = v.sl + 32 * c;
        const uint32_t j = base + (uint32_t)vl;
        hw[c] = j < H ? (T)pw[j] : (T)0;
        if (gathered) continue;
        const int64_t lrank = (int64_t)anchor - 1 - vl;  // virtual lane e: light rank anchor-1-e
        lw[c] = (T)0;
        if (lrank >= 0) {
            if constexpr (DIRECT) {
                lx[c] = light_pos(idx, ls.lpos, (uint32_t)lrank, L, H);
                lw[c] = (T)ls.w[lx[c]];
            } else {
                lw[c] = (T)pw[N - 1u - (uint32_t)lrank];
            }
        }
    }
    vexcl_scan<T>(hw[0], hw[1], v);  // HP(base + e) - HP(base)
    vexcl_scan<T>(lw[0], lw[1], v);  // sum of light ranks [anchor - e, anchor)
    T lp[2], hp[2];
    bool pr[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int vl = v.sl + 32 * c;
        const uint32_t j = base + (uint32_t)vl;
        hp[c] = hp_base + hw[c];
        const int64_t i = (int64_t)n - (int64_t)j;
        int64_t e64 = (int64_t)anchor - i;
        e64 = e64 < 0 ? 0 : (e64 > kWave - 1 ? kWave - 1 : e64);
        const T suf = vshfl<T>(lw[0], lw[1], (int)e64, v);
        lp[c] = (T)0;
        if (j > hi) {
            pr[c] = false;
        } else if (j < lo || i > (int64_t)anchor) {
            pr[c] = true;
        } else {
            lp[c] = lp_anchor - suf;
            pr[c] = lp[c] + hp[c] <= target;
        }
    }
    const uint64_t bal = vballot(pr[0], pr[1], v);
    int t = bal ? 63 - __builtin_clzll(bal) : (int)(lo - base);
    if ((uint32_t)t + base < lo) t = (int)(lo - base);
    const T lp_t = vshfl<T>(lp[0], lp[1], t, v);
    const T hp_t = vshfl<T>(hp[0], hp[1], t, v);
    const uint32_t js = base + (uint32_t)t;
    const uint32_t is = n - js;  // in [anchor - 63, anchor]
    uint32_t xs = 0;
    if constexpr (DIRECT) {
        int64_t sel = (int64_t)anchor - 1 - (int64_t)is;
        sel = sel < 0 ? 0 : (sel > kWave - 1 ? kWave - 1 : sel);
        const uint32_t xl = vshfl<uint32_t>(lx[0], lx[1], (int)sel, v);
        xs = is == anchor ? (anchor < L ? ls.lpos[anchor / kSample] : N) : xl;
    }
    return Coarse{is, js, (double)lp_t, (double)hp_t, xs, 0u};
}

// A half-wave takes `run` consecutive boundaries, each after the first
// searched in the bracket its predecessor's split gives.  The launcher sizes
// the run so that the grid is about one block per resident slot of the
// device (8 blocks of 256 threads per CU): on the 2^26 Zipf build (65 537
// boundaries, 256 CUs) that is 4, measured fastest against 1, 2, 3, 5, 6 and
// 8 (shorter runs leave a second partial wave of blocks, longer ones idle
// CUs and lengthen the chain of dependent searches).

template <int PREC, bool DIRECT>
__global__ __launch_bounds__(256) void k_coarse(const float* __restrict__ pw,
                                                const double* __restrict__ samp_h,
                                                const double* __restrict__ samp_l, uint32_t N,
                                                const Params* __restrict__ prm, uint32_t nbound,
                                                uint32_t tile, Coarse* __restrict__ out,
                                                const uint32_t* __restrict__ idx, LightSrc ls,
                                                uint32_t run) {
    using T = typename RealT<PREC>::T;
    if (prm->error) return;
    const uint32_t b0 = (blockIdx.x * 8u + threadIdx.x / 32u) * run;  // this half-wave's first
    uint32_t jb_lo = 0u, jb_hi = 0xffffffffu;
    for (uint32_t b = b0; b < b0 + run && b < nbound; ++b) {  // uniform per half-wave
        const uint32_t n = (uint32_t)min((uint64_t)b * tile, (uint64_t)N);  // N < 2^32 <= b * tile
        const Coarse c = coarse_split_half<PREC, DIRECT>(pw, samp_h, samp_l, N, prm->heavy_count,
                                                         (T)prm->mean, (T)prm->light_sum, n, idx,
                                                         ls, prm->pivot, jb_lo, jb_hi);
        if ((lane_id() & 31) == 0) out[b] = c;
        jb_lo = c.j;
        jb_hi = c.j + tile;
    }
}

// Segmented form: boundary p of the flattened boundary list belongs to segment
// bound_seg[p] (a list the batch plan derives from the offsets).
template <int PREC, bool DIRECT>
__global__ __launch_bounds__(256) void k_seg_coarse(
    const float* __restrict__ pw, const double* __restrict__ samp_h,
    const double* __restrict__ samp_l, const uint64_t* __restrict__ seg_off,
    const uint64_t* __restrict__ seg_samp0, const uint32_t* __restrict__ seg_b0,
    const uint32_t* __restrict__ bound_seg, uint32_t nbound, const Params* __restrict__ seg_prm,
    uint32_t tile, Coarse* __restrict__ out, const uint32_t* __restrict__ idx,
    const float* __restrict__ w, const uint32_t* __restrict__ lpos,
    const uint64_t* __restrict__ seg_lpos0) {
    using T = typename RealT<PREC>::T;
    const uint32_t p = blockIdx.x * 8u + threadIdx.x / 32u;  // one boundary per half-wave
    if (p >= nbound) return;  // the whole half-wave exits together
    const uint32_t g = bound_seg[p];
    const uint64_t off = seg_off[g];
    const uint32_t N = (uint32_t)(seg_off[g + 1] - off);
    const uint32_t n = min((p - seg_b0[g]) * tile, N);
    const uint64_t so = seg_samp0[g];
    const LightSrc ls = DIRECT ? LightSrc{w + off, lpos + seg_lpos0[g]} : LightSrc{};
    const Coarse c = coarse_split_half<PREC, DIRECT>(pw + off, samp_h + so, samp_l + so, N,
                                                     seg_prm[g].heavy_count, (T)seg_prm[g].mean,
                                                     (T)seg_prm[g].light_sum, n,
                                                     DIRECT ? idx + off : nullptr, ls,
                                                     seg_prm[g].pivot);
    if ((lane_id() & 31) == 0) out[p] = c;
}

// quant of a float weight w >= 0 (+-0 included) without FP64 arithmetic, the
// value of merge_quant((double)w, qs): w * 2^qs < 2^63 is exact in FP32 (an
// exponent shift; a result below 2^-126 rounds to 0 either way), rintf rounds
// it to the nearest even integer as the FP64 path does, and the
// integer-valued float converts to int64 as its 24-bit significand shifted
// left (y >= 2^23) or by the exact 32-bit conversion (y < 2^23).
__device__ __forceinline__ int64_t quant_w(float w, int qs) {
    const float y = rintf(ldexpf(w, qs));
    const uint32_t b = __float_as_uint(y) & 0x7fffffffu;
    const int ex = (int)(b >> 23) - 150;  // y = significand * 2^ex
    const uint64_t sig = (uint64_t)((b & 0x7fffffu) | 0x800000u);
    return ex >= 0 ? (int64_t)(sig << ex) : (int64_t)(uint32_t)fabsf(y);
}

// LDS of one merge block, slot e = the e-th item of the block: lights first
// (slot e < nl is light i0 + e), then heavies (slot nl + q is heavy j0 + q,
// q <= nh; the weight of heavy nh is not needed).  A thread owns kPer
// consecutive slots and reads and writes them as 16-byte vectors, so the
// arrays are 16-byte aligned and need no bank swizzle.
template <int TILE, int BLOCK>
struct MergeShared {
    __align__(16) uint64_t key[TILE + 4];  // running key sums S(e); the entries afterwards
    __align__(16) uint32_t idx[TILE + 4];
    __align__(16) float w[TILE + 4];
    uint64_t wave_l[BLOCK / kWave];
    int64_t dq;   // quant(D), D the block's split offset
    double dres;  // residual offset (see merge_core)
};

// kPer consecutive words at p, p aligned to their total size (up to 16 bytes).
template <int K>
__device__ __forceinline__ void lds_load(const float* p, float (&v)[K]) {
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int i = 0; i < K; i += 4) {
            const float4 x = *reinterpret_cast<const float4*>(p + i);
            v[i] = x.x;
            v[i + 1] = x.y;
            v[i + 2] = x.z;
            v[i + 3] = x.w;
        }
    } else {
        static_assert(K == 2, "merge: per-thread slots");
        const float2 x = *reinterpret_cast<const float2*>(p);
        v[0] = x.x;
        v[1] = x.y;
    }
}
template <int K>
__device__ __forceinline__ void lds_store(uint64_t* p, const uint64_t (&v)[K]) {
#pragma unroll
    for (int i = 0; i < K; i += 2)
        *reinterpret_cast<ulonglong2*>(p + i) = make_ulonglong2(v[i], v[i + 1]);
}
template <int K>
__device__ __forceinline__ void lds_store(uint2* p, const uint2 (&v)[K]) {
#pragma unroll
    for (int i = 0; i < K; i += 2)
        *reinterpret_cast<uint4*>(p + i) = make_uint4(v[i].x, v[i].y, v[i + 1].x, v[i + 1].y);
}

// Indices and weights of one block's slots, in registers between their load
// and their store to LDS: slot e < nl is light i0 + e, slot nl <= e <= ne is
// heavy j0 + e - nl (the weight of slot ne is not needed).  Every load is
// issued (at position 0 for the slots beyond ne), so that the staging has no
// branches.
template <int TILE, int BLOCK>
struct MergeStage {
    static constexpr int kPer = TILE / BLOCK;
    uint32_t gi[kPer + 1];
    float gw[kPer + 1];

    __device__ __forceinline__ void issue(const float* __restrict__ pw,
                                          const uint32_t* __restrict__ idx, uint32_t N,
                                          const Coarse& c0, const Coarse& c1) {
        const uint32_t t = threadIdx.x;
        const uint32_t nl = c1.i - c0.i, ne = nl + (c1.j - c0.j);
#pragma unroll
        for (int k = 0; k <= kPer; ++k) {
            const uint32_t e = t + (uint32_t)BLOCK * k;
            uint32_t pos = e < nl ? N - 1u - (c0.i + e) : c0.j + (e - nl);
            pos = e <= ne ? pos : 0u;
            gi[k] = idx[pos];
            gw[k] = pw[pos];
        }
    }
    __device__ __forceinline__ void store(MergeShared<TILE, BLOCK>& s, uint32_t ne) const {
        const uint32_t t = threadIdx.x;
#pragma unroll
        for (int k = 0; k <= kPer; ++k) {
            const uint32_t e = t + (uint32_t)BLOCK * k;
            s.idx[e <= ne ? e : (uint32_t)TILE + 1u] = gi[k];
            s.w[e < ne ? e : (uint32_t)TILE] = gw[k];
        }
    }
};

// Direct-form staging of one block (the same slots as MergeStage): heavies
// j0..j1 from idx/pw; the nl lights i0..i1-1 are the lights (!(w > pivot), the
// partition's test) of the element range [x0, x1) = [c0.xi, c1.xi), in
// element order.  A range of at most kRounds * BLOCK elements (every block
// of a distribution whose heavies do not cluster) is loaded whole: each wave
// takes a contiguous share of it (a multiple of 64 elements, the waves in
// order) in rounds of 64 consecutive elements, counts its lights by ballots,
// and one exclusive scan of the per-wave counts gives every light its slot; a longer range (a run of
// heavies between two of the block's lights) locates each light by
// light_pos.  s_cnt holds BLOCK / kWave counts.  Ends with the slots
// written, not synchronised.
template <int TILE, int BLOCK>
__device__ __forceinline__ void stage_direct(MergeShared<TILE, BLOCK>& s,
                                             uint32_t* __restrict__ s_cnt,
                                             const float* __restrict__ pw,
                                             const uint32_t* __restrict__ idx,
                                             const LightSrc& ls, float pivot, uint32_t N,
                                             uint32_t H, const Coarse& c0, const Coarse& c1) {
    constexpr int kPer = TILE / BLOCK;
    constexpr int kRounds = 2 * kPer;
    constexpr int kWaves = BLOCK / kWave;
    const uint32_t t = threadIdx.x;
    const int lane = lane_id();
    const int wid = (int)t / kWave;
    const uint32_t L = N - H;
    const uint32_t nl = c1.i - c0.i, nh = c1.j - c0.j;
    const uint32_t x1 = min(c1.xi, N), x0 = min(c0.xi, x1), span = x1 - x0;
    const bool dense = span <= (uint32_t)(kRounds * BLOCK);
    // Rounds past the heavies or past the range are skipped as a whole (the
    // conditions are block-uniform): a block usually holds a few heavies and
    // about TILE elements of range.
    // heavy h in [0, nh] goes to slot nl + h (the weight of heavy nh is not needed)
    uint32_t gi[kPer + 1];
    float gw[kPer + 1];
#pragma unroll
    for (int k = 0; k <= kPer; ++k) {
        if ((uint32_t)BLOCK * k <= nh) {
            const uint32_t h = t + (uint32_t)BLOCK * k;
            const uint32_t pos = h <= nh ? c0.j + h : 0u;
            gi[k] = idx[pos];
            gw[k] = pw[pos];
        }
    }
    // wave wid compacts range offsets [wid Q, (wid+1) Q) in rounds of kWave
    // consecutive elements (Q: a multiple of kWave, at most kRounds rounds)
    const uint32_t Q = ((span + kWaves - 1) / kWaves + kWave - 1) & ~(uint32_t)(kWave - 1);
    const uint32_t ow = (uint32_t)wid * Q + (uint32_t)lane;
    float lw[kRounds];
    if (dense) {
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            if ((uint32_t)kWave * k < Q) {  // wave-uniform
                const uint32_t o = ow + (uint32_t)kWave * k;  // offset in the range
                lw[k] = o < span ? ls.w[x0 + o] : 0.0f;
            }
        }
    }
#pragma unroll
    for (int k = 0; k <= kPer; ++k) {
        if ((uint32_t)BLOCK * k <= nh) {
            const uint32_t h = t + (uint32_t)BLOCK * k;
            s.idx[h <= nh ? nl + h : (uint32_t)TILE + 1u] = gi[k];
            s.w[h < nh ? nl + h : (uint32_t)TILE] = gw[k];
        }
    }
    if (dense) {
        uint64_t bal[kRounds];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            bal[k] = 0;
            if ((uint32_t)kWave * k < Q) {
                const uint32_t o = ow + (uint32_t)kWave * k;
                bal[k] = __ballot(o < span && !(lw[k] > pivot));
                cnt += (uint32_t)__popcll(bal[k]);
            }
        }
        if (lane == 0) s_cnt[wid] = cnt;
        __syncthreads();
        uint32_t e_round = 0;  // lights before this round of this wave
        for (int q = 0; q < kWaves; ++q) e_round += q < wid ? s_cnt[q] : 0u;
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            if ((bal[k] >> lane) & 1ull) {
                const uint32_t e = e_round + (uint32_t)__popcll(bal[k] & below);
                if (e < nl) {
                    s.idx[e] = x0 + ow + (uint32_t)kWave * k;
                    s.w[e] = lw[k];
                }
            }
            e_round += (uint32_t)__popcll(bal[k]);
        }
    } else {
        for (uint32_t e = t; e < nl; e += BLOCK) {
            const uint32_t x = light_pos(idx, ls.lpos, c0.i + e, L, H);
            s.idx[e] = x;
            s.w[e] = ls.w[x];
        }
    }
}

// The build scalars the merge uses (from Params).
struct MergeConst {
    double mean, inv_mean;
    int64_t mean_q;
    int qs;
};

// Keys and entries of one block whose slots are staged in s (block
// synchronised, s.dq and s.dres set): leaves the entry of slot e at
// ((uint2*)s.key)[e], e < ne, and ends synchronised.
//
// Local keys are exact 64-bit fixed-point running sums of non-negative
// increments (w_q = quant(w), mean_q = quant(mean)):
//     a_r = sum_{r' < r}  (mean_q - w_q(light i0 + r'))      r in [0, nl]
//     b_q = sum_{q' <= q} (w_q(heavy j0 + q') - mean_q)       q in [0, nh)
// so A and B are sorted whatever the rounding of the prefix sums.  With the
// split offset D = X(i0) - Y(j0) = n0*mean - LP(i0) - HP(j0) (FP64), heavy q
// goes before light r iff Y(j0+q+1) <= X(i0+r), i.e. B[q] = b_q - quant(D) <=
// a_r = A[r]; its residual is Y(j0+q+1) - X(i0+r) + mean = (B[q] - A[r]) 2^-qs
// + dres with dres = quant(D) 2^-qs - D + mean.  The sweep of the merged
// order is never walked: an entry depends only on how many items of the other
// kind precede its item,
//     light r:  alias = heavy q(r),  q(r) = #{q < nh : B[q] <= A[r]}
//               (q(r) == nh: heavy j0 + nh, current at the block's end),
//     heavy q:  residual against light r(q) = #{r < nl : A[r] < B[q]}
//               (A[nl] = S(nl) when every light precedes it), alias heavy q+1,
// both found by binary search over the sorted keys (a light continues from
// its predecessor's q, so it usually needs one comparison).
template <int PREC, int TILE, int BLOCK>
__device__ __forceinline__ void merge_core(MergeShared<TILE, BLOCK>& s, const MergeConst& mc,
                                           uint32_t nl, uint32_t nh) {
    constexpr int kPer = TILE / BLOCK;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    const uint32_t ne = nl + nh;
    const int qs = mc.qs;

    // ---- keys: one running sum over all slots, S(e) = sum_{e' < e} inc(e')
    // (lights first: A[r] = S(r), B[q] = S(nl + q + 1) - S(nl) - quant(D)).
    // Every increment is below 2^62 and S below 2^63 + a few ulps, so S is
    // exact in uint64.
    const uint32_t e0 = (uint32_t)(kPer * t);
    float w[kPer];
    lds_load<kPer>(&s.w[e0], w);
    uint64_t ex[kPer];  // S(e0 + k) once the thread's offset is added
    uint64_t run = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t e = e0 + k;
        uint64_t inc = 0;
        if (e < ne) {
            const uint64_t d = (uint64_t)quant_w(w[k], qs) - (uint64_t)mc.mean_q;
            inc = e < nl ? (uint64_t)0 - d : d;
        }
        ex[k] = run;
        run += inc;
    }
    const uint64_t incl = wave_incl_scan_u64(run);
    if (lane == kWave - 1) s.wave_l[wid] = incl;
    __syncthreads();
    uint64_t off = incl - run;
    for (int q = 0; q < BLOCK / kWave; ++q)
        if (q < wid) off += s.wave_l[q];
#pragma unroll
    for (int k = 0; k < kPer; ++k) ex[k] += off;
    lds_store<kPer>(&s.key[e0], ex);
    if (e0 + (uint32_t)kPer == (uint32_t)TILE) s.key[TILE] = off + run;  // S(TILE), a full block
    __syncthreads();

    // ---- entries -------------------------------------------------------
    // p = residual / mean: FP32 mode divides as the reference does; the FP64
    // modes multiply by the FP64 reciprocal (within 1.5 ulp of the FP64
    // quotient before the rounding to float)
    auto ratio = [&](double x) -> float {
        if constexpr (PREC == kFP32) return (float)((float)x / (float)mc.mean);
        else return (float)(x * mc.inv_mean);
    };
    const uint64_t kb = s.key[nl] + (uint64_t)s.dq;  // B[q] = S(nl + q + 1) - kb
    auto A = [&](uint32_t r) -> int64_t { return (int64_t)s.key[r]; };  // r in [0, nl]
    auto B = [&](uint32_t q) -> int64_t { return (int64_t)(s.key[nl + 1u + q] - kb); };
    constexpr int64_t kInf = INT64_MAX;  // B[nh]: above every A[r], r < nl
    // first B[q] > a for q in [lo, nh] (nh if none)
    auto q_above = [&](uint32_t lo, int64_t a) -> uint32_t {
        uint32_t hi = nh;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (B(mid) <= a) lo = mid + 1u;
            else hi = mid;
        }
        return lo;
    };
    uint2 ent[kPer];
    if (e0 < nl) {
        uint32_t q = q_above(0u, (int64_t)ex[0]);
        int64_t bq = q < nh ? B(q) : kInf;
        uint32_t alias = s.idx[nl + q];  // re-read only when q moves
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (e0 + (uint32_t)k < nl) {
                const int64_t a = (int64_t)ex[k];
                if (k > 0 && bq <= a) {  // heavies between this light and the last
                    q = q_above(q + 1u, a);
                    bq = q < nh ? B(q) : kInf;
                    alias = s.idx[nl + q];
                }
                const float p = fminf(ratio((double)w[k]), 1.0f);
                ent[k] = make_uint2(__float_as_uint(p), alias);
            }
        }
    }
    // heavies: heavy q by thread q mod BLOCK
    uint2 hent[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t q = (uint32_t)t + (uint32_t)BLOCK * j;
        if ((uint32_t)BLOCK * j < nh && q < nh) {
            const int64_t bv = B(q);
            uint32_t lo = 0, hi = nl;  // r(q): first A[r] >= bv, r in [0, nl]
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (A(mid) < bv) lo = mid + 1u;
                else hi = mid;
            }
            const double resid = ldexp((double)(bv - A(lo)), -qs) + s.dres;
            const float p = fminf(fmaxf(ratio(resid), 0.0f), 1.0f);
            hent[j] = make_uint2(__float_as_uint(p), s.idx[nl + q + 1u]);
        }
    }
    __syncthreads();  // keys no longer read: the buffer now takes the entries
    uint2* sent = reinterpret_cast<uint2*>(s.key);
    if (e0 + (uint32_t)kPer <= nl) {
        lds_store<kPer>(&sent[e0], ent);
    } else {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (e0 + (uint32_t)k < nl) sent[e0 + k] = ent[k];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t q = (uint32_t)t + (uint32_t)BLOCK * j;
        if ((uint32_t)BLOCK * j < nh && q < nh) sent[nl + q] = hent[j];
    }
    __syncthreads();
}

// One block of TILE buckets [n0, n1) between boundary splits c0 and c1, BLOCK
// threads, TILE / BLOCK slots per thread.  Rows are written with consecutive
// lanes on consecutive slots (consecutive light indices).  The last heavy of
// all gets (1, itself), as in pack.hpp:78-81.
template <int PREC, int TILE, int BLOCK, bool DIRECT = false>
__device__ __forceinline__ void mergepack_block(const float* __restrict__ pw,
                                                const uint32_t* __restrict__ idx, uint32_t N,
                                                const Params* __restrict__ prm, uint32_t n0,
                                                uint32_t n1, Coarse c0, Coarse c1,
                                                AliasEntry* __restrict__ table,
                                                const LightSrc& ls = LightSrc{}) {
    static_assert(TILE % BLOCK == 0 && BLOCK % kWave == 0 && TILE % kSample == 0, "merge shape");
    __shared__ MergeShared<TILE, BLOCK> s;
    const uint32_t t = threadIdx.x;
    const uint32_t H = prm->heavy_count;
    const uint32_t L = N - H;
    if (H == 0 || L == 0) {  // all weights equal the mean up to its rounding: identity
        for (uint32_t e = t; e < n1 - n0; e += BLOCK) table[n0 + e] = AliasEntry{1.0f, n0 + e};
        return;
    }
    const MergeConst mc{prm->mean, prm->inv_mean, prm->mean_q, prm->merge_qs};
    const uint32_t nl = c1.i - c0.i;  // lights packed here
    const uint32_t nh = c1.j - c0.j;  // heavies packed here (heavy c1.j is current at n1)
    const uint32_t ne = nl + nh;      // == n1 - n0 (- 1 in the last block)
    if constexpr (DIRECT) {
        __shared__ uint32_t s_cnt[BLOCK / kWave];
        stage_direct<TILE, BLOCK>(s, s_cnt, pw, idx, ls, prm->pivot, N, H, c0, c1);
    } else {
        MergeStage<TILE, BLOCK> st;
        st.issue(pw, idx, N, c0, c1);
        st.store(s, ne);
    }
    if (t == 0) {  // the block's split offset (after the staging loads are issued)
        const double D = (double)n0 * mc.mean - (c0.lp + c0.hp);
        const int64_t dq = merge_quant(D, mc.qs);
        s.dq = dq;
        s.dres = ldexp((double)dq, -mc.qs) - D + mc.mean;
    }
    __syncthreads();
    merge_core<PREC, TILE, BLOCK>(s, mc, nl, nh);
    uint2* tab = reinterpret_cast<uint2*>(table);
    const uint2* sent = reinterpret_cast<const uint2*>(s.key);
    for (uint32_t e = t; e < ne; e += BLOCK) {
        const uint32_t x = s.idx[e];
        if (!DIRECT || x < N) tab[x] = sent[e];  // direct: every slot is < N by construction
    }
    if (n1 == N && t == 0) {
        const uint32_t h = s.idx[ne];  // heavy H-1
        table[h] = AliasEntry{1.0f, h};
    }
}

template <int PREC, int TILE, int BLOCK, bool DIRECT>
__global__ __launch_bounds__(BLOCK) void k_mergepack(const float* __restrict__ pw,
                                                     const uint32_t* __restrict__ idx, uint32_t N,
                                                     const Params* __restrict__ prm,
                                                     const Coarse* __restrict__ bounds,
                                                     AliasEntry* __restrict__ table,
                                                     LightSrc ls) {
    if (prm->error) return;  // an earlier wait of the build gave up: no table
    const uint32_t b = blockIdx.x;
    const uint32_t n0 = b * (uint32_t)TILE;  // < N
    const uint32_t n1 = (uint32_t)min((uint64_t)n0 + TILE, (uint64_t)N);
    mergepack_block<PREC, TILE, BLOCK, DIRECT>(pw, idx, N, prm, n0, n1, bounds[b], bounds[b + 1],
                                               table, ls);
}

template <int PREC, int TILE, int BLOCK, bool DIRECT>
__global__ __launch_bounds__(BLOCK) void k_seg_mergepack(const float* __restrict__ pw,
                                                         const uint32_t* __restrict__ idx,
                                                         const SegItem* __restrict__ mitem_d,
                                                         const Params* __restrict__ seg_prm,
                                                         const Coarse* __restrict__ bounds,
                                                         AliasEntry* __restrict__ table,
                                                         const float* __restrict__ w,
                                                         const uint32_t* __restrict__ lpos,
                                                         const uint64_t* __restrict__ seg_lpos0) {
    const SegItem it = mitem_d[blockIdx.x];
    const uint32_t N = it.n;
    const uint32_t n0 = it.blk * (uint32_t)TILE;
    const uint32_t n1 = min(n0 + (uint32_t)TILE, N);
    const Coarse* bd = bounds + it.bnd;
    const LightSrc ls = DIRECT ? LightSrc{w + it.off, lpos + seg_lpos0[it.seg]} : LightSrc{};
    mergepack_block<PREC, TILE, BLOCK, DIRECT>(pw + it.off, idx + it.off, N, seg_prm + it.seg, n0,
                                               n1, bd[0], bd[1], table + it.off, ls);
}

// ---- boundary repair ---------------------------------------------------------
// Splits found independently per boundary are monotone in the boundary when
// the key sequences X and Y are; rounding in large prefix sums (FP32 mode at
// large N, or weights within rounding of the mean) can break that.  This pass
// makes every block range valid.  With m_b the buckets before boundary b:
//     j'_b  = max_{b' <= b} j_b'              (heavy ranks non-decreasing)
//     i''_b = max_{b' <= b} (m_b' - j'_b')    (light ranks non-decreasing)
//     j''_b = m_b - i''_b
// keeps i'' + j'' = m_b (so no block grows), j'' non-decreasing (i'' - i'
// grows by at most m_{b+1} - m_b per step), and the ranges of the raw splits
// (j'' in [max(0, m_b - L), min(m_b, H - 1)]).  Exact keys are monotone
// already and then nothing changes.  The last boundary (m = N) is fixed.
// segmented form: one wave per segment (a batch holds thousands of short
// segments; a segment has at most 2^23 / kMinMergeTile + 1 boundaries, i.e. at
// most 33 rounds of kRepairThreads * kRepairPer)
constexpr int kRepairThreads = 64;
constexpr int kRepairPer = 4;

__device__ __forceinline__ uint32_t wave_incl_max_u(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = shfl_up(v, d);
        if (l >= d) v = v < o ? o : v;
    }
    return v;
}

// A repaired boundary gets the prefixes LP(i), HP(j) of its new split: the
// nearest sample below plus at most kSample - 1 weights (as in k_coarse), and
// in the direct form the element index of light i.
template <bool DIRECT>
__device__ __forceinline__ void repair_anchor(Coarse& c, const float* __restrict__ pw,
                                              const double* __restrict__ samp_h,
                                              const double* __restrict__ samp_l, uint32_t N,
                                              uint32_t H, double light_sum,
                                              const uint32_t* __restrict__ idx,
                                              const LightSrc& ls) {
    const uint32_t L = N - H;
    double lp = 0.0, hp = 0.0;
    if (c.i == L) {
        lp = light_sum;
    } else if (c.i > 0) {
        const uint32_t m = c.i / kSample;
        lp = m ? samp_l[m - 1u] : 0.0;
        for (uint32_t r = m * kSample; r < c.i; ++r)
            lp += DIRECT ? (double)ls.w[light_pos(idx, ls.lpos, r, L, H)]
                         : (double)pw[N - 1u - r];
    }
    if constexpr (DIRECT) c.xi = c.i < L ? light_pos(idx, ls.lpos, c.i, L, H) : N;
    if (c.j > 0) {
        const uint32_t m = c.j / kSample;
        hp = m ? samp_h[m - 1u] : 0.0;
        for (uint32_t q = m * kSample; q < c.j; ++q) hp += (double)pw[q];
    }
    c.lp = lp;
    c.hp = hp;
}

// Segmented form: one block (kRepairThreads) per segment, boundaries
// [seg_b0[g], seg_b0[g + 1]).  The raw split j is kept in the record's pad
// word between the passes.
template <bool DIRECT>
__global__ __launch_bounds__(kRepairThreads) void k_repair_bounds(
    Coarse* __restrict__ bounds, const uint32_t* __restrict__ seg_b0,
    const uint64_t* __restrict__ seg_off, const uint64_t* __restrict__ seg_samp0,
    const Params* __restrict__ seg_prm, const float* __restrict__ pw,
    const double* __restrict__ samp_h, const double* __restrict__ samp_l, uint32_t tile,
    const uint32_t* __restrict__ idx, const float* __restrict__ w,
    const uint32_t* __restrict__ lpos, const uint64_t* __restrict__ seg_lpos0) {
    __shared__ uint32_t s_w[kRepairThreads / kWave];
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    const uint32_t g = blockIdx.x;
    const uint32_t b0 = seg_b0[g];
    const uint32_t b1 = seg_b0[g + 1];
    const uint32_t M = b1 - b0 - 1u;  // all but the last boundary
    Coarse* bd = bounds + b0;
    const uint64_t off = seg_off[g];
    const uint32_t N = (uint32_t)(seg_off[g + 1] - off);
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < M; c0 += (uint32_t)(kRepairThreads * kRepairPer)) {
            const uint32_t base = c0 + (uint32_t)t * kRepairPer;
            uint32_t v[kRepairPer];
            uint32_t m = 0;
#pragma unroll
            for (int k = 0; k < kRepairPer; ++k) {
                const uint32_t b = base + k;
                v[k] = 0;
                if (b < M) v[k] = pass == 0 ? bd[b].j : b * tile - bd[b].j;
                m = m < v[k] ? v[k] : m;
            }
            const uint32_t incl = wave_incl_max_u(m);
            uint32_t run = shfl_up(incl, 1);
            if (lane == 0) run = 0;
            if (lane == kWave - 1) s_w[wid] = incl;
            __syncthreads();
            uint32_t chunk = carry;
            for (int q = 0; q < kRepairThreads / kWave; ++q) {
                if (q < wid) run = run < s_w[q] ? s_w[q] : run;
                chunk = chunk < s_w[q] ? s_w[q] : chunk;
            }
            run = run < carry ? carry : run;
#pragma unroll
            for (int k = 0; k < kRepairPer; ++k) {
                const uint32_t b = base + k;
                run = run < v[k] ? v[k] : run;
                if (b < M) {
                    if (pass == 0) {
                        bd[b].aux = bd[b].j;
                        bd[b].j = run;
                    } else {
                        Coarse c = bd[b];
                        c.i = run;
                        c.j = b * tile - run;
                        if (c.j != c.aux)
                            repair_anchor<DIRECT>(
                                c, pw + off, samp_h + seg_samp0[g], samp_l + seg_samp0[g], N,
                                seg_prm[g].heavy_count, seg_prm[g].light_sum,
                                DIRECT ? idx + off : nullptr,
                                DIRECT ? LightSrc{w + off, lpos + seg_lpos0[g]} : LightSrc{});
                        bd[b] = c;
                    }
                }
            }
            carry = chunk;
            __syncthreads();  // s_w is rewritten by the next chunk; pass 1 reads pass 0's writes
        }
    }
}

// Single-distribution form: block g repairs boundaries
// [g * kRepairSpan, (g + 1) * kRepairSpan) and chains to its predecessors by
// decoupled look-back, once per pass (max is exact and order-free, so the
// result does not depend on the look-back timing).
constexpr int kRepairLbThreads = 256;
constexpr int kRepairLbPer = (int)kRepairSpan / kRepairLbThreads;

struct MaxOps {
    static __device__ __forceinline__ uint32_t zero() { return 0u; }
    static __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) { return a < b ? b : a; }
    static __device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = __shfl_xor(v, d, kWave);
            v = v < o ? o : v;
        }
        return v;
    }
    static __device__ __forceinline__ bool read(const uint64_t* r, uint32_t& v) {
        const uint64_t x = ld_word(r);
        v = (uint32_t)x;
        return (x >> 63) == 0;
    }
    static __device__ __forceinline__ void publish(uint64_t* r, const uint32_t& v) {
        st_word(r, (uint64_t)v);
    }
};

// In-place inclusive max-scan of the block's items (thread t holds items
// kRepairLbPer*t ..); returns the block maximum.  Ends synchronised.
__device__ __forceinline__ uint32_t block_incl_max(uint32_t (&v)[kRepairLbPer], uint32_t* s_w) {
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kRepairLbPer; ++k) v[k] = m = m < v[k] ? v[k] : m;
    const uint32_t incl = wave_incl_max_u(m);
    uint32_t run = shfl_up(incl, 1);
    if (lane == 0) run = 0;
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    uint32_t total = 0;
    for (int q = 0; q < kRepairLbThreads / kWave; ++q) {
        if (q < wid) run = run < s_w[q] ? s_w[q] : run;
        total = total < s_w[q] ? s_w[q] : total;
    }
#pragma unroll
    for (int k = 0; k < kRepairLbPer; ++k) v[k] = v[k] < run ? run : v[k];
    __syncthreads();
    return total;
}

// Exclusive max over the blocks before blk (records `recs`), executed by wave 0;
// result broadcast through *s_ex.
__device__ __forceinline__ void repair_lookback(uint64_t* recs, uint32_t blk, uint32_t agg,
                                                uint32_t* s_ex, Params* prm) {
    if (threadIdx.x >= (unsigned)kWave) return;
    const int lane = lane_id();
    uint64_t* rec = recs + (uint64_t)blk * kStateWords;
    uint32_t ex = 0;
    if (blk == 0) {
        if (lane == 0) publish_inclusive<uint32_t, MaxOps>(rec, agg);
    } else {
        if (lane == 0) publish_aggregate<uint32_t, MaxOps>(rec, agg);
        ex = lookback<uint32_t, MaxOps>(recs, blk, prm->wait_polls, &prm->error);
        if (lane == 0) publish_inclusive<uint32_t, MaxOps>(rec, MaxOps::add(ex, agg));
    }
    if (lane == 0) *s_ex = ex;
}

template <bool DIRECT>
__global__ __launch_bounds__(kRepairLbThreads) void k_repair_bounds_lb(
    Coarse* __restrict__ bounds, uint32_t nbound, uint32_t N, Params* __restrict__ prm,
    uint64_t* __restrict__ recs, const float* __restrict__ pw, const double* __restrict__ samp_h,
    const double* __restrict__ samp_l, uint32_t tile, const uint32_t* __restrict__ idx,
    LightSrc ls) {
    __shared__ uint32_t s_w[kRepairLbThreads / kWave];
    __shared__ uint32_t s_blk, s_ex;
    if (prm->error) return;
    if (threadIdx.x == 0) s_blk = atomicAdd(&prm->repair_ticket, 1u);  // predecessors first
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t M = nbound - 1u;  // all but the last boundary
    const uint32_t nrb = (M + kRepairSpan - 1u) / kRepairSpan;
    const uint32_t base = blk * kRepairSpan + (uint32_t)threadIdx.x * kRepairLbPer;
    // pass 0: j' = running max of j
    uint32_t v[kRepairLbPer], raw_j[kRepairLbPer];
#pragma unroll
    for (int k = 0; k < kRepairLbPer; ++k) raw_j[k] = v[k] = base + k < M ? bounds[base + k].j : 0u;
    uint32_t agg = block_incl_max(v, s_w);
    repair_lookback(recs, blk, agg, &s_ex, prm);
    __syncthreads();
    uint32_t ex = s_ex;
    // pass 1: i'' = running max of m_b - j'_b
#pragma unroll
    for (int k = 0; k < kRepairLbPer; ++k) {
        const uint32_t jp = v[k] < ex ? ex : v[k];
        v[k] = base + k < M ? (base + k) * tile - jp : 0u;
    }
    agg = block_incl_max(v, s_w);
    repair_lookback(recs + (uint64_t)nrb * kStateWords, blk, agg, &s_ex, prm);
    __syncthreads();
    ex = s_ex;
#pragma unroll
    for (int k = 0; k < kRepairLbPer; ++k) {
        const uint32_t b = base + k;
        if (b < M) {
            const uint32_t i = v[k] < ex ? ex : v[k];
            const uint32_t j = b * tile - i;
            if (j != raw_j[k]) {
                Coarse c = bounds[b];
                c.i = i;
                c.j = j;
                repair_anchor<DIRECT>(c, pw, samp_h, samp_l, N, prm->heavy_count,
                                      prm->light_sum, idx, ls);
                bounds[b] = c;
            }
        }
    }
}

hipError_t launch_repair_bounds(Coarse* bounds, uint64_t n, MergeShape ms, Params* prm,
                                uint64_t* states, const float* pw, const double* samp_h,
                                const double* samp_l, hipStream_t s, const uint32_t* idx,
                                LightSrc ls) {
    const uint32_t nbound = (uint32_t)merge_blocks_for(n, ms.tile) + 1u;
    const uint32_t grid = (uint32_t)repair_blocks_for(n, ms.tile);
    uint64_t* recs = states + repair_records_offset(n);
    if (ls.lpos)
        k_repair_bounds_lb<true><<<grid, kRepairLbThreads, 0, s>>>(
            bounds, nbound, (uint32_t)n, prm, recs, pw, samp_h, samp_l, ms.tile, idx, ls);
    else
        k_repair_bounds_lb<false><<<grid, kRepairLbThreads, 0, s>>>(
            bounds, nbound, (uint32_t)n, prm, recs, pw, samp_h, samp_l, ms.tile, idx, ls);
    return hipGetLastError();
}

hipError_t launch_coarse(const float* pw, uint64_t n, MergeShape ms, int prec, const Params* prm,
                         const double* samp_h, const double* samp_l, Coarse* bounds,
                         hipStream_t s, const uint32_t* idx, LightSrc ls) {
    const uint32_t N = (uint32_t)n;
    const uint32_t nbound = (uint32_t)merge_blocks_for(n, ms.tile) + 1u;
    static const uint32_t slots = [] {  // resident 256-thread blocks: 8 per CU
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        return 8u * (uint32_t)cus;
    }();
    const uint32_t half_waves = 8u * slots;  // 8 half-waves a block
    const uint32_t run = std::min(8u, std::max(1u, (nbound + half_waves / 2u) / half_waves));
    const uint32_t runs = (nbound + run - 1u) / run;
    const uint32_t cgrid = (runs + 7u) / 8u;  // 8 runs of boundaries per 256-thread block
    // the split arithmetic is FP32 in FP32 mode and FP64 in all others
#define WRS_COARSE(P, D)                                                                    \
    k_coarse<P, D><<<cgrid, 256, 0, s>>>(pw, samp_h, samp_l, N, prm, nbound, ms.tile, bounds, \
                                         idx, ls, run)
    if (prec == kFP32) {
        if (ls.lpos) WRS_COARSE(kFP32, true);
        else WRS_COARSE(kFP32, false);
    } else {
        if (ls.lpos) WRS_COARSE(kFP64Carry, true);
        else WRS_COARSE(kFP64Carry, false);
    }
#undef WRS_COARSE
    return hipGetLastError();
}

// The merge shapes instantiated: (buckets per block, buckets per thread).
#define WRS_MERGE_SHAPES(X) \
    X(1024, 2) X(1024, 4) X(1024, 8) X(2048, 2) X(2048, 4) X(2048, 8) X(4096, 4) X(4096, 8)

bool merge_shape_supported(MergeShape ms) {
#define WRS_SHAPE_OK(T, P) if (ms.tile == (T) && ms.per == (P)) return true;
    WRS_MERGE_SHAPES(WRS_SHAPE_OK)
#undef WRS_SHAPE_OK
    return false;
}

namespace {
template <int P, int T, int B>
hipError_t launch_mp(const float* pw, uint32_t N, const Params* prm, const uint32_t* idx,
                     const Coarse* bounds, AliasEntry* table, hipStream_t s, LightSrc ls) {
    const uint32_t grid = (uint32_t)merge_blocks_for(N, T);
    if (ls.lpos)
        k_mergepack<P, T, B, true><<<grid, B, 0, s>>>(pw, idx, N, prm, bounds, table, ls);
    else
        k_mergepack<P, T, B, false><<<grid, B, 0, s>>>(pw, idx, N, prm, bounds, table, ls);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_mergepack(const float* pw, uint64_t n, MergeShape ms, int prec,
                            const Params* prm, const uint32_t* idx, const Coarse* bounds,
                            AliasEntry* table, hipStream_t s, LightSrc ls) {
    const uint32_t N = (uint32_t)n;
    // the merge keys are exact in every mode; only FP32 mode's p differs
#define WRS_MP(T, P)                                                                       \
    if (ms.tile == (T) && ms.per == (P))                                                   \
        return prec == kFP32                                                               \
                   ? launch_mp<kFP32, T, T / P>(pw, N, prm, idx, bounds, table, s, ls)      \
                   : launch_mp<kFP64Carry, T, T / P>(pw, N, prm, idx, bounds, table, s, ls);
    WRS_MERGE_SHAPES(WRS_MP)
#undef WRS_MP
    return hipErrorInvalidValue;
}

hipError_t launch_seg_mergepack(const float* pw, const uint32_t* idx, const double* samp_h,
                                const double* samp_l, const uint64_t* seg_off,
                                const uint64_t* seg_samp0, const uint32_t* seg_b0, uint32_t nseg,
                                uint32_t nbound, const uint32_t* bound_seg,
                                const SegItem* mitem_d, uint32_t mitems, MergeShape ms, int prec,
                                const Params* seg_prm, Coarse* bounds, AliasEntry* table,
                                hipStream_t s, const float* w, const uint32_t* lpos,
                                const uint64_t* seg_lpos0) {
    if (!merge_shape_supported(ms)) return hipErrorInvalidValue;
    const bool direct = lpos != nullptr;
    if (direct && (w == nullptr || seg_lpos0 == nullptr)) return hipErrorInvalidValue;
    const uint32_t cgrid = (nbound + 7u) / 8u;  // 8 boundaries per 256-thread block
#define WRS_SEG_COARSE(P, D)                                                                      \
    k_seg_coarse<P, D><<<cgrid, 256, 0, s>>>(pw, samp_h, samp_l, seg_off, seg_samp0, seg_b0,      \
                                             bound_seg, nbound, seg_prm, ms.tile, bounds, idx, w, \
                                             lpos, seg_lpos0)
    if (prec == kFP32) {
        if (direct) WRS_SEG_COARSE(kFP32, true);
        else WRS_SEG_COARSE(kFP32, false);
    } else {
        if (direct) WRS_SEG_COARSE(kFP64Carry, true);
        else WRS_SEG_COARSE(kFP64Carry, false);
    }
#undef WRS_SEG_COARSE
    if (direct)
        k_repair_bounds<true><<<nseg, kRepairThreads, 0, s>>>(bounds, seg_b0, seg_off, seg_samp0,
                                                              seg_prm, pw, samp_h, samp_l, ms.tile,
                                                              idx, w, lpos, seg_lpos0);
    else
        k_repair_bounds<false><<<nseg, kRepairThreads, 0, s>>>(bounds, seg_b0, seg_off, seg_samp0,
                                                               seg_prm, pw, samp_h, samp_l,
                                                               ms.tile, idx, w, lpos, seg_lpos0);
#define WRS_SMP(T, P)                                                                           \
    if (ms.tile == (T) && ms.per == (P)) {                                                      \
        if (prec == kFP32) {                                                                    \
            if (direct)                                                                         \
                k_seg_mergepack<kFP32, T, T / P, true><<<mitems, T / P, 0, s>>>(                \
                    pw, idx, mitem_d, seg_prm, bounds, table, w, lpos, seg_lpos0);              \
            else                                                                                \
                k_seg_mergepack<kFP32, T, T / P, false><<<mitems, T / P, 0, s>>>(               \
                    pw, idx, mitem_d, seg_prm, bounds, table, w, lpos, seg_lpos0);              \
        } else {                                                                                \
            if (direct)                                                                         \
                k_seg_mergepack<kFP64Carry, T, T / P, true><<<mitems, T / P, 0, s>>>(           \
                    pw, idx, mitem_d, seg_prm, bounds, table, w, lpos, seg_lpos0);              \
            else                                                                                \
                k_seg_mergepack<kFP64Carry, T, T / P, false><<<mitems, T / P, 0, s>>>(          \
                    pw, idx, mitem_d, seg_prm, bounds, table, w, lpos, seg_lpos0);              \
        }                                                                                       \
    }
    WRS_MERGE_SHAPES(WRS_SMP)
#undef WRS_SMP
    return hipGetLastError();
}

}  // namespace wrsk
