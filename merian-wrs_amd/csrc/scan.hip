// scan.hip -- inclusive prefix sum (the ITS build, and the standalone scan).
//
// Semantics: out[i] = w[0] + ... + w[i] (reference DecoupledPrefixSum,
// src/device/prefix_sum/decoupled/shader.comp:321-333, CPU form
// src/host/reference/prefix_sum.hpp:13-34); with `normalise` the prefix is
// divided by the FP64 total, giving a CDF.
//
// Two forms over the same tile machinery as the prefix-partition (4096-element
// tiles, LDS transpose to 16 consecutive elements per thread, wave64 scan,
// tail-safe loads and stores; the reference kernel has no tail check,
// shader.comp:70-92,181-197):
//   - small n: one pass with decoupled look-back (dynamic tile ticket), after
//     the mean pass that supplies the total and resets the records;
//   - large n: reduce-then-scan (tile sums, a spine scan of them, an emit pass).
// In-tile sums use the mode's tile arithmetic (tilesum.hpp).  The carries are
// exact fixed-point sums in the FP64 modes (so their bits do not depend on
// look-back timing) and an FP32 sum in FP32 mode.  Outputs are staged back
// through LDS and written as coalesced float4 rows.
#include "device_common.hpp"
#include "internal.hpp"
#include "lookback.hpp"
#include "tileload.hpp"
#include "tilesum.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {
constexpr int kPad = kTile + kTile / 16;

template <int PREC>
__device__ __forceinline__ FX sc_add(FX a, FX b) {
    if constexpr (PREC == kFP32) {
        return FX{(uint64_t)__float_as_uint(__uint_as_float((uint32_t)a.lo) +
                                            __uint_as_float((uint32_t)b.lo)),
                  0};
    } else {
        return fx_add(a, b);
    }
}
template <int PREC>
__device__ __forceinline__ FX sc_wave_sum(FX v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        FX o;
        o.lo = __shfl_xor(v.lo, d, kWave);
        o.hi = __shfl_xor(v.hi, d, kWave);
        v = sc_add<PREC>(v, o);
    }
    return v;
}
__device__ __forceinline__ bool sc_read(const uint64_t* rec, FX& v) {
    v.lo = ld_word(rec + 0);
    v.hi = ld_word(rec + 1);
    return ((v.lo | v.hi) >> 63) == 0;
}
__device__ __forceinline__ void sc_publish(uint64_t* rec, FX v) {
    st_word(rec + 0, v.lo);
    st_word(rec + 1, v.hi);
}
template <int PREC>
struct ScanOps {
    static __device__ __forceinline__ FX zero() { return fx_zero(); }
    static __device__ __forceinline__ FX add(FX a, FX b) { return sc_add<PREC>(a, b); }
    static __device__ __forceinline__ FX wave_sum(FX v) { return sc_wave_sum<PREC>(v); }
    static __device__ __forceinline__ bool read(const uint64_t* r, FX& v) { return sc_read(r, v); }
    static __device__ __forceinline__ void publish(uint64_t* r, const FX& v) { sc_publish(r, v); }
};

// LDS of one tile: the padded transpose buffer and the per-wave totals.
template <int PREC>
struct ScanShared {
    float buf[kPad];
    typename Acc1<PREC>::T wave[kBlock / kWave];
};

// Loads tile [base, base + m) into registers (thread t holds the 16
// consecutive elements 16t..16t+15, through an LDS transpose of coalesced
// float4 rows, tileload.hpp) and scans it: thr_excl is the sum of the tile's elements before
// thread t's, tile_agg the tile's sum.  Every kernel of this file computes a
// tile's sums through this function, so the single-pass and the
// reduce-then-scan forms give identical in-tile sums.
template <int PREC>
__device__ __forceinline__ void scan_tile(const float* __restrict__ w, uint32_t base, uint32_t m,
                                          ScanShared<PREC>& s, float (&v)[kItems],
                                          typename Acc1<PREC>::T& thr_excl,
                                          typename Acc1<PREC>::T& tile_agg) {
    using A1 = Acc1<PREC>;
    using R = typename A1::T;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    load_tile_lds<true>(w + base, m, s.buf);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = s.buf[17 * t + k];
    const uint32_t e0 = (uint32_t)(kItems * t);

    R mine = A1::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (e0 + k < m) mine = A1::add1(mine, v[k]);
    R incl = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const R o = A1::shfl_up(incl, d);
        if (lane >= d) incl = A1::add(o, incl);
    }
    R lane_excl = A1::shfl_up(incl, 1);
    if (lane == 0) lane_excl = A1::zero();
    if (lane == kWave - 1) s.wave[wid] = incl;
    __syncthreads();
    R wave_excl = A1::zero();
    for (int q = 0; q < wid; ++q) wave_excl = A1::add(wave_excl, s.wave[q]);
    tile_agg = s.wave[0];
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q) tile_agg = A1::add(tile_agg, s.wave[q]);
    thr_excl = A1::add(wave_excl, lane_excl);
}

// Writes the tile's inclusive prefixes: carry c (FP64 modes) or cf (FP32
// mode) plus the in-tile running sum, times scale; hi + lo float pairs in the
// modes with a low word.  Staged through LDS so that rows leave as float4.
template <int PREC>
__device__ __forceinline__ void scan_tile_store(uint32_t base, uint32_t m, bool full,
                                                const float (&v)[kItems],
                                                typename Acc1<PREC>::T thr_excl, double c,
                                                float cf, double scale, ScanShared<PREC>& s,
                                                float* __restrict__ hi_out,
                                                float* __restrict__ lo_out) {
    using A1 = Acc1<PREC>;
    const int t = threadIdx.x;
    float lo_keep[kItems];
    {
        typename A1::T run = thr_excl;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            float hi, lo = 0.0f;
            run = A1::add1(run, v[k]);
            if constexpr (prec_has_lo(PREC)) {
                const double P = (c + A1::value(run)) * scale;
                hi = (float)P;
                lo = (float)(P - (double)hi);
            } else if constexpr (PREC == kFP64Carry) {
                hi = (float)((c + (double)run) * scale);
            } else {
                hi = (float)((double)(cf + run) * scale);
            }
            lo_keep[k] = lo;
            s.buf[17 * t + k] = hi;
        }
    }
    __syncthreads();
    if (full) {
        float4* o4 = reinterpret_cast<float4*>(hi_out + base);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = 4 * (t + kBlock * j);
            const int q = e + e / 16;
            o4[t + kBlock * j] = make_float4(s.buf[q], s.buf[q + 1], s.buf[q + 2], s.buf[q + 3]);
        }
    } else {
        for (uint32_t e = t; e < m; e += kBlock) hi_out[base + e] = s.buf[e + e / 16];
    }
    if (prec_has_lo(PREC) && lo_out != nullptr) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; ++k) s.buf[17 * t + k] = lo_keep[k];
        __syncthreads();
        if (full) {
            float4* o4 = reinterpret_cast<float4*>(lo_out + base);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = 4 * (t + kBlock * j);
                const int q = e + e / 16;
                o4[t + kBlock * j] = make_float4(s.buf[q], s.buf[q + 1], s.buf[q + 2], s.buf[q + 3]);
            }
        } else {
            for (uint32_t e = t; e < m; e += kBlock) lo_out[base + e] = s.buf[e + e / 16];
        }
    }
}

// Tile aggregate as a carry word pair: FX at the build's shift in the FP64
// modes, the float's bits in word .lo in FP32 mode.
template <int PREC>
__device__ __forceinline__ FX agg_to_carry(typename Acc1<PREC>::T agg, int shift) {
    if constexpr (PREC == kFP32) return FX{(uint64_t)__float_as_uint(agg), 0};
    else return fx_from_double(Acc1<PREC>::value(agg), shift);
}

}  // namespace

// ---- single pass with decoupled look-back (small n) ------------------------

template <int PREC, bool VEC>
__global__ __launch_bounds__(kBlock) void k_scan(const float* __restrict__ w, uint32_t n,
                                                 int normalise, Params* __restrict__ prm,
                                                 uint64_t* __restrict__ states,
                                                 float* __restrict__ hi_out,
                                                 float* __restrict__ lo_out) {
    __shared__ ScanShared<PREC> s;
    __shared__ uint32_t s_tile;
    __shared__ double s_carry;
    __shared__ float s_carry_f;
    using R = typename Acc1<PREC>::T;

    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    if (prm->error) return;  // the total's wait gave up
    if (t == 0) s_tile = atomicAdd(&prm->ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const int shift = prm->shift;
    const double scale = normalise ? prm->inv_total : 1.0;
    const bool full = VEC && m == (uint32_t)kTile;

    float v[kItems];
    R thr_excl, tile_agg;
    scan_tile<PREC>(w, base, m, s, v, thr_excl, tile_agg);

    uint64_t* rec = states + (uint64_t)tile * kStateWords;
    if (wid == 0) {
        const FX agg = agg_to_carry<PREC>(tile_agg, shift);
        FX excl = fx_zero();
        if (tile == 0) {
            if (lane == 0) publish_inclusive<FX, ScanOps<PREC>>(rec, agg);
        } else {
            if (lane == 0) publish_aggregate<FX, ScanOps<PREC>>(rec, agg);
            excl = lookback<FX, ScanOps<PREC>>(states, tile, prm->wait_polls, &prm->error);
            if (lane == 0) publish_inclusive<FX, ScanOps<PREC>>(rec, sc_add<PREC>(excl, agg));
        }
        if (lane == 0) {
            if constexpr (PREC == kFP32) s_carry_f = __uint_as_float((uint32_t)excl.lo);
            else s_carry = fx_to_double(excl, shift);
        }
    }
    __syncthreads();
    scan_tile_store<PREC>(base, m, full, v, thr_excl, s_carry, s_carry_f, scale, s, hi_out, lo_out);
}

// ---- reduce-then-scan (large n) ---------------------------------------------
//
// A single-pass chained scan resolves about one look-back round trip (a
// device-coherent access through the L2 of another XCD) per 64 tiles, which
// at millions of elements costs more than one extra streaming read of the
// input.  The large form therefore runs three kernels:
//   k_scan_tilesum  tile sums (dense, one word per tile); also resets the
//                   spine's look-back records and ticket;
//   k_scan_spine    exclusive prefixes of the tile sums, 1024 tiles per block,
//                   look-back between blocks; the last block writes the total;
//   k_scan_emit     re-reads each tile, adds its carry, writes the prefixes.
// In the FP64 modes the spine adds the tile sums as 315-bit fixed-point words
// (five 63-bit limbs) at the fixed scale 2^150.  Every tile sum is a multiple
// of 2^-149 (a float is, and so is a sum of floats held in a double: a double
// below 2^-96 that is such a sum is exact, and above it the double's own unit
// is at least 2^-148) and below 2^141, and a scan has at most 2^20 tiles, so
// every carry is the exact sum of the tile sums before it, whatever order the
// look-back combines them in; no pass has to find the scale first.
//
// Layout in the state buffer (t = tiles): tile sums at [0, t), exclusive
// carries (as doubles, FP32 mode: float bits) at [t, 2t), spine records of
// kStateWords words at [16t, 16t + 16 * spine blocks).

constexpr int kSpineThreads = 256;
constexpr int kSpinePer = 4;
static_assert(kSpineThreads * kSpinePer == (int)kSpineTiles, "spine block shape");
constexpr int kLimbs = 5;
constexpr int kFxBias = 150;  // limb word k holds bits [63k, 63k + 63) of x * 2^150

struct FX5 {
    uint64_t w[kLimbs];
};
__device__ __forceinline__ FX5 fx5_zero() {
    FX5 r;
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) r.w[k] = 0;
    return r;
}
// exact for a non-negative multiple of 2^-150 below 2^165
__device__ __forceinline__ FX5 fx5_from_double(double d) {
    FX5 r = fx5_zero();
    if (!(d > 0.0)) return r;
    const uint64_t bits = (uint64_t)__double_as_longlong(d);
    const int e = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1ull);
    int p;
    if (e == 0) {
        p = -1074 + kFxBias;
    } else {
        m |= 1ull << 52;
        p = e - 1075 + kFxBias;
    }
    if (p < 0) {  // bits below 2^-150 (none for the sums of floats this file forms)
        m = -p >= 64 ? 0 : m >> (-p);
        p = 0;
    }
    const int k = p / 63;
    const unsigned __int128 v = (unsigned __int128)m << (p - 63 * k);
    const uint64_t lo = (uint64_t)v & kMask63, hi = (uint64_t)(v >> 63);
#pragma unroll
    for (int j = 0; j < kLimbs; ++j) r.w[j] = j == k ? lo : (j == k + 1 ? hi : 0);
    return r;
}
__device__ __forceinline__ FX5 fx5_add(const FX5& a, const FX5& b) {
    FX5 r;
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) {
        const uint64_t x = a.w[k] + b.w[k] + c;
        r.w[k] = x & kMask63;
        c = x >> 63;
    }
    return r;
}
// value as a double (Horner from the top limb; a fixed function of the limbs)
__device__ __forceinline__ double fx5_to_double(const FX5& v) {
    double r = (double)v.w[kLimbs - 1];
#pragma unroll
    for (int k = kLimbs - 2; k >= 0; --k) r = r * 0x1.0p63 + (double)v.w[k];
    return ldexp(r, -kFxBias);
}

// Spine carry: FX5 in the FP64 modes, an FP32 sum (bits in limb 0) in FP32 mode.
template <int PREC>
__device__ __forceinline__ FX5 sp_add(const FX5& a, const FX5& b) {
    if constexpr (PREC == kFP32) {
        FX5 r = fx5_zero();
        r.w[0] = (uint64_t)__float_as_uint(__uint_as_float((uint32_t)a.w[0]) +
                                           __uint_as_float((uint32_t)b.w[0]));
        return r;
    } else {
        return fx5_add(a, b);
    }
}
__device__ __forceinline__ FX5 fx5_shfl_up(const FX5& v, int d) {
    FX5 r;
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) r.w[k] = __shfl_up(v.w[k], (unsigned)d, kWave);
    return r;
}
__device__ __forceinline__ FX5 fx5_shfl_xor(const FX5& v, int d) {
    FX5 r;
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) r.w[k] = __shfl_xor(v.w[k], d, kWave);
    return r;
}
template <int PREC>
struct SpineOps {
    static __device__ __forceinline__ FX5 zero() { return fx5_zero(); }
    static __device__ __forceinline__ FX5 add(const FX5& a, const FX5& b) { return sp_add<PREC>(a, b); }
    static __device__ __forceinline__ FX5 wave_sum(FX5 v) {
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) v = sp_add<PREC>(v, fx5_shfl_xor(v, d));
        return v;
    }
    static __device__ __forceinline__ bool read(const uint64_t* r, FX5& v) {
        uint64_t any = 0;
#pragma unroll
        for (int k = 0; k < kLimbs; ++k) {
            v.w[k] = ld_word(r + k);
            any |= v.w[k];
        }
        return (any >> 63) == 0;
    }
    static __device__ __forceinline__ void publish(uint64_t* r, const FX5& v) {
#pragma unroll
        for (int k = 0; k < kLimbs; ++k) st_word(r + k, v.w[k]);
    }
};
static_assert(kLimbs <= 7, "a look-back record half holds 7 payload words");

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_scan_tilesum(const float* __restrict__ w, uint32_t n,
                                                         Params* __restrict__ prm,
                                                         uint64_t* __restrict__ states) {
    __shared__ ScanShared<PREC> s;
    using R = typename Acc1<PREC>::T;
    const uint32_t tile = blockIdx.x;
    const uint32_t tiles = gridDim.x;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    // the spine kernel that follows uses a ticket and sentinel-reset records
    const uint32_t rec_words = ((tiles + kSpineTiles - 1) / kSpineTiles) * (uint32_t)kStateWords;
    uint64_t* recs = states + (uint64_t)tiles * kStateWords;
    for (uint32_t i = tile * kBlock + threadIdx.x; i < rec_words; i += tiles * kBlock)
        recs[i] = kSentinel;
    if (tile == 0 && threadIdx.x == 0) {  // a fresh build for the kernels that follow
        prm->ticket = 0;
        prm->error = 0;
        prm->wait_polls = kDefaultWaitPolls;
        prm->hooks = 0;
    }
    float v[kItems];
    R thr_excl, tile_agg;
    scan_tile<PREC>(w, base, m, s, v, thr_excl, tile_agg);
    if (threadIdx.x == 0) {
        if constexpr (PREC == kFP32) states[tile] = (uint64_t)__float_as_uint(tile_agg);
        else states[tile] = (uint64_t)__double_as_longlong(Acc1<PREC>::value(tile_agg));
    }
}

template <int PREC>
__global__ __launch_bounds__(kSpineThreads) void k_scan_spine(uint64_t* __restrict__ states,
                                                              uint32_t tiles, uint32_t n,
                                                              Params* __restrict__ prm) {
    __shared__ FX5 s_wave[kSpineThreads / kWave];
    __shared__ FX5 s_excl;
    __shared__ uint32_t s_blk;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    if (t == 0) s_blk = atomicAdd(&prm->ticket, 1u);  // predecessors are scheduled first
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t nblk = (tiles + kSpineTiles - 1) / kSpineTiles;
    const uint64_t* sums = states;
    uint64_t* excl_out = states + tiles;
    uint64_t* recs = states + (uint64_t)tiles * kStateWords;

    const uint32_t b = blk * kSpineTiles + (uint32_t)t * kSpinePer;
    FX5 a[kSpinePer];
#pragma unroll
    for (int k = 0; k < kSpinePer; ++k) {
        a[k] = fx5_zero();
        if (b + k < tiles) {
            const uint64_t x = sums[b + k];
            if constexpr (PREC == kFP32) a[k].w[0] = x;
            else a[k] = fx5_from_double(__longlong_as_double((long long)x));
        }
    }
    FX5 mine = a[0];
#pragma unroll
    for (int k = 1; k < kSpinePer; ++k) mine = sp_add<PREC>(mine, a[k]);
    FX5 incl = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const FX5 o = fx5_shfl_up(incl, d);
        if (lane >= d) incl = sp_add<PREC>(o, incl);
    }
    FX5 lane_excl = fx5_shfl_up(incl, 1);
    if (lane == 0) lane_excl = fx5_zero();
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        FX5 blk_agg = s_wave[0];
#pragma unroll
        for (int q = 1; q < kSpineThreads / kWave; ++q) blk_agg = sp_add<PREC>(blk_agg, s_wave[q]);
        uint64_t* rec = recs + (uint64_t)blk * kStateWords;
        FX5 ex = fx5_zero();
        if (blk == 0) {
            if (lane == 0) publish_inclusive<FX5, SpineOps<PREC>>(rec, blk_agg);
        } else {
            if (lane == 0) publish_aggregate<FX5, SpineOps<PREC>>(rec, blk_agg);
            ex = lookback<FX5, SpineOps<PREC>>(recs, blk, prm->wait_polls, &prm->error);
            if (lane == 0) publish_inclusive<FX5, SpineOps<PREC>>(rec, sp_add<PREC>(ex, blk_agg));
        }
        if (lane == 0) {
            s_excl = ex;
            if (blk == nblk - 1) {
                const FX5 tot = sp_add<PREC>(ex, blk_agg);
                if constexpr (PREC == kFP32) {
                    const float total = __uint_as_float((uint32_t)tot.w[0]);
                    prm->total = (double)total;
                    prm->mean = (double)(total / (float)n);
                    prm->inv_total = total > 0.0f ? (double)(1.0f / total) : 0.0;
                } else {
                    const double total = fx5_to_double(tot);
                    prm->total = total;
                    prm->mean = total / (double)n;
                    prm->inv_total = total > 0.0 ? 1.0 / total : 0.0;
                }
            }
        }
    }
    __syncthreads();
    FX5 run = s_excl;
    for (int q = 0; q < wid; ++q) run = sp_add<PREC>(run, s_wave[q]);
    run = sp_add<PREC>(run, lane_excl);
#pragma unroll
    for (int k = 0; k < kSpinePer; ++k) {
        if (b + k < tiles) {
            if constexpr (PREC == kFP32)
                excl_out[b + k] = run.w[0];
            else
                excl_out[b + k] = (uint64_t)__double_as_longlong(fx5_to_double(run));
        }
        run = sp_add<PREC>(run, a[k]);
    }
}

template <int PREC, bool VEC>
__global__ __launch_bounds__(kBlock) void k_scan_emit(const float* __restrict__ w, uint32_t n,
                                                      int normalise,
                                                      const Params* __restrict__ prm,
                                                      const uint64_t* __restrict__ states,
                                                      float* __restrict__ hi_out,
                                                      float* __restrict__ lo_out) {
    __shared__ ScanShared<PREC> s;
    using R = typename Acc1<PREC>::T;
    if (prm->error) return;  // the spine's look-back gave up
    const uint32_t tile = blockIdx.x;
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);  // n + kTile - 1 may exceed 2^32
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const bool full = VEC && m == (uint32_t)kTile;
    const double scale = normalise ? prm->inv_total : 1.0;
    const uint64_t cw = states[tiles + tile];
    const double c = PREC == kFP32 ? 0.0 : __longlong_as_double((long long)cw);
    const float cf = PREC == kFP32 ? __uint_as_float((uint32_t)cw) : 0.0f;
    float v[kItems];
    R thr_excl, tile_agg;
    scan_tile<PREC>(w, base, m, s, v, thr_excl, tile_agg);
    scan_tile_store<PREC>(base, m, full, v, thr_excl, c, cf, scale, s, hi_out, lo_out);
}

namespace {
template <int P, bool V>
void launch_rts_t(const float* w, uint32_t n, int normalise, Params* prm, uint64_t* states,
                  float* hi, float* lo, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    const uint32_t nblk = (tiles + kSpineTiles - 1) / kSpineTiles;
    k_scan_tilesum<P><<<tiles, kBlock, 0, s>>>(w, n, prm, states);
    k_scan_spine<P><<<nblk, kSpineThreads, 0, s>>>(states, tiles, n, prm);
    k_scan_emit<P, V><<<tiles, kBlock, 0, s>>>(w, n, normalise, prm, states, hi, lo);
}
template <int P>
void launch_rts_p(bool vec, const float* w, uint32_t n, int normalise, Params* prm,
                  uint64_t* states, float* hi, float* lo, hipStream_t s) {
    if (vec) launch_rts_t<P, true>(w, n, normalise, prm, states, hi, lo, s);
    else launch_rts_t<P, false>(w, n, normalise, prm, states, hi, lo, s);
}
}  // namespace

hipError_t launch_scan_rts(const float* w, uint64_t n, int prec, int normalise, Params* prm,
                           uint64_t* states, float* hi, float* lo, hipStream_t s) {
    // VEC: 16-byte aligned outputs (stored as float4 rows); inputs load as rows at any alignment
    const bool vec = ((reinterpret_cast<uintptr_t>(hi) |
                       (lo ? reinterpret_cast<uintptr_t>(lo) : 0)) & 15u) == 0;
    const uint32_t nn = (uint32_t)n;
    switch (prec) {
        case kFP32: launch_rts_p<kFP32>(vec, w, nn, normalise, prm, states, hi, lo, s); break;
        case kFP64Carry: launch_rts_p<kFP64Carry>(vec, w, nn, normalise, prm, states, hi, lo, s); break;
        case kComp: launch_rts_p<kComp>(vec, w, nn, normalise, prm, states, hi, lo, s); break;
        default: launch_rts_p<kFP64>(vec, w, nn, normalise, prm, states, hi, lo, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_scan(const float* w, uint64_t n, int prec, int normalise, Params* prm,
                       uint64_t* states, float* hi, float* lo, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    // VEC: 16-byte aligned outputs (stored as float4 rows); inputs load as rows at any alignment
    const bool vec = ((reinterpret_cast<uintptr_t>(hi) |
                       (lo ? reinterpret_cast<uintptr_t>(lo) : 0)) & 15u) == 0;
    const uint32_t nn = (uint32_t)n;
#define WRS_LAUNCH_SCAN(P, V) \
    k_scan<P, V><<<tiles, kBlock, 0, s>>>(w, nn, normalise, prm, states, hi, lo)
    if (prec == kFP64) {
        if (vec) WRS_LAUNCH_SCAN(kFP64, true);
        else WRS_LAUNCH_SCAN(kFP64, false);
    } else if (prec == kComp) {
        if (vec) WRS_LAUNCH_SCAN(kComp, true);
        else WRS_LAUNCH_SCAN(kComp, false);
    } else if (prec == kFP64Carry) {
        if (vec) WRS_LAUNCH_SCAN(kFP64Carry, true);
        else WRS_LAUNCH_SCAN(kFP64Carry, false);
    } else {
        if (vec) WRS_LAUNCH_SCAN(kFP32, true);
        else WRS_LAUNCH_SCAN(kFP32, false);
    }
#undef WRS_LAUNCH_SCAN
    return hipGetLastError();
}

}  // namespace wrsk
