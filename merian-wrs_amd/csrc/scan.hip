// scan.hip -- single-pass inclusive prefix sum with decoupled look-back
// (the ITS build, and the standalone scan).
//
// Semantics: out[i] = w[0] + ... + w[i] (reference DecoupledPrefixSum,
// src/device/prefix_sum/decoupled/shader.comp:321-333, CPU form
// src/host/reference/prefix_sum.hpp:13-34); with `normalise` the prefix is
// divided by the FP64 total of the preceding reduce pass, giving a CDF.
//
// Same tile machinery as the prefix-partition: 4096-element tiles, dynamic
// ticket, LDS transpose to 16 consecutive elements per thread, wave64 scan,
// tail-safe loads and stores (the reference kernel has no tail check,
// shader.comp:70-92,181-197).  In-tile sums are FP64 in the default mode.  The carry is the 126-bit fixed-point word pair
// (exact, so its bits do not depend on look-back timing) in the FP64 modes and
// an FP32 sum in FP32 mode.  The output is staged back through LDS and written
// as coalesced float4 rows.
#include "device_common.hpp"
#include "internal.hpp"
#include "lookback.hpp"
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
}  // namespace

template <int PREC, bool VEC>
__global__ __launch_bounds__(kBlock) void k_scan(const float* __restrict__ w, uint32_t n,
                                                 int normalise, Params* __restrict__ prm,
                                                 uint64_t* __restrict__ states,
                                                 float* __restrict__ hi_out,
                                                 float* __restrict__ lo_out) {
    __shared__ float s_buf[kPad];
    __shared__ uint32_t s_tile;
    // in-tile accumulator in the mode's tile arithmetic (tilesum.hpp)
    using A1 = Acc1<PREC>;
    using R = typename A1::T;
    __shared__ R s_wave[kBlock / kWave];
    __shared__ double s_carry;
    __shared__ float s_carry_f;

    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    if (t == 0) s_tile = atomicAdd(&prm->ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const int shift = prm->shift;
    const double scale = normalise ? prm->inv_total : 1.0;

    const bool full = VEC && m == (uint32_t)kTile;
    if (full) {
        const float4* w4 = reinterpret_cast<const float4*>(w + base);
        float4 v4[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v4[j] = ld_nt4(w4 + t + kBlock * j);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = 4 * (t + kBlock * j);
            const int q = e + e / 16;
            s_buf[q + 0] = v4[j].x;
            s_buf[q + 1] = v4[j].y;
            s_buf[q + 2] = v4[j].z;
            s_buf[q + 3] = v4[j].w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint32_t e = (uint32_t)t + (uint32_t)kBlock * j;
            s_buf[e + e / 16] = e < m ? w[base + e] : 0.0f;
        }
    }
    __syncthreads();
    float v[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = s_buf[17 * t + k];
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
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    R wave_excl = A1::zero();
    for (int q = 0; q < wid; ++q) wave_excl = A1::add(wave_excl, s_wave[q]);
    R tile_agg = s_wave[0];
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q) tile_agg = A1::add(tile_agg, s_wave[q]);
    const R thr_excl = A1::add(wave_excl, lane_excl);

    uint64_t* rec = states + (uint64_t)tile * kStateWords;
    if (wid == 0) {
        FX agg;
        if constexpr (PREC == kFP32) agg = FX{(uint64_t)__float_as_uint(tile_agg), 0};
        else agg = fx_from_double(A1::value(tile_agg), shift);
        FX excl = fx_zero();
        if (tile == 0) {
            if (lane == 0) publish_inclusive<FX, ScanOps<PREC>>(rec, agg);
        } else {
            if (lane == 0) publish_aggregate<FX, ScanOps<PREC>>(rec, agg);
            excl = lookback<FX, ScanOps<PREC>>(states, tile);
            if (lane == 0) publish_inclusive<FX, ScanOps<PREC>>(rec, sc_add<PREC>(excl, agg));
        }
        if (lane == 0) {
            if constexpr (PREC == kFP32) s_carry_f = __uint_as_float((uint32_t)excl.lo);
            else s_carry = fx_to_double(excl, shift);
        }
    }
    __syncthreads();

    float lo_keep[kItems];
    {
        R run = thr_excl;
        const double c = s_carry;
        const float cf = s_carry_f;
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
            s_buf[17 * t + k] = hi;
        }
    }
    __syncthreads();
    if (full) {
        float4* o4 = reinterpret_cast<float4*>(hi_out + base);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = 4 * (t + kBlock * j);
            const int q = e + e / 16;
            o4[t + kBlock * j] = make_float4(s_buf[q], s_buf[q + 1], s_buf[q + 2], s_buf[q + 3]);
        }
    } else {
        for (uint32_t e = t; e < m; e += kBlock) hi_out[base + e] = s_buf[e + e / 16];
    }
    if (prec_has_lo(PREC) && lo_out != nullptr) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; ++k) s_buf[17 * t + k] = lo_keep[k];
        __syncthreads();
        if (full) {
            float4* o4 = reinterpret_cast<float4*>(lo_out + base);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = 4 * (t + kBlock * j);
                const int q = e + e / 16;
                o4[t + kBlock * j] = make_float4(s_buf[q], s_buf[q + 1], s_buf[q + 2], s_buf[q + 3]);
            }
        } else {
            for (uint32_t e = t; e < m; e += kBlock) lo_out[base + e] = s_buf[e + e / 16];
        }
    }
}

hipError_t launch_scan(const float* w, uint64_t n, int prec, int normalise, Params* prm,
                       uint64_t* states, float* hi, float* lo, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    const bool vec = ((reinterpret_cast<uintptr_t>(w) | reinterpret_cast<uintptr_t>(hi) |
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
