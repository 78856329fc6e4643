// device_common.hpp -- building blocks shared by the CDNA4 (gfx950) kernels.
//
//  * Philox4x32-10 (Random123, Salmon et al. SC'11), counter-based, so every
//    sample / element draws its own counter and the stream is independent of
//    the launch shape.
//  * error-free transformation two_sum and the double-float ("compensated
//    FP32") combine used by the compensated tile sums.
//  * a 126-bit fixed-point accumulator (two 63-bit limbs) for the inter-tile
//    look-back carry: integer addition is associative, so the carry bits do
//    not depend on the order in which the look-back happened to find its
//    predecessors' aggregates.
//  * self-validating 64-bit look-back words: every published word starts as
//    the sentinel ~0 and is written exactly once per launch with a value whose
//    top bit is clear, so a reader that sees all words of a record non-sentinel
//    has the final record without any release/acquire fence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wrsdev {

constexpr int kWave = 64;
constexpr uint64_t kSentinel = ~0ull;
constexpr uint64_t kMask63 = (1ull << 63) - 1ull;

// ------------------------------------------------------------- Philox ----
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

// counter = (lo(t), hi(t), c2, c3), key = (lo(seed), hi(seed))
__device__ __forceinline__ U4 philox_at(uint64_t seed, uint64_t t, uint32_t c2, uint32_t c3) {
    return philox4x32_10((uint32_t)t, (uint32_t)(t >> 32), c2, c3, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
}

constexpr uint32_t kTagGen = 0x47454E31u;  // weight generators
constexpr uint32_t kTagFeistel = 0x46535431u;
constexpr uint32_t kTagGroup = 0x47525031u;  // windowed samplers: one stream per group

__device__ __forceinline__ float u24(uint32_t x) { return (float)(x >> 8) * 0x1.0p-24f; }
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * 0x1.0p-53;
}
__device__ __forceinline__ double u53_open0(uint32_t hi, uint32_t lo) {
    return (double)(((((uint64_t)hi << 32) | lo) >> 11) + 1ull) * 0x1.0p-53;
}

// ------------------------------------------------- compensated FP32 ----
// two_sum: s + e == a + b exactly (Knuth, TAOCP vol. 2, 4.2.2).  Built with
// -ffp-contract=off and no reassociation, so the operation order below is what
// executes.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

struct DF {  // unevaluated sum hi + lo
    float hi, lo;
};

__device__ __forceinline__ DF df_add(DF a, DF b) {
    float s, e;
    two_sum(a.hi, b.hi, s, e);
    return DF{s, (a.lo + b.lo) + e};
}
__device__ __forceinline__ DF df_add1(DF a, float x) {
    float s, e;
    two_sum(a.hi, x, s, e);
    return DF{s, a.lo + e};
}
__device__ __forceinline__ double df_value(DF a) { return (double)a.hi + (double)a.lo; }

// ---------------------------------------------------- 126-bit fixed point ----
struct FX {  // value = hi * 2^63 + lo, lo < 2^63, hi < 2^63
    uint64_t lo, hi;
};
__device__ __forceinline__ FX fx_zero() { return FX{0, 0}; }
__device__ __forceinline__ FX fx_add(FX a, FX b) {
    const uint64_t l = a.lo + b.lo;
    return FX{l & kMask63, a.hi + b.hi + (l >> 63)};
}
// floor(d * 2^shift) for finite d >= 0 whose scaled value is < 2^126.
__device__ __forceinline__ FX fx_from_double(double d, int shift) {
    if (!(d > 0.0)) return fx_zero();
    const uint64_t bits = (uint64_t)__double_as_longlong(d);
    const int e = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1ull);
    int ex;
    if (e == 0) {
        ex = -1074;
    } else {
        m |= (1ull << 52);
        ex = e - 1075;
    }
    const int sh = ex + shift;
    if (sh >= 0) {
        const unsigned __int128 v = (unsigned __int128)m << sh;
        return FX{(uint64_t)v & kMask63, (uint64_t)(v >> 63)};
    }
    if (sh <= -64) return fx_zero();
    return FX{(m >> (-sh)) & kMask63, 0};
}
__device__ __forceinline__ double fx_to_double(FX v, int shift) {
    const double r = (double)v.hi * 0x1.0p63 + (double)v.lo;
    return ldexp(r, -shift);
}

// ------------------------------------------------------ agent-scope words ----
__device__ __forceinline__ void st_word(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_word(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------- streaming (NT) access ----
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt4(const float4* p) {
    const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_nt_u4(uint4* p, uint4 v) {
    u32x4_t x;
    x.x = v.x;
    x.y = v.y;
    x.z = v.z;
    x.w = v.w;
    __builtin_nontemporal_store(x, reinterpret_cast<u32x4_t*>(p));
}
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_nt2(const float2* p) {
    const f32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt4(float4* p, float4 v) {
    f32x4_t x;
    x.x = v.x;
    x.y = v.y;
    x.z = v.z;
    x.w = v.w;
    __builtin_nontemporal_store(x, reinterpret_cast<f32x4_t*>(p));
}

// Quantisation of the merge keys: x -> rint(x * 2^qs) as a 64-bit integer.
__device__ __forceinline__ int64_t merge_quant(double x, int qs) {
    return __double2ll_rn(ldexp(x, qs));
}

// ------------------------------------------------------------ wave ops ----
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// One DPP move (update_dpp with old = 0): lanes whose source is outside the
// row, or whose row is masked off, receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, true);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float v) {
    return __uint_as_float(dpp_u32<CTRL, ROW_MASK>(__float_as_uint(v)));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)v);
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// v + (the value of v DPP-moved by CTRL): both halves move, and lanes whose
// source is outside the row, or whose row is masked off, add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t add_dpp_u64(uint64_t v) {
    const uint32_t vl = (uint32_t)v, vh = (uint32_t)(v >> 32);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)vl, CTRL, ROW_MASK, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)vh, CTRL, ROW_MASK, 0xf, true);
    unsigned c0, c1;
    const uint32_t rl = __builtin_addc(vl, lo, 0u, &c0);
    const uint32_t rh = __builtin_addc(vh, hi, c0, &c1);
    return ((uint64_t)rh << 32) | rl;
}

// Inclusive scan of a 32-bit integer over the 64 lanes, by the DPP steps of
// wave_incl_scan_u64 below.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    v += dpp_u32<0x111, 0xf>(v);  // row_shr:1
    v += dpp_u32<0x112, 0xf>(v);  // row_shr:2
    v += dpp_u32<0x114, 0xf>(v);  // row_shr:4
    v += dpp_u32<0x118, 0xf>(v);  // row_shr:8
    v += dpp_u32<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_u32<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Inclusive scan of a 64-bit integer over the 64 lanes (modular addition, so
// the association does not matter): shifts by 1, 2, 4, 8 lanes within each
// row of 16, then the total of row 0 (2) added to row 1 (3) by row_bcast:15,
// then the total of rows 0-1 added to rows 2-3 by row_bcast:31.
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    v = add_dpp_u64<0x111, 0xf>(v);  // row_shr:1
    v = add_dpp_u64<0x112, 0xf>(v);  // row_shr:2
    v = add_dpp_u64<0x114, 0xf>(v);  // row_shr:4
    v = add_dpp_u64<0x118, 0xf>(v);  // row_shr:8
    v = add_dpp_u64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = add_dpp_u64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

template <class T>
__device__ __forceinline__ T shfl_up(T v, int d) {
    return __shfl_up(v, (unsigned)d, kWave);
}
template <class T>
__device__ __forceinline__ T shfl(T v, int src) {
    return __shfl(v, src, kWave);
}

// Inclusive wave scan, Hillis-Steele over 64 lanes: lane l receives
// v_0 (+) ... (+) v_l with the earlier lanes always the left operand, so the
// result is a fixed function of the lane values.
__device__ __forceinline__ float wave_incl_f(float v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float n = shfl_up(v, d);
        if (l >= d) v = n + v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_incl_u(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t n = shfl_up(v, d);
        if (l >= d) v = n + v;
    }
    return v;
}

// Fixed-order wave sum (butterfly) of a double: identical bits on all lanes.
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v = v + __shfl_xor(v, d, kWave);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v = v + __shfl_xor(v, d, kWave);
    return v;
}
__device__ __forceinline__ FX wave_sum_fx(FX v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        FX o;
        o.lo = __shfl_xor(v.lo, d, kWave);
        o.hi = __shfl_xor(v.hi, d, kWave);
        v = fx_add(v, o);
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Largest float f with (double)f <= m (round toward -inf): for float w,
// (double)w > m  <=>  w > pivot_rd(m).
__device__ __forceinline__ float pivot_rd(double m) {
    float f = (float)m;
    if ((double)f > m) f = nextafterf(f, -__builtin_huge_valf());
    return f;
}

}  // namespace wrsdev
