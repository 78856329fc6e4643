// tilesum.hpp -- the running-sum type of the in-tile scans, by build arithmetic.
//
//   kFP32, kFP64Carry  FP32 sums (the reference kernels' arithmetic)
//   kComp              compensated FP32: a double-float hi + lo kept exact by
//                      two_sum at every add (error ~2^-48 relative instead of
//                      2^-24 times the term count)
//   kFP64              FP64 sums: a float weight converts exactly; an FP64
//                      add holds its SIMD twice as long as an FP32 add on
//                      gfx950 (4 cycles a wave64 instruction against 2), but
//                      one of them replaces the seven FP32 operations of a
//                      compensated add
//
// All operations are in a fixed order, so a tile's sums are a function of its
// weights only.  pin() is an empty register barrier: it makes a value exist at
// that point of the program (an optimisation fence for the scheduler, no
// instruction).
#pragma once

#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {

template <int PREC>
struct Acc1 {
    using T = float;
    static __device__ __forceinline__ T zero() { return 0.0f; }
    static __device__ __forceinline__ T add1(T a, float x) { return a + x; }
    static __device__ __forceinline__ T add(T a, T b) { return a + b; }
    static __device__ __forceinline__ double value(T a) { return (double)a; }
    static __device__ __forceinline__ T shfl_up(T a, int d) { return wrsdev::shfl_up(a, d); }
    static __device__ __forceinline__ T sel(bool c, T a, T b) { return c ? a : b; }
    template <int CTRL, int RM>
    static __device__ __forceinline__ T dpp(T a) { return wrsdev::dpp_f32<CTRL, RM>(a); }
    static __device__ __forceinline__ void pin(T& a) { asm volatile("" : "+v"(a)); }
};

template <>
struct Acc1<kComp> {
    using T = wrsdev::DF;
    static __device__ __forceinline__ T zero() { return T{0.0f, 0.0f}; }
    static __device__ __forceinline__ T add1(T a, float x) { return wrsdev::df_add1(a, x); }
    static __device__ __forceinline__ T add(T a, T b) { return wrsdev::df_add(a, b); }
    static __device__ __forceinline__ double value(T a) { return wrsdev::df_value(a); }
    static __device__ __forceinline__ T shfl_up(T a, int d) {
        return T{wrsdev::shfl_up(a.hi, d), wrsdev::shfl_up(a.lo, d)};
    }
    static __device__ __forceinline__ T sel(bool c, T a, T b) {
        return T{c ? a.hi : b.hi, c ? a.lo : b.lo};
    }
    template <int CTRL, int RM>
    static __device__ __forceinline__ T dpp(T a) {
        return T{wrsdev::dpp_f32<CTRL, RM>(a.hi), wrsdev::dpp_f32<CTRL, RM>(a.lo)};
    }
    static __device__ __forceinline__ void pin(T& a) { asm volatile("" : "+v"(a.hi), "+v"(a.lo)); }
};

template <>
struct Acc1<kFP64> {
    using T = double;
    static __device__ __forceinline__ T zero() { return 0.0; }
    static __device__ __forceinline__ T add1(T a, float x) { return a + (double)x; }
    static __device__ __forceinline__ T add(T a, T b) { return a + b; }
    static __device__ __forceinline__ double value(T a) { return a; }
    static __device__ __forceinline__ T shfl_up(T a, int d) { return wrsdev::shfl_up(a, d); }
    static __device__ __forceinline__ T sel(bool c, T a, T b) { return c ? a : b; }
    template <int CTRL, int RM>
    static __device__ __forceinline__ T dpp(T a) {
        return __longlong_as_double(
            (long long)wrsdev::dpp_u64<CTRL, RM>((uint64_t)__double_as_longlong(a)));
    }
    static __device__ __forceinline__ void pin(T& a) { asm volatile("" : "+v"(a)); }
};

// Adds x to h if heavy, else to l, with one addition and selects (no branch;
// the same bits as adding to the chosen sum).
template <int PREC>
__device__ __forceinline__ void add_split(typename Acc1<PREC>::T& h, typename Acc1<PREC>::T& l,
                                          bool heavy, float x) {
    using A1 = Acc1<PREC>;
    const typename A1::T a = A1::add1(A1::sel(heavy, h, l), x);
    h = A1::sel(heavy, a, h);
    l = A1::sel(heavy, l, a);
}

}  // namespace wrsk
