// partition.hip -- single-pass heavy/light prefix-partition with decoupled
// look-back (pass 2 of a PSA build).
//
// Semantics follow the reference kernel src/device/prefix_partition/decoupled/
// shader.comp:313-362 and its CPU counterpart src/host/reference/partition.hpp:35-53:
//   heavy <=> w > pivot (strict), stable order;
//   idx[0..H)        heavy element indices, ascending;
//   idx[N-1-r]       the r-th light element index (lights stored reversed);
//   prefix[same pos] inclusive prefix of the heavy (resp. light) stream;
//   heavy_count = H.
//
// MI355X design (one 256-thread block = 4 wave64s per 4096-element tile):
//   1. dynamic tile ticket (atomicAdd) so that every predecessor of a tile has
//      been scheduled before it spins on it;
//   2. coalesced float4 loads, transposed through LDS (row stride 17 words, so
//      both the transpose writes and the per-thread reads avoid bank
//      conflicts) so that each thread owns 16 consecutive elements;
//   3. per-thread sequential (compensated) sums, wave64 shuffle scan, then a
//      4-entry cross-wave scan in LDS, all in a fixed order;
//   4. the tile aggregate is published as self-validating 64-bit words in HBM
//      (126-bit fixed point for the two weight streams, exact integer count);
//      wave 0 looks back up to 64 predecessors per step (one per lane);
//   5. outputs are staged in LDS in final order (heavies by rank, lights by
//      reversed rank) and written as two contiguous coalesced runs.
#include "device_common.hpp"
#include "internal.hpp"
#include "lookback.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

constexpr int kPad = kTile + kTile / 16;  // LDS words of the padded transpose buffer

// Look-back record payload (lookback.hpp): heavy sum (2 words), light sum
// (2 words), heavy count; aggregate at words 0..4, inclusive prefix at 8..12.
// FP32 mode keeps each sum in one float (word .lo; .hi = 0).
struct Carry {
    FX h, l;
    uint64_t c;
};

__device__ __forceinline__ Carry carry_zero() { return Carry{fx_zero(), fx_zero(), 0}; }

template <int PREC>
__device__ __forceinline__ Carry carry_add(Carry a, Carry b) {
    Carry r;
    if constexpr (PREC == kFP32) {
        r.h.lo = (uint64_t)__float_as_uint(__uint_as_float((uint32_t)a.h.lo) +
                                           __uint_as_float((uint32_t)b.h.lo));
        r.l.lo = (uint64_t)__float_as_uint(__uint_as_float((uint32_t)a.l.lo) +
                                           __uint_as_float((uint32_t)b.l.lo));
        r.h.hi = r.l.hi = 0;
    } else {
        r.h = fx_add(a.h, b.h);
        r.l = fx_add(a.l, b.l);
    }
    r.c = a.c + b.c;
    return r;
}

template <int PREC>
__device__ __forceinline__ Carry wave_sum_carry(Carry v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        Carry o;
        o.h.lo = __shfl_xor(v.h.lo, d, kWave);
        o.h.hi = __shfl_xor(v.h.hi, d, kWave);
        o.l.lo = __shfl_xor(v.l.lo, d, kWave);
        o.l.hi = __shfl_xor(v.l.hi, d, kWave);
        o.c = __shfl_xor(v.c, d, kWave);
        v = carry_add<PREC>(v, o);
    }
    return v;
}

__device__ __forceinline__ void publish_rec(uint64_t* rec, const Carry& v) {
    st_word(rec + 0, v.h.lo);
    st_word(rec + 1, v.h.hi);
    st_word(rec + 2, v.l.lo);
    st_word(rec + 3, v.l.hi);
    st_word(rec + 4, v.c);
}

// Returns true if all five words are published.
__device__ __forceinline__ bool read_rec(const uint64_t* rec, Carry& v) {
    v.h.lo = ld_word(rec + 0);
    v.h.hi = ld_word(rec + 1);
    v.l.lo = ld_word(rec + 2);
    v.l.hi = ld_word(rec + 3);
    v.c = ld_word(rec + 4);
    return ((v.h.lo | v.h.hi | v.l.lo | v.l.hi | v.c) >> 63) == 0;
}

template <int PREC>
struct CarryOps {
    static __device__ __forceinline__ Carry zero() { return carry_zero(); }
    static __device__ __forceinline__ Carry add(Carry a, Carry b) { return carry_add<PREC>(a, b); }
    static __device__ __forceinline__ Carry wave_sum(Carry v) { return wave_sum_carry<PREC>(v); }
    static __device__ __forceinline__ bool read(const uint64_t* r, Carry& v) { return read_rec(r, v); }
    static __device__ __forceinline__ void publish(uint64_t* r, const Carry& v) { publish_rec(r, v); }
};

// In-tile per-thread accumulator: compensated (DF) or plain float.
template <int PREC>
struct Acc {
    DF h, l;
    uint32_t c;
};

template <int PREC>
__device__ __forceinline__ Acc<PREC> acc_add(Acc<PREC> a, Acc<PREC> b) {
    Acc<PREC> r;
    if constexpr (PREC == kComp) {
        r.h = df_add(a.h, b.h);
        r.l = df_add(a.l, b.l);
    } else {
        r.h = DF{a.h.hi + b.h.hi, 0.0f};
        r.l = DF{a.l.hi + b.l.hi, 0.0f};
    }
    r.c = a.c + b.c;
    return r;
}

template <int PREC>
__device__ __forceinline__ Acc<PREC> acc_shfl_up(Acc<PREC> v, int d) {
    Acc<PREC> r;
    r.h.hi = shfl_up(v.h.hi, d);
    r.l.hi = shfl_up(v.l.hi, d);
    if constexpr (PREC == kComp) {
        r.h.lo = shfl_up(v.h.lo, d);
        r.l.lo = shfl_up(v.l.lo, d);
    } else {
        r.h.lo = 0.0f;
        r.l.lo = 0.0f;
    }
    r.c = shfl_up(v.c, d);
    return r;
}

template <int PREC>
__device__ __forceinline__ double stream_value(DF v) {
    if constexpr (PREC == kComp) return df_value(v);
    else return (double)v.hi;
}

}  // namespace

// One 4096-element tile of the distribution w[0..n): `tile` is its index,
// `states` the look-back records of tile 0 onward, prm its build scalars.
// Used by the single-distribution kernel and by the segmented (batched) one.
template <int PREC, bool VEC>
__device__ __forceinline__ void partition_tile(const float* __restrict__ w, uint32_t n,
                                               uint32_t tile, Params* __restrict__ prm,
                                               uint64_t* __restrict__ states,
                                               uint32_t* __restrict__ idx_out,
                                               float* __restrict__ hi_out,
                                               float* __restrict__ lo_out,
                                               double* __restrict__ samp_h,
                                               double* __restrict__ samp_l) {
    __shared__ float s_buf[kPad];
    __shared__ Acc<PREC> s_wave[kBlock / kWave];
    __shared__ double s_carry_h, s_carry_l, s_tot_h, s_tot_l;
    __shared__ float s_carry_hf, s_carry_lf;
    __shared__ uint32_t s_carry_c;

    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const float pivot = prm->pivot;
    const int shift = prm->shift;

    // ---- 2. load + transpose -------------------------------------------
    if (VEC && m == (uint32_t)kTile) {
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

    // ---- 3. in-tile scan -----------------------------------------------
    Acc<PREC> mine{DF{0.0f, 0.0f}, DF{0.0f, 0.0f}, 0u};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (e0 + k < m) {
            const bool heavy = v[k] > pivot;
            if constexpr (PREC == kComp) {
                if (heavy) mine.h = df_add1(mine.h, v[k]);
                else mine.l = df_add1(mine.l, v[k]);
            } else {
                if (heavy) mine.h.hi += v[k];
                else mine.l.hi += v[k];
            }
            mine.c += heavy ? 1u : 0u;
        }
    }
    Acc<PREC> incl = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const Acc<PREC> o = acc_shfl_up<PREC>(incl, d);
        if (lane >= d) incl = acc_add<PREC>(o, incl);
    }
    Acc<PREC> lane_excl = acc_shfl_up<PREC>(incl, 1);
    if (lane == 0) lane_excl = Acc<PREC>{DF{0.0f, 0.0f}, DF{0.0f, 0.0f}, 0u};
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    Acc<PREC> wave_excl{DF{0.0f, 0.0f}, DF{0.0f, 0.0f}, 0u};
    for (int q = 0; q < wid; ++q) wave_excl = acc_add<PREC>(wave_excl, s_wave[q]);
    Acc<PREC> tile_agg = s_wave[0];
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q) tile_agg = acc_add<PREC>(tile_agg, s_wave[q]);
    const Acc<PREC> thr_excl = acc_add<PREC>(wave_excl, lane_excl);

    // ---- 4. publish + decoupled look-back ------------------------------
    uint64_t* rec = states + (uint64_t)tile * kStateWords;
    if (wid == 0) {
        Carry agg;
        if constexpr (PREC == kFP32) {
            agg.h = FX{(uint64_t)__float_as_uint(tile_agg.h.hi), 0};
            agg.l = FX{(uint64_t)__float_as_uint(tile_agg.l.hi), 0};
        } else {
            agg.h = fx_from_double(stream_value<PREC>(tile_agg.h), shift);
            agg.l = fx_from_double(stream_value<PREC>(tile_agg.l), shift);
        }
        agg.c = tile_agg.c;
        if (tile == 0) {
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, agg);
        } else {
            if (lane == 0) publish_aggregate<Carry, CarryOps<PREC>>(rec, agg);
            const Carry excl = lookback<Carry, CarryOps<PREC>>(states, tile);
            if (lane == 0) {
                const Carry inc = carry_add<PREC>(excl, agg);
                publish_inclusive<Carry, CarryOps<PREC>>(rec, inc);
                if constexpr (PREC == kFP32) {
                    s_carry_hf = __uint_as_float((uint32_t)excl.h.lo);
                    s_carry_lf = __uint_as_float((uint32_t)excl.l.lo);
                    s_tot_h = (double)__uint_as_float((uint32_t)inc.h.lo);
                    s_tot_l = (double)__uint_as_float((uint32_t)inc.l.lo);
                } else {
                    s_carry_h = fx_to_double(excl.h, shift);
                    s_carry_l = fx_to_double(excl.l, shift);
                    s_tot_h = fx_to_double(inc.h, shift);
                    s_tot_l = fx_to_double(inc.l, shift);
                }
                s_carry_c = (uint32_t)excl.c;
            }
        }
        if (tile == 0 && lane == 0) {
            s_carry_h = 0.0;
            s_carry_l = 0.0;
            s_carry_hf = 0.0f;
            s_carry_lf = 0.0f;
            s_carry_c = 0;
            if constexpr (PREC == kFP32) {
                s_tot_h = (double)__uint_as_float((uint32_t)agg.h.lo);
                s_tot_l = (double)__uint_as_float((uint32_t)agg.l.lo);
            } else {
                s_tot_h = fx_to_double(agg.h, shift);
                s_tot_l = fx_to_double(agg.l, shift);
            }
        }
    }
    __syncthreads();

    // ---- 5. outputs, staged in LDS in final order -----------------------
    const uint32_t Hx = s_carry_c;           // heavies before this tile
    const uint32_t Lx = base - Hx;           // lights before this tile
    const uint32_t hc = tile_agg.c;          // heavies in this tile
    uint32_t* s_u = reinterpret_cast<uint32_t*>(s_buf);

    uint32_t slot[kItems];
    {
        uint32_t hr = thr_excl.c;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t e = e0 + k;
            const bool heavy = v[k] > pivot;
            slot[k] = heavy ? hr : (m - 1u - (e - hr));
            hr += heavy ? 1u : 0u;
        }
    }
    // positions of LDS slot s: heavy run then light run
    auto out_pos = [&](uint32_t s) -> uint32_t { return s < hc ? Hx + s : n - m - Lx + s; };

    // round A: indices
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (e0 + k < m) s_u[slot[k]] = base + e0 + k;
    __syncthreads();
    for (uint32_t s = t; s < m; s += kBlock) idx_out[out_pos(s)] = s_u[s];
    __syncthreads();

    // round B: the inclusive prefix of every element (FP64 in the FP64 modes).
    // Ranks r with (r+1) % kSample == 0 publish their prefix as a sample; the
    // full hi (and lo) prefix arrays are written only when requested.
    float hi_keep[kItems], lo_keep[kItems];
    {
        DF rh = thr_excl.h, rl = thr_excl.l;
        const double ch = s_carry_h, cl = s_carry_l;
        const float chf = s_carry_hf, clf = s_carry_lf;
        uint32_t hr = Hx + thr_excl.c;                    // global heavy rank of the next heavy
        uint32_t lr = Lx + (e0 - thr_excl.c);             // global light rank of the next light
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const bool heavy = v[k] > pivot;
            double P;
            float hi, lo = 0.0f;
            if constexpr (PREC == kComp) {
                if (heavy) rh = df_add1(rh, v[k]);
                else rl = df_add1(rl, v[k]);
                P = (heavy ? ch : cl) + df_value(heavy ? rh : rl);
                hi = (float)P;
                lo = (float)(P - (double)hi);
            } else if constexpr (PREC == kFP64Carry) {
                if (heavy) rh.hi += v[k];
                else rl.hi += v[k];
                P = (heavy ? ch : cl) + (double)(heavy ? rh.hi : rl.hi);
                hi = (float)P;
            } else {
                if (heavy) rh.hi += v[k];
                else rl.hi += v[k];
                hi = (heavy ? chf : clf) + (heavy ? rh.hi : rl.hi);
                P = (double)hi;
            }
            hi_keep[k] = hi;
            lo_keep[k] = lo;
            if (e0 + k < m && samp_h != nullptr) {
                const uint32_t r1 = (heavy ? hr : lr) + 1u;  // prefix covers ranks < r1
                if ((r1 & (kSample - 1)) == 0) (heavy ? samp_h : samp_l)[r1 / kSample - 1u] = P;
            }
            hr += heavy ? 1u : 0u;
            lr += heavy ? 0u : 1u;
        }
    }
    if (hi_out != nullptr) {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (e0 + k < m) s_buf[slot[k]] = hi_keep[k];
        __syncthreads();
        for (uint32_t s = t; s < m; s += kBlock) hi_out[out_pos(s)] = s_buf[s];
        if constexpr (PREC == kComp) {
            if (lo_out != nullptr) {
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kItems; ++k)
                    if (e0 + k < m) s_buf[slot[k]] = lo_keep[k];
                __syncthreads();
                for (uint32_t s = t; s < m; s += kBlock) lo_out[out_pos(s)] = s_buf[s];
            }
        }
    }
    if (t == 0 && base + m == n) {
        prm->heavy_count = Hx + hc;
        prm->heavy_sum = s_tot_h;
        prm->light_sum = s_tot_l;
    }
}

template <int PREC, bool VEC>
__global__ __launch_bounds__(kBlock) void k_partition(const float* __restrict__ w, uint32_t n,
                                                      Params* __restrict__ prm,
                                                      uint64_t* __restrict__ states,
                                                      uint32_t* __restrict__ idx_out,
                                                      float* __restrict__ hi_out,
                                                      float* __restrict__ lo_out,
                                                      double* __restrict__ samp_h,
                                                      double* __restrict__ samp_l) {
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(&prm->ticket, 1u);
    __syncthreads();
    partition_tile<PREC, VEC>(w, n, s_tile, prm, states, idx_out, hi_out, lo_out, samp_h, samp_l);
}

// Segmented form: one work item per (segment, tile), items of a segment are
// consecutive and handed out in ticket order, so a tile's predecessors in its
// segment always hold smaller tickets.  Segment outputs use segment-local
// indices and occupy the segment's own range of idx/hi/lo.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_seg_partition(
    const float* __restrict__ w, const uint64_t* __restrict__ seg_off,
    const uint32_t* __restrict__ item_seg, const uint32_t* __restrict__ item_tile,
    const uint32_t* __restrict__ seg_item0, Params* __restrict__ seg_prm,
    uint32_t* __restrict__ ticket, uint64_t* __restrict__ states, uint32_t* __restrict__ idx,
    float* __restrict__ hi, float* __restrict__ lo, double* __restrict__ samp_h,
    double* __restrict__ samp_l, const uint64_t* __restrict__ seg_samp0) {
    __shared__ uint32_t s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    const uint32_t sg = item_seg[item];
    const uint64_t off = seg_off[sg];
    const uint32_t n = (uint32_t)(seg_off[sg + 1] - off);
    const uint64_t so = seg_samp0[sg];
    partition_tile<PREC, false>(w + off, n, item_tile[item], seg_prm + sg,
                                states + (uint64_t)seg_item0[sg] * kStateWords, idx + off,
                                hi ? hi + off : nullptr, lo ? lo + off : nullptr, samp_h + so,
                                samp_l + so);
}

hipError_t launch_partition(const float* w, uint64_t n, int prec, Params* prm, uint64_t* states,
                            uint32_t* idx, float* hi, float* lo, double* samp_h, double* samp_l,
                            hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    const bool vec = (reinterpret_cast<uintptr_t>(w) & 15u) == 0;
    const uint32_t nn = (uint32_t)n;
#define WRS_LAUNCH_PART(P, V) \
    k_partition<P, V><<<tiles, kBlock, 0, s>>>(w, nn, prm, states, idx, hi, lo, samp_h, samp_l)
    if (prec == kComp) {
        if (vec) WRS_LAUNCH_PART(kComp, true);
        else WRS_LAUNCH_PART(kComp, false);
    } else if (prec == kFP64Carry) {
        if (vec) WRS_LAUNCH_PART(kFP64Carry, true);
        else WRS_LAUNCH_PART(kFP64Carry, false);
    } else {
        if (vec) WRS_LAUNCH_PART(kFP32, true);
        else WRS_LAUNCH_PART(kFP32, false);
    }
#undef WRS_LAUNCH_PART
    return hipGetLastError();
}


hipError_t launch_seg_partition(const float* w, const uint64_t* seg_off, const uint32_t* item_seg,
                                const uint32_t* item_tile, const uint32_t* seg_item0,
                                uint32_t items, int prec, Params* seg_prm, uint32_t* ticket,
                                uint64_t* states, uint32_t* idx, float* hi, float* lo,
                                double* samp_h, double* samp_l, const uint64_t* seg_samp0,
                                hipStream_t s) {
#define WRS_LAUNCH_SEG(P)                                                                 \
    k_seg_partition<P><<<items, kBlock, 0, s>>>(w, seg_off, item_seg, item_tile, seg_item0, \
                                                seg_prm, ticket, states, idx, hi, lo,       \
                                                samp_h, samp_l, seg_samp0)
    if (prec == kComp) WRS_LAUNCH_SEG(kComp);
    else if (prec == kFP64Carry) WRS_LAUNCH_SEG(kFP64Carry);
    else WRS_LAUNCH_SEG(kFP32);
#undef WRS_LAUNCH_SEG
    return hipGetLastError();
}

}  // namespace wrsk
