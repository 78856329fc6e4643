// reduce.hip -- deterministic total weight (pass 1 of a build).
//
// Replaces the reference's AtomicMean (src/device/mean/atomic/shader.comp:69-82),
// whose float atomicAdd makes the pivot change from run to run.  Here every
// element is accumulated in FP64 (exact conversion of the float), in an order
// fixed by N alone: block b sums tiles b, b+G, b+2G, ...; thread t of a tile
// sums its four float4 chunks; waves reduce by a fixed butterfly and the
// finalize kernel sums the G partials in a fixed tree.  The same launch also
// resets the look-back records of the kernel that follows it.
#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {
using namespace wrsdev;

template <bool FP64, bool VEC>
__global__ __launch_bounds__(kBlock) void k_reduce(const float* __restrict__ w, uint64_t n,
                                                   double* __restrict__ partials,
                                                   uint64_t* __restrict__ states,
                                                   uint64_t state_words) {
    using Acc = typename std::conditional<FP64, double, float>::type;
    __shared__ Acc s_wave[kBlock / kWave];
    const int t = threadIdx.x;
    const uint32_t G = gridDim.x;
    const uint64_t tiles = (n + kTile - 1) / kTile;
    Acc acc = 0;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += G) {
        const uint64_t base = tile * kTile;
        if (VEC && base + kTile <= n) {
            const float4* w4 = reinterpret_cast<const float4*>(w + base);
            float4 v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = w4[t + kBlock * j];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc += (Acc)v[j].x;
                acc += (Acc)v[j].y;
                acc += (Acc)v[j].z;
                acc += (Acc)v[j].w;
            }
        } else {
#pragma unroll 4
            for (int j = 0; j < kItems; ++j) {
                const uint64_t e = base + (uint64_t)t + (uint64_t)kBlock * j;
                if (e < n) acc += (Acc)w[e];
            }
        }
    }
    Acc ws;
    if constexpr (FP64) ws = wave_sum_d(acc);
    else ws = wave_sum_f(acc);
    if (lane_id() == 0) s_wave[t / kWave] = ws;
    // reset the look-back records for the next kernel on the stream
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + t; i < state_words; i += (uint64_t)G * kBlock)
        states[i] = kSentinel;
    __syncthreads();
    if (t == 0) {
        Acc b = s_wave[0];
        for (int k = 1; k < kBlock / kWave; ++k) b = b + s_wave[k];
        partials[blockIdx.x] = (double)b;
    }
}

template <bool FP64>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials,
                                                     uint32_t G, uint64_t n,
                                                     int has_pivot, double pivot,
                                                     Params* __restrict__ prm) {
    using Acc = typename std::conditional<FP64, double, float>::type;
    __shared__ Acc s_wave[kBlock / kWave];
    const int t = threadIdx.x;
    Acc acc = 0;
    for (uint32_t i = t; i < G; i += kBlock) acc = acc + (Acc)partials[i];
    Acc ws;
    if constexpr (FP64) ws = wave_sum_d(acc);
    else ws = wave_sum_f(acc);
    if (lane_id() == 0) s_wave[t / kWave] = ws;
    __syncthreads();
    if (t == 0) {
        Acc total = s_wave[0];
        for (int k = 1; k < kBlock / kWave; ++k) total = total + s_wave[k];
        Params p;
        if constexpr (FP64) {
            p.total = total;
            p.mean = total / (double)n;
            p.pivot = pivot_rd(p.mean);
            p.shift = total > 0.0 ? 110 - ilogb(total) : 0;
            p.inv_total = total > 0.0 ? 1.0 / total : 0.0;
        } else {
            const float mean = total / (float)n;  // psa_alias_table.hpp:39, in FP32
            p.total = (double)total;
            p.mean = (double)mean;
            p.pivot = mean;
            p.shift = 0;
            p.inv_total = total > 0.0f ? (double)(1.0f / total) : 0.0;
        }
        if (has_pivot) {
            p.mean = pivot;
            p.pivot = FP64 ? pivot_rd(p.mean) : (float)p.mean;
        }
        p.ticket = 0;
        p.heavy_count = 0;
        p.reserved[0] = p.reserved[1] = p.reserved[2] = 0;
        *prm = p;
    }
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t words, uint64_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

hipError_t launch_reduce(const float* w, uint64_t n, int prec, double* partials, uint32_t G,
                         uint64_t* states, uint64_t state_words, hipStream_t s) {
    const bool vec = (reinterpret_cast<uintptr_t>(w) & 15u) == 0;
    const bool f64 = prec != kFP32;
    if (f64 && vec)
        k_reduce<true, true><<<G, kBlock, 0, s>>>(w, n, partials, states, state_words);
    else if (f64)
        k_reduce<true, false><<<G, kBlock, 0, s>>>(w, n, partials, states, state_words);
    else if (vec)
        k_reduce<false, true><<<G, kBlock, 0, s>>>(w, n, partials, states, state_words);
    else
        k_reduce<false, false><<<G, kBlock, 0, s>>>(w, n, partials, states, state_words);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* partials, uint32_t G, uint64_t n, int prec,
                           int has_pivot, double pivot, Params* prm, hipStream_t s) {
    if (prec != kFP32)
        k_finalize<true><<<1, kBlock, 0, s>>>(partials, G, n, has_pivot, pivot, prm);
    else
        k_finalize<false><<<1, kBlock, 0, s>>>(partials, G, n, has_pivot, pivot, prm);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t words, uint64_t value, hipStream_t s) {
    if (words == 0) return hipSuccess;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill_u64<<<(uint32_t)blocks, 256, 0, s>>>(p, words, value);
    return hipGetLastError();
}

}  // namespace wrsk
