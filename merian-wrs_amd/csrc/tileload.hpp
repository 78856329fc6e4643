// tileload.hpp -- loading one tile of weights into LDS as 16-byte rows,
// whatever the alignment of its first element.
//
// The weights of a tile (or of a segment of a batch) may start at any 4-byte
// boundary; 4-byte loads would halve the achieved bandwidth.  Instead the
// float4 rows of the 16-byte-aligned address at or below the tile's start are
// loaded (one more row when the start is not aligned) and each element is
// stored to its LDS slot.  The words read outside [p, p + m) lie in the same
// 16-byte row as a word inside it, hence in the same memory page, and are
// discarded.
#pragma once

#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {

// LDS slot of tile element e: with PAD one word of padding after every 16
// elements (the transpose layout of the scan kernels, thread t then reads
// elements 16t..16t+15 from slots 17t..17t+15 without bank conflicts).
template <bool PAD>
__device__ __forceinline__ uint32_t tile_slot(uint32_t e) {
    return PAD ? e + e / 16u : e;
}

// Stores elements [0, m) of p (m <= kTile) to buf[tile_slot(e)] and zero to
// the slots of [m, kTile).  Called by all kBlock threads; the caller
// synchronises before reading buf.
template <bool PAD>
__device__ __forceinline__ void load_tile_lds(const float* __restrict__ p, uint32_t m,
                                              float* __restrict__ buf) {
    constexpr int kRows = kTile / 4 / kBlock;  // full rows per thread
    const int t = threadIdx.x;
    const uint32_t a = (uint32_t)((reinterpret_cast<uintptr_t>(p) >> 2) & 3u);
    const float4* q = reinterpret_cast<const float4*>(p - a);
    const uint32_t rows = (a + m + 3u) >> 2;  // rows holding an element of [0, m)
    float4 r[kRows + 1];
#pragma unroll
    for (int j = 0; j <= kRows; ++j) {
        const uint32_t row = (uint32_t)t + (uint32_t)kBlock * j;
        r[j] = row < rows ? wrsdev::ld_nt4(q + row) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int j = 0; j <= kRows; ++j) {
        const int e0 = 4 * (t + kBlock * j) - (int)a;
        const float x[4] = {r[j].x, r[j].y, r[j].z, r[j].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e = e0 + c;
            if (e >= 0 && e < (int)m) buf[tile_slot<PAD>((uint32_t)e)] = x[c];
        }
    }
    if (m < (uint32_t)kTile)
        for (uint32_t e = m + (uint32_t)t; e < (uint32_t)kTile; e += kBlock) buf[tile_slot<PAD>(e)] = 0.0f;
}

}  // namespace wrsk
