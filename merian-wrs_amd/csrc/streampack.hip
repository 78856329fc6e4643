// streampack.hip -- PSA split + pack streamed in element order (the pack form
// WRS_PACK_STREAM).
//
// The reference packs each split interval with a serial sweep
// (src/host/reference/pack.hpp:14-83, pack/scalar/float.comp:60-180): with
// heavy j current and residual r, bucket i + j goes to heavy j if r <= mean
// (p = r/mean, alias = heavy j+1), else to light i (p = w_i/mean, alias =
// heavy j).  With LP(i), HP(j) the sums of the first i lights / j heavies,
// "r <= mean" is Y(j+1) <= X(i) for
//
//     X(i) = i mean - LP(i)   (non-decreasing),   Y(j) = HP(j) - j mean   (increasing),
//
// so the sweep is the merge of the lights keyed X with the heavies keyed
// Y(j+1), heavy first on ties, and every row of the table follows from a count
// in that merge:
//
//     light i:   alias heavy q(i),  q(i) = #{q < H-1 : Y(q+1) <= X(i)}
//     heavy q:   (Y(q+1) - X(r) + mean) / mean,  alias heavy q+1,
//                r = r(q) = #{i < L : X(i) < Y(q+1)}
//     heavy H-1: (1, itself)  (pack.hpp:78-81)
//
// This file evaluates the keys EXACTLY: with K = 53 - ilogb(mean) the FP64
// mean times 2^K is an integer below 2^54, a light weight times 2^K is an
// integer below 2^54 (rounded to the nearest integer below mean 2^-30, where
// the float's own bits run out), and every X and Y is an integer sum below
// 2^86.  Exact integer keys are consistent between kernels and tiles whatever
// the order of the additions, so the table can be produced in element order,
// each tile of 4096 weights on its own:
//
//   k_sp_agg      per tile: heavy count, exact light and heavy key sums (the
//                 tile records the partition's tile scan chains), the tile's
//                 heavies (index, weight) into its slot range
//   (tile scan)   k_part_tilescan<.., true> (partition.hip): the exclusive
//                 prefix of every tile's record
//   k_sp_heavy    per tile: Y(q+1) and the element index of each of its
//                 heavies q, in heavy order (hkey, hidx; every 64th key also
//                 in hsmp)
//   k_sp_qrange   per tile t: qa[t] = #{q < H-1 : Y(q+1) <= X(r0(t))}, r0(t)
//                 the lights before the tile -- the alias range of its lights
//                 is [qa[t], qa[t+1]], and heavies q in [qa[t], qa[t+1]) have
//                 their crossing r(q) among its lights
//   k_sp_rows     per tile: the rows of the heavies whose crossing lies among
//                 its lights, into hrow[q] (tiles without any skip their read)
//   k_sp_table    per tile: every row of the tile -- a light's alias from the
//                 crossings before it, its p = w/mean; heavies from hrow --
//                 written in element order, as whole 128-byte lines
//
// The light p is w/mean as in the merge pack; a heavy's p is its exact
// residual rounded once to FP64 and divided by the mean (FP64 modes) or in
// FP32 (FP32 mode), clamped to [0, 1].
#include "device_common.hpp"
#include "internal.hpp"
#include "tileload.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

using u128 = unsigned __int128;
template <bool V>
struct Flag {
    static constexpr bool value = V;
};

constexpr int kWaves = kBlock / kWave;       // waves of the heavy-range search blocks
// Geometry of a streaming pass with B threads per tile: B / 64 waves, each
// thread kG = 4096 / (4 B) float4 groups, rows of 4 B elements, kS (row, wave)
// segments.
template <int B>
struct Geo {
    static constexpr int kW = B / kWave;
    static constexpr int kG = kTile / 4 / B;
    static constexpr uint32_t kStride = 4u * B;
    static constexpr int kS = kW * kG;
};
constexpr int kRowsBlock = 512;   // rows pass
constexpr int kTableBlock = 256;  // table pass
constexpr uint32_t kKeyCache = 1024;         // alias indices staged in LDS per tile (table pass)
constexpr uint32_t kRowCache = 512;          // heavy rows staged in LDS per tile (table pass)

// Tile layout of the streaming passes: B threads per tile, thread t holds the
// float4 groups t + B j (j < 4096 / (4 B)), i.e. elements 4 B j + 4 t + c
// (c < 4) -- the loads and the row stores are coalesced as they are, and the
// scans run per row of groups (the rows pass: B = 512, two groups a thread,
// registers low enough for 8 waves per SIMD; the table pass: B = 256, four).

// rint(w 2^K) for a float w >= 0 with w 2^K < 2^64: the scaling is exact in
// FP32, rintf rounds below 2^23 (where the float's bits run out), and the
// integer-valued float converts as its significand shifted left (y >= 2^23)
// or by the exact 32-bit conversion (y < 2^23).
__device__ __forceinline__ uint64_t qkey_round(float w, int K) {
    const float y = rintf(ldexpf(w, K));
    const uint32_t b = __float_as_uint(y) & 0x7fffffffu;
    const int ex = (int)(b >> 23) - 150;
    const uint64_t sig = (uint64_t)((b & 0x7fffffu) | 0x800000u);
    return ex >= 0 ? sig << (ex & 63) : (uint64_t)(uint32_t)y;
}
// The same value as qkey_round for w 2^K < 2^64: a normal float whose scaled
// value is an integer (exponent field e with e - 150 + K >= 0) is its
// significand shifted left; the others (zero, subnormal, small weights) take
// the rounding route, behind a wave-uniform test so that a wave without such
// a weight skips it.
__device__ __forceinline__ bool qkey_exact(float w, int K) {
    const int e = (int)(__float_as_uint(w) >> 23);
    return e != 0 && e - 150 + K >= 0;
}
// The same test over many weights: the smallest exponent field among them,
// against the smallest exact one.
__device__ __forceinline__ uint32_t qkey_emin(int K) { return (uint32_t)max(1, 150 - K); }
__device__ __forceinline__ uint64_t qkey(float w, int K, bool wave_slow) {
    const uint32_t b = __float_as_uint(w);
    const int e = (int)(b >> 23);
    const uint64_t sig = (uint64_t)((b & 0x7fffffu) | 0x800000u);
    uint64_t q = sig << ((e - 150 + K) & 63);
    if (wave_slow) {
        if (!qkey_exact(w, K)) q = qkey_round(w, K);
    }
    return q;
}

__device__ __forceinline__ FX fx_of(u128 v) {
    return FX{(uint64_t)v & kMask63, (uint64_t)(v >> 63)};
}
__device__ __forceinline__ u128 u128_of(uint64_t lo63, uint64_t hi) {
    return ((u128)hi << 63) | (u128)lo63;
}
// nearest double of v
__device__ __forceinline__ double u128_to_double(u128 v) {
    const uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    if (hi == 0) return (double)lo;
    const int sh = 64 - __clzll((long long)hi);  // 1..64: v >> sh fits 64 bits
    const u128 t = v >> sh;
    const uint64_t sticky = (v & (((u128)1 << sh) - 1)) != 0 ? 1u : 0u;
    return ldexp((double)((uint64_t)t | sticky), sh);
}

template <bool F32>
__device__ __forceinline__ float ratio(double x, const Params* prm) {
    if constexpr (F32) return (float)((float)x / (float)prm->mean);
    else return (float)(x * prm->inv_mean);
}

// The float4 groups of this thread (zero past m; scalar loads where w is not
// 16-byte aligned or a group straddles m).  A whole tile of aligned weights
// (all but the last tile) takes a block-uniform branch to plain float4 loads.
template <int B>
__device__ __forceinline__ void nat_load(const float* __restrict__ w, uint32_t base, uint32_t m,
                                         bool aligned, float (&x)[4 * Geo<B>::kG]) {
    const uint32_t t = threadIdx.x;
    if (aligned && m == (uint32_t)kTile) {
#pragma unroll
        for (int j = 0; j < Geo<B>::kG; ++j) {
            const float4 r = ld_nt4(reinterpret_cast<const float4*>(w + base + Geo<B>::kStride * j + 4u * t));
            x[4 * j] = r.x;
            x[4 * j + 1] = r.y;
            x[4 * j + 2] = r.z;
            x[4 * j + 3] = r.w;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < Geo<B>::kG; ++j) {
        const uint32_t e0 = Geo<B>::kStride * j + 4u * t;
        if (aligned && e0 + 3u < m) {
            const float4 r = ld_nt4(reinterpret_cast<const float4*>(w + base + e0));
            x[4 * j] = r.x;
            x[4 * j + 1] = r.y;
            x[4 * j + 2] = r.z;
            x[4 * j + 3] = r.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) x[4 * j + c] = e0 + c < m ? w[base + e0 + c] : 0.0f;
        }
    }
}

// Per-row wave scans of the thread's per-group light counts and key sums:
// lane 63 of each wave leaves the wave totals in s_c / s_s[Geo<B>::kW * j +
// wid]; excl_c / excl_s are the thread's exclusive prefixes within its wave.
template <int B>
struct NatScan {
    uint32_t excl_c[Geo<B>::kG];
    uint64_t excl_s[Geo<B>::kG];
};
template <int B>
__device__ __forceinline__ void nat_scan(const uint32_t (&lc)[Geo<B>::kG],
                                         const uint64_t (&ls)[Geo<B>::kG], uint32_t* s_c,
                                         uint64_t* s_s, NatScan<B>& o) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < Geo<B>::kG; ++j) {
        const uint32_t ci = wave_incl_scan_u32(lc[j]);
        const uint64_t si = wave_incl_scan_u64(ls[j]);  // a wave's row: < 64 * 4 * 2^54
        if (lane == kWave - 1) {
            s_c[Geo<B>::kW * j + wid] = ci;
            s_s[Geo<B>::kW * j + wid] = si;
        }
        o.excl_c[j] = ci - lc[j];
        o.excl_s[j] = si - ls[j];
    }
}
__device__ __forceinline__ bool le128(u128 a, u128 b) {
    const uint64_t ah = (uint64_t)(a >> 64), bh = (uint64_t)(b >> 64);
    return ah < bh || (ah == bh && (uint64_t)a <= (uint64_t)b);
}

// X(r0): the key before the first light of tile t (t == tiles: X(L)).
__device__ __forceinline__ u128 tile_key(const uint64_t* __restrict__ states, uint32_t n,
                                         uint32_t tiles, uint32_t t, uint64_t mk) {
    const uint64_t* ex;
    u128 lp;
    uint32_t h0;
    if (t < tiles) {
        ex = states + (uint64_t)kExclBase * tiles + (uint64_t)t * kCarryWords;
        lp = u128_of(ex[2], ex[3]);
        h0 = (uint32_t)ex[4];
    } else {
        ex = states + (uint64_t)kExclBase * tiles + (uint64_t)(tiles - 1u) * kCarryWords;
        const uint64_t* ag = states + (uint64_t)(tiles - 1u) * kCarryWords;
        lp = u128_of(ex[2], ex[3]) + u128_of(ag[2], ag[3]);
        h0 = (uint32_t)(ex[4] + ag[4]);
    }
    const uint64_t r0 = min((uint64_t)t * kTile, (uint64_t)n) - h0;  // lights before the tile
    return (u128)r0 * (u128)mk - lp;
}

}  // namespace

// The light key sum of 16 weights by the rounding route (waves holding a
// weight that is zero, subnormal or below mean 2^-30 only).
__device__ __attribute__((noinline)) uint64_t light_keys_round(const float* v, float pivot, int K) {
    uint64_t ls = 0;
    for (int k = 0; k < kItems; ++k) ls += v[k] > pivot ? 0ull : qkey_round(v[k], K);
    return ls;
}

// ---- pass 2a: tile records and heavy slots ---------------------------------
// 256 threads, each owning 16 consecutive elements of the tile (transposed
// through LDS, tileload.hpp): per element only the light key is summed; the
// thread's heavies -- few -- are taken one set bit at a time and their keys
// summed as key / 2^30 (a heavy weight times 2^K is an integer multiple of
// 2^30 below 2^86: its significand shifted left, exactly).  A wave with at
// most 64 heavies lists their tile offsets in LDS in slot order, so that its
// lanes store the slots with one coalesced store each after the barrier
// (others store them per thread).
constexpr int kPadAgg = kTile + kTile / 16;
__global__ __launch_bounds__(kBlock) void k_sp_agg(const float* __restrict__ w, uint32_t n,
                                                   const Params* __restrict__ prm,
                                                   uint64_t* __restrict__ states,
                                                   uint32_t* __restrict__ slot_idx,
                                                   float* __restrict__ slot_w,
                                                   uint32_t* __restrict__ seg) {
    __shared__ float s_buf[kPadAgg];
    __shared__ uint32_t s_stage[kWaves * kWave];  // per wave: tile offsets of its heavies
    __shared__ uint32_t s_c[kWaves];
    __shared__ uint64_t s_l[kWaves], s_hl[kWaves], s_hh[kWaves];
    if (prm->error) return;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    load_tile_lds<true>(w + base, m, s_buf);
    __syncthreads();
    const float pivot = prm->pivot;
    const int K = prm->key_k;
    const uint32_t e0 = (uint32_t)(kItems * t);
    const uint32_t nvalid = e0 >= m ? 0u : min((uint32_t)kItems, m - e0);
    uint32_t hm = 0, emn = 255u;
    uint64_t ls = 0;
    // four elements at a time (one LDS wait each), the key of every element
    // computed and the lights' kept by selects
#pragma unroll
    for (int k0 = 0; k0 < kItems; k0 += 4) {
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = s_buf[17 * t + k0 + c];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool heavy = v[c] > pivot;  // weights past m are zero: light, key 0
            const uint64_t q = qkey(v[c], K, false);
            ls += heavy ? 0ull : q;
            emn = min(emn, heavy ? 255u : __float_as_uint(v[c]) >> 23);
            hm |= (heavy ? 1u : 0u) << (k0 + c);
        }
        asm volatile("" : "+v"(ls));  // a batch at a time: few registers
    }
    hm &= (1u << nvalid) - 1u;
    if (__any(emn < qkey_emin(K))) ls = light_keys_round(s_buf + 17 * t, pivot, K);
    const uint32_t c = (uint32_t)__popc(hm);
    const uint32_t ci = wave_incl_scan_u32(c);
    const uint32_t cw = (uint32_t)__builtin_amdgcn_readlane((int)ci, kWave - 1);  // the wave's heavies
    const bool listed = cw <= (uint32_t)kWave;  // wave-uniform
    uint64_t hs = 0;
    {
        uint32_t i = wid * kWave + ci - c;
        for (uint32_t mk = hm; mk; mk &= mk - 1u) {
            const int k = __builtin_ctz(mk);
            const uint32_t b = __float_as_uint(s_buf[17 * t + k]);
            hs += (uint64_t)((b & 0x7fffffu) | 0x800000u) << (((int)(b >> 23) - 180 + K) & 63);
            if (listed) s_stage[i++] = e0 + (uint32_t)k;
        }
    }
    const uint64_t lt = wave_incl_scan_u64(ls);  // < 64 * 16 * 2^54
    const uint64_t hl = wave_incl_scan_u64(hs & 0xffffffffull);
    const uint64_t hh = wave_incl_scan_u64(hs >> 32);
    if (lane == kWave - 1) {
        s_c[wid] = ci;
        s_l[wid] = lt;
        s_hl[wid] = hl;
        s_hh[wid] = hh;
    }
    __syncthreads();
    uint32_t slot0 = 0, ctile = 0;  // heavies of the tile before this wave, in all
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
        slot0 += q < wid ? s_c[q] : 0u;
        ctile += s_c[q];
    }
    // heavies before each segment of 256 elements (16 threads)
    if ((t & 15) == 0) seg[(uint64_t)tile * kSegWords + (uint32_t)(t >> 4)] = slot0 + ci - c;
    if (listed) {
        if ((uint32_t)lane < cw) {
            const uint32_t off = s_stage[wid * kWave + lane];
            slot_idx[base + slot0 + lane] = base + off;
            slot_w[base + slot0 + lane] = s_buf[off + (off >> 4)];
        }
    } else {
        uint32_t slot = slot0 + ci - c;
        for (uint32_t mk = hm; mk; mk &= mk - 1u) {
            const int k = __builtin_ctz(mk);
            slot_idx[base + slot] = base + e0 + (uint32_t)k;
            slot_w[base + slot] = s_buf[17 * t + k];
            ++slot;
        }
    }
    if (t == 0) {
        u128 L = 0, Hl = 0, Hh = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) {
            L += (u128)s_l[q];
            Hl += (u128)s_hl[q];
            Hh += (u128)s_hh[q];
        }
        const FX h = fx_of((Hl + (Hh << 32)) << 30), l = fx_of(L);
        uint64_t* rec = states + (uint64_t)tile * kCarryWords;
        rec[0] = h.lo;
        rec[1] = h.hi;
        rec[2] = l.lo;
        rec[3] = l.hi;
        rec[4] = ctile;
    }
}

// ---- pass 3a: the heavies' keys in heavy order ------------------------------
// One wave per two consecutive tiles (all waves fit on the device at once),
// both tiles' first slots loaded up front.  A heavy weight times 2^K is its
// significand shifted left, a multiple of 2^30 below 2^86 (the key k_sp_agg
// sums): the wave scans key / 2^30, below 2^57, so that 64 of them fit one
// 64-bit scan.  Every 64th key is also written to hsmp, the first levels of
// the range search.
constexpr uint32_t kHeavyTiles = 2;
__global__ __launch_bounds__(kBlock) void k_sp_heavy(uint32_t n, const Params* __restrict__ prm,
                                                    const uint64_t* __restrict__ states,
                                                    const uint32_t* __restrict__ slot_idx,
                                                    const float* __restrict__ slot_w,
                                                    u128* __restrict__ hkey,
                                                    uint32_t* __restrict__ hidx,
                                                    u128* __restrict__ hsmp) {
    if (prm->error) return;
    const uint32_t H = prm->heavy_count;
    if (H == 0 || H == n) return;  // identity table
    const int lane = lane_id();
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
    const uint32_t t0 = (blockIdx.x * (uint32_t)kWaves + threadIdx.x / kWave) * kHeavyTiles;
    if (t0 >= tiles) return;  // the whole wave
    // the first two chunks of each tile's slots (the slot range [base, base +
    // kTile) is allocated whatever the tile's count)
    constexpr int kPre = 2;
    float pv[kHeavyTiles][kPre];
    uint32_t pid[kHeavyTiles][kPre];
#pragma unroll
    for (uint32_t tt = 0; tt < kHeavyTiles; ++tt) {
        const uint32_t base = (t0 + tt) * (uint32_t)kTile;
#pragma unroll
        for (int r = 0; r < kPre; ++r) {
            const uint32_t k = (uint32_t)(kWave * r + lane);
            const bool in = t0 + tt < tiles && base + k < n;
            pv[tt][r] = in ? slot_w[base + k] : 0.0f;
            pid[tt][r] = in ? slot_idx[base + k] : 0u;
        }
    }
    const uint64_t mk = prm->mean_k;
    const int K = prm->key_k;
#pragma unroll
    for (uint32_t tt = 0; tt < kHeavyTiles; ++tt) {
        const uint32_t tile = t0 + tt;
        if (tile >= tiles) break;  // the whole wave
        const uint32_t base = tile * (uint32_t)kTile;
        const uint64_t* ag = states + (uint64_t)tile * kCarryWords;
        const uint64_t* ex = states + (uint64_t)kExclBase * tiles + (uint64_t)tile * kCarryWords;
        const uint32_t ct = (uint32_t)ag[4];
        const uint32_t h0 = (uint32_t)ex[4];
        u128 run = u128_of(ex[0], ex[1]);  // HP(h0)
        for (uint32_t r = 0; r < ct; r += kWave) {  // wave-uniform
            const uint32_t k = r + (uint32_t)lane;
            const bool act = k < ct;
            float v;
            uint32_t id;
            if (r < (uint32_t)(kPre * kWave)) {
                v = act ? (r == 0 ? pv[tt][0] : pv[tt][1]) : 0.0f;
                id = act ? (r == 0 ? pid[tt][0] : pid[tt][1]) : 0u;
            } else {
                v = act ? slot_w[base + k] : 0.0f;
                id = act ? slot_idx[base + k] : 0u;
            }
            const uint32_t b = __float_as_uint(v);
            const uint64_t q = act ? (uint64_t)((b & 0x7fffffu) | 0x800000u) << (((int)(b >> 23) - 180 + K) & 63)
                                   : 0ull;  // key / 2^30
            const uint64_t sc = wave_incl_scan_u64(q);  // < 64 * 2^57
            if (act) {
                const uint32_t qi = h0 + k;
                const u128 y = run + ((u128)sc << 30) - (u128)(qi + 1u) * (u128)mk;  // Y(qi + 1)
                hkey[qi] = y;
                hidx[qi] = id;
                if ((qi & 63u) == 0) hsmp[qi >> 6] = y;  // every 64th key: the range search's index
            }
            run += (u128)__shfl(sc, kWave - 1, kWave) << 30;
        }
    }
}

// ---- pass 3b: the alias range of every tile ----------------------------------
// One wave per tile: qa[t] = first q in [0, H-1) with Y(q+1) > X(r0(t)) (H-1
// if none); qa[tiles] = H-1.  A 64-ary search over every 64th key (hsmp, a
// small array the caches hold; its first round's probes are the same for
// every tile and load with the tile's record) finds the 64 keys that hold the
// answer, and one load of them (consecutive) and a ballot finish it.
__global__ __launch_bounds__(kBlock) void k_sp_qrange(uint32_t n, const Params* __restrict__ prm,
                                                      const uint64_t* __restrict__ states,
                                                      const u128* __restrict__ hkey,
                                                      const u128* __restrict__ hsmp,
                                                      uint32_t* __restrict__ qa) {
    if (prm->error) return;
    const int lane = lane_id();
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
    const uint32_t t = blockIdx.x * (uint32_t)kWaves + threadIdx.x / kWave;
    if (t > tiles) return;  // the whole wave
    const uint32_t H = prm->heavy_count;
    if (H == 0 || H == n) {
        if (lane == 0) qa[t] = 0;
        return;
    }
    const uint32_t hq = H - 1u;
    if (t == tiles || hq == 0) {
        if (lane == 0) qa[t] = hq;
        return;
    }
    // samples S[i] = Y(64 i + 1), i < ns: i* = first i with S[i] > x (ns if none)
    const uint32_t ns = (hq + 63u) / 64u;
    const uint32_t step0 = (ns + (uint32_t)kWave - 1u) / (uint32_t)kWave;
    const bool in0 = (uint64_t)lane * step0 < ns;
    const u128 y0 = in0 ? hsmp[(uint32_t)lane * step0] : (u128)0;
    const u128 x0 = tile_key(states, n, tiles, t, prm->mean_k);
    uint32_t lo = 0, hi = ns;  // i* lies in [lo, hi]
    bool first = true;
    while (lo < hi) {          // wave-uniform
        const uint32_t R = hi - lo;
        const uint32_t step = (R + (uint32_t)kWave - 1u) / (uint32_t)kWave;
        const uint32_t p = lo + (uint32_t)lane * step;
        const bool in = (uint64_t)lane * step < R;
        const bool pr = in && (first ? y0 : hsmp[p]) > x0;
        first = false;
        const uint64_t b = __ballot(pr);
        const uint64_t bi = __ballot(in);
        if (b) {
            const int f = __builtin_ctzll(b);
            const uint32_t pf = lo + (uint32_t)f * step;
            lo = f == 0 ? lo : lo + (uint32_t)(f - 1) * step + 1u;
            hi = pf;
        } else {
            const int last = 63 - __builtin_clzll(bi);
            lo = lo + (uint32_t)last * step + 1u;
        }
    }
    // S[i* - 1] <= x < S[i*]: the answer lies in [64 (i* - 1) + 1, min(64 i*, hq)];
    // heavies q = b0 .. b0 + cnt - 1 are checked, min(64 i*, hq) if none passes
    uint32_t ans = 0;
    if (lo > 0) {
        const uint32_t b0 = 64u * (lo - 1u) + 1u;
        const uint32_t e = min(64u * lo, hq);  // the answer if no earlier key passes
        const uint32_t cnt = e > b0 ? e - b0 : 0u;  // <= 63 keys before e
        const bool in = (uint32_t)lane < cnt;
        const bool pr = in && hkey[b0 + (uint32_t)lane] > x0;
        const uint64_t b = __ballot(pr);
        ans = b ? b0 + (uint32_t)__builtin_ctzll(b) : e;
    }
    if (lane == 0) qa[t] = ans;
}

// ---- pass 3c: rows of the heavies that cross among a tile's lights -----------
// Each heavy q the tile owns gets its row (p, heavy q+1) in hrow[q] and the
// tile offset of the light it crosses at in cross[q] (0xffff past the last
// light): the table pass counts those crossings instead of comparing keys.
//
// The tile's key after every element, K(e) (X after the lights up to e), is
// non-decreasing; it is kept in LDS relative to the start of its segment of
// 256 elements (one wave's share of a row of groups), where it is below
// 2^62, with the segments' exact start keys beside it.  Heavy q crosses at the
// first element e with K(e) >= Y(q+1) -- a light, since K grows only at
// lights -- found by one thread per key with a search over the segment
// starts and one over the segment.
template <bool F32, int B>
__global__ __launch_bounds__(B) void k_sp_rows(const float* __restrict__ w, uint32_t n,
                                               const Params* __restrict__ prm,
                                               const uint64_t* __restrict__ states,
                                               const u128* __restrict__ hkey,
                                               const uint32_t* __restrict__ hidx,
                                               const uint32_t* __restrict__ qa,
                                               AliasEntry* __restrict__ hrow,
                                               uint16_t* __restrict__ cross,
                                               uint32_t* __restrict__ seg) {
    using G = Geo<B>;
    constexpr uint32_t kSeg = kTile / G::kS;  // elements per segment (256)
    static_assert(G::kS < kWave, "segment prefixes are scanned by one wave");
    __shared__ uint64_t s_k[kTile];
    __shared__ u128 s_xs[G::kS + 1];
    __shared__ uint32_t s_hist[G::kS + 1];
    __shared__ uint32_t s_c[G::kS];
    __shared__ uint64_t s_s[G::kS];
    if (prm->error) return;
    const uint32_t H = prm->heavy_count;
    if (H == 0 || H == n) return;
    const uint32_t tile = blockIdx.x;
    const uint32_t q0 = qa[tile], q1 = qa[tile + 1u];
    if (q0 >= q1) return;  // no crossing among this tile's lights: its weights are not read
    const int t = threadIdx.x;
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const int K = prm->key_k;
    const uint64_t mk = prm->mean_k;
    const float pivot = prm->pivot;
    const bool aligned = (reinterpret_cast<uintptr_t>(w) & 15u) == 0;
    float x[4 * G::kG];
    nat_load<B>(w, base, m, aligned, x);
    const uint32_t kc = q1 - q0;
    // the first chunk's keys and alias indices, loaded with the tile
    u128 y0 = 0;
    uint32_t i0 = 0;
    if ((uint32_t)t < kc) {
        y0 = hkey[q0 + t];
        i0 = hidx[q0 + 1u + t];
    }
    uint32_t lc[G::kG];
    uint64_t ls[G::kG];
    uint32_t lm = 0, emn = 255u;
#pragma unroll
    for (int k = 0; k < 4 * G::kG; ++k) emn = min(emn, x[k] > pivot ? 255u : __float_as_uint(x[k]) >> 23);
    const bool slow = __any(emn < qkey_emin(K));
#pragma unroll
    for (int j = 0; j < G::kG; ++j) {
        const uint32_t e0 = G::kStride * j + 4u * (uint32_t)t;
        lc[j] = 0;
        ls[j] = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool light = !(x[4 * j + c] > pivot) && e0 + c < m;
            lm |= (light ? 1u : 0u) << (4 * j + c);
            lc[j] += light ? 1u : 0u;
            ls[j] += light ? qkey(x[4 * j + c], K, slow) : 0ull;
        }
    }
    NatScan<B> ns;
    nat_scan<B>(lc, ls, s_c, s_s, ns);
    // K(e) relative to the segment start: (lights so far) mean_k - (their keys)
#pragma unroll
    for (int j = 0; j < G::kG; ++j) {
        uint64_t cnt = ns.excl_c[j], sum = ns.excl_s[j];
        uint64_t kr[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if ((lm >> (4 * j + c)) & 1u) {
                ++cnt;
                sum += qkey(x[4 * j + c], K, slow);
            }
            kr[c] = cnt * mk - sum;
        }
        uint64_t* d = &s_k[G::kStride * j + 4u * (uint32_t)t];
        reinterpret_cast<ulonglong2*>(d)[0] = make_ulonglong2(kr[0], kr[1]);
        reinterpret_cast<ulonglong2*>(d)[1] = make_ulonglong2(kr[2], kr[3]);
    }
    if (t <= G::kS) s_hist[t] = 0;  // crossings per segment (G::kS: past the last light)
    const u128 x0 = tile_key(states, n, tiles, tile, mk);
    __syncthreads();
    if (t < kWave) {  // wave 0: segment start keys; s_xs[kS]: the key after the tile's lights
        // exclusive scans over lanes 0 .. kS of the segment light counts and
        // key sums (a segment's sum is below 2^62: two 32-bit halves)
        const uint32_t c = t < G::kS ? s_c[t] : 0u;
        const uint64_t sv = t < G::kS ? s_s[t] : 0ull;
        const uint32_t ce = wave_incl_scan_u32(c) - c;
        const uint64_t sl = wave_incl_scan_u64(sv & 0xffffffffull) - (sv & 0xffffffffull);
        const uint64_t sh = wave_incl_scan_u64(sv >> 32) - (sv >> 32);
        if (t <= G::kS) s_xs[t] = x0 + (u128)ce * mk - ((u128)sl + ((u128)sh << 32));
    }
    __syncthreads();
    const u128 xend = s_xs[G::kS];
    for (uint32_t a0 = 0; a0 < kc; a0 += B) {  // block-uniform
        const uint32_t a = a0 + (uint32_t)t;
        uint32_t e = 0xffffu;
        if (a < kc) {
            const u128 y = a0 == 0 ? y0 : hkey[q0 + a];
            u128 xa;  // X after the crossing light (X(L) past the last light)
            if (!le128(y, xend)) {  // past the last light (the last tile only): r(q) = L
                xa = xend;
            } else {
                // last segment whose start key is below y, then the first element
                // of it whose key reaches y
                uint32_t lo = 0, hi = G::kS - 1u;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1u) >> 1;
                    if (le128(y, s_xs[mid])) hi = mid - 1u;
                    else lo = mid;
                }
                const u128 xs = s_xs[lo];
                const uint64_t yr = (uint64_t)(y - xs);  // (0, segment span]: below 2^62
                uint32_t b0 = lo * kSeg, b1 = b0 + kSeg - 1u;
                while (b0 < b1) {
                    const uint32_t mid = (b0 + b1) >> 1;
                    if (s_k[mid] >= yr) b1 = mid;
                    else b0 = mid + 1u;
                }
                e = b0;
                xa = xs + (u128)s_k[b0];
            }
            const double r = ldexp(u128_to_double(y + (u128)mk - xa), -K);
            const float p = fminf(fmaxf(ratio<F32>(r, prm), 0.0f), 1.0f);
            hrow[q0 + a] = AliasEntry{p, a0 == 0 ? i0 : hidx[q0 + 1u + a]};
            cross[q0 + a] = (uint16_t)e;
        }
        if (a < kc) atomicAdd(&s_hist[min(e / kSeg, (uint32_t)G::kS)], 1u);
    }
    // crossings before each segment boundary 256 s (s <= kS)
    __syncthreads();
    if (t < kWave) {  // wave 0: exclusive scan over lanes 0 .. kS
        const uint32_t h = t <= G::kS ? s_hist[t] : 0u;
        const uint32_t c = wave_incl_scan_u32(h) - h;
        if (t <= G::kS) seg[(uint64_t)tile * kSegWords + 32u + (uint32_t)t] = c;
    }
}

// ---- pass 4: the table, in element order -------------------------------------
// A light's alias is heavy q0 + (the crossings at lights before it in the
// tile); a heavy's row comes from hrow.  No keys are compared here.  A tile
// whose alias range and heavy rows fit in LDS (every tile of a spread
// distribution) stages them and counts its crossings per element into an LDS
// histogram, so that wave scans give every element both its heavy rank and
// its crossing count on top of the segment prefixes of the aggregate pass
// (heavies) and the rows pass (crossings).  Other tiles read the indices and
// rows from memory and count crossings in their segment's slice of the sorted
// crossing list.
//
// Layout: wave w of row j owns segment s = kW j + w (elements [256 s, 256 s +
// 256)), and its lane l the pairs of elements at 256 s + 2 l (half 0) and
// 256 s + 128 + 2 l (half 1): a wave's 16-byte row stores then cover 1 KiB of
// the table without gaps (a lane holding four consecutive rows writes them
// as two 16-byte stores that each cover half of every line, measurably
// slower), and its 8-byte weight loads 512 bytes.
template <bool F32, int B>
__global__ __launch_bounds__(B) void k_sp_table(const float* __restrict__ w, uint32_t n,
                                                const Params* __restrict__ prm,
                                                const uint64_t* __restrict__ states,
                                                const uint32_t* __restrict__ hidx,
                                                const uint16_t* __restrict__ cross,
                                                const uint32_t* __restrict__ qa,
                                                const AliasEntry* __restrict__ hrow,
                                                const uint32_t* __restrict__ seg,
                                                AliasEntry* __restrict__ table) {
    using G = Geo<B>;
    constexpr uint32_t kSeg = kTile / G::kS;  // elements per segment (256)
    static_assert(kSeg == 4u * kWave, "a segment is one wave's share of a row");
    constexpr uint32_t kHalf = kSeg / 2u;
    __shared__ uint32_t s_hist[kTile / 2];      // crossings at each element, two 16-bit counts a word
    __shared__ uint32_t s_idx[kKeyCache + 1];   // hidx[q0 .. q0 + kc]
    __shared__ AliasEntry s_row[kRowCache + 1]; // hrow[h0 ..]
    if (prm->error) return;
    const int t = threadIdx.x;
    const uint32_t lane = (uint32_t)t & (kWave - 1u);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const uint32_t H = prm->heavy_count;
    // tile offset of this thread's pair (j, h)
    auto pos = [&](int j, int h) {
        return kSeg * (uint32_t)(G::kW * j + wid) + kHalf * (uint32_t)h + 2u * lane;
    };
    // the rows of a pair: FULL (a whole tile into a 16-byte aligned table, a
    // block-uniform choice made once per tile, so that the two paths stay
    // apart in the code) one 16-byte nontemporal store (the table is not read
    // again by the build; streaming it past the caches measured 156 -> 144 us
    // for this pass); else a plain 16-byte store where the pair is whole and
    // aligned, or per row
    const bool tab16 = (reinterpret_cast<uintptr_t>(table) & 15u) == 0;
    const bool full16 = tab16 && m == (uint32_t)kTile;
    auto store = [&](uint32_t e0, uint2 r0, uint2 r1, auto full) {
        if (decltype(full)::value) {
            st_nt_u4(reinterpret_cast<uint4*>(table + base + e0), make_uint4(r0.x, r0.y, r1.x, r1.y));
        } else if (tab16 && e0 + 1u < m) {
            *reinterpret_cast<uint4*>(table + base + e0) = make_uint4(r0.x, r0.y, r1.x, r1.y);
        } else {
            uint2* o = reinterpret_cast<uint2*>(table + base + e0);
            if (e0 < m) o[0] = r0;
            if (e0 + 1u < m) o[1] = r1;
        }
    };
    if (H == 0 || H == n) {  // all weights equal the mean up to its rounding: identity
#pragma unroll
        for (int j = 0; j < G::kG; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t e0 = pos(j, h);
                store(e0, make_uint2(__float_as_uint(1.0f), base + e0),
                      make_uint2(__float_as_uint(1.0f), base + e0 + 1u), Flag<false>{});
            }
        }
        return;
    }
    const float pivot = prm->pivot;
    // qa is non-decreasing for valid weights (exact keys); a range that is
    // not (NaN weights) is taken as empty rather than read out of bounds
    const uint32_t q0 = qa[tile], q1 = max(qa[tile + 1u], q0);
    const uint64_t* ag = states + (uint64_t)tile * kCarryWords;
    const uint64_t* ex = states + (uint64_t)kExclBase * tiles + (uint64_t)tile * kCarryWords;
    const uint32_t h0 = (uint32_t)ex[4];
    const uint32_t ct = (uint32_t)ag[4];
    const uint32_t* hseg = seg + (uint64_t)tile * kSegWords;
    const uint32_t* cseg = hseg + 32u;
    // x[4 j + 2 h + c]: element pos(j, h) + c (zero past m)
    float x[4 * G::kG];
    if ((reinterpret_cast<uintptr_t>(w) & 7u) == 0 && m == (uint32_t)kTile) {  // block-uniform
#pragma unroll
        for (int j = 0; j < G::kG; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float2 v = ld_nt2(reinterpret_cast<const float2*>(w + base + pos(j, h)));
                x[4 * j + 2 * h] = v.x;
                x[4 * j + 2 * h + 1] = v.y;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < G::kG; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t e0 = pos(j, h);
                x[4 * j + 2 * h] = e0 < m ? w[base + e0] : 0.0f;
                x[4 * j + 2 * h + 1] = e0 + 1u < m ? w[base + e0 + 1u] : 0.0f;
            }
        }
    }
    const uint32_t kc = q1 - q0;
    // the segment prefixes of this wave's segments (wave-uniform; the crossing
    // words are written for tiles with crossings only)
    uint32_t hpre[G::kG], clo[G::kG], chi[G::kG];
#pragma unroll
    for (int j = 0; j < G::kG; ++j) {
        hpre[j] = hseg[G::kW * j + wid];
        clo[j] = kc ? cseg[G::kW * j + wid] : 0u;
        chi[j] = kc ? cseg[G::kW * j + wid + 1] : 0u;
    }
    // heavy flags (bit c) and counts of the two pairs of segment j
    auto heavies = [&](int j, uint32_t (&hm)[2], uint32_t (&hc)[2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t e0 = pos(j, h);
            const bool a = x[4 * j + 2 * h] > pivot && e0 < m;
            const bool b = x[4 * j + 2 * h + 1] > pivot && e0 + 1u < m;
            hm[h] = (a ? 1u : 0u) | (b ? 2u : 0u);
            hc[h] = (a ? 1u : 0u) + (b ? 1u : 0u);
        }
    };
    const bool cached = kc <= kKeyCache && ct <= kRowCache;
    constexpr int kPer = (int)(kKeyCache / B);
    constexpr int kRowPer = (int)(kRowCache / B);
    // the light p: w / mean rounded once to float
    auto light_p = [&](float v) { return fminf(ratio<F32>((double)v, prm), 1.0f); };
    if (cached) {  // block-uniform
        uint32_t cv[kPer], iv[kPer];
        AliasEntry rv[kRowPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = (uint32_t)t + B * r;
            cv[r] = i < kc ? (uint32_t)cross[q0 + i] : 0xffffu;
            iv[r] = i < kc ? hidx[q0 + i] : 0u;
        }
        const uint32_t ilast = t == 0 ? hidx[q1] : 0u;  // alias of the tile's last lights
#pragma unroll
        for (int r = 0; r < kRowPer; ++r) {
            const uint32_t i = (uint32_t)t + B * r;
            if (i < ct) rv[r] = hrow[h0 + i];
        }
        if (kc) {
            uint4* h4 = reinterpret_cast<uint4*>(s_hist);
            for (uint32_t k = (uint32_t)t; k < kTile / 8u; k += B) h4[k] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = (uint32_t)t + B * r;
            if (i < kc) s_idx[i] = iv[r];
            // 0xffff: past the tile's last light (the last tile only)
            if (cv[r] < kTile) atomicAdd(&s_hist[cv[r] >> 1], 1u << (16u * (cv[r] & 1u)));
        }
        if (t == 0) s_idx[kc] = ilast;
#pragma unroll
        for (int r = 0; r < kRowPer; ++r) {
            const uint32_t i = (uint32_t)t + B * r;
            if (i < ct) s_row[i] = rv[r];
        }
        __syncthreads();
        auto body = [&](auto full) {
#pragma unroll
            for (int j = 0; j < G::kG; ++j) {
                uint32_t hm[2], hc[2], na[2], v[2];
                heavies(j, hm, hc);
#pragma unroll
                for (int h = 0; h < 2; ++h) {  // crossings at the pair's first element, at both
                    uint32_t nt = 0;
                    na[h] = 0;
                    if (kc) {
                        const uint32_t hw = s_hist[pos(j, h) >> 1];
                        na[h] = hw & 0xffffu;
                        nt = na[h] + (hw >> 16);
                    }
                    // heavies (<= 128 a half) in the low 12 bits, crossings
                    // (<= kKeyCache) above
                    v[h] = hc[h] | (nt << 12);
                }
                const uint32_t s0 = wave_incl_scan_u32(v[0]);
                const uint32_t s1 = wave_incl_scan_u32(v[1]);
                const uint32_t tot0 = (uint32_t)__builtin_amdgcn_readlane((int)s0, kWave - 1);
                const uint32_t exv[2] = {s0 - v[0], tot0 + s1 - v[1]};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t e0 = pos(j, h);
                    uint32_t hr = hpre[j] + (exv[h] & 0xfffu);
                    uint32_t cx = clo[j] + (exv[h] >> 12);
                    uint2 r[2];
                    // both rows read and one selected, without branches: s_row[hr]
                    // for a light past the tile's last heavy (hr = ct) is read and unused
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const bool hv = (hm[h] >> c) & 1u;
                        const AliasEntry e = s_row[hr];
                        const uint32_t ia = s_idx[cx];
                        const float lp = light_p(x[4 * j + 2 * h + c]);
                        const bool last = h0 + hr + 1u == H;  // the last heavy: (1, itself)
                        const float p = hv ? (last ? 1.0f : e.p) : lp;
                        const uint32_t a = hv ? (last ? base + e0 + (uint32_t)c : e.a) : ia;
                        r[c] = make_uint2(__float_as_uint(p), a);
                        hr += hv ? 1u : 0u;
                        if (c == 0) cx += na[h];
                    }
                    store(e0, r[0], r[1], full);
                }
            }
        };
        if (full16)
            body(Flag<true>{});
        else
            body(Flag<false>{});
        return;
    }
    // crossings before element e: from the segment's slice [clo, chi) of the
    // sorted list -- a short slice by one wave-uniform pass over it, a long one
    // by a search and a walk
    constexpr uint32_t kShortSlice = 32;
    const uint16_t* cr = cross + q0;
    const uint32_t* ix = hidx + q0;
    const AliasEntry* hr_src = hrow + h0;
#pragma unroll
    for (int j = 0; j < G::kG; ++j) {
        uint32_t hm[2], hc[2];
        heavies(j, hm, hc);
        const uint32_t v = hc[0] | (hc[1] << 16);  // heavies of a half: <= 128
        const uint32_t sc = wave_incl_scan_u32(v);
        const uint32_t tot0 = (uint32_t)__builtin_amdgcn_readlane((int)sc, kWave - 1) & 0xffffu;
        const uint32_t hx[2] = {(sc - v) & 0xffffu, tot0 + ((sc - v) >> 16)};
        const uint32_t lo = min(clo[j], kc), hi = max(min(chi[j], kc), lo);
        uint32_t cxa[2], cxb[2];  // crossings before the pair's first, second element
        if (hi - lo <= kShortSlice) {  // wave-uniform
            const uint32_t e0 = pos(j, 0), e1 = pos(j, 1);
            uint32_t b0 = 0, q0n = 0, b1 = 0, q1n = 0;
            for (uint32_t k = lo; k < hi; ++k) {
                const uint32_t cv = cr[k];
                b0 += cv < e0 ? 1u : 0u;
                q0n += cv == e0 ? 1u : 0u;
                b1 += cv < e1 ? 1u : 0u;
                q1n += cv == e1 ? 1u : 0u;
            }
            cxa[0] = lo + b0;
            cxb[0] = cxa[0] + q0n;
            cxa[1] = lo + b1;
            cxb[1] = cxa[1] + q1n;
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t e0 = pos(j, h);
                uint32_t a = lo, b = hi;  // first crossing at or past e0
                while (a < b) {
                    const uint32_t mid = (a + b) >> 1;
                    if (cr[mid] < e0) a = mid + 1u;
                    else b = mid;
                }
                cxa[h] = a;
                while (a < hi && cr[a] < e0 + 1u) ++a;
                cxb[h] = a;
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t e0 = pos(j, h);
            uint32_t hr = hpre[j] + hx[h];
            uint2 r[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if ((hm[h] >> c) & 1u) {
                    if (h0 + hr + 1u == H) {
                        r[c] = make_uint2(__float_as_uint(1.0f), base + e0 + (uint32_t)c);
                    } else {
                        const AliasEntry e = hr_src[hr];
                        r[c] = make_uint2(__float_as_uint(e.p), e.a);
                    }
                    ++hr;
                } else {  // a light (or past m: not stored)
                    r[c] = make_uint2(__float_as_uint(light_p(x[4 * j + 2 * h + c])),
                                      ix[c == 0 ? cxa[h] : cxb[h]]);
                }
            }
            store(e0, r[0], r[1], Flag<false>{});
        }
    }
}

// ---- launchers -----------------------------------------------------------------
hipError_t launch_stream_agg(const float* w, uint64_t n, const Params* prm, uint64_t* states,
                             const StreamBufs& b, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    k_sp_agg<<<tiles, kBlock, 0, s>>>(w, (uint32_t)n, prm, states, b.slot_idx, b.slot_w, b.seg);
    return hipGetLastError();
}

hipError_t launch_stream_pack(const float* w, uint64_t n, int prec, const Params* prm,
                              const uint64_t* states, const StreamBufs& b, AliasEntry* table,
                              hipStream_t s, hipEvent_t ev_keys, hipEvent_t ev_rows) {
    const uint32_t nn = (uint32_t)n;
    const uint32_t tiles = (uint32_t)tiles_for(n);
    u128* hkey = static_cast<u128*>(b.hkey);
    u128* hsmp = static_cast<u128*>(b.hsmp);
    const uint32_t hwaves = (tiles + kHeavyTiles - 1u) / kHeavyTiles;
    k_sp_heavy<<<(hwaves + kWaves - 1u) / kWaves, kBlock, 0, s>>>(nn, prm, states, b.slot_idx,
                                                                   b.slot_w, hkey, b.hidx, hsmp);
    k_sp_qrange<<<(tiles + 1u + kWaves - 1u) / kWaves, kBlock, 0, s>>>(nn, prm, states, hkey, hsmp,
                                                                       b.qa);
    if (ev_keys) (void)hipEventRecord(ev_keys, s);
    if (prec == kFP32) {
        k_sp_rows<true, kRowsBlock><<<tiles, kRowsBlock, 0, s>>>(w, nn, prm, states, hkey, b.hidx,
                                                                 b.qa, b.hrow, b.cross, b.seg);
        if (ev_rows) (void)hipEventRecord(ev_rows, s);
        k_sp_table<true, kTableBlock><<<tiles, kTableBlock, 0, s>>>(w, nn, prm, states, b.hidx,
                                                                    b.cross, b.qa, b.hrow, b.seg,
                                                                    table);
    } else {
        k_sp_rows<false, kRowsBlock><<<tiles, kRowsBlock, 0, s>>>(w, nn, prm, states, hkey, b.hidx,
                                                                  b.qa, b.hrow, b.cross, b.seg);
        if (ev_rows) (void)hipEventRecord(ev_rows, s);
        k_sp_table<false, kTableBlock><<<tiles, kTableBlock, 0, s>>>(w, nn, prm, states, b.hidx,
                                                                     b.cross, b.qa, b.hrow, b.seg,
                                                                     table);
    }
    return hipGetLastError();
}

}  // namespace wrsk
