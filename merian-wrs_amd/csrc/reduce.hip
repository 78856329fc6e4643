// reduce.hip -- deterministic total weight (pass 1 of a build).
//
// Replaces the reference's AtomicMean (src/device/mean/atomic/shader.comp:69-82),
// whose float atomicAdd makes the pivot change from run to run.  Here every
// element is accumulated in FP64 (exact conversion of the float) in an order
// fixed by N alone: block b sums tiles b, b+G, b+2G, ...; within a tile thread t
// sums the elements 4(t+256j)+q, j = 0..3, q = 0..3 in that order (the same order
// whether the rows come straight from memory or through LDS, so the result does
// not depend on pointer alignment); waves reduce by a fixed butterfly and the
// last block to finish sums the G partials in a fixed tree.  The same launch
// also resets the look-back records of the kernel that follows it.
//
// The segmented (batched) kernels run exactly the same per-tile and finalize
// code per segment, so a batched build yields the same bits as building each
// segment on its own.
#include "device_common.hpp"
#include "internal.hpp"
#include "tileload.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

template <bool FP64>
using AccT = typename std::conditional<FP64, double, float>::type;

// Adds tile `tile` of w[0..n) into acc in the fixed per-thread order.  A full
// tile at a 16-byte aligned address is loaded straight into registers; any
// other tile (the tail, or a segment of a batch, which may start anywhere) is
// staged in LDS as rows (tileload.hpp) and read back as the same float4 rows,
// so the order -- and the sum -- is the same either way.  Block-uniform.
template <bool FP64>
__device__ __forceinline__ void reduce_tile(const float* __restrict__ w, uint64_t n, uint64_t tile,
                                            AccT<FP64>& acc) {
    using Acc = AccT<FP64>;
    const int t = threadIdx.x;
    const uint64_t base = tile * kTile;
    const float* p = w + base;
    float4 v[4];
    if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0 && base + kTile <= n) {
        const float4* w4 = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = ld_nt4(w4 + t + kBlock * j);
    } else {
        __shared__ float4 s_rows[kTile / 4];
        const uint32_t m = (uint32_t)(n - base < (uint64_t)kTile ? n - base : (uint64_t)kTile);
        __syncthreads();  // the previous tile of a grid-stride loop has been read
        load_tile_lds<false>(p, m, reinterpret_cast<float*>(s_rows));
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = s_rows[t + kBlock * j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        acc += (Acc)v[j].x;
        acc += (Acc)v[j].y;
        acc += (Acc)v[j].z;
        acc += (Acc)v[j].w;
    }
}

// Block sum of one value per thread: wave butterflies, then waves in order.
template <bool FP64>
__device__ __forceinline__ AccT<FP64> block_sum(AccT<FP64> acc) {
    using Acc = AccT<FP64>;
    __shared__ Acc s_wave[kBlock / kWave];
    Acc ws;
    if constexpr (FP64) ws = wave_sum_d(acc);
    else ws = wave_sum_f(acc);
    if (lane_id() == 0) s_wave[threadIdx.x / kWave] = ws;
    __syncthreads();
    Acc b = s_wave[0];
    for (int k = 1; k < kBlock / kWave; ++k) b = b + s_wave[k];
    return b;
}

// Derives the build scalars from this thread's running sum of the partials
// i = t, t + kBlock, ... (in that order); thread 0 writes *prm.
template <bool FP64>
__device__ __forceinline__ void finalize_acc(AccT<FP64> acc, uint64_t n, int has_pivot,
                                             double pivot, Params* __restrict__ prm,
                                             uint32_t error = 0u,
                                             uint32_t wait_polls = kDefaultWaitPolls,
                                             uint32_t hooks = 0u) {
    using Acc = AccT<FP64>;
    const int t = threadIdx.x;
    const Acc total = block_sum<FP64>(acc);
    if (t == 0) {
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
        p.heavy_sum = 0.0;
        p.light_sum = 0.0;
        p.repair_ticket = 0;
        // merge-key scale: every key and offset of a merge block stays below
        // 2^62 in magnitude
        p.merge_qs = p.total > 0.0 ? 61 - ilogb(p.total) : 0;
        p.mean_q = p.total > 0.0 ? merge_quant(p.mean, p.merge_qs) : 0;
        p.inv_mean = p.mean > 0.0 ? 1.0 / p.mean : 0.0;
        // exact keys of the streamed pack: the mean (53 significant bits)
        // times 2^key_k is an integer below 2^54
        p.key_k = p.mean > 0.0 ? 53 - ilogb(p.mean) : 0;
        p.key_pad = 0;
        p.mean_k = p.mean > 0.0 ? (uint64_t)ldexp(p.mean, p.key_k) : 0;
        p.error = error;
        p.wait_polls = wait_polls;
        p.hooks = hooks;
        p.epoch = prm->epoch + 1u;  // the next build's tag
        *prm = p;
    }
}

// Sums G partials in a fixed order and derives the build scalars.
template <bool FP64>
__device__ __forceinline__ void finalize_block(const double* __restrict__ partials, uint32_t G,
                                               uint64_t n, int has_pivot, double pivot,
                                               Params* __restrict__ prm) {
    AccT<FP64> acc = 0;
    for (uint32_t i = threadIdx.x; i < G; i += kBlock) acc = acc + (AccT<FP64>)partials[i];
    finalize_acc<FP64>(acc, n, has_pivot, pivot, prm);
}

}  // namespace

// Block totals, then the build scalars by the last block to finish.  red holds
// kMaxPartials total words, the sentinel ~0 between launches (a total is a
// non-negative double or a NaN from a float, never all ones), and the arrival
// counters, zero between launches.  A total word carries the parity of the
// build's epoch (Params::epoch, which every build's finalize advances) in its
// sign bit, which a non-negative total leaves free: a store of an earlier build
// that lands late -- after that build's reader gave up and restored the
// sentinel -- has the other parity, and the next reader keeps waiting for its
// own build's word instead of summing the stale one.  Every block stores its total with an
// agent-coherent store and then counts itself in its group's counter (groups
// of kRedGroup blocks); the last of a group counts the group in the top
// counter, and the last group's last block -- which therefore runs after every
// other block has issued its total -- reads the totals, sums them in the
// fixed order of the segmented finalize (same bits), writes *prm, and restores
// the sentinels and zero counters for the next launch.  No block ever waits
// for another to make progress: the reader only waits until stores already
// issued become visible, a wait bounded by `wait_polls` polls per word whose
// expiry (a machine fault, or a hook of the tests) sets kErrReduceWait.  The
// split counters keep the same-address atomics to kRedGroup per word.
template <bool FP64>
__global__ __launch_bounds__(kBlock) void k_reduce_final(const float* __restrict__ w, uint64_t n,
                                                         uint64_t* __restrict__ red,
                                                         uint64_t* __restrict__ states,
                                                         uint64_t state_words, int has_pivot,
                                                         double pivot, Params* __restrict__ prm,
                                                         uint32_t wait_polls, uint32_t hooks) {
    __shared__ uint32_t s_last;
    const int t = threadIdx.x;
    const uint32_t G = gridDim.x;
    const uint64_t tiles = (n + kTile - 1) / kTile;
    AccT<FP64> acc = 0;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += G) reduce_tile<FP64>(w, n, tile, acc);
    // reset the look-back records for the next kernel on the stream
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + t; i < state_words; i += (uint64_t)G * kBlock)
        states[i] = kSentinel;
    const AccT<FP64> b = block_sum<FP64>(acc);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(red + kMaxPartials);
    const uint64_t tag = (uint64_t)(prm->epoch & 1u) << 63;  // read before any block finalizes
    constexpr uint64_t kSign = 1ull << 63;
    if (t == 0) {
        const uint64_t bits = ((uint64_t)__double_as_longlong((double)b) & ~kSign) | tag;
        if (blockIdx.x == 0 && (hooks & kHookStaleTotal))
            st_word(red, bits ^ kSign);  // the previous build's tag
        else if (!(blockIdx.x == 0 && (hooks & kHookSkipPartial)))
            st_word(red + blockIdx.x, bits);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the store is issued before the count
        const uint32_t g = blockIdx.x / kRedGroup;
        const uint32_t gsize = min(kRedGroup, G - g * kRedGroup);
        uint32_t last = 0;
        if (atomicAdd(&cnt[g], 1u) == gsize - 1u)
            last = atomicAdd(&cnt[kRedCounters - 1], 1u) == (G + kRedGroup - 1) / kRedGroup - 1u;
        s_last = last;
    }
    __syncthreads();
    if (!s_last) return;
    // this thread's words i = t + kBlock k, all loads in flight at once
    constexpr int kPer = (int)(kMaxPartials / kBlock);
    uint64_t x[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = (uint32_t)t + (uint32_t)kBlock * k;
        x[k] = i < G ? ld_word(red + i) : 0ull;
    }
    // a word of this build: not the sentinel, and this build's tag
    auto mine = [&](uint64_t v) { return v != kSentinel && (v & kSign) == tag; };
    bool missing = false;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = (uint32_t)t + (uint32_t)kBlock * k;
        for (uint32_t spin = 0; i < G && !mine(x[k]) && spin < wait_polls; ++spin) {
            __builtin_amdgcn_s_sleep(1);
            x[k] = ld_word(red + i);
        }
        missing |= i < G && !mine(x[k]);
    }
    const bool failed = __syncthreads_or(missing) != 0;
    AccT<FP64> sum = 0;  // the order of finalize_block
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if ((uint32_t)t + (uint32_t)kBlock * k < G)
            sum = sum + (AccT<FP64>)__longlong_as_double((long long)(x[k] & ~kSign));
    finalize_acc<FP64>(sum, n, has_pivot, pivot, prm, failed ? kErrReduceWait : 0u, wait_polls,
                       hooks);
    // every thread restores only the words it read itself
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = (uint32_t)t + (uint32_t)kBlock * k;
        if (i < G) st_word(red + i, kSentinel);
    }
    if (t < (int)kRedCounters) cnt[t] = 0u;
}

// Segmented: one block per (segment, tile) item; every segment has at most
// kMaxPartials tiles, so the single-distribution kernel would also give each
// tile its own block -- identical partials.
template <bool FP64>
__global__ __launch_bounds__(kBlock) void k_seg_reduce(const float* __restrict__ w,
                                                       const SegItem* __restrict__ item_d,
                                                       double* __restrict__ partials,
                                                       uint64_t* __restrict__ states) {
    const uint32_t item = blockIdx.x;
    const SegItem it = item_d[item];
    AccT<FP64> acc = 0;
    reduce_tile<FP64>(w + it.off, it.n, it.blk, acc);
    if (threadIdx.x < kStateWords) states[(uint64_t)item * kStateWords + threadIdx.x] = kSentinel;
    const AccT<FP64> b = block_sum<FP64>(acc);
    if (threadIdx.x == 0) partials[item] = (double)b;
}

template <bool FP64>
__global__ __launch_bounds__(kBlock) void k_seg_finalize(const uint64_t* __restrict__ seg_off,
                                                         const uint32_t* __restrict__ seg_item0,
                                                         const double* __restrict__ partials,
                                                         Params* __restrict__ seg_prm,
                                                         uint32_t* __restrict__ ticket) {
    const uint32_t sg = blockIdx.x;
    const uint32_t i0 = seg_item0[sg];
    const uint32_t G = seg_item0[sg + 1] - i0;
    finalize_block<FP64>(partials + i0, G, seg_off[sg + 1] - seg_off[sg], 0, 0.0, seg_prm + sg);
    if (sg == 0 && threadIdx.x == 0) *ticket = 0;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t words, uint64_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

hipError_t launch_reduce_final(const float* w, uint64_t n, int prec, uint64_t* red, uint32_t G,
                               uint64_t* states, uint64_t state_words, int has_pivot,
                               double pivot, Params* prm, hipStream_t s, uint32_t wait_polls,
                               uint32_t hooks) {
    if (G == 0 || G > kMaxPartials) return hipErrorInvalidValue;
    if (prec != kFP32)
        k_reduce_final<true><<<G, kBlock, 0, s>>>(w, n, red, states, state_words, has_pivot,
                                                  pivot, prm, wait_polls, hooks);
    else
        k_reduce_final<false><<<G, kBlock, 0, s>>>(w, n, red, states, state_words, has_pivot,
                                                   pivot, prm, wait_polls, hooks);
    return hipGetLastError();
}

hipError_t launch_seg_reduce(const float* w, const SegItem* item_d, uint32_t items, int prec,
                             double* partials, uint64_t* states, hipStream_t s) {
    if (prec != kFP32)
        k_seg_reduce<true><<<items, kBlock, 0, s>>>(w, item_d, partials, states);
    else
        k_seg_reduce<false><<<items, kBlock, 0, s>>>(w, item_d, partials, states);
    return hipGetLastError();
}

hipError_t launch_seg_finalize(const uint64_t* seg_off, const uint32_t* seg_item0, uint32_t nseg,
                               int prec, const double* partials, Params* seg_prm,
                               uint32_t* ticket, hipStream_t s) {
    if (prec != kFP32)
        k_seg_finalize<true><<<nseg, kBlock, 0, s>>>(seg_off, seg_item0, partials, seg_prm, ticket);
    else
        k_seg_finalize<false><<<nseg, kBlock, 0, s>>>(seg_off, seg_item0, partials, seg_prm, ticket);
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
