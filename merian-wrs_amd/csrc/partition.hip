// partition.hip -- heavy/light prefix-partition (pass 2 of a PSA build).
//
// Semantics follow the reference kernel src/device/prefix_partition/decoupled/
// shader.comp:313-362 and its CPU counterpart src/host/reference/partition.hpp:35-53:
//   heavy <=> w > pivot (strict), stable order;
//   idx[0..H)        heavy element indices, ascending;
//   idx[N-1-r]       the r-th light element index (lights stored reversed);
//   prefix[same pos] inclusive prefix of the heavy (resp. light) stream;
//   heavy_count = H.
//
// MI355X design.  A tile is 4096 elements handled by one 256-thread block
// (4 wave64s):
//   * coalesced float4 rows at any input alignment (tileload.hpp), transposed
//     through LDS (row stride 17 words, so both the transpose writes and the
//     per-thread reads avoid bank conflicts) so that each thread owns 16
//     consecutive elements;
//   * per-thread sequential sums in the mode's tile arithmetic (tilesum.hpp),
//     wave64 shuffle scan, then a 4-entry cross-wave scan in LDS, all in a
//     fixed order;
//   * tile carries are 126-bit fixed-point sums (exact, so their bits do not
//     depend on the order in which tiles are combined) plus an exact count;
//   * outputs are staged in LDS in final order (heavies by rank, lights by
//     reversed rank) and written as two contiguous coalesced runs.
//
// Two forms, for a single distribution and per segment of a batch (the tuner
// picks one per problem):
//   * reduce-then-scan: the aggregate of every tile, a scan of the
//     aggregates, then an emit pass that re-reads the tile.  The 8 XCDs of
//     the chip have separate L2 caches, so a single-pass chained scan pays a
//     device-coherent round trip per look-back step on every tile; one extra
//     streaming read of the weights is cheaper than that at large N;
//   * single pass with decoupled look-back (the reference's form): one kernel,
//     better where the chain of tiles is short.
//
// All forms compute identical tile aggregates and carries, so they give the
// same bits, and a segment built in a batch is bit-identical to the same
// distribution built alone.
#include "device_common.hpp"
#include "internal.hpp"
#include "lookback.hpp"
#include "tileload.hpp"
#include "tilesum.hpp"

namespace wrsk {
using namespace wrsdev;

namespace {

constexpr int kPad = kTile + kTile / 16;  // LDS words of the padded transpose buffer

// Tile carry: heavy sum (2 words), light sum (2 words), heavy count.  In the
// look-back records the aggregate sits at words 0..4 and the inclusive prefix
// at 8..12 of the tile's record.  FP32 mode keeps each sum in one float (word
// .lo; .hi = 0).
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

__device__ __forceinline__ Carry carry_shfl_up(Carry v, int d) {
    Carry o;
    o.h.lo = __shfl_up(v.h.lo, (unsigned)d, kWave);
    o.h.hi = __shfl_up(v.h.hi, (unsigned)d, kWave);
    o.l.lo = __shfl_up(v.l.lo, (unsigned)d, kWave);
    o.l.hi = __shfl_up(v.l.hi, (unsigned)d, kWave);
    o.c = __shfl_up(v.c, (unsigned)d, kWave);
    return o;
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

// Plain (kernel-ordered) record access for the reduce-then-scan path.
__device__ __forceinline__ Carry load_carry(const uint64_t* rec) {
    return Carry{FX{rec[0], rec[1]}, FX{rec[2], rec[3]}, rec[4]};
}
__device__ __forceinline__ void store_carry(uint64_t* rec, const Carry& v) {
    rec[0] = v.h.lo;
    rec[1] = v.h.hi;
    rec[2] = v.l.lo;
    rec[3] = v.l.hi;
    rec[4] = v.c;
}

template <int PREC>
struct CarryOps {
    static __device__ __forceinline__ Carry zero() { return carry_zero(); }
    static __device__ __forceinline__ Carry add(Carry a, Carry b) { return carry_add<PREC>(a, b); }
    static __device__ __forceinline__ Carry wave_sum(Carry v) { return wave_sum_carry<PREC>(v); }
    static __device__ __forceinline__ bool read(const uint64_t* r, Carry& v) { return read_rec(r, v); }
    static __device__ __forceinline__ void publish(uint64_t* r, const Carry& v) { publish_rec(r, v); }
};

// In-tile accumulator: the heavy and light running sums in the mode's
// arithmetic (tilesum.hpp) and the heavy count.
template <int PREC>
struct Acc {
    typename Acc1<PREC>::T h, l;
    uint32_t c;
};

template <int PREC>
__device__ __forceinline__ Acc<PREC> acc_zero() {
    return Acc<PREC>{Acc1<PREC>::zero(), Acc1<PREC>::zero(), 0u};
}

template <int PREC>
__device__ __forceinline__ Acc<PREC> acc_add(Acc<PREC> a, Acc<PREC> b) {
    return Acc<PREC>{Acc1<PREC>::add(a.h, b.h), Acc1<PREC>::add(a.l, b.l), a.c + b.c};
}

template <int PREC>
__device__ __forceinline__ Acc<PREC> acc_shfl_up(Acc<PREC> v, int d) {
    return Acc<PREC>{Acc1<PREC>::shfl_up(v.h, d), Acc1<PREC>::shfl_up(v.l, d), shfl_up(v.c, d)};
}

// v + (v DPP-moved by CTRL), the moved value the left operand; lanes that
// receive nothing add zero (an exact no-op on the sums).
template <int PREC, int CTRL, int RM>
__device__ __forceinline__ Acc<PREC> acc_add_dpp(const Acc<PREC>& v) {
    using A1 = Acc1<PREC>;
    const Acc<PREC> o{A1::template dpp<CTRL, RM>(v.h), A1::template dpp<CTRL, RM>(v.l),
                      dpp_u32<CTRL, RM>(v.c)};
    return acc_add<PREC>(o, v);
}

// Tile aggregate as a carry word set.
template <int PREC>
__device__ __forceinline__ Carry to_carry(const Acc<PREC>& a, int shift) {
    Carry r;
    if constexpr (PREC == kFP32) {
        r.h = FX{(uint64_t)__float_as_uint(a.h), 0};
        r.l = FX{(uint64_t)__float_as_uint(a.l), 0};
    } else {
        r.h = fx_from_double(Acc1<PREC>::value(a.h), shift);
        r.l = fx_from_double(Acc1<PREC>::value(a.l), shift);
    }
    r.c = a.c;
    return r;
}

template <int PREC>
__device__ __forceinline__ double carry_value(const FX& v, int shift) {
    if constexpr (PREC == kFP32) return (double)__uint_as_float((uint32_t)v.lo);
    else return fx_to_double(v, shift);
}

// LDS of one tile block.
template <int PREC>
struct TileShared {
    float buf[kPad];
    Acc<PREC> wave[kBlock / kWave];
    uint32_t cnt[kBlock / kWave];
};

// Loads the m (<= kTile) elements of w[base..) so that thread t owns elements
// 16t..16t+15 (zero beyond m).  Ends with s.buf still in use (the caller
// synchronises before reusing it).
template <int PREC>
__device__ __forceinline__ void load_thread_elems(const float* __restrict__ w, uint32_t base,
                                                  uint32_t m, TileShared<PREC>& s,
                                                  float (&v)[kItems]) {
    const int t = threadIdx.x;
    load_tile_lds<true>(w + base, m, s.buf);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = s.buf[17 * t + k];
}

// The thread's heavy and light sums of its elements (heavy: bit k of hmask) in
// element order, in the mode's tile arithmetic (tilesum.hpp), with the running
// sums captured after element kh (heavy) and kl (light); -1 captures nothing.
template <int PREC>
__device__ __forceinline__ Acc<PREC> run_sums(const float (&v)[kItems], uint32_t hmask, int kh,
                                              int kl, typename Acc1<PREC>::T& hcap,
                                              typename Acc1<PREC>::T& lcap) {
    using A1 = Acc1<PREC>;
    Acc<PREC> r = acc_zero<PREC>();
    hcap = A1::zero();
    lcap = A1::zero();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const bool heavy = (hmask >> k) & 1u;
        add_split<PREC>(r.h, r.l, heavy, v[k]);
        hcap = A1::sel(k == kh, r.h, hcap);
        lcap = A1::sel(k == kl, r.l, lcap);
        // taken here: otherwise the compiler keeps all 16 running sums live
        // to select one at the end (126 instead of 66 VGPRs in the emit)
        A1::pin(hcap);
        A1::pin(lcap);
    }
    r.c = (uint32_t)__popc(hmask);
    return r;
}

// Heavy / light masks of the thread's elements (only elements < m count).
__device__ __forceinline__ void classify(const float (&v)[kItems], float pivot, uint32_t e0,
                                         uint32_t m, uint32_t& hmask, uint32_t& lmask) {
    uint32_t h = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) h |= (v[k] > pivot ? 1u : 0u) << k;
    const uint32_t nvalid = e0 >= m ? 0u : min((uint32_t)kItems, m - e0);
    const uint32_t valid = (1u << nvalid) - 1u;  // nvalid <= 16
    hmask = h & valid;
    lmask = ~h & valid;
}

// Position of the j-th (from 0) set bit of a 16-bit mask holding more than j bits.
__device__ __forceinline__ int nth_bit(uint32_t mask, uint32_t j) {
    int pos = 0;
    uint32_t c = (uint32_t)__popc(mask & 0xffu);
    if (j >= c) { j -= c; mask >>= 8; pos += 8; }
    c = (uint32_t)__popc(mask & 0xfu);
    if (j >= c) { j -= c; mask >>= 4; pos += 4; }
    c = (uint32_t)__popc(mask & 0x3u);
    if (j >= c) { j -= c; mask >>= 2; pos += 2; }
    c = mask & 1u;
    if (j >= c) pos += 1;
    return pos;
}

// Block-wide exclusive heavy count of this thread (c = its own count) and
// the tile's count.  Synchronises the block.
template <int PREC>
__device__ __forceinline__ void count_scan(uint32_t c, TileShared<PREC>& s, uint32_t& excl,
                                           uint32_t& total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t incl = wave_incl_scan_u32(c);
    if (lane == kWave - 1) s.cnt[wid] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q) {
        before += q < wid ? s.cnt[q] : 0u;
        all += s.cnt[q];
    }
    excl = before + incl - c;
    total = all;
}

// Tile total from the per-wave totals in s.wave (block synchronised), the
// waves in order; with wave_excl the total of the waves before wave wid (the
// prefixes of the same chain of additions).
template <int PREC>
__device__ __forceinline__ Acc<PREC> combine_waves(const TileShared<PREC>& s, int wid,
                                                   Acc<PREC>* wave_excl) {
    Acc<PREC> c = s.wave[0];
    if (wave_excl) *wave_excl = wid == 0 ? acc_zero<PREC>() : c;
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q) {
        c = acc_add<PREC>(c, s.wave[q]);
        if (wave_excl && wid == q + 1) *wave_excl = c;
    }
    return c;
}

// Inclusive wave scan of the per-thread sums by DPP row steps (within rows of
// 16 by shifts of 1, 2, 4, 8, then row 2's total into row 3 and rows 0-1's
// into rows 2-3), the moved value the left operand.  Lane 63 receives the
// wave total by the same additions, in the same order, as the 64-lane
// Hillis-Steele scan, so tile totals built on it have the bits of tile_total.
template <int PREC>
__device__ __forceinline__ Acc<PREC> wave_scan_dpp(Acc<PREC> x) {
    x = acc_add_dpp<PREC, 0x111, 0xf>(x);  // row_shr:1
    x = acc_add_dpp<PREC, 0x112, 0xf>(x);  // row_shr:2
    x = acc_add_dpp<PREC, 0x114, 0xf>(x);  // row_shr:4
    x = acc_add_dpp<PREC, 0x118, 0xf>(x);  // row_shr:8
    x = acc_add_dpp<PREC, 0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
    x = acc_add_dpp<PREC, 0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
    return x;
}

// The tile total alone (the aggregate passes).  Ends with the block
// synchronised and s.buf free.
template <int PREC>
__device__ __forceinline__ Acc<PREC> tile_total(const float* __restrict__ w, uint32_t base,
                                                uint32_t m, float pivot, TileShared<PREC>& s) {
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    float v[kItems];
    load_thread_elems<PREC>(w, base, m, s, v);
    // the sums of run_sums (elements past m are zero: adding them changes no
    // bits, whatever their class), the count of the elements below m only
    const uint32_t e0 = (uint32_t)(kItems * t);
    Acc<PREC> x = acc_zero<PREC>();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const bool heavy = v[k] > pivot;
        add_split<PREC>(x.h, x.l, heavy, v[k]);
        x.c += heavy && e0 + k < m ? 1u : 0u;
    }
    x = wave_scan_dpp<PREC>(x);
    if (lane == kWave - 1) s.wave[wid] = x;
    __syncthreads();
    return combine_waves<PREC>(s, wid, nullptr);
}

// Per-thread state of a tile pass.
template <int PREC>
struct TileScan {
    using T = typename Acc1<PREC>::T;
    uint32_t hmask, lmask;  // bit k: element e0 + k (< m) is heavy / light
    uint32_t c_excl;        // heavies of the tile before this thread
    Acc<PREC> excl;         // sums (and count) of the tile before this thread
    Acc<PREC> agg;          // the tile's sums and count
    // sampled ranks: this thread's heavy and light whose inclusive prefix is a
    // sample (rank + 1 a multiple of kSample) and its light of rank a multiple
    // of kSample (local element, -1: none), and the ranks themselves
    int kh, kl, kp;
    uint32_t rh, rl, rp;
    T hcap, lcap;           // the thread's running sums through kh / kl
};

// Ranks of the thread's sampled elements, given the heavies (Hx) and lights
// (Lx) before the tile.
template <int PREC>
__device__ __forceinline__ void sample_ranks(TileScan<PREC>& ts, uint32_t Hx, uint32_t Lx,
                                             uint32_t e0) {
    const uint32_t nh = (uint32_t)__popc(ts.hmask), nl = (uint32_t)__popc(ts.lmask);
    const uint32_t hr0 = Hx + ts.c_excl;         // rank of the thread's first heavy
    const uint32_t lr0 = Lx + (e0 - ts.c_excl);  // rank of its first light (if any)
    const uint32_t jh = (kSample - 1u - hr0) & (kSample - 1u);
    const uint32_t jl = (kSample - 1u - lr0) & (kSample - 1u);
    const uint32_t jp = (0u - lr0) & (kSample - 1u);
    ts.kh = jh < nh ? nth_bit(ts.hmask, jh) : -1;
    ts.kl = jl < nl ? nth_bit(ts.lmask, jl) : -1;
    ts.kp = jp < nl ? nth_bit(ts.lmask, jp) : -1;
    ts.rh = hr0 + jh;
    ts.rl = lr0 + jl;
    ts.rp = lr0 + jp;
}

// One tile pass: loads the tile, classifies, scans the heavy counts, then (with
// `carry_known`, the heavies before the tile in Hx) sums the thread's elements
// once in the tile arithmetic, capturing the running sums at its sampled ranks;
// finally the block scan of the sums (ts.excl, ts.agg).  Without
// `carry_known` (the look-back forms, whose carry follows from the aggregate)
// nothing is captured: call recapture() once the carry is known.  Ends with
// the block synchronised and s.buf free.
template <int PREC>
__device__ __forceinline__ void tile_pass(const float* __restrict__ w, uint32_t base, uint32_t m,
                                          float pivot, bool carry_known, uint32_t Hx,
                                          TileShared<PREC>& s, float (&v)[kItems],
                                          TileScan<PREC>& ts) {
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    const uint32_t e0 = (uint32_t)(kItems * t);
    load_thread_elems<PREC>(w, base, m, s, v);
    classify(v, pivot, e0, m, ts.hmask, ts.lmask);
    uint32_t c_tile;
    count_scan<PREC>((uint32_t)__popc(ts.hmask), s, ts.c_excl, c_tile);
    ts.kh = ts.kl = ts.kp = -1;
    if (carry_known) sample_ranks<PREC>(ts, Hx, base - Hx, e0);
    const Acc<PREC> mine = run_sums<PREC>(v, ts.hmask, ts.kh, ts.kl, ts.hcap, ts.lcap);
    const Acc<PREC> incl = wave_scan_dpp<PREC>(mine);
    using A1 = Acc1<PREC>;
    Acc<PREC> lane_excl{A1::template dpp<0x138, 0xf>(incl.h), A1::template dpp<0x138, 0xf>(incl.l),
                        0u};  // wave_shr:1, lane 0 receives zero
    if (lane == kWave - 1) s.wave[wid] = incl;
    __syncthreads();
    Acc<PREC> wave_excl;
    ts.agg = combine_waves<PREC>(s, wid, &wave_excl);
    ts.agg.c = c_tile;
    ts.excl = acc_add<PREC>(wave_excl, lane_excl);
    ts.excl.c = ts.c_excl;
}

// The captures of a pass run without a known carry.
template <int PREC>
__device__ __forceinline__ void recapture(const float (&v)[kItems], uint32_t Hx, uint32_t base,
                                          TileScan<PREC>& ts) {
    sample_ranks<PREC>(ts, Hx, base - Hx, (uint32_t)(kItems * threadIdx.x));
    (void)run_sums<PREC>(v, ts.hmask, ts.kh, ts.kl, ts.hcap, ts.lcap);
}

// Carries of one tile: the sums and count of everything before it.
struct TileCarry {
    double h, l;   // FP64 modes
    float hf, lf;  // FP32 mode
    uint32_t c;
};

template <int PREC>
__device__ __forceinline__ TileCarry tile_carry(const Carry& excl, int shift) {
    TileCarry r;
    if constexpr (PREC == kFP32) {
        r.hf = __uint_as_float((uint32_t)excl.h.lo);
        r.lf = __uint_as_float((uint32_t)excl.l.lo);
        r.h = r.l = 0.0;
    } else {
        r.h = fx_to_double(excl.h, shift);
        r.l = fx_to_double(excl.l, shift);
        r.hf = r.lf = 0.0f;
    }
    r.c = (uint32_t)excl.c;
    return r;
}

// Inclusive prefix of the heavy (light) stream at an element of this thread
// whose thread-local running sum is `run`: the tile carry plus (the sum of the
// thread's predecessors in the tile plus `run`), FP64 in the FP64-carry modes.
template <int PREC>
__device__ __forceinline__ double prefix_at(bool heavy, const TileCarry& cy,
                                            const TileScan<PREC>& ts,
                                            const typename Acc1<PREC>::T& run) {
    using A1 = Acc1<PREC>;
    const typename A1::T part = A1::add(heavy ? ts.excl.h : ts.excl.l, run);
    if constexpr (PREC == kFP32) return (double)((heavy ? cy.hf : cy.lf) + part);
    else return (heavy ? cy.h : cy.l) + A1::value(part);
}

// Tiles of the direct form with at most this many heavies write them straight
// from the registers (a few store instructions, mostly with few lanes active);
// denser tiles stage them in LDS and write them as a run of consecutive
// words: on a 2^26 build, 94 vs 103 us for Zipf(1.1) weights (2.7% heavy) and
// 267 vs 119 us for U(0, 1) weights (half heavy) with the other choice.
constexpr uint32_t kRegHeavies = 512;

// Writes the tile's outputs given its carries (block-uniform): the partition
// indices, the weights in partition order (pw_out, when non-null), the prefix
// samples at ranks r with (r+1) % kSample == 0 (from the pass's captures),
// and with FULL the hi (and lo) prefix arrays of every element.  With DIRECT
// only the heavies' indices and weights are written (their positions [0, H)
// of idx/pw), and for every
// light rank r with r % kSample == 0 its element index goes to lpos[r /
// kSample]: the merge pack then gathers the lights from the weights
// themselves.  Expects s.buf free (the block synchronised after its last use).
template <int PREC, bool FULL, bool DIRECT = false>
__device__ __forceinline__ void tile_emit(uint32_t n, uint32_t base, uint32_t m,
                                          const float (&v)[kItems], const TileScan<PREC>& ts,
                                          const TileCarry& cy, TileShared<PREC>& s,
                                          uint32_t* __restrict__ idx_out,
                                          float* __restrict__ pw_out,
                                          float* __restrict__ hi_out, float* __restrict__ lo_out,
                                          double* __restrict__ samp_h,
                                          double* __restrict__ samp_l,
                                          uint32_t* __restrict__ lpos = nullptr) {
    static_assert(!(FULL && DIRECT), "the direct form writes no full prefix arrays");
    const int t = threadIdx.x;
    const uint32_t e0 = (uint32_t)(kItems * t);
    const uint32_t hc = ts.agg.c;   // heavies of the tile
    const uint32_t Hx = cy.c;       // heavies before this tile
    const uint32_t Lx = base - Hx;  // lights before this tile
    uint32_t* s_u = reinterpret_cast<uint32_t*>(s.buf);

    // prefix samples (at most one heavy and one light per thread)
    if (samp_h != nullptr) {
        if (ts.kh >= 0) samp_h[(ts.rh + 1u) / kSample - 1u] = prefix_at<PREC>(true, cy, ts, ts.hcap);
        if (ts.kl >= 0) samp_l[(ts.rl + 1u) / kSample - 1u] = prefix_at<PREC>(false, cy, ts, ts.lcap);
    }
    if constexpr (DIRECT) {
        if (ts.kp >= 0) lpos[ts.rp / kSample] = base + e0 + (uint32_t)ts.kp;
        if (hc <= kRegHeavies) {  // few heavies: straight from the registers
            uint32_t hr = Hx + ts.c_excl;  // global rank of the next heavy
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                if ((ts.hmask >> k) & 1u) {
                    idx_out[hr] = base + e0 + k;
                    pw_out[hr] = v[k];
                }
                hr += (ts.hmask >> k) & 1u;
            }
            return;
        }
    }

    // LDS slot of each element: heavy run (by tile-local heavy rank), then the
    // light run reversed
    uint32_t slot[kItems];
    {
        uint32_t hr = ts.c_excl;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t e = e0 + k;
            const bool heavy = (ts.hmask >> k) & 1u;
            slot[k] = heavy ? hr : (m - 1u - (e - hr));
            hr += heavy ? 1u : 0u;
        }
    }
    // positions of LDS slot q: heavy run then light run
    auto out_pos = [&](uint32_t q) -> uint32_t { return q < hc ? Hx + q : n - m - Lx + q; };
    // the direct form writes the heavy run only: slots [0, hc)
    const uint32_t q_end = DIRECT ? hc : m;
    if (DIRECT && hc == 0) return;
    // round A: indices
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (e0 + k < m) s_u[slot[k]] = base + e0 + k;
    __syncthreads();
    for (uint32_t q = t; q < q_end; q += kBlock) idx_out[out_pos(q)] = s_u[q];
    // round A': weights
    if (pw_out != nullptr) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (e0 + k < m) s.buf[slot[k]] = v[k];
        __syncthreads();
        for (uint32_t q = t; q < q_end; q += kBlock) pw_out[out_pos(q)] = s.buf[q];
    }
    if constexpr (FULL) {
        // round B: the inclusive prefix of every element (the samples above
        // are the values of the same formula at the sampled elements)
        float hi_keep[kItems], lo_keep[kItems];
        using A1 = Acc1<PREC>;
        typename A1::T rh = A1::zero(), rl = A1::zero();
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const bool heavy = (ts.hmask >> k) & 1u;
            add_split<PREC>(rh, rl, heavy, v[k]);
            const double P = prefix_at<PREC>(heavy, cy, ts, heavy ? rh : rl);
            const float hi = (float)P;
            hi_keep[k] = hi;
            lo_keep[k] = prec_has_lo(PREC) ? (float)(P - (double)hi) : 0.0f;
        }
        if (hi_out != nullptr) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kItems; ++k)
                if (e0 + k < m) s.buf[slot[k]] = hi_keep[k];
            __syncthreads();
            for (uint32_t q = t; q < m; q += kBlock) hi_out[out_pos(q)] = s.buf[q];
            if constexpr (prec_has_lo(PREC)) {
                if (lo_out != nullptr) {
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < kItems; ++k)
                        if (e0 + k < m) s.buf[slot[k]] = lo_keep[k];
                    __syncthreads();
                    for (uint32_t q = t; q < m; q += kBlock) lo_out[out_pos(q)] = s.buf[q];
                }
            }
        }
    }
}

}  // namespace

// ---- single distribution: reduce-then-scan ---------------------------------


// Pass 2a: aggregate of every tile.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_part_agg(const float* __restrict__ w, uint32_t n,
                                                     const Params* __restrict__ prm,
                                                     uint64_t* __restrict__ states) {
    __shared__ TileShared<PREC> s;
    if (prm->error) return;  // an earlier wait of the build gave up
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const Acc<PREC> tile_agg = tile_total<PREC>(w, base, m, prm->pivot, s);
    if (threadIdx.x == 0)
        store_carry(states + (uint64_t)tile * kCarryWords, to_carry<PREC>(tile_agg, prm->shift));
}

// Pass 2b: exclusive scan of the tile aggregates.  Block g scans tiles
// [g*kSpineTiles, (g+1)*kSpineTiles) (thread t takes kSpinePer consecutive
// tiles, all loaded at once) and chains to its predecessors by decoupled
// look-back over per-block records; every aggregate already exists, so the
// chain resolves in about one round trip.  The last block writes the totals
// into prm.  The FP64 modes add exactly, so the result does not depend on the
// look-back order.
constexpr int kSpineThreads = 256;
constexpr int kSpinePer = 4;
static_assert(kSpineThreads * kSpinePer == (int)kSpineTiles, "spine block shape");

// KEYS: the aggregates are the streamed pack's exact key sums (scale
// 2^key_k) instead of the partition's carries (scale 2^shift).
template <int PREC, bool KEYS = false>
__global__ __launch_bounds__(kSpineThreads) void k_part_tilescan(uint64_t* __restrict__ states,
                                                                 uint32_t tiles,
                                                                 Params* __restrict__ prm) {
    __shared__ Carry s_wave[kSpineThreads / kWave];
    __shared__ Carry s_excl;
    __shared__ uint32_t s_blk;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wid = t / kWave;
    if (prm->error) return;
    if (t == 0) s_blk = atomicAdd(&prm->ticket, 1u);  // predecessors are scheduled first
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t nblk = (tiles + kSpineTiles - 1) / kSpineTiles;
    const uint64_t* agg = states;
    uint64_t* excl = states + (uint64_t)kExclBase * tiles;
    uint64_t* recs = states + (uint64_t)tiles * kStateWords;  // block records after the tiles'

    const uint32_t b = blk * kSpineTiles + (uint32_t)t * kSpinePer;
    Carry a[kSpinePer];
#pragma unroll
    for (int k = 0; k < kSpinePer; ++k)
        a[k] = b + k < tiles ? load_carry(agg + (uint64_t)(b + k) * kCarryWords) : carry_zero();
    Carry mine = a[0];
#pragma unroll
    for (int k = 1; k < kSpinePer; ++k) mine = carry_add<PREC>(mine, a[k]);
    Carry incl = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const Carry o = carry_shfl_up(incl, d);
        if (lane >= d) incl = carry_add<PREC>(o, incl);
    }
    Carry lane_excl = carry_shfl_up(incl, 1);
    if (lane == 0) lane_excl = carry_zero();
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        Carry blk_agg = s_wave[0];
#pragma unroll
        for (int q = 1; q < kSpineThreads / kWave; ++q) blk_agg = carry_add<PREC>(blk_agg, s_wave[q]);
        uint64_t* rec = recs + (uint64_t)blk * kStateWords;
        Carry ex = carry_zero();
        if (blk == 0) {
            if (lane == 0 && !(prm->hooks & kHookSkipSpine))
                publish_inclusive<Carry, CarryOps<PREC>>(rec, blk_agg);
        } else {
            if (lane == 0) publish_aggregate<Carry, CarryOps<PREC>>(rec, blk_agg);
            ex = lookback<Carry, CarryOps<PREC>>(recs, blk, prm->wait_polls, &prm->error);
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, carry_add<PREC>(ex, blk_agg));
        }
        if (lane == 0) {
            s_excl = ex;
            if (blk == nblk - 1) {
                const Carry tot = carry_add<PREC>(ex, blk_agg);
                const int shift = KEYS ? prm->key_k : prm->shift;
                prm->heavy_count = (uint32_t)tot.c;
                prm->heavy_sum = carry_value<PREC>(tot.h, shift);
                prm->light_sum = carry_value<PREC>(tot.l, shift);
            }
        }
    }
    __syncthreads();
    Carry run = s_excl;
    for (int q = 0; q < wid; ++q) run = carry_add<PREC>(run, s_wave[q]);
    run = carry_add<PREC>(run, lane_excl);
#pragma unroll
    for (int k = 0; k < kSpinePer; ++k) {
        if (b + k < tiles) store_carry(excl + (uint64_t)(b + k) * kCarryWords, run);
        run = carry_add<PREC>(run, a[k]);
    }
}

// Pass 2c: outputs of every tile from its exclusive prefix.
template <int PREC, bool FULL, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_part_emit(const float* __restrict__ w, uint32_t n,
                                                      const Params* __restrict__ prm,
                                                      const uint64_t* __restrict__ states,
                                                      uint32_t* __restrict__ idx_out,
                                                      float* __restrict__ pw_out,
                                                      float* __restrict__ hi_out,
                                                      float* __restrict__ lo_out,
                                                      double* __restrict__ samp_h,
                                                      double* __restrict__ samp_l,
                                                      uint32_t* __restrict__ lpos) {
    __shared__ TileShared<PREC> s;
    if (prm->error) return;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const float pivot = prm->pivot;
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);  // n + kTile - 1 may exceed 2^32
    const Carry excl =
        load_carry(states + (uint64_t)kExclBase * tiles + (uint64_t)tile * kCarryWords);
    const TileCarry cy = tile_carry<PREC>(excl, prm->shift);
    float v[kItems];
    TileScan<PREC> ts;
    tile_pass<PREC>(w, base, m, pivot, true, cy.c, s, v, ts);
    tile_emit<PREC, FULL, DIRECT>(n, base, m, v, ts, cy, s, idx_out, pw_out, hi_out, lo_out,
                                  samp_h, samp_l, lpos);
}

// ---- single distribution: single pass with decoupled look-back --------------
// The reference's form (prefix_partition/decoupled/shader.comp:96-362): tiles
// taken in ticket order, each publishes its aggregate, looks back over its
// predecessors' records for its exclusive prefix and emits.  One kernel
// instead of three, which pays where the chain of tiles is short; the tuner
// chooses between the forms.  Records and ticket are reset by the reduce and
// finalize passes.
template <int PREC, bool FULL, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_part_lookback(
    const float* __restrict__ w, uint32_t n, Params* __restrict__ prm,
    uint64_t* __restrict__ states, uint32_t* __restrict__ idx_out, float* __restrict__ pw_out,
    float* __restrict__ hi_out, float* __restrict__ lo_out, double* __restrict__ samp_h,
    double* __restrict__ samp_l, uint32_t* __restrict__ lpos) {
    __shared__ TileShared<PREC> s;
    __shared__ uint32_t s_tile;
    __shared__ Carry s_excl;
    if (prm->error) return;
    if (threadIdx.x == 0) s_tile = atomicAdd(&prm->ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const float pivot = prm->pivot;
    const int shift = prm->shift;
    float v[kItems];
    TileScan<PREC> ts;
    tile_pass<PREC>(w, base, m, pivot, false, 0u, s, v, ts);

    const int lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x < kWave) {
        uint64_t* rec = states + (uint64_t)tile * kStateWords;
        const Carry agg = to_carry<PREC>(ts.agg, shift);
        Carry excl = carry_zero();
        if (tile == 0) {
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, agg);
        } else {
            if (lane == 0) publish_aggregate<Carry, CarryOps<PREC>>(rec, agg);
            excl = lookback<Carry, CarryOps<PREC>>(states, tile, prm->wait_polls, &prm->error);
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, carry_add<PREC>(excl, agg));
        }
        if (lane == 0) {
            s_excl = excl;
            if (base + m == n) {
                const Carry tot = carry_add<PREC>(excl, agg);
                prm->heavy_count = (uint32_t)tot.c;
                prm->heavy_sum = carry_value<PREC>(tot.h, shift);
                prm->light_sum = carry_value<PREC>(tot.l, shift);
            }
        }
    }
    __syncthreads();
    const TileCarry cy = tile_carry<PREC>(s_excl, shift);
    recapture<PREC>(v, cy.c, base, ts);
    tile_emit<PREC, FULL, DIRECT>(n, base, m, v, ts, cy, s, idx_out, pw_out, hi_out, lo_out,
                                  samp_h, samp_l, lpos);
}

// ---- segmented: single pass with decoupled look-back -----------------------

// One work item per (segment, tile), items of a segment are consecutive and
// handed out in ticket order, so a tile's predecessors in its segment always
// hold smaller tickets.  Segment outputs use segment-local indices and occupy
// the segment's own range of idx/pw.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_seg_partition(
    const float* __restrict__ w, const SegItem* __restrict__ item_d,
    const uint32_t* __restrict__ seg_item0, Params* __restrict__ seg_prm,
    uint32_t* __restrict__ ticket, uint64_t* __restrict__ states, uint32_t* __restrict__ idx,
    float* __restrict__ pw, double* __restrict__ samp_h, double* __restrict__ samp_l,
    const uint64_t* __restrict__ seg_samp0) {
    __shared__ TileShared<PREC> s;
    __shared__ uint32_t s_item;
    __shared__ Carry s_excl;
    if (threadIdx.x == 0) s_item = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    const SegItem it = item_d[item];
    const uint32_t sg = it.seg;
    const uint64_t off = it.off;
    const uint32_t n = it.n;
    const uint64_t so = seg_samp0[sg];
    const uint32_t tile = it.blk;
    Params* prm = seg_prm + sg;
    uint64_t* seg_states = states + (uint64_t)seg_item0[sg] * kStateWords;

    const uint32_t base = tile * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const float pivot = prm->pivot;
    const int shift = prm->shift;
    float v[kItems];
    TileScan<PREC> ts;
    tile_pass<PREC>(w + off, base, m, pivot, false, 0u, s, v, ts);

    const int lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x < kWave) {
        uint64_t* rec = seg_states + (uint64_t)tile * kStateWords;
        const Carry agg = to_carry<PREC>(ts.agg, shift);
        Carry excl = carry_zero();
        if (tile == 0) {
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, agg);
        } else {
            if (lane == 0) publish_aggregate<Carry, CarryOps<PREC>>(rec, agg);
            excl = lookback<Carry, CarryOps<PREC>>(seg_states, tile, prm->wait_polls, &prm->error);
            if (lane == 0) publish_inclusive<Carry, CarryOps<PREC>>(rec, carry_add<PREC>(excl, agg));
        }
        if (lane == 0) {
            s_excl = excl;
            if (base + m == n) {
                const Carry tot = carry_add<PREC>(excl, agg);
                prm->heavy_count = (uint32_t)tot.c;
                prm->heavy_sum = carry_value<PREC>(tot.h, shift);
                prm->light_sum = carry_value<PREC>(tot.l, shift);
            }
        }
    }
    __syncthreads();
    const TileCarry cy = tile_carry<PREC>(s_excl, shift);
    recapture<PREC>(v, cy.c, base, ts);
    tile_emit<PREC, false>(n, base, m, v, ts, cy, s, idx + off, pw + off, nullptr, nullptr,
                           samp_h + so, samp_l + so);
}

// ---- segmented: reduce-then-scan ------------------------------------------
// The same three steps as the single-distribution form, per segment: the
// aggregate of every (segment, tile) item (dense records, item order), an
// exclusive scan of each segment's aggregates by one wave, then the emit.
// No block waits on another, which at batch scale beats the single pass above
// (whose blocks hold their CU during the look-back round trip) by more than
// the extra read of the weights costs.  The FP64-mode carries are exact, so
// both forms -- and the single-distribution build -- give the same bits.

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_seg_part_agg(const float* __restrict__ w,
                                                         const SegItem* __restrict__ item_d,
                                                         const Params* __restrict__ seg_prm,
                                                         uint64_t* __restrict__ states) {
    __shared__ TileShared<PREC> s;
    const uint32_t item = blockIdx.x;
    const SegItem it = item_d[item];
    const uint64_t off = it.off;
    const uint32_t n = it.n;
    const uint32_t base = it.blk * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const Params* prm = seg_prm + it.seg;
    const Acc<PREC> tile_agg = tile_total<PREC>(w + off, base, m, prm->pivot, s);
    if (threadIdx.x == 0)
        store_carry(states + (uint64_t)item * kCarryWords, to_carry<PREC>(tile_agg, prm->shift));
}

// One wave per segment: exclusive prefixes of the segment's item aggregates,
// kSegSpinePer consecutive items per lane per round; the segment totals go to
// its Params.
constexpr int kSegSpinePer = 4;

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_seg_part_spine(const uint32_t* __restrict__ seg_item0,
                                                           uint32_t nseg, uint32_t items,
                                                           Params* __restrict__ seg_prm,
                                                           uint64_t* __restrict__ states) {
    const uint32_t sg = blockIdx.x * (uint32_t)(kBlock / kWave) + threadIdx.x / kWave;
    if (sg >= nseg) return;  // whole wave
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t i0 = seg_item0[sg], i1 = seg_item0[sg + 1];
    const uint64_t* agg = states;
    uint64_t* excl = states + (uint64_t)kExclBase * items;
    Carry run = carry_zero();
    for (uint32_t c0 = i0; c0 < i1; c0 += (uint32_t)(kWave * kSegSpinePer)) {
        const uint32_t b = c0 + (uint32_t)lane * kSegSpinePer;
        Carry a[kSegSpinePer];
#pragma unroll
        for (int k = 0; k < kSegSpinePer; ++k)
            a[k] = b + k < i1 ? load_carry(agg + (uint64_t)(b + k) * kCarryWords) : carry_zero();
        Carry mine = a[0];
#pragma unroll
        for (int k = 1; k < kSegSpinePer; ++k) mine = carry_add<PREC>(mine, a[k]);
        Carry incl = mine;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const Carry o = carry_shfl_up(incl, d);
            if (lane >= d) incl = carry_add<PREC>(o, incl);
        }
        Carry lex = carry_shfl_up(incl, 1);
        if (lane == 0) lex = carry_zero();
        Carry r = carry_add<PREC>(run, lex);
#pragma unroll
        for (int k = 0; k < kSegSpinePer; ++k) {
            if (b + k < i1) store_carry(excl + (uint64_t)(b + k) * kCarryWords, r);
            r = carry_add<PREC>(r, a[k]);
        }
        Carry chunk;
        chunk.h.lo = __shfl(incl.h.lo, kWave - 1, kWave);
        chunk.h.hi = __shfl(incl.h.hi, kWave - 1, kWave);
        chunk.l.lo = __shfl(incl.l.lo, kWave - 1, kWave);
        chunk.l.hi = __shfl(incl.l.hi, kWave - 1, kWave);
        chunk.c = __shfl(incl.c, kWave - 1, kWave);
        run = carry_add<PREC>(run, chunk);
    }
    if (lane == 0) {
        Params* prm = seg_prm + sg;
        const int shift = prm->shift;
        prm->heavy_count = (uint32_t)run.c;
        prm->heavy_sum = carry_value<PREC>(run.h, shift);
        prm->light_sum = carry_value<PREC>(run.l, shift);
    }
}

// DIRECT: the direct pack form (heavies' idx + pw and the segment's light
// positions at lpos + seg_lpos0[segment], as the single-distribution emit).
template <int PREC, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_seg_part_emit(
    const float* __restrict__ w, const SegItem* __restrict__ item_d, uint32_t items,
    const Params* __restrict__ seg_prm, const uint64_t* __restrict__ states,
    uint32_t* __restrict__ idx, float* __restrict__ pw, double* __restrict__ samp_h,
    double* __restrict__ samp_l, const uint64_t* __restrict__ seg_samp0,
    uint32_t* __restrict__ lpos, const uint64_t* __restrict__ seg_lpos0) {
    __shared__ TileShared<PREC> s;
    const uint32_t item = blockIdx.x;
    const SegItem it = item_d[item];
    const uint32_t sg = it.seg;
    const uint64_t off = it.off;
    const uint32_t n = it.n;
    const uint64_t so = seg_samp0[sg];
    const uint32_t base = it.blk * (uint32_t)kTile;
    const uint32_t m = min((uint32_t)kTile, n - base);
    const Params* prm = seg_prm + sg;
    const float pivot = prm->pivot;
    const Carry excl = load_carry(states + (uint64_t)kExclBase * items + (uint64_t)item * kCarryWords);
    const TileCarry cy = tile_carry<PREC>(excl, prm->shift);
    float v[kItems];
    TileScan<PREC> ts;
    tile_pass<PREC>(w + off, base, m, pivot, true, cy.c, s, v, ts);
    tile_emit<PREC, false, DIRECT>(n, base, m, v, ts, cy, s, idx + off, pw + off, nullptr,
                                   nullptr, samp_h + so, samp_l + so,
                                   DIRECT ? lpos + seg_lpos0[sg] : nullptr);
}

namespace {

template <int P>
void launch_part_t(const float* w, uint32_t n, int lookback, Params* prm, uint64_t* states,
                   uint32_t* idx, float* pw, float* hi, float* lo, double* samp_h,
                   double* samp_l, uint32_t* lpos, hipStream_t s, hipEvent_t ev_agg,
                   hipEvent_t ev_scan) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    if (lookback) {  // one kernel: the aggregate and tile-scan phases are empty
        if (ev_agg) (void)hipEventRecord(ev_agg, s);
        if (ev_scan) (void)hipEventRecord(ev_scan, s);
        if (hi)
            k_part_lookback<P, true, false><<<tiles, kBlock, 0, s>>>(
                w, n, prm, states, idx, pw, hi, lo, samp_h, samp_l, nullptr);
        else if (lpos)
            k_part_lookback<P, false, true><<<tiles, kBlock, 0, s>>>(
                w, n, prm, states, idx, pw, nullptr, nullptr, samp_h, samp_l, lpos);
        else
            k_part_lookback<P, false, false><<<tiles, kBlock, 0, s>>>(
                w, n, prm, states, idx, pw, nullptr, nullptr, samp_h, samp_l, nullptr);
        return;
    }
    k_part_agg<P><<<tiles, kBlock, 0, s>>>(w, n, prm, states);
    if (ev_agg) (void)hipEventRecord(ev_agg, s);
    k_part_tilescan<P><<<(tiles + kSpineTiles - 1) / kSpineTiles, kSpineThreads, 0, s>>>(
        states, tiles, prm);
    if (ev_scan) (void)hipEventRecord(ev_scan, s);
    if (hi)
        k_part_emit<P, true, false><<<tiles, kBlock, 0, s>>>(w, n, prm, states, idx, pw, hi, lo,
                                                             samp_h, samp_l, nullptr);
    else if (lpos)
        k_part_emit<P, false, true><<<tiles, kBlock, 0, s>>>(w, n, prm, states, idx, pw, nullptr,
                                                             nullptr, samp_h, samp_l, lpos);
    else
        k_part_emit<P, false, false><<<tiles, kBlock, 0, s>>>(w, n, prm, states, idx, pw, nullptr,
                                                              nullptr, samp_h, samp_l, nullptr);
}

}  // namespace

hipError_t launch_partition(const float* w, uint64_t n, int prec, int lookback, Params* prm,
                            uint64_t* states, uint32_t* idx, float* pw, float* hi, float* lo,
                            double* samp_h, double* samp_l, uint32_t* lpos, hipStream_t s,
                            hipEvent_t ev_agg, hipEvent_t ev_scan) {
    const uint32_t nn = (uint32_t)n;
#define WRS_PART(P)                                                                      \
    launch_part_t<P>(w, nn, lookback, prm, states, idx, pw, hi, lo, samp_h, samp_l, lpos, s, \
                     ev_agg, ev_scan)
    switch (prec) {
        case kFP32: WRS_PART(kFP32); break;
        case kFP64Carry: WRS_PART(kFP64Carry); break;
        case kComp: WRS_PART(kComp); break;
        default: WRS_PART(kFP64); break;
    }
#undef WRS_PART
    return hipGetLastError();
}

hipError_t launch_key_tilescan(uint64_t* states, uint64_t n, Params* prm, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    k_part_tilescan<kFP64Carry, true><<<(tiles + kSpineTiles - 1) / kSpineTiles, kSpineThreads, 0,
                                          s>>>(states, tiles, prm);
    return hipGetLastError();
}

hipError_t launch_seg_partition(const float* w, const SegItem* item_d, const uint32_t* seg_item0,
                                uint32_t nseg,
                                uint32_t items, int prec, int lookback, Params* seg_prm,
                                uint32_t* ticket, uint64_t* states, uint32_t* idx, float* pw,
                                double* samp_h, double* samp_l, const uint64_t* seg_samp0,
                                hipStream_t s, uint32_t* lpos, const uint64_t* seg_lpos0) {
    if (lpos && lookback) return hipErrorInvalidValue;  // direct: reduce-then-scan only
    const uint32_t spine_blocks = (nseg + kBlock / kWave - 1) / (kBlock / kWave);
#define WRS_LAUNCH_SEG(P)                                                                        \
    if (lookback) {                                                                              \
        k_seg_partition<P><<<items, kBlock, 0, s>>>(w, item_d, seg_item0,                        \
                                                    seg_prm, ticket, states, idx, pw, samp_h,    \
                                                    samp_l, seg_samp0);                          \
    } else {                                                                                     \
        k_seg_part_agg<P><<<items, kBlock, 0, s>>>(w, item_d, seg_prm, states);                  \
        k_seg_part_spine<P><<<spine_blocks, kBlock, 0, s>>>(seg_item0, nseg, items, seg_prm,     \
                                                            states);                             \
        if (lpos)                                                                                \
            k_seg_part_emit<P, true><<<items, kBlock, 0, s>>>(w, item_d, items, seg_prm, states, \
                                                              idx, pw, samp_h, samp_l, seg_samp0, \
                                                              lpos, seg_lpos0);                   \
        else                                                                                     \
            k_seg_part_emit<P, false><<<items, kBlock, 0, s>>>(w, item_d, items, seg_prm, states, \
                                                               idx, pw, samp_h, samp_l,           \
                                                               seg_samp0, nullptr, nullptr);      \
    }
    switch (prec) {
        case kFP32: WRS_LAUNCH_SEG(kFP32); break;
        case kFP64Carry: WRS_LAUNCH_SEG(kFP64Carry); break;
        case kComp: WRS_LAUNCH_SEG(kComp); break;
        default: WRS_LAUNCH_SEG(kFP64); break;
    }
#undef WRS_LAUNCH_SEG
    return hipGetLastError();
}

}  // namespace wrsk
