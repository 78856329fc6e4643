// internal.hpp -- device-resident build parameters, tile geometry and the
// host-side launchers each kernel translation unit exports.  Not part of the
// public C ABI (include/wrs.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wrsk {

// Arithmetic of a build (the C ABI's wrs_precision values).
enum Prec : int {
    kFP32 = 0,       // FP32 everywhere, FP32 look-back carry (the reference GPU's arithmetic)
    kFP64Carry = 1,  // FP32 in-tile scan, exact fixed-point carry read out in FP64, FP64 mean/split
    kComp = 2,       // compensated FP32 in-tile sums (two_sum hi+lo) + the kFP64Carry carry and split
    kFP64 = 3,       // FP64 in-tile sums + the kFP64Carry carry and split
};
// the modes whose prefixes are exported as a hi + lo float pair
constexpr bool prec_has_lo(int prec) { return prec >= kComp; }

// Per-build scalars, resident in device memory so that a whole build can be
// recorded on a stream (or captured in a graph) without host round trips.
struct Params {
    double total;          // sum of the weights (FP32 mode: an FP32 value)
    double mean;           // total / N: pivot and bucket size (FP32 mode: FP32 value)
    float pivot;           // heavy <=> w > pivot (pivot = mean rounded toward -inf)
    int32_t shift;         // fixed-point scale: carry words hold floor(x * 2^shift)
    uint32_t ticket;       // dynamic tile ticket of the look-back kernels
    uint32_t heavy_count;  // H, written by the partition's last tile
    double inv_total;      // 1 / total (ITS normalisation)
    double heavy_sum;      // sum of the heavy weights (written by the partition's last tile)
    double light_sum;      // sum of the light weights (idem)
    uint32_t repair_ticket;  // dynamic block ticket of the boundary repair
    int32_t merge_qs;        // merge keys: x -> rint(x * 2^merge_qs), 61 - ilogb(total)
    double inv_mean;         // 1 / mean (FP64 division)
    int64_t mean_q;          // rint(mean * 2^merge_qs)
    uint32_t error;          // kErr* bits: a bounded wait of this build gave up
    uint32_t wait_polls;     // polls of a bounded wait before it gives up
    uint32_t hooks;          // kHook* failure injection (tests)
    uint32_t epoch;          // builds finalized with these Params (the reduce's word tags)
    int32_t key_k;           // exact keys of the streamed pack: x -> x * 2^key_k,
                             // 53 - ilogb(mean), so that the mean is an integer
    uint32_t key_pad;
    uint64_t mean_k;         // mean * 2^key_k (exact, in [2^53, 2^54))
};
static_assert(sizeof(Params) == 112, "Params layout");

// Params::error bits.  A build whose error is not zero has invalid outputs;
// the kernels after the failing one return at once (no table is written).
constexpr uint32_t kErrReduceWait = 1u;    // the total's last block waited in vain for a partial
constexpr uint32_t kErrLookbackWait = 2u;  // a look-back waited in vain for a predecessor
// default bound of the waits: about a second of polling, far beyond any
// legitimate wait (a predecessor's record or an issued store)
constexpr uint32_t kDefaultWaitPolls = 1u << 22;
// failure injection for tests of the error path (the plan's test hooks)
constexpr uint32_t kHookSkipPartial = 1u;  // reduce block 0 does not publish its partial
constexpr uint32_t kHookSkipSpine = 2u;    // the tile-scan spine's block 0 publishes nothing
constexpr uint32_t kHookStaleTotal = 4u;   // reduce block 0 publishes its total under the
                                           // previous build's tag (a late store of that build)


struct AliasEntry {
    float p;
    uint32_t a;
};

// Exact split at a block boundary: i lights and j heavies precede bucket n,
// lp = LP(i), hp = HP(j) (prefix sums of the first i lights / j heavies).
struct Coarse {  // split (i lights, j heavies) at a merge-block boundary
    uint32_t i, j;
    double lp, hp;  // LP(i), HP(j)
    uint32_t xi;    // direct form: element index of light i (N when i == L)
    uint32_t aux;   // scratch of the segmented boundary repair
};
static_assert(sizeof(Coarse) == 32, "Coarse layout");

// One work item of a segmented launch: a (segment, tile) or a (segment, merge
// block) pair, with what its kernel needs before its first data load, so that
// the data loads wait for one lookup only.
struct SegItem {
    uint64_t off;  // first weight of the segment
    uint32_t n;    // segment length
    uint32_t seg;  // segment index
    uint32_t blk;  // tile (or merge block) within the segment
    uint32_t bnd;  // merge items: index of the block's first boundary
};
static_assert(sizeof(SegItem) == 24, "SegItem layout");

struct SplitRec {  // debug output of the split stage
    uint32_t i;
    uint32_t j;
    double spill;
};

constexpr int kBlock = 256;                 // threads per look-back block (4 waves)
constexpr int kItems = 16;                  // consecutive elements per thread
constexpr int kTile = kBlock * kItems;      // 4096 elements per tile
constexpr int kStateWords = 16;             // 128-byte look-back record per tile
constexpr uint32_t kMaxPartials = 2048;     // reduce blocks
// The reduce's words (wrs::Workspace::red): kMaxPartials block totals (the
// sentinel ~0 between launches), then kRedCounters 32-bit arrival counters
// (zero between launches).
constexpr uint32_t kRedGroup = 32;  // blocks per first-level arrival counter
constexpr uint32_t kRedCounters = kMaxPartials / kRedGroup + 1;
constexpr uint32_t kRedWords = kMaxPartials + (kRedCounters + 1) / 2;
constexpr int kSplitBlock = 256;            // debug split kernel block size
constexpr int kPackLanes = 16;              // lanes sharing one interval in the subgroup pack
constexpr uint32_t kSample = 64;            // partition prefix sample stride (ranks)
// Merge-pack geometry: `tile` buckets per block (a multiple of kSample),
// `per` buckets per thread (tile / per threads).  Supported shapes:
// tile 1024 or 2048 with per 2, 4 or 8; tile 4096 with per 4 or 8.
struct MergeShape {
    uint32_t tile = 1024;
    uint32_t per = 4;
};
constexpr uint32_t kMinMergeTile = 1024;
bool merge_shape_supported(MergeShape ms);

inline uint64_t samples_for(uint64_t n) { return n / kSample + 1; }
inline uint64_t merge_blocks_for(uint64_t n, uint32_t tile) { return (n + tile - 1) / tile; }

inline uint64_t tiles_for(uint64_t n) { return (n + kTile - 1) / kTile; }
// tile aggregates scanned per block by the partition's spine scan
constexpr uint32_t kSpineTiles = 1024;
inline uint64_t spine_blocks_for(uint64_t n) { return (tiles_for(n) + kSpineTiles - 1) / kSpineTiles; }
// boundaries per block of the boundary repair
constexpr uint32_t kRepairSpan = 1024;
inline uint64_t repair_blocks_for(uint64_t n, uint32_t tile) {
    return (merge_blocks_for(n, tile) + kRepairSpan - 1) / kRepairSpan;
}
// state buffer of a single-distribution build: one record per tile, one per
// spine block, then two per repair block (one per max-scan pass, sized for
// the smallest merge tile); all reset to the sentinel by the reduce pass
inline uint64_t repair_records_offset(uint64_t n) {
    return (tiles_for(n) + spine_blocks_for(n)) * kStateWords;
}
// Dense record layout of the reduce-then-scan partition inside the state
// buffer (tiles * kStateWords words): the aggregate of tile i at words
// [5i, 5i+5), its exclusive prefix at [kExclBase * tiles + 5i, ...); each as
// heavy sum (two 63-bit limbs), light sum (two limbs), heavy count.
constexpr uint32_t kCarryWords = 5;
constexpr uint32_t kExclBase = 8;
static_assert(kExclBase + kCarryWords <= kStateWords, "carry records exceed the state buffer");
inline uint64_t state_words_for(uint64_t n) {
    return repair_records_offset(n) + 2 * repair_blocks_for(n, kMinMergeTile) * kStateWords;
}
inline uint32_t partials_for(uint64_t n) {
    const uint64_t t = tiles_for(n);
    return (uint32_t)(t < kMaxPartials ? (t ? t : 1) : kMaxPartials);
}

// ---- launchers (each returns the launch's hipError_t) ----
// Total, mean, pivot and the other build scalars into *prm, the look-back
// records reset; red: kMaxPartials words, all ~0 before the first launch and
// restored by every launch (wrs::Workspace allocates them so).
hipError_t launch_reduce_final(const float* w, uint64_t n, int prec, uint64_t* red, uint32_t G,
                               uint64_t* states, uint64_t state_words, int has_pivot,
                               double pivot, Params* prm, hipStream_t s,
                               uint32_t wait_polls = kDefaultWaitPolls, uint32_t hooks = 0);
// idx is always written; pw (the weights in partition order), hi/lo (full
// prefix) and the prefix samples only when non-null.  ev_agg / ev_scan, when
// non-null, are recorded after the tile-aggregate and tile-scan kernels.
// lookback != 0: the single-pass look-back form (one kernel), else
// reduce-then-scan (aggregate, tile scan, emit).  lpos non-null (and hi
// null): the direct form -- idx/pw get only the heavies ([0, H)) and
// lpos[r / kSample] the element index of every light rank r % kSample == 0.
hipError_t launch_partition(const float* w, uint64_t n, int prec, int lookback, Params* prm,
                            uint64_t* states, uint32_t* idx, float* pw, float* hi, float* lo,
                            double* samp_h, double* samp_l, uint32_t* lpos, hipStream_t s,
                            hipEvent_t ev_agg = nullptr, hipEvent_t ev_scan = nullptr);
// exact splits at every kMergeTile buckets, then the LDS merge-path pack
// The lights of a direct-form build: the partition arrays idx/pw hold the
// heavies only, lights are read from the weights w and located through
// lpos[k] = element index of light rank k * kSample.  lpos null: the classic
// form (lights at idx/pw[N - 1 - rank]).
struct LightSrc {
    const float* w = nullptr;
    const uint32_t* lpos = nullptr;
};
inline uint64_t lightpos_for(uint64_t n) { return n / kSample + 1; }

// ---- the streamed pack form (streampack.hip) ----
// exclusive scan of the exact key-sum aggregates that k_sp_agg wrote
hipError_t launch_key_tilescan(uint64_t* states, uint64_t n, Params* prm, hipStream_t s);
// Buffers of a streamed build of n weights: slot_idx / slot_w hold tile t's
// heavies at [t * kTile, t * kTile + c_t) (n entries each); hkey, hidx, hrow,
// cross one entry per heavy (at most n); qa tiles_for(n) + 1 entries.
struct StreamBufs {
    uint32_t* slot_idx = nullptr;
    float* slot_w = nullptr;
    void* hkey = nullptr;      // 16-byte keys
    uint32_t* hidx = nullptr;
    AliasEntry* hrow = nullptr;
    uint16_t* cross = nullptr; // tile offset of the light a heavy's key crosses at
    uint32_t* qa = nullptr;
    uint32_t* seg = nullptr;   // kSegWords per tile: heavies before each 256-element
                               // segment, crossings before each segment boundary
    void* hsmp = nullptr;      // 16-byte keys of heavies 0, 64, 128, ... (the range search's index)
};
constexpr uint32_t kSegWords = 64;  // [0, 16) heavies, [32, 49) crossings
// the aggregate pass (after the reduce), then -- after launch_key_tilescan --
// the heavy keys, the tiles' alias ranges, the heavies' rows and the table;
// ev_keys / ev_rows (may be null) recorded after the heavy keys and ranges
// and after the rows
hipError_t launch_stream_agg(const float* w, uint64_t n, const Params* prm, uint64_t* states,
                             const StreamBufs& b, hipStream_t s);
hipError_t launch_stream_pack(const float* w, uint64_t n, int prec, const Params* prm,
                              const uint64_t* states, const StreamBufs& b, AliasEntry* table,
                              hipStream_t s, hipEvent_t ev_keys = nullptr,
                              hipEvent_t ev_rows = nullptr);

hipError_t launch_coarse(const float* pw, uint64_t n, MergeShape ms, int prec, const Params* prm,
                         const double* samp_h, const double* samp_l, Coarse* bounds,
                         hipStream_t s, const uint32_t* idx = nullptr, LightSrc ls = {});
// makes the boundary splits monotone (no-op for exact keys); states are the
// build's look-back records (state_words_for(n) words, reset by the reduce)
hipError_t launch_repair_bounds(Coarse* bounds, uint64_t n, MergeShape ms, Params* prm,
                                uint64_t* states, const float* pw, const double* samp_h,
                                const double* samp_l, hipStream_t s,
                                const uint32_t* idx = nullptr, LightSrc ls = {});
hipError_t launch_mergepack(const float* pw, uint64_t n, MergeShape ms, int prec,
                            const Params* prm, const uint32_t* idx, const Coarse* bounds,
                            AliasEntry* table, hipStream_t s, LightSrc ls = {});
// the K split points S(k), k = 1..K, of the reference's splitK layout (debug output)
hipError_t launch_splits(uint64_t n, uint32_t K, int prec, const Params* prm, const float* hi,
                         const float* lo, SplitRec* splits, hipStream_t s);
// the sweep pack forms: serial (split records from launch_splits, then one
// thread per split interval) and inline (splits and sweeps in one kernel);
// idx, pw: the classic partition arrays; hi, lo: the full prefix arrays
hipError_t launch_pack_serial(uint64_t n, uint32_t K, int prec, const Params* prm,
                              const SplitRec* splits, const uint32_t* idx, const float* pw,
                              AliasEntry* table, hipStream_t s);
// subgroup form: the serial form's split records, kPackLanes lanes per interval
hipError_t launch_pack_subgroup(uint64_t n, uint32_t K, int prec, const Params* prm,
                                const SplitRec* splits, const uint32_t* idx, const float* pw,
                                AliasEntry* table, hipStream_t s);
hipError_t launch_splitpack_inline(uint64_t n, uint32_t K, int prec, const Params* prm,
                                   const float* hi, const float* lo, const uint32_t* idx,
                                   const float* pw, AliasEntry* table, hipStream_t s);
// reduce-then-scan form for large n: needs no total beforehand (it writes
// prm->total, mean and inv_total itself) and no reset pass; states as for a
// PSA build of n
hipError_t launch_scan_rts(const float* w, uint64_t n, int prec, int normalise, Params* prm,
                           uint64_t* states, float* hi, float* lo, hipStream_t s);
// tiles up to which the single-pass look-back scan is used
constexpr uint32_t kScanSinglePassTiles = 64;
hipError_t launch_scan(const float* w, uint64_t n, int prec, int normalise, Params* prm,
                       uint64_t* states, float* hi, float* lo, hipStream_t s);
hipError_t launch_alias_sample(const AliasEntry* table, uint64_t n, uint64_t s_count,
                               uint64_t seed, uint64_t counter_base, uint32_t* out, hipStream_t s);
hipError_t launch_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                             uint64_t s_count, uint64_t seed, uint64_t counter_base,
                             uint32_t* out, hipStream_t s);
// the same samples with the line indexes in a static device array (no
// allocation); n <= kItsNoAllocMaxN
constexpr uint64_t kItsNoAllocMaxN = 1ull << 25;
hipError_t launch_its_sample_noalloc(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                     uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                     uint32_t* out, hipStream_t s);
// windowed samplers (groups of 64 consecutive counters share a window);
// s_count >= 1, window >= 1
hipError_t launch_alias_sample_window(const AliasEntry* table, uint64_t n, uint64_t s_count,
                                      uint64_t seed, uint64_t counter_base, uint32_t window,
                                      uint32_t* out, hipStream_t s);
hipError_t launch_its_sample_window(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                    uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                    uint64_t window, int parts, uint32_t* out, hipStream_t s);
// guide[0..m] of the normalised CDF (m >= 1), and guided ITS samples
hipError_t launch_cutpoint_build(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                 uint32_t* guide, uint64_t m, hipStream_t s);
hipError_t launch_cutpoint_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                  const uint32_t* guide, uint64_t m, uint64_t s_count,
                                  uint64_t seed, uint64_t counter_base, uint32_t* out,
                                  hipStream_t s);
hipError_t launch_philox(uint64_t seed, uint32_t stream_word, uint64_t base, uint64_t n,
                         uint32_t* out4, hipStream_t s);
hipError_t launch_generate(int kind, double a, double b, uint64_t seed, uint64_t n, float* out,
                           hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t words, uint64_t value, hipStream_t s);
// sample statistics: hist[0..n] (hist[n] counts samples >= n); stats over
// tiles of 4096 bins with one 16-word record per tile in recs
hipError_t launch_histogram(const uint32_t* samples, uint64_t s_count, uint64_t n,
                            uint64_t* hist, int clear, hipStream_t s);
uint64_t stats_tiles_for(uint64_t n);
hipError_t launch_stats(const uint64_t* hist, const float* w, const float* cdf_hi,
                        const float* cdf_lo, uint64_t n, uint64_t s_count, double min_expected,
                        const Params* prm, uint64_t* recs, double* out, hipStream_t s);

// segmented (batched) forms: items are (segment, tile) pairs, points are the
// flattened split points of all segments
hipError_t launch_seg_reduce(const float* w, const SegItem* item_d, uint32_t items, int prec,
                             double* partials, uint64_t* states, hipStream_t s);
hipError_t launch_seg_finalize(const uint64_t* seg_off, const uint32_t* seg_item0, uint32_t nseg,
                               int prec, const double* partials, Params* seg_prm,
                               uint32_t* ticket, hipStream_t s);
// lookback != 0: one pass with decoupled look-back per segment (states reset
// by the segmented reduce, ticket by the segmented finalize); else
// reduce-then-scan (aggregate, per-segment spine, emit; no reset needed)
hipError_t launch_seg_partition(const float* w, const SegItem* item_d, const uint32_t* seg_item0,
                                uint32_t nseg,
                                uint32_t items, int prec, int lookback, Params* seg_prm,
                                uint32_t* ticket, uint64_t* states, uint32_t* idx, float* pw,
                                double* samp_h, double* samp_l, const uint64_t* seg_samp0,
                                hipStream_t s, uint32_t* lpos = nullptr,
                                const uint64_t* seg_lpos0 = nullptr);
// lpos non-null: the direct pack form (segment g's light positions at
// lpos + seg_lpos0[g]; the reduce-then-scan partition only)
hipError_t launch_seg_mergepack(const float* pw, const uint32_t* idx, const double* samp_h,
                                const double* samp_l, const uint64_t* seg_off,
                                const uint64_t* seg_samp0, const uint32_t* seg_b0, uint32_t nseg,
                                uint32_t nbound, const uint32_t* bound_seg,
                                const SegItem* mitem_d, uint32_t mitems, MergeShape ms, int prec,
                                const Params* seg_prm, Coarse* bounds, AliasEntry* table,
                                hipStream_t s, const float* w = nullptr,
                                const uint32_t* lpos = nullptr,
                                const uint64_t* seg_lpos0 = nullptr);
hipError_t launch_seg_sample(const AliasEntry* tables, const uint64_t* seg_off, uint32_t nseg,
                             uint32_t per_segment, uint64_t seed, uint64_t counter_base,
                             uint32_t* out, hipStream_t s);
hipError_t launch_copy_f32(float* dst, const float* src, uint64_t n, hipStream_t s);

}  // namespace wrsk
