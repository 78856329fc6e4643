// lookback.hpp -- decoupled look-back over per-tile records in HBM, shared by
// the prefix-partition and the scan kernels.
//
// Record of tile t (kStateWords 64-bit words, all reset to the sentinel ~0 by
// the reduce pass that precedes every look-back kernel):
//   [0..7]   aggregate payload      [8..14]  inclusive-prefix payload
//   [15]     status: kStAgg once the aggregate is published, kStInc once the
//            inclusive prefix is.
// Payload words are self-validating (top bit clear once written), so the
// status word needs no release fence: it is a hint that tells a reader which
// payload to fetch, and a reader that finds a payload word still holding the
// sentinel simply polls again.  Polling touches one 8-byte word per
// predecessor; payloads are read once, only for the predecessors consumed.
//
// The reference polls its state words in subgroups of PARALLEL_LOOKBACK_DEPTH
// = 32 (src/device/prefix_partition/decoupled/shader.comp:206-311); here one
// wave64 polls 64 predecessors per round trip.
#pragma once

#include "device_common.hpp"
#include "internal.hpp"

namespace wrsdev {

constexpr uint64_t kStAgg = 1;
constexpr uint64_t kStInc = 2;
constexpr int kStatusWord = 15;

// Ops must provide: static P zero(); static P add(P, P); static P wave_sum(P);
// static bool read(const uint64_t* words, P&); static void publish(uint64_t*, const P&).
template <class P, class Ops>
__device__ __forceinline__ void publish_aggregate(uint64_t* rec, const P& agg) {
    Ops::publish(rec + 0, agg);
    st_word(rec + kStatusWord, kStAgg);
}
template <class P, class Ops>
__device__ __forceinline__ void publish_inclusive(uint64_t* rec, const P& inc) {
    Ops::publish(rec + 8, inc);
    st_word(rec + kStatusWord, kStInc);
}

// Exclusive prefix of tile `tile` (> 0); executed by one whole wave, the
// result is identical in every lane.  After `polls` consecutive polls without
// progress it gives up: it sets kErrLookbackWait in *err and returns what it
// has (the build's outputs are then invalid and its later kernels return at
// once).  A predecessor publishes within microseconds of starting, so the
// bound (kDefaultWaitPolls, about a second) is only reached if one never
// starts or never publishes.
template <class P, class Ops>
__device__ __forceinline__ P lookback(const uint64_t* states, uint32_t tile, uint32_t polls,
                                      uint32_t* err) {
    const int lane = lane_id();
    P excl = Ops::zero();
    int64_t look = (int64_t)tile - 1;
    uint32_t idle = 0;
    for (;;) {
        if (idle > polls) {
            if (lane == 0) atomicOr(err, wrsk::kErrLookbackWait);
            return excl;
        }
        const int64_t tt = look - lane;
        const uint64_t stw = tt < 0 ? kStInc : ld_word(states + (uint64_t)tt * wrsk::kStateWords + kStatusWord);
        const int st = stw == kStInc ? 2 : (stw == kStAgg ? 1 : 0);
        const uint64_t b_inc = __ballot(st == 2);
        const uint64_t b_nr = __ballot(st == 0);
        const int fi = b_inc ? __builtin_ctzll(b_inc) : kWave;
        const int fn = b_nr ? __builtin_ctzll(b_nr) : kWave;
        const bool finish = fi < fn;
        const int cnt = finish ? fi + 1 : fn;
        if (cnt == 0) {
            __builtin_amdgcn_s_sleep(2);
            ++idle;
            continue;
        }
        P pv = Ops::zero();
        bool ok = true;
        if (lane < cnt && tt >= 0)
            ok = Ops::read(states + (uint64_t)tt * wrsk::kStateWords + (st == 2 ? 8 : 0), pv);
        if (!__all(ok)) {  // a payload is not visible yet
            __builtin_amdgcn_s_sleep(1);
            ++idle;
            continue;
        }
        excl = Ops::add(excl, Ops::wave_sum(lane < cnt ? pv : Ops::zero()));
        if (finish) return excl;
        look -= cnt;
        idle = 0;
    }
}

}  // namespace wrsdev
