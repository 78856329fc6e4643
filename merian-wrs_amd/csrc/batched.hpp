// batched.hpp -- host side of the segmented (batched) PSA build: thousands of
// independent distributions of uneven sizes, laid out back to back (CSR
// offsets), built by segmented launches whose work lists (segment, tile) and
// (segment, merge block) are derived from the offsets when the plan is created
// (the merge lists again when the tuning changes the merge tile).  Each
// segment is built with exactly the arithmetic and order of a
// single-distribution build, so the tables are bit-identical to building the
// segments one by one with the same merge tile.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/wrs.h"
#include "internal.hpp"

namespace wrsb {

struct BatchPlan {
    uint32_t split_size = 32;
    int prec = 2;
    wrs_tuning tuning{};         // partition: 1 single pass with look-back, 0 reduce-then-scan
    wrsk::MergeShape shape;      // merge shape of the tuning
    uint32_t nseg = 0;
    uint32_t items = 0;          // (segment, tile) work items of reduce and partition
    uint32_t mitems = 0;         // (segment, merge block) work items of the merge pack
    uint32_t nbound = 0;         // block-boundary splits over all segments
    uint64_t total_n = 0;        // weights over all segments
    std::vector<uint64_t> offsets;  // host copy, nseg + 1
    // device
    uint64_t* seg_off = nullptr;     // [nseg+1]
    uint32_t* seg_item0 = nullptr;   // [nseg+1]
    uint32_t* seg_b0 = nullptr;      // [nseg+1] first boundary of each segment
    uint64_t* seg_samp0 = nullptr;   // [nseg+1] first prefix sample of each segment
    wrsk::SegItem* item_d = nullptr;   // [items] (segment, tile) items
    wrsk::SegItem* mitem_d = nullptr;  // [mitems] (segment, merge block) items
    uint32_t* bound_seg = nullptr;   // [nbound] segment of each block boundary
    wrsk::Params* seg_prm = nullptr; // [nseg]
    double* partials = nullptr;      // [items]
    uint64_t* states = nullptr;      // [items * kStateWords]
    uint32_t* ticket = nullptr;      // [1]
    uint32_t* idx = nullptr;         // [total_n]
    float* pw = nullptr;             // [total_n] weights in partition order
    double* samp_h = nullptr;        // [seg_samp0[nseg]]
    double* samp_l = nullptr;        // [seg_samp0[nseg]]
    wrsk::Coarse* bounds = nullptr;  // [nbound at the smallest merge tile]
    uint32_t* lpos = nullptr;        // [seg_lpos0[nseg]] direct form: light positions
    uint64_t* seg_lpos0 = nullptr;   // [nseg+1] first light position of each segment
    uint32_t mitems_cap = 0, nbound_cap = 0;

    void create(uint32_t split_size, int prec, const wrs_tuning& tuning,
                const uint64_t* offsets_host, uint32_t nseg);
    // changes the launch geometry; rebuilds the merge work lists when the
    // merge tile changes (synchronous uploads: not while a build is in flight)
    void set_tuning(const wrs_tuning& tuning);
    void destroy();
    void build(const float* weights, wrsk::AliasEntry* tables, hipStream_t s) const;
};

}  // namespace wrsb
