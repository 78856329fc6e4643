// batched.cpp -- BatchPlan: geometry of the segmented build and its launches.
#include "batched.hpp"

#include "../../include/wrs.h"
#include "wrs_ops.hpp"

namespace wrsb {

namespace {
template <class T>
T* dalloc(uint64_t count) {
    void* p = nullptr;
    wrs::check_hip(hipMalloc(&p, (count ? count : 1) * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}
template <class T>
void upload(T* dst, const std::vector<T>& src) {
    wrs::check_hip(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice),
                   "upload");
}
}  // namespace

namespace {
// merge work lists of a merge tile: (segment, block) items and the first
// boundary of every segment
void merge_lists(const std::vector<uint64_t>& offsets, uint32_t nseg, uint32_t tile,
                 std::vector<wrsk::SegItem>& mitem, std::vector<uint32_t>& b0,
                 std::vector<uint32_t>& bseg, uint64_t& nbound) {
    mitem.clear();
    bseg.clear();
    b0.assign(nseg + 1, 0);
    uint64_t bnd = 0;
    for (uint32_t g = 0; g < nseg; ++g) {
        const uint64_t n = offsets[g + 1] - offsets[g];
        const uint64_t blocks = wrsk::merge_blocks_for(n, tile);
        b0[g] = (uint32_t)bnd;
        for (uint64_t b = 0; b < blocks; ++b)
            mitem.push_back(wrsk::SegItem{offsets[g], (uint32_t)n, g, (uint32_t)b,
                                          (uint32_t)(bnd + b)});
        for (uint64_t b = 0; b <= blocks; ++b) bseg.push_back(g);
        bnd += blocks + 1;
    }
    b0[nseg] = (uint32_t)bnd;
    nbound = bnd;
}
}  // namespace

void BatchPlan::create(uint32_t split_size_, int prec_, const wrs_tuning& tuning_,
                       const uint64_t* offsets_host, uint32_t nseg_) {
    split_size = split_size_ ? split_size_ : 32u;
    prec = prec_;
    nseg = nseg_;
    if (prec < wrsk::kFP32 || prec > wrsk::kFP64) throw wrs::Error(WRS_EINVAL, "unknown precision");
    wrs::require_tuning(tuning_);
    offsets.assign(offsets_host, offsets_host + nseg + 1);
    if (offsets[0] != 0) throw wrs::Error(WRS_EINVAL, "offsets[0] must be 0");
    std::vector<wrsk::SegItem> h_items;
    std::vector<uint32_t> h_seg_item0(nseg + 1);
    std::vector<uint64_t> h_seg_samp0(nseg + 1);
    std::vector<uint64_t> h_seg_lpos0(nseg + 1);
    uint64_t items64 = 0, samp = 0, lp = 0;
    for (uint32_t g = 0; g < nseg; ++g) {
        if (offsets[g + 1] <= offsets[g])
            throw wrs::Error(WRS_EINVAL, "segments must be non-empty and offsets increasing");
        const uint64_t n = offsets[g + 1] - offsets[g];
        const uint64_t tiles = wrsk::tiles_for(n);
        if (tiles > wrsk::kMaxPartials)
            throw wrs::Error(WRS_EINVAL, "segment larger than 2^23 weights");
        h_seg_item0[g] = (uint32_t)items64;
        for (uint64_t t = 0; t < tiles; ++t) {
            h_items.push_back(wrsk::SegItem{offsets[g], (uint32_t)n, g, (uint32_t)t, 0u});
        }
        items64 += tiles;
        h_seg_samp0[g] = samp;
        samp += wrsk::samples_for(n);
        h_seg_lpos0[g] = lp;
        lp += wrsk::lightpos_for(n);
    }
    // capacity of the merge lists: the smallest merge tile has the most blocks
    std::vector<wrsk::SegItem> mi_;
    std::vector<uint32_t> b0_, bs_;
    uint64_t bnd_cap = 0;
    merge_lists(offsets, nseg, wrsk::kMinMergeTile, mi_, b0_, bs_, bnd_cap);
    if (items64 >= (1ull << 31) || bnd_cap >= (1ull << 31)) throw wrs::Error(WRS_EINVAL, "too many tiles");
    h_seg_item0[nseg] = (uint32_t)items64;
    h_seg_samp0[nseg] = samp;
    h_seg_lpos0[nseg] = lp;
    items = (uint32_t)items64;
    mitems_cap = (uint32_t)mi_.size();
    nbound_cap = (uint32_t)bnd_cap;
    total_n = offsets[nseg];

    seg_off = dalloc<uint64_t>(nseg + 1);
    seg_item0 = dalloc<uint32_t>(nseg + 1);
    seg_b0 = dalloc<uint32_t>(nseg + 1);
    seg_samp0 = dalloc<uint64_t>(nseg + 1);
    item_d = dalloc<wrsk::SegItem>(items);
    mitem_d = dalloc<wrsk::SegItem>(mitems_cap);
    bound_seg = dalloc<uint32_t>(nbound_cap);
    seg_prm = dalloc<wrsk::Params>(nseg);
    partials = dalloc<double>(items);
    states = dalloc<uint64_t>((uint64_t)items * wrsk::kStateWords);
    ticket = dalloc<uint32_t>(1);
    idx = dalloc<uint32_t>(total_n);
    pw = dalloc<float>(total_n);
    samp_h = dalloc<double>(samp);
    samp_l = dalloc<double>(samp);
    bounds = dalloc<wrsk::Coarse>(nbound_cap);
    lpos = dalloc<uint32_t>(lp);
    seg_lpos0 = dalloc<uint64_t>(nseg + 1);
    upload(seg_off, offsets);
    upload(seg_lpos0, h_seg_lpos0);
    upload(seg_item0, h_seg_item0);
    upload(seg_samp0, h_seg_samp0);
    upload(item_d, h_items);
    shape.tile = 0;  // forces the merge lists of the first tuning
    set_tuning(tuning_);
}

void BatchPlan::set_tuning(const wrs_tuning& t) {
    wrs::require_tuning(t);
    if (t.direct == WRS_PACK_SERIAL || t.direct == WRS_PACK_INLINE ||
        t.direct == WRS_PACK_SUBGROUP || t.direct == WRS_PACK_STREAM)
        throw wrs::Error(WRS_EINVAL, "batch plans build the merge pack forms only");
    const wrsk::MergeShape ms = wrs::merge_shape_of(t);
    if (ms.tile != shape.tile) {
        std::vector<wrsk::SegItem> h_mitem;
        std::vector<uint32_t> h_b0, h_bseg;
        uint64_t bnd = 0;
        merge_lists(offsets, nseg, ms.tile, h_mitem, h_b0, h_bseg, bnd);
        upload(mitem_d, h_mitem);
        upload(seg_b0, h_b0);
        upload(bound_seg, h_bseg);
        mitems = (uint32_t)h_mitem.size();
        nbound = (uint32_t)bnd;
    }
    tuning = t;
    shape = ms;
}

void BatchPlan::destroy() {
    void* ptrs[] = {seg_off,  seg_item0, seg_b0, seg_samp0, item_d, mitem_d,
                    bound_seg, seg_prm,  partials, states,  ticket, idx,
                    pw,       samp_h,    samp_l, bounds,    lpos,   seg_lpos0};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    seg_off = nullptr;
    seg_item0 = nullptr;
    seg_b0 = nullptr;
    seg_samp0 = nullptr;
    item_d = nullptr;
    mitem_d = nullptr;
    bound_seg = nullptr;
    seg_prm = nullptr;
    partials = nullptr;
    states = nullptr;
    ticket = nullptr;
    idx = nullptr;
    pw = nullptr;
    samp_h = nullptr;
    samp_l = nullptr;
    bounds = nullptr;
    lpos = nullptr;
    seg_lpos0 = nullptr;
}

// The direct pack form (tuning.direct) runs with the reduce-then-scan
// partition; with the single-pass partition the classic form is built (the
// same tables either way).
void BatchPlan::build(const float* w, wrsk::AliasEntry* tables, hipStream_t s) const {
    const bool direct = tuning.direct != WRS_PACK_CLASSIC && tuning.partition == 0;
    wrs::check_hip(wrsk::launch_seg_reduce(w, item_d, items, prec, partials,
                                           states, s),
                   "segmented reduce launch");
    wrs::check_hip(wrsk::launch_seg_finalize(seg_off, seg_item0, nseg, prec, partials, seg_prm,
                                             ticket, s),
                   "segmented finalize launch");
    wrs::check_hip(wrsk::launch_seg_partition(w, item_d, seg_item0, nseg,
                                              items, prec, (int)tuning.partition, seg_prm,
                                              ticket, states, idx, pw, samp_h, samp_l, seg_samp0,
                                              s, direct ? lpos : nullptr, seg_lpos0),
                   "segmented partition launch");
    wrs::check_hip(wrsk::launch_seg_mergepack(pw, idx, samp_h, samp_l, seg_off, seg_samp0, seg_b0,
                                              nseg, nbound, bound_seg, mitem_d, mitems, shape,
                                              prec, seg_prm, bounds, tables, s, w,
                                              direct ? lpos : nullptr, seg_lpos0),
                   "segmented merge pack launch");
}

}  // namespace wrsb
