// wrs_host.cpp -- the C++ operator layer (wrs_ops.hpp) and the C ABI
// (include/wrs.h, include/wrs_runtime.h) on top of the HIP kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/wrs.h"
#include "../../include/wrs_runtime.h"
#include "batched.hpp"
#include "internal.hpp"
#include "wrs_ops.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace wrs {

void set_last_error(const std::string& msg) { g_last_error = msg; }

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::string msg = std::string(what) + ": " + hipGetErrorName(e) + " (" +
                          hipGetErrorString(e) + ")";
        throw Error(e == hipErrorOutOfMemory ? WRS_ENOMEM : WRS_EHIP, msg);
    }
}

template <class T>
static T* dev_alloc(uint64_t count) {
    void* p = nullptr;
    if (count == 0) return nullptr;
    check_hip(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

Workspace Workspace::allocate(uint64_t max_n, bool partition_buffers) {
    Workspace ws;
    ws.max_n = max_n;
    try {
        ws.params = dev_alloc<wrsk::Params>(1);
        ws.red = dev_alloc<uint64_t>(wrsk::kRedWords);
        // block totals at the sentinel, arrival counters at zero (each launch
        // of the reduce restores both)
        check_hip(hipMemset(ws.red, 0xff, wrsk::kMaxPartials * sizeof(uint64_t)), "hipMemset");
        check_hip(hipMemset(ws.red + wrsk::kMaxPartials, 0,
                            (wrsk::kRedWords - wrsk::kMaxPartials) * sizeof(uint64_t)),
                  "hipMemset");
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        ws.state_words = wrsk::state_words_for(max_n);
        ws.states = dev_alloc<uint64_t>(ws.state_words);
        if (partition_buffers) {
            ws.indices = dev_alloc<uint32_t>(max_n);
            ws.pweights = dev_alloc<float>(max_n);
            ws.samp_h = dev_alloc<double>(wrsk::samples_for(max_n));
            ws.samp_l = dev_alloc<double>(wrsk::samples_for(max_n));
            ws.bounds = dev_alloc<wrsk::Coarse>(wrsk::merge_blocks_for(max_n, wrsk::kMinMergeTile) + 2);
            ws.lightpos = dev_alloc<uint32_t>(wrsk::lightpos_for(max_n));
        }
    } catch (...) {
        ws.release();
        throw;
    }
    return ws;
}

void Workspace::ensure_prefix(bool lo) {
    if (!prefix_hi) prefix_hi = dev_alloc<float>(max_n);
    if (lo && !prefix_lo) prefix_lo = dev_alloc<float>(max_n);
}

void Workspace::ensure_splits(uint64_t K) {
    if (split_cap >= K) return;
    if (splits) (void)hipFree(splits);
    splits = nullptr;
    split_cap = 0;
    splits = dev_alloc<wrsk::SplitRec>(K);
    split_cap = K;
}

void Workspace::ensure_stream() {
    if (skey) return;
    try {
        skey = dev_alloc<uint64_t>(2 * max_n);
        shidx = dev_alloc<uint32_t>(max_n);
        shrow = dev_alloc<wrsk::AliasEntry>(max_n);
        scross = dev_alloc<uint16_t>(max_n);
        sqa = dev_alloc<uint32_t>(wrsk::tiles_for(max_n) + 2);
        sseg = dev_alloc<uint32_t>(wrsk::kSegWords * wrsk::tiles_for(max_n));
        shsmp = dev_alloc<uint64_t>(2 * (max_n / 64 + 2));
    } catch (...) {
        void* ptrs[] = {skey, shidx, shrow, scross, sqa, sseg, shsmp};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        skey = nullptr;
        shidx = nullptr;
        shrow = nullptr;
        scross = nullptr;
        sqa = nullptr;
        sseg = nullptr;
        shsmp = nullptr;
        throw;
    }
}

void Workspace::release() {
    void* ptrs[] = {params,    red,    states,   indices, pweights, prefix_hi, prefix_lo,
                    samp_h,    samp_l, bounds,   lightpos, splits,  skey,      shidx,
                    shrow,     scross, sqa,      sseg,     shsmp};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    params = nullptr;
    red = nullptr;
    states = nullptr;
    indices = nullptr;
    pweights = nullptr;
    prefix_hi = nullptr;
    prefix_lo = nullptr;
    samp_h = nullptr;
    samp_l = nullptr;
    bounds = nullptr;
    lightpos = nullptr;
    splits = nullptr;
    split_cap = 0;
    skey = nullptr;
    shidx = nullptr;
    shrow = nullptr;
    scross = nullptr;
    sqa = nullptr;
    sseg = nullptr;
    shsmp = nullptr;
}

static void require(bool cond, const char* msg) {
    if (!cond) throw Error(WRS_EINVAL, msg);
}

wrsk::MergeShape merge_shape_of(const wrs_tuning& t) {
    wrsk::MergeShape ms;
    if (t.merge_tile) ms.tile = t.merge_tile;
    if (t.merge_per) ms.per = t.merge_per;
    return ms;
}

void require_tuning(const wrs_tuning& t) {
    require(wrsk::merge_shape_supported(merge_shape_of(t)),
            "unsupported merge shape (merge_tile 1024/2048 with merge_per 2/4/8, or 4096 with 4/8)");
    require(t.partition <= 1u, "partition must be 0 or 1");
    require(t.direct <= WRS_PACK_STREAM, "direct must be a WRS_PACK_* value");
}

void Mean::run(hipStream_t s, const float* weights, uint64_t N, Workspace& ws, double pivot) const {
    require(weights != nullptr, "weights is null");
    require(N > 0 && N <= ws.max_n, "N must be in [1, max_n]");
    const uint32_t G = wrsk::partials_for(N);
    const uint64_t words = wrsk::state_words_for(N);
    const bool has_pivot = !std::isnan(pivot);
    check_hip(wrsk::launch_reduce_final(weights, N, (int)m_prec, ws.red, G, ws.states, words,
                                        has_pivot ? 1 : 0, has_pivot ? pivot : 0.0, ws.params, s,
                                        ws.wait_polls, ws.hooks),
              "reduce launch");
}

void PrefixPartition::run(hipStream_t s, const Buffers& b, uint64_t N, Workspace& ws) const {
    require(b.elements && b.partitionIndices && b.partitionPrefix, "null partition buffer");
    require(N > 0 && N <= ws.max_n, "N must be in [1, max_n]");
    check_hip(wrsk::launch_partition(b.elements, N, (int)m_prec, 0, ws.params, ws.states,
                                     b.partitionIndices, nullptr, b.partitionPrefix,
                                     wrsk::prec_has_lo(m_prec) ? b.partitionPrefixLo : nullptr,
                                     nullptr, nullptr, nullptr, s),
              "partition launch");
}

void PrefixSum::run(hipStream_t s, const Buffers& b, uint64_t N, Workspace& ws) const {
    require(b.elements && b.prefix, "null scan buffer");
    require(N > 0 && N <= ws.max_n, "N must be in [1, max_n]");
    float* lo = wrsk::prec_has_lo(m_prec) ? b.prefixLo : nullptr;
    if (wrsk::tiles_for(N) <= wrsk::kScanSinglePassTiles) {
        // short look-back chain: mean pass (total, shift, record reset) + one scan pass
        Mean(m_prec).run(s, b.elements, N, ws);
        check_hip(wrsk::launch_scan(b.elements, N, (int)m_prec, m_normalise ? 1 : 0, ws.params,
                                    ws.states, b.prefix, lo, s),
                  "scan launch");
    } else {
        check_hip(wrsk::launch_scan_rts(b.elements, N, (int)m_prec, m_normalise ? 1 : 0,
                                        ws.params, ws.states, b.prefix, lo, s),
                  "scan launch");
    }
}

void PSA::run(hipStream_t s, const Buffers& b, uint64_t N, const wrs_psa_debug* dbg,
              hipEvent_t* ev) const {
    require(b.weights && b.aliasTable && b.internals, "null PSA buffer");
    Workspace& ws = *b.internals;
    require(N > 0 && N <= ws.max_n, "N must be in [1, max_n]");
    require(ws.indices && ws.samp_h && ws.bounds, "workspace has no partition buffers");
    const wrs_precision prec = m_cfg.precision;
    // full prefix arrays only for a debug build (parity tests); the build
    // itself needs the indices and the sparse prefix samples
    // the sweep forms (the reference's serial, subgroup and inline split +
    // pack) read the full prefix arrays and the classic partition arrays
    const uint32_t form = dbg ? WRS_PACK_CLASSIC : m_cfg.tuning.direct;
    const bool records = form == WRS_PACK_SERIAL || form == WRS_PACK_SUBGROUP;
    const bool sweep = records || form == WRS_PACK_INLINE;
    const bool full_prefix = sweep || (dbg && (dbg->prefix_hi || dbg->prefix_lo || dbg->splits));
    if (full_prefix) ws.ensure_prefix(wrsk::prec_has_lo(prec));
    if (records) ws.ensure_splits(splitCount(N));
    float* hi = full_prefix ? ws.prefix_hi : nullptr;
    float* lo = full_prefix && wrsk::prec_has_lo(prec) ? ws.prefix_lo : nullptr;
    // the direct form (partition arrays hold the heavies only) for plain builds;
    // a debug build reads the full partition arrays
    wrsk::LightSrc ls;
    if ((form == WRS_PACK_DEFAULT || form == WRS_PACK_DIRECT) && !dbg)
        ls = wrsk::LightSrc{b.weights, ws.lightpos};
    const uint32_t G = wrsk::partials_for(N);
    const uint64_t words = wrsk::state_words_for(N);
    if (form == WRS_PACK_STREAM && !dbg) ws.ensure_stream();
    if (ev) check_hip(hipEventRecord(ev[0], s), "event");
    check_hip(wrsk::launch_reduce_final(b.weights, N, (int)prec, ws.red, G, ws.states, words, 0,
                                        0.0, ws.params, s, ws.wait_polls, ws.hooks),
              "reduce launch");
    // the finalize runs inside the reduce (its last block): an empty phase
    if (ev) check_hip(hipEventRecord(ev[1], s), "event");
    if (ev) check_hip(hipEventRecord(ev[2], s), "event");
    if (form == WRS_PACK_STREAM && !dbg) {
        // phases: tile aggregate, tile scan, heavy keys and ranges, heavy rows, table
        const wrsk::StreamBufs sb{ws.indices, ws.pweights, ws.skey, ws.shidx,
                                  ws.shrow,    ws.scross,   ws.sqa, ws.sseg, ws.shsmp};
        check_hip(wrsk::launch_stream_agg(b.weights, N, ws.params, ws.states, sb, s),
                  "stream aggregate launch");
        if (ev) check_hip(hipEventRecord(ev[3], s), "event");
        check_hip(wrsk::launch_key_tilescan(ws.states, N, ws.params, s), "tile scan launch");
        if (ev) check_hip(hipEventRecord(ev[4], s), "event");
        check_hip(wrsk::launch_stream_pack(b.weights, N, (int)prec, ws.params, ws.states, sb,
                                           reinterpret_cast<wrsk::AliasEntry*>(b.aliasTable), s,
                                           ev ? ev[5] : nullptr, ev ? ev[6] : nullptr),
                  "stream pack launch");
        if (ev) check_hip(hipEventRecord(ev[7], s), "event");
        return;
    }
    check_hip(wrsk::launch_partition(b.weights, N, (int)prec, (int)m_cfg.tuning.partition,
                                     ws.params, ws.states, ws.indices,
                                     ws.pweights, hi, lo, ws.samp_h, ws.samp_l,
                                     ls.lpos ? ws.lightpos : nullptr, s,
                                     ev ? ev[3] : nullptr, ev ? ev[4] : nullptr),
              "partition launch");
    if (ev) check_hip(hipEventRecord(ev[5], s), "event");
    if (sweep) {
        const uint32_t K = splitCount(N);
        auto* table = reinterpret_cast<wrsk::AliasEntry*>(b.aliasTable);
        if (records) {
            check_hip(wrsk::launch_splits(N, K, (int)prec, ws.params, hi, lo, ws.splits, s),
                      "splits launch");
            if (ev) check_hip(hipEventRecord(ev[6], s), "event");
            if (form == WRS_PACK_SERIAL)
                check_hip(wrsk::launch_pack_serial(N, K, (int)prec, ws.params, ws.splits,
                                                   ws.indices, ws.pweights, table, s),
                          "serial pack launch");
            else
                check_hip(wrsk::launch_pack_subgroup(N, K, (int)prec, ws.params, ws.splits,
                                                     ws.indices, ws.pweights, table, s),
                          "subgroup pack launch");
        } else {
            if (ev) check_hip(hipEventRecord(ev[6], s), "event");
            check_hip(wrsk::launch_splitpack_inline(N, K, (int)prec, ws.params, hi, lo,
                                                    ws.indices, ws.pweights, table, s),
                      "inline split pack launch");
        }
        if (ev) check_hip(hipEventRecord(ev[7], s), "event");
        return;
    }
    const wrsk::MergeShape ms = merge_shape_of(m_cfg.tuning);
    check_hip(wrsk::launch_coarse(ws.pweights, N, ms, (int)prec, ws.params, ws.samp_h, ws.samp_l,
                                  ws.bounds, s, ws.indices, ls),
              "coarse split launch");
    check_hip(wrsk::launch_repair_bounds(ws.bounds, N, ms, ws.params, ws.states, ws.pweights,
                                         ws.samp_h, ws.samp_l, s, ws.indices, ls),
              "boundary repair launch");
    if (ev) check_hip(hipEventRecord(ev[6], s), "event");
    check_hip(wrsk::launch_mergepack(ws.pweights, N, ms, (int)prec, ws.params, ws.indices,
                                     ws.bounds, reinterpret_cast<wrsk::AliasEntry*>(b.aliasTable),
                                     s, ls),
              "merge pack launch");
    if (ev) check_hip(hipEventRecord(ev[7], s), "event");
    if (dbg) {
        if (dbg->splits)
            check_hip(wrsk::launch_splits(N, splitCount(N), (int)prec, ws.params, ws.prefix_hi,
                                          wrsk::prec_has_lo(prec) ? ws.prefix_lo : nullptr,
                                          reinterpret_cast<wrsk::SplitRec*>(dbg->splits), s),
                      "splits launch");
        if (dbg->mean)
            check_hip(hipMemcpyAsync(dbg->mean, &ws.params->mean, sizeof(double),
                                     hipMemcpyDeviceToDevice, s), "copy mean");
        if (dbg->heavy_count)
            check_hip(hipMemcpyAsync(dbg->heavy_count, &ws.params->heavy_count, sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, s), "copy H");
        if (dbg->indices)
            check_hip(hipMemcpyAsync(dbg->indices, ws.indices, N * sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, s), "copy indices");
        if (dbg->prefix_hi)
            check_hip(hipMemcpyAsync(dbg->prefix_hi, ws.prefix_hi, N * sizeof(float),
                                     hipMemcpyDeviceToDevice, s), "copy prefix");
        if (dbg->prefix_lo && lo)
            check_hip(hipMemcpyAsync(dbg->prefix_lo, lo, N * sizeof(float),
                                     hipMemcpyDeviceToDevice, s), "copy prefix lo");
    }
}

void SampleAliasTable::run(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed,
                           uint64_t counterBase) const {
    require(b.aliasTable && b.samples, "null sampling buffer");
    require(N > 0 && N < (1ull << 32), "N must be in [1, 2^32)");
    if (S == 0) return;
    if (m_cfg.cooperativeSampleSize)
        check_hip(wrsk::launch_alias_sample_window(
                      reinterpret_cast<const wrsk::AliasEntry*>(b.aliasTable), N, S, seed,
                      counterBase, m_cfg.cooperativeSampleSize, b.samples, s),
                  "windowed alias sample launch");
    else
        check_hip(wrsk::launch_alias_sample(reinterpret_cast<const wrsk::AliasEntry*>(b.aliasTable),
                                            N, S, seed, counterBase, b.samples, s),
                  "alias sample launch");
}

void InverseTransformSampling::run(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S,
                                   uint64_t seed, uint64_t counterBase) const {
    require(b.cmf && b.samples, "null sampling buffer");
    require(N > 0 && N < (1ull << 32), "N must be in [1, 2^32)");
    if (S == 0) return;
    if (m_cfg.cooperativeSamplingSize)
        check_hip(wrsk::launch_its_sample_window(b.cmf, b.cmfLo, N, S, seed, counterBase,
                                                 m_cfg.cooperativeSamplingSize,
                                                 m_cfg.pArraySearch ? 1 : 0, b.samples, s),
                  "windowed its sample launch");
    else
        check_hip(wrsk::launch_its_sample(b.cmf, b.cmfLo, N, S, seed, counterBase, b.samples, s),
                  "its sample launch");
}

static void require_cutpoint(const float* hi, uint64_t N, const uint32_t* guide, uint64_t m) {
    require(hi && guide, "null cutpoint buffer");
    require(N > 0 && N < (1ull << 32), "N must be in [1, 2^32)");
    require(m > 0 && m < (1ull << 32), "guide size must be in [1, 2^32)");
}

void Cutpoint::build(hipStream_t s, const Buffers& b, uint64_t N) const {
    require(b.internals != nullptr, "null workspace");
    m_scan.run(s, PrefixSum::Buffers{b.weights, b.cdf, b.cdfLo}, N, *b.internals);
    const uint64_t m = guideSize(N);
    require_cutpoint(b.cdf, N, b.guide, m);
    const float* lo = wrsk::prec_has_lo(m_cfg.precision) ? b.cdfLo : nullptr;
    check_hip(wrsk::launch_cutpoint_build(b.cdf, lo, N, b.guide, m, s), "guide launch");
}

void Cutpoint::sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed,
                      uint64_t counterBase) const {
    const uint64_t m = guideSize(N);
    require_cutpoint(b.cdf, N, b.guide, m);
    require(b.samples || S == 0, "null sampling buffer");
    if (S == 0) return;
    const float* lo = wrsk::prec_has_lo(m_cfg.precision) ? b.cdfLo : nullptr;
    check_hip(wrsk::launch_cutpoint_sample(b.cdf, lo, N, b.guide, m, S, seed, counterBase,
                                           b.samples, s),
              "cutpoint sample launch");
}

void AliasTable::build(hipStream_t s, const Buffers& b, uint64_t N) const {
    m_psa.run(s, PSA::Buffers{b.weights, b.aliasTable, b.internals}, N);
}
void AliasTable::sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S,
                        uint64_t seed) const {
    m_sampler.run(s, SampleAliasTable::Buffers{b.aliasTable, b.samples}, N, S, seed);
}
void ITS::build(hipStream_t s, const Buffers& b, uint64_t N) const {
    m_scan.run(s, PrefixSum::Buffers{b.weights, b.cdf, b.cdfLo}, N, *b.internals);
}
void ITS::sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed) const {
    m_sampler.run(s, InverseTransformSampling::Buffers{b.cdf, b.cdfLo, b.samples}, N, S, seed);
}

WRS::WRS(WRSConfig cfg)
    : m_method(std::holds_alternative<PSAConfig>(cfg)
                   ? std::variant<AliasTable, ITS>(AliasTable(std::get<PSAConfig>(cfg)))
                   : std::variant<AliasTable, ITS>(ITS(std::get<wrs_precision>(cfg)))) {}

void WRS::build(hipStream_t s, const Buffers& b, uint64_t N) const {
    if (auto* a = std::get_if<AliasTable>(&m_method))
        a->build(s, AliasTable::Buffers{b.weights, b.aliasTable, b.samples, b.internals}, N);
    else
        std::get<ITS>(m_method).build(s, ITS::Buffers{b.weights, b.cdf, b.cdfLo, b.samples, b.internals},
                                      N);
}
void WRS::sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed) const {
    if (auto* a = std::get_if<AliasTable>(&m_method))
        a->sample(s, AliasTable::Buffers{b.weights, b.aliasTable, b.samples, b.internals}, N, S,
                  seed);
    else
        std::get<ITS>(m_method).sample(
            s, ITS::Buffers{b.weights, b.cdf, b.cdfLo, b.samples, b.internals}, N, S, seed);
}

}  // namespace wrs

// ===================================================================== C ABI
struct wrs_plan {
    wrs::PSAConfig cfg;
    wrs::Workspace ws;
    bool timing = false;
    hipEvent_t ev[WRS_PHASE_COUNT + 1] = {};  // brackets the build's kernels
};

namespace {

template <class F>
wrs_status guarded(F&& f) {
    try {
        f();
        return WRS_OK;
    } catch (const wrs::Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return WRS_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return WRS_EHIP;
    }
}

wrs::PSAConfig config_of(const wrs_config* cfg) {
    wrs::PSAConfig c;
    if (cfg) {
        c.splitSize = cfg->split_size ? cfg->split_size : 32u;
        c.precision = cfg->precision;
        c.tuning = cfg->tuning;
    }
    wrs::require_tuning(c.tuning);
    return c;
}

void require_precision(wrs_precision p) {
    if ((int)p < (int)WRS_PREC_FP32 || (int)p > (int)WRS_PREC_FP64)
        throw wrs::Error(WRS_EINVAL, "unknown precision");
}

void require_plan(const wrs_plan* p) {
    if (!p) throw wrs::Error(WRS_EINVAL, "plan is null");
}

hipStream_t hs(wrs_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

wrs_status wrs_plan_create(const wrs_config* cfg, uint64_t max_n, wrs_plan** out) {
    return guarded([&] {
        if (!out) throw wrs::Error(WRS_EINVAL, "out is null");
        *out = nullptr;
        if (max_n == 0 || max_n >= (1ull << 32))
            throw wrs::Error(WRS_EINVAL, "max_n must be in [1, 2^32)");
        wrs::PSAConfig c = config_of(cfg);
        require_precision(c.precision);
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
            throw wrs::Error(WRS_ENOTSUP, "no HIP device");
        wrs_plan* p = new wrs_plan;
        p->cfg = c;
        try {
            p->ws = wrs::Workspace::allocate(max_n, true);
            for (auto& e : p->ev) wrs::check_hip(hipEventCreate(&e), "hipEventCreate");
        } catch (...) {
            p->ws.release();
            for (auto& e : p->ev)
                if (e) (void)hipEventDestroy(e);
            delete p;
            throw;
        }
        *out = p;
    });
}

wrs_status wrs_plan_destroy(wrs_plan* plan) {
    return guarded([&] {
        if (!plan) return;
        plan->ws.release();
        for (auto& e : plan->ev)
            if (e) (void)hipEventDestroy(e);
        delete plan;
    });
}

wrs_status wrs_psa_build(wrs_plan* plan, const float* weights, uint64_t n, wrs_alias_entry* table,
                         wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        wrs::PSA(plan->cfg).run(hs(stream), wrs::PSA::Buffers{weights, table, &plan->ws}, n,
                                nullptr, plan->timing ? plan->ev : nullptr);
    });
}

wrs_status wrs_psa_build_debug(wrs_plan* plan, const float* weights, uint64_t n,
                               wrs_alias_entry* table, const wrs_psa_debug* dbg,
                               wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        wrs::PSA(plan->cfg).run(hs(stream), wrs::PSA::Buffers{weights, table, &plan->ws}, n, dbg,
                                plan->timing ? plan->ev : nullptr);
    });
}

wrs_status wrs_mean(wrs_plan* plan, const float* weights, uint64_t n, double* out2,
                    wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        if (!out2) throw wrs::Error(WRS_EINVAL, "out is null");
        wrs::Mean(plan->cfg.precision).run(hs(stream), weights, n, plan->ws);
        wrs::check_hip(hipMemcpyAsync(out2, plan->ws.params, 2 * sizeof(double),
                                      hipMemcpyDeviceToDevice, hs(stream)),
                       "copy mean");
    });
}

wrs_status wrs_prefix_partition(wrs_plan* plan, const float* weights, uint64_t n, double pivot,
                                uint32_t* indices, float* prefix_hi, float* prefix_lo,
                                uint32_t* heavy_count, wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        if (!heavy_count) throw wrs::Error(WRS_EINVAL, "heavy_count is null");
        wrs::Mean(plan->cfg.precision).run(hs(stream), weights, n, plan->ws, pivot);
        wrs::PrefixPartition(plan->cfg.precision)
            .run(hs(stream), wrs::PrefixPartition::Buffers{weights, indices, prefix_hi, prefix_lo},
                 n, plan->ws);
        wrs::check_hip(hipMemcpyAsync(heavy_count, &plan->ws.params->heavy_count, sizeof(uint32_t),
                                      hipMemcpyDeviceToDevice, hs(stream)),
                       "copy H");
    });
}

wrs_status wrs_scan_inclusive(wrs_plan* plan, const float* in, uint64_t n, float* out_hi,
                              float* out_lo, wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        wrs::PrefixSum(plan->cfg.precision, false)
            .run(hs(stream), wrs::PrefixSum::Buffers{in, out_hi, out_lo}, n, plan->ws);
    });
}

wrs_status wrs_its_build(wrs_plan* plan, const float* weights, uint64_t n, float* cdf_hi,
                         float* cdf_lo, wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        wrs::PrefixSum(plan->cfg.precision, true)
            .run(hs(stream), wrs::PrefixSum::Buffers{weights, cdf_hi, cdf_lo}, n, plan->ws);
    });
}

wrs_status wrs_alias_sample(const wrs_alias_entry* table, uint64_t n, uint64_t s, uint64_t seed,
                            uint64_t counter_base, uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        wrs::SampleAliasTable().run(hs(stream), wrs::SampleAliasTable::Buffers{table, out}, n, s,
                                    seed, counter_base);
    });
}

wrs_status wrs_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n, uint64_t s,
                          uint64_t seed, uint64_t counter_base, uint32_t* out,
                          wrs_stream_t stream) {
    return guarded([&] {
        wrs::InverseTransformSampling().run(
            hs(stream), wrs::InverseTransformSampling::Buffers{cdf_hi, cdf_lo, out}, n, s, seed,
            counter_base);
    });
}

wrs_status wrs_alias_sample_window(const wrs_alias_entry* table, uint64_t n, uint64_t s,
                                   uint64_t seed, uint64_t counter_base, uint32_t window,
                                   uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        if (window == 0) throw wrs::Error(WRS_EINVAL, "window must be >= 1");
        wrs::SampleAliasTable(wrs::SampleAliasTableConfig{window})
            .run(hs(stream), wrs::SampleAliasTable::Buffers{table, out}, n, s, seed, counter_base);
    });
}

wrs_status wrs_its_sample_window(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                 uint64_t s, uint64_t seed, uint64_t counter_base,
                                 uint32_t window, int parts, uint32_t* out,
                                 wrs_stream_t stream) {
    return guarded([&] {
        if (window == 0) throw wrs::Error(WRS_EINVAL, "window must be >= 1");
        wrs::InverseTransformSampling(wrs::InverseTransformSamplingConfig{window, parts != 0})
            .run(hs(stream), wrs::InverseTransformSampling::Buffers{cdf_hi, cdf_lo, out}, n, s,
                 seed, counter_base);
    });
}

static_assert(WRS_ITS_NOALLOC_MAX_N == wrsk::kItsNoAllocMaxN, "static ITS index limit");

wrs_status wrs_its_sample_noalloc(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                  uint64_t s, uint64_t seed, uint64_t counter_base,
                                  uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        wrs::require(cdf_hi && out, "null sampling buffer");
        wrs::require(n > 0 && n <= WRS_ITS_NOALLOC_MAX_N,
                     "n must be in [1, WRS_ITS_NOALLOC_MAX_N = 2^25] for the static index");
        if (s == 0) return;
        wrs::check_hip(wrsk::launch_its_sample_noalloc(cdf_hi, cdf_lo, n, s, seed, counter_base,
                                                       out, hs(stream)),
                       "its sample (static index) launch");
    });
}

wrs_status wrs_cutpoint_build(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                              uint32_t* guide, uint64_t m, wrs_stream_t stream) {
    return guarded([&] {
        wrs::require_cutpoint(cdf_hi, n, guide, m);
        wrs::check_hip(wrsk::launch_cutpoint_build(cdf_hi, cdf_lo, n, guide, m, hs(stream)),
                       "guide launch");
    });
}

wrs_status wrs_cutpoint_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                               const uint32_t* guide, uint64_t m, uint64_t s, uint64_t seed,
                               uint64_t counter_base, uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        wrs::require_cutpoint(cdf_hi, n, guide, m);
        if (!out && s) throw wrs::Error(WRS_EINVAL, "out is null");
        if (s == 0) return;
        wrs::check_hip(wrsk::launch_cutpoint_sample(cdf_hi, cdf_lo, n, guide, m, s, seed,
                                                    counter_base, out, hs(stream)),
                       "cutpoint sample launch");
    });
}

wrs_status wrs_histogram(const uint32_t* samples, uint64_t s, uint64_t n, uint64_t* hist,
                         int clear, wrs_stream_t stream) {
    return guarded([&] {
        if (!hist || (!samples && s)) throw wrs::Error(WRS_EINVAL, "null histogram buffer");
        if (n == 0 || n >= (1ull << 32)) throw wrs::Error(WRS_EINVAL, "n must be in [1, 2^32)");
        wrs::check_hip(wrsk::launch_histogram(samples, s, n, hist, clear, hs(stream)),
                       "histogram launch");
    });
}

wrs_status wrs_sample_stats(wrs_plan* plan, const float* weights, uint64_t n, const uint64_t* hist,
                            uint64_t s, double min_expected, double* out6, wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        if (!weights || !hist || !out6) throw wrs::Error(WRS_EINVAL, "null statistics buffer");
        if (n == 0 || n > plan->ws.max_n) throw wrs::Error(WRS_EINVAL, "n must be in [1, max_n]");
        if (s == 0) throw wrs::Error(WRS_EINVAL, "s must be positive");
        wrs::Workspace& ws = plan->ws;
        ws.ensure_prefix(wrsk::prec_has_lo(plan->cfg.precision));
        const bool lo = wrsk::prec_has_lo(plan->cfg.precision) && ws.prefix_lo;
        wrs::PrefixSum(plan->cfg.precision, true)
            .run(hs(stream), wrs::PrefixSum::Buffers{weights, ws.prefix_hi, lo ? ws.prefix_lo : nullptr},
                 n, ws);
        wrs::check_hip(wrsk::launch_stats(hist, weights, ws.prefix_hi, lo ? ws.prefix_lo : nullptr,
                                          n, s, min_expected, ws.params, ws.states, out6,
                                          hs(stream)),
                       "statistics launch");
    });
}

wrs_status wrs_philox(uint64_t seed, uint32_t stream_word, uint64_t base, uint64_t n,
                      uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        if (!out && n) throw wrs::Error(WRS_EINVAL, "out is null");
        if (n == 0) return;
        wrs::check_hip(wrsk::launch_philox(seed, stream_word, base, n, out, hs(stream)),
                       "philox launch");
    });
}

wrs_status wrs_generate(wrs_generator kind, double a, double b, uint64_t seed, uint64_t n,
                        float* out, wrs_stream_t stream) {
    return guarded([&] {
        if (!out && n) throw wrs::Error(WRS_EINVAL, "out is null");
        if ((int)kind < 0 || (int)kind > 3) throw wrs::Error(WRS_EINVAL, "unknown generator");
        if (n == 0) return;
        wrs::check_hip(wrsk::launch_generate((int)kind, a, b, seed, n, out, hs(stream)),
                       "generate launch");
    });
}

// ------------------------------------------------------------- batched ----
struct wrs_batch_plan {
    wrsb::BatchPlan impl;
};

wrs_status wrs_batch_plan_create(const wrs_config* cfg, const uint64_t* offsets_host,
                                 uint32_t nseg, wrs_batch_plan** out) {
    return guarded([&] {
        if (!out || !offsets_host || nseg == 0)
            throw wrs::Error(WRS_EINVAL, "bad batch plan arguments");
        *out = nullptr;
        const wrs::PSAConfig c = config_of(cfg);
        require_precision(c.precision);
        wrs_batch_plan* p = new wrs_batch_plan;
        try {
            p->impl.create(c.splitSize, (int)c.precision, c.tuning, offsets_host, nseg);
        } catch (...) {
            p->impl.destroy();
            delete p;
            throw;
        }
        *out = p;
    });
}

wrs_status wrs_batch_plan_destroy(wrs_batch_plan* plan) {
    return guarded([&] {
        if (!plan) return;
        plan->impl.destroy();
        delete plan;
    });
}

wrs_status wrs_psa_build_batched(wrs_batch_plan* plan, const float* weights,
                                 wrs_alias_entry* tables, wrs_stream_t stream) {
    return guarded([&] {
        if (!plan || !weights || !tables)
            throw wrs::Error(WRS_EINVAL, "null batched build argument");
        plan->impl.build(weights, reinterpret_cast<wrsk::AliasEntry*>(tables), hs(stream));
    });
}

wrs_status wrs_alias_sample_batched(const wrs_alias_entry* tables, const uint64_t* offsets_dev,
                                    uint32_t nseg, uint32_t per_segment, uint64_t seed,
                                    uint64_t counter_base, uint32_t* out, wrs_stream_t stream) {
    return guarded([&] {
        if (!tables || !offsets_dev || !out || nseg == 0)
            throw wrs::Error(WRS_EINVAL, "null batched sample argument");
        if (per_segment == 0) return;
        wrs::check_hip(wrsk::launch_seg_sample(reinterpret_cast<const wrsk::AliasEntry*>(tables),
                                                   offsets_dev, nseg, per_segment, seed,
                                                   counter_base, out, hs(stream)),
                       "batched sample launch");
    });
}

// ------------------------------------------------------- introspection ----
const char* wrs_last_error(void) { return g_last_error.c_str(); }
const char* wrs_version(void) { return "merian-wrs_amd 0.1 (gfx950)"; }

wrs_status wrs_plan_set_tuning(wrs_plan* plan, const wrs_tuning* tuning) {
    return guarded([&] {
        require_plan(plan);
        if (!tuning) throw wrs::Error(WRS_EINVAL, "tuning is null");
        wrs::require_tuning(*tuning);
        plan->cfg.tuning = *tuning;
    });
}

wrs_status wrs_plan_set_timing(wrs_plan* plan, int enable) {
    return guarded([&] {
        require_plan(plan);
        plan->timing = enable != 0;
    });
}

wrs_status wrs_plan_phase_times(wrs_plan* plan, float* ms, uint32_t count) {
    return guarded([&] {
        require_plan(plan);
        if (!ms && count) throw wrs::Error(WRS_EINVAL, "ms is null");
        wrs::check_hip(hipEventSynchronize(plan->ev[WRS_PHASE_COUNT]), "event sync");
        for (uint32_t k = 0; k < count && k < WRS_PHASE_COUNT; ++k)
            wrs::check_hip(hipEventElapsedTime(&ms[k], plan->ev[k], plan->ev[k + 1]),
                           "event elapsed");
    });
}

wrs_status wrs_plan_status(wrs_plan* plan, wrs_stream_t stream) {
    wrs_status st = WRS_OK;
    const wrs_status call = guarded([&] {
        require_plan(plan);
        wrs::check_hip(hipStreamSynchronize(hs(stream)), "stream sync");
        uint32_t err = 0;
        wrs::check_hip(hipMemcpy(&err, &plan->ws.params->error, sizeof(err),
                                 hipMemcpyDeviceToHost),
                       "copy error word");
        if (err) {
            st = WRS_EWAIT;
            g_last_error = std::string("a bounded wait of the last build gave up (") +
                           ((err & wrsk::kErrReduceWait) ? "total" : "") +
                           ((err & wrsk::kErrReduceWait) && (err & wrsk::kErrLookbackWait) ? ", " : "") +
                           ((err & wrsk::kErrLookbackWait) ? "look-back" : "") +
                           "): its outputs are invalid";
        }
    });
    return call != WRS_OK ? call : st;
}

wrs_status wrs_plan_set_test_hooks(wrs_plan* plan, uint32_t wait_polls, uint32_t hooks) {
    return guarded([&] {
        require_plan(plan);
        plan->ws.wait_polls = wait_polls ? wait_polls : wrsk::kDefaultWaitPolls;
        plan->ws.hooks = hooks;
    });
}

wrs_status wrs_batch_plan_status(wrs_batch_plan* plan, wrs_stream_t stream) {
    wrs_status st = WRS_OK;
    const wrs_status call = guarded([&] {
        if (!plan) throw wrs::Error(WRS_EINVAL, "plan is null");
        wrs::check_hip(hipStreamSynchronize(hs(stream)), "stream sync");
        std::vector<wrsk::Params> prm(plan->impl.nseg);
        wrs::check_hip(hipMemcpy(prm.data(), plan->impl.seg_prm,
                                 prm.size() * sizeof(wrsk::Params), hipMemcpyDeviceToHost),
                       "copy segment parameters");
        for (uint32_t g = 0; g < plan->impl.nseg; ++g)
            if (prm[g].error) {
                st = WRS_EWAIT;
                g_last_error = "a look-back wait of segment " + std::to_string(g) +
                               " gave up: the batch's outputs are invalid";
                break;
            }
    });
    return call != WRS_OK ? call : st;
}

// ------------------------------------------------------------- runtime ----
wrs_status wrs_rt_device_count(int* count) {
    return guarded([&] {
        if (!count) throw wrs::Error(WRS_EINVAL, "count is null");
        *count = 0;
        if (hipGetDeviceCount(count) != hipSuccess) *count = 0;
    });
}

wrs_status wrs_rt_device_info(char* name, size_t len, char* arch, size_t arch_len, int* cu_count,
                              uint64_t* hbm_bytes) {
    return guarded([&] {
        hipDeviceProp_t prop;
        wrs::check_hip(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        if (name && len) {
            std::strncpy(name, prop.name, len - 1);
            name[len - 1] = 0;
        }
        if (arch && arch_len) {
            std::strncpy(arch, prop.gcnArchName, arch_len - 1);
            arch[arch_len - 1] = 0;
        }
        if (cu_count) *cu_count = prop.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    });
}

wrs_status wrs_rt_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) throw wrs::Error(WRS_EINVAL, "ptr is null");
        *ptr = nullptr;
        wrs::check_hip(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
    });
}
wrs_status wrs_rt_free(void* ptr) {
    return guarded([&] {
        if (ptr) wrs::check_hip(hipFree(ptr), "hipFree");
    });
}
wrs_status wrs_rt_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guarded([&] { wrs::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "h2d"); });
}
wrs_status wrs_rt_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guarded([&] { wrs::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "d2h"); });
}
wrs_status wrs_rt_memcpy_dtod(void* dst, const void* src, size_t bytes, wrs_stream_t stream) {
    return guarded([&] {
        wrs::check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, hs(stream)), "d2d");
    });
}
wrs_status wrs_rt_memset(void* dst, int value, size_t bytes, wrs_stream_t stream) {
    return guarded([&] { wrs::check_hip(hipMemsetAsync(dst, value, bytes, hs(stream)), "memset"); });
}
wrs_status wrs_rt_stream_create(wrs_stream_t* stream) {
    return guarded([&] {
        hipStream_t s = nullptr;
        wrs::check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream create");
        *stream = reinterpret_cast<wrs_stream_t>(s);
    });
}
wrs_status wrs_rt_stream_destroy(wrs_stream_t stream) {
    return guarded([&] {
        if (stream) wrs::check_hip(hipStreamDestroy(hs(stream)), "stream destroy");
    });
}
wrs_status wrs_rt_stream_sync(wrs_stream_t stream) {
    return guarded([&] { wrs::check_hip(hipStreamSynchronize(hs(stream)), "stream sync"); });
}
wrs_status wrs_rt_device_sync(void) {
    return guarded([&] { wrs::check_hip(hipDeviceSynchronize(), "device sync"); });
}
wrs_status wrs_rt_event_create(wrs_event_t* ev) {
    return guarded([&] {
        hipEvent_t e = nullptr;
        wrs::check_hip(hipEventCreate(&e), "event create");
        *ev = reinterpret_cast<wrs_event_t>(e);
    });
}
wrs_status wrs_rt_event_destroy(wrs_event_t ev) {
    return guarded([&] {
        if (ev) wrs::check_hip(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)), "event destroy");
    });
}
wrs_status wrs_rt_event_record(wrs_event_t ev, wrs_stream_t stream) {
    return guarded([&] {
        wrs::check_hip(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), hs(stream)), "event record");
    });
}
wrs_status wrs_rt_event_elapsed(float* ms, wrs_event_t start, wrs_event_t stop) {
    return guarded([&] {
        wrs::check_hip(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)), "event sync");
        wrs::check_hip(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                           reinterpret_cast<hipEvent_t>(stop)),
                       "event elapsed");
    });
}
wrs_status wrs_rt_copy_f32(float* dst, const float* src, uint64_t n, wrs_stream_t stream) {
    return guarded([&] {
        if (n == 0) return;
        wrs::check_hip(wrsk::launch_copy_f32(dst, src, n, hs(stream)), "copy launch");
    });
}

wrs_status wrs_rt_psa_build_sample_each(wrs_plan* plan, const float* weights,
                                        const uint64_t* offsets_host, uint32_t nseg,
                                        wrs_alias_entry* tables, uint32_t per_segment,
                                        uint64_t seed, uint64_t counter_base, uint32_t* samples,
                                        wrs_stream_t stream) {
    return guarded([&] {
        require_plan(plan);
        if (!weights || !offsets_host || !tables) throw wrs::Error(WRS_EINVAL, "null argument");
        const wrs::PSA psa(plan->cfg);
        const wrs::SampleAliasTable sampler;
        for (uint32_t g = 0; g < nseg; ++g) {
            const uint64_t off = offsets_host[g];
            const uint64_t n = offsets_host[g + 1] - off;
            psa.run(hs(stream), wrs::PSA::Buffers{weights + off, tables + off, &plan->ws}, n);
            if (per_segment && samples)
                sampler.run(hs(stream),
                            wrs::SampleAliasTable::Buffers{tables + off,
                                                           samples + (uint64_t)g * per_segment},
                            n, per_segment, seed, counter_base + (uint64_t)g * per_segment);
        }
    });
}

struct wrs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

wrs_status wrs_rt_graph_begin(wrs_stream_t stream) {
    return guarded([&] {
        if (!stream) throw wrs::Error(WRS_EINVAL, "graph capture needs a non-default stream");
        wrs::check_hip(hipStreamBeginCapture(hs(stream), hipStreamCaptureModeThreadLocal),
                       "begin capture");
    });
}
wrs_status wrs_rt_graph_end(wrs_stream_t stream, wrs_graph** out) {
    return guarded([&] {
        wrs_graph* g = new wrs_graph;
        try {
            wrs::check_hip(hipStreamEndCapture(hs(stream), &g->graph), "end capture");
            wrs::check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0),
                           "graph instantiate");
        } catch (...) {
            if (g->graph) (void)hipGraphDestroy(g->graph);
            delete g;
            throw;
        }
        *out = g;
    });
}
wrs_status wrs_rt_graph_launch(wrs_graph* graph, wrs_stream_t stream) {
    return guarded([&] {
        if (!graph) throw wrs::Error(WRS_EINVAL, "graph is null");
        wrs::check_hip(hipGraphLaunch(graph->exec, hs(stream)), "graph launch");
    });
}
wrs_status wrs_rt_graph_destroy(wrs_graph* graph) {
    return guarded([&] {
        if (!graph) return;
        if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
        if (graph->graph) (void)hipGraphDestroy(graph->graph);
        delete graph;
    });
}

}  // extern "C"
