// tune.cpp -- the autotuner: picks the launch geometry (wrs_tuning) of a PSA
// build or of a batched build by timing candidates on the caller's data.
//
// Search: coordinate descent over the parts of a build.
//   1. the prefix-partition form (reduce-then-scan or single pass with
//      decoupled look-back, per segment in a batch), at the default merge shape;
//   2. the pack form (classic or direct) with the form of step 1 (a batch builds
//      the direct form with the reduce-then-scan partition only, so there the
//      direct candidate takes that partition form); a single build also times
//      the reference's serial and inline split + pack at the plan's split size;
//   3. every supported merge shape (buckets per block x buckets per thread)
//      with the forms of steps 1 and 2 (none when a sweep form won: it has no
//      merge shape);
//   4. the two fastest candidates again, each after kWarm builds of its own,
//      and the faster of those two.
// A candidate's time is the median of kReps builds (HIP events on the
// caller's stream) after one warm-up build; kWarm builds of the first
// candidate run untimed before the search, so that the first candidate is not
// timed on an idle (down-clocked) device.  The partition form only changes
// the partition kernels and the merge shape only the split/merge kernels, so
// the steps interact little beyond the shared memory bandwidth.
//
// Results are cached per process under a key made of the device (name, CU
// count), the precision and the problem shape; with WRS_TUNE_CACHE set they
// also persist in that file, one "key<TAB>tile per partition direct" line each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wrs.h"
#include "batched.hpp"
#include "internal.hpp"
#include "wrs_ops.hpp"

namespace wrst {

namespace {

constexpr int kReps = 9;
constexpr int kWarm = 20;

// the merge shapes tried, default first
const wrsk::MergeShape kShapes[] = {{1024, 4}, {1024, 2}, {1024, 8}, {2048, 2},
                                    {2048, 4}, {2048, 8}, {4096, 4}, {4096, 8}};

std::mutex g_mu;
std::map<std::string, wrs_tuning> g_cache;
bool g_file_loaded = false;

const char* cache_path() {
    const char* p = std::getenv("WRS_TUNE_CACHE");
    return (p && *p) ? p : nullptr;
}

void load_file_locked() {
    if (g_file_loaded) return;
    g_file_loaded = true;
    const char* path = cache_path();
    if (!path) return;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos) continue;
        std::istringstream vals(line.substr(tab + 1));
        wrs_tuning t{};
        if (vals >> t.merge_tile >> t.merge_per >> t.partition) {
            if (!(vals >> t.direct)) continue;
            try {
                wrs::require_tuning(t);
            } catch (...) {
                continue;  // a line this build does not support
            }
            g_cache[line.substr(0, tab)] = t;
        }
    }
}

bool cache_get(const std::string& key, wrs_tuning& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_file_locked();
    auto it = g_cache.find(key);
    if (it == g_cache.end()) return false;
    out = it->second;
    return true;
}

void cache_put(const std::string& key, const wrs_tuning& t) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_file_locked();
    g_cache[key] = t;
    if (const char* path = cache_path()) {
        std::ofstream out(path, std::ios::app);
        if (out)
            out << key << '\t' << t.merge_tile << ' ' << t.merge_per << ' ' << t.partition << ' '
                << t.direct << '\n';
    }
}

std::string device_key() {
    int dev = 0;
    wrs::check_hip(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    wrs::check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::ostringstream k;
    // "t5": the tuning encoding (WRS_PACK_* pack forms, sweep and streamed
    // forms included); lines of other encodings in a cache file do not match
    k << "t5/" << prop.gcnArchName << "/cu" << prop.multiProcessorCount;
    return k.str();
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        wrs::check_hip(hipEventCreate(&a), "hipEventCreate");
        wrs::check_hip(hipEventCreate(&b), "hipEventCreate");
    }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// median time (ms) of kReps runs of fn on stream s, after one warm-up
template <class F>
float time_median(F&& fn, hipStream_t s, Events& ev, int warm = 1) {
    for (int r = 0; r < warm; ++r) fn();
    wrs::check_hip(hipStreamSynchronize(s), "tune sync");
    std::vector<float> t(kReps);
    for (int r = 0; r < kReps; ++r) {
        wrs::check_hip(hipEventRecord(ev.a, s), "event");
        fn();
        wrs::check_hip(hipEventRecord(ev.b, s), "event");
        wrs::check_hip(hipEventSynchronize(ev.b), "event sync");
        wrs::check_hip(hipEventElapsedTime(&t[r], ev.a, ev.b), "event time");
    }
    std::sort(t.begin(), t.end());
    return t[kReps / 2];
}

struct Log {
    wrs_tune_entry* log;
    uint32_t cap;
    uint32_t n = 0;
    void add(const wrs_tuning& t, float ms) {
        if (log && n < cap) log[n] = wrs_tune_entry{t, ms};
        ++n;
    }
};

wrs_tuning with_shape(wrs_tuning t, const wrsk::MergeShape& ms) {
    t.merge_tile = ms.tile;
    t.merge_per = ms.per;
    return t;
}

}  // namespace

// Coordinate descent over a measuring function: the partition form at the
// default merge shape, the pack form, then every other merge shape with the
// better forms.
template <class M>
wrs_tuning descend(M&& measure_raw, bool direct_rts_only) {
    // every candidate timed, for the confirmation round
    std::vector<std::pair<float, wrs_tuning>> seen;
    auto measure = [&](const wrs_tuning& t, int warm = 1) {
        const float ms = measure_raw(t, warm);
        seen.emplace_back(ms, t);
        return ms;
    };
    wrs_tuning t0 = with_shape(wrs_tuning{}, kShapes[0]);
    t0.direct = WRS_PACK_CLASSIC;
    wrs_tuning best = t0;
    float best_ms = measure(t0, kWarm);
    wrs_tuning t1 = t0;
    t1.partition = 1;
    const float ms1 = measure(t1);
    if (ms1 < best_ms) {
        best_ms = ms1;
        best = t1;
    }
    {
        wrs_tuning t2 = best;
        t2.direct = WRS_PACK_DIRECT;
        if (direct_rts_only) t2.partition = 0;
        const float ms2 = measure(t2);
        if (ms2 < best_ms) {
            best_ms = ms2;
            best = t2;
        }
    }
    bool sweep = false;
    if (!direct_rts_only) {  // the reference's sweep forms and the streamed form (single builds)
        for (const uint32_t f :
             {WRS_PACK_SERIAL, WRS_PACK_INLINE, WRS_PACK_SUBGROUP, WRS_PACK_STREAM}) {
            wrs_tuning t3 = best;
            t3.direct = f;
            const float ms3 = measure(t3);
            if (ms3 < best_ms) {
                best_ms = ms3;
                best = t3;
            }
        }
        // the merge shape does not apply to them (nor to the streamed form)
        sweep = best.direct >= WRS_PACK_SERIAL;
    }
    if (!sweep) {
        const wrs_tuning form = best;
        for (size_t i = 1; i < sizeof(kShapes) / sizeof(kShapes[0]); ++i) {
            const wrs_tuning t = with_shape(form, kShapes[i]);
            const float ms = measure(t);
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    }
    // confirmation: the two fastest candidates again, each after kWarm builds
    // of its own (a candidate timed right after a slower one can read slow:
    // the first builds after other work run below the steady rate)
    std::stable_sort(seen.begin(), seen.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    if (seen.size() >= 2) {
        const wrs_tuning a = seen[0].second, b = seen[1].second;
        const float ma = measure_raw(a, kWarm);
        const float mb = measure_raw(b, kWarm);
        best = mb < ma ? b : a;
    }
    return best;
}

// PSA build of one distribution of n weights.
wrs_tuning tune_psa(const wrs::PSAConfig& base, const float* w, uint64_t n, uint32_t flags,
                    hipStream_t s, Log& log) {
    std::ostringstream key;
    key << device_key() << "/p" << (int)base.precision << "/psa/n" << n << "/s" << base.splitSize;
    wrs_tuning best{};
    if (!(flags & WRS_TUNE_NO_CACHE) && cache_get(key.str(), best)) return best;

    wrs::Workspace ws = wrs::Workspace::allocate(n, true);
    wrs_alias_entry* table = nullptr;
    try {
        wrs::check_hip(hipMalloc(&table, n * sizeof(wrs_alias_entry)), "hipMalloc");
        Events ev;
        best = descend([&](const wrs_tuning& t, int warm = 1) {
            wrs::PSAConfig c = base;
            c.tuning = t;
            const wrs::PSA psa(c);
            const float ms = time_median(
                [&] { psa.run(s, wrs::PSA::Buffers{w, table, &ws}, n); }, s, ev, warm);
            log.add(t, ms);
            return ms;
        }, false);
    } catch (...) {
        if (table) (void)hipFree(table);
        ws.release();
        throw;
    }
    (void)hipFree(table);
    ws.release();
    cache_put(key.str(), best);
    return best;
}

// Batched build: the partition form, the pack form, then the merge shapes.
wrs_tuning tune_batched(const wrs::PSAConfig& base, const uint64_t* offsets, uint32_t nseg,
                        const float* w, uint32_t flags, hipStream_t s, Log& log) {
    // problem shape: segment count, total, histogram of floor(log2(size))
    uint32_t hist[33] = {};
    for (uint32_t g = 0; g < nseg; ++g) {
        const uint64_t sz = offsets[g + 1] - offsets[g];
        int b = 0;
        while (b < 32 && (sz >> (b + 1)) != 0) ++b;
        ++hist[b];
    }
    std::ostringstream key;
    key << device_key() << "/p" << (int)base.precision << "/batch/s" << nseg << "/n"
        << offsets[nseg] << "/h";
    for (int b = 0; b < 33; ++b)
        if (hist[b]) key << b << ':' << hist[b] << ',';
    wrs_tuning best{};
    if (!(flags & WRS_TUNE_NO_CACHE) && cache_get(key.str(), best)) return best;

    wrsb::BatchPlan bp;
    wrsk::AliasEntry* tables = nullptr;
    try {
        bp.create(base.splitSize, (int)base.precision, wrs_tuning{}, offsets, nseg);
        wrs::check_hip(hipMalloc(&tables, bp.total_n * sizeof(wrsk::AliasEntry)), "hipMalloc");
        Events ev;
        best = descend([&](const wrs_tuning& t, int warm = 1) {
            bp.set_tuning(t);
            const float ms = time_median([&] { bp.build(w, tables, s); }, s, ev, warm);
            log.add(t, ms);
            return ms;
        }, true);
    } catch (...) {
        if (tables) (void)hipFree(tables);
        bp.destroy();
        throw;
    }
    (void)hipFree(tables);
    bp.destroy();
    cache_put(key.str(), best);
    return best;
}

}  // namespace wrst

// ------------------------------------------------------------------- C ABI --
namespace {

template <class F>
wrs_status tune_guarded(F&& f) {
    try {
        f();
        return WRS_OK;
    } catch (const wrs::Error& e) {
        wrs::set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        wrs::set_last_error("host allocation failed");
        return WRS_ENOMEM;
    } catch (const std::exception& e) {
        wrs::set_last_error(e.what());
        return WRS_EHIP;
    }
}

wrs::PSAConfig tune_base(const wrs_config* cfg) {
    wrs::PSAConfig c;
    if (cfg) {
        c.splitSize = cfg->split_size ? cfg->split_size : 32u;
        c.precision = cfg->precision;
    }
    if ((int)c.precision < (int)WRS_PREC_FP32 || (int)c.precision > (int)WRS_PREC_FP64)
        throw wrs::Error(WRS_EINVAL, "unknown precision");
    return c;
}
}  // namespace

extern "C" {

wrs_status wrs_autotune_psa(const wrs_config* cfg, const float* weights, uint64_t n,
                            uint32_t flags, wrs_tuning* best, wrs_tune_entry* log,
                            uint32_t log_cap, uint32_t* log_len, wrs_stream_t stream) {
    return tune_guarded([&] {
        if (!weights || !best) throw wrs::Error(WRS_EINVAL, "weights and best must be non-null");
        if (n == 0 || n >= (1ull << 32)) throw wrs::Error(WRS_EINVAL, "n must be in [1, 2^32)");
        wrst::Log lg{log, log_cap};
        *best = wrst::tune_psa(tune_base(cfg), weights, n, flags,
                               reinterpret_cast<hipStream_t>(stream), lg);
        if (log_len) *log_len = lg.n;
    });
}

wrs_status wrs_autotune_batched(const wrs_config* cfg, const uint64_t* offsets_host,
                                uint32_t nseg, const float* weights, uint32_t flags,
                                wrs_tuning* best, wrs_tune_entry* log, uint32_t log_cap,
                                uint32_t* log_len, wrs_stream_t stream) {
    return tune_guarded([&] {
        if (!weights || !best || !offsets_host || nseg == 0)
            throw wrs::Error(WRS_EINVAL, "bad autotune arguments");
        wrst::Log lg{log, log_cap};
        *best = wrst::tune_batched(tune_base(cfg), offsets_host, nseg, weights, flags,
                                   reinterpret_cast<hipStream_t>(stream), lg);
        if (log_len) *log_len = lg.n;
    });
}

}  // extern "C"
