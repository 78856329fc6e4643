// wrs_ops.hpp -- C++ host layer above the kernels, mirroring the reference's
// operator classes (same names, argument meaning and build/sample split) with a
// HIP stream in place of the Vulkan command buffer:
//
//   reference (src/device/...)                       here
//   Mean<float>::run            mean/Mean.hpp        wrs::Mean::run
//   PrefixPartition<float>::run prefix_partition/    wrs::PrefixPartition::run
//   PrefixSum<float>::run       prefix_sum/          wrs::PrefixSum::run
//   PSA::run                    wrs/alias/psa/PSA.hpp wrs::PSA::run
//   SampleAliasTable::run       wrs/alias/sampling/  wrs::SampleAliasTable::run
//   InverseTransformSampling::run wrs/its/sampling/  wrs::InverseTransformSampling::run
//   AliasTable / ITS / WRS      wrs/alias, wrs/its, wrs/WRS.hpp  same names
//
// Buffers are plain device pointers owned by the caller; each class's
// Buffers::allocate sizes the internal workspace (look-back records, partition
// indices and prefixes) the way the reference's Buffers::allocate does.
// Failures throw wrs::Error carrying a wrs_status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <variant>

#include "../../include/wrs.h"
#include "internal.hpp"

namespace wrs {

class Error : public std::runtime_error {
  public:
    Error(wrs_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    wrs_status status;
};

void check_hip(hipError_t e, const char* what);
// sets the calling thread's wrs_last_error() message
void set_last_error(const std::string& msg);

// Device workspace shared by the build operators.  All sizes derive from max_n.
struct Workspace {
    uint64_t max_n = 0;
    wrsk::Params* params = nullptr;    // [1]
    uint64_t* red = nullptr;           // [kRedWords] reduce block totals (~0 between builds)
                                       // and arrival counters (0 between builds)
    uint32_t wait_polls = wrsk::kDefaultWaitPolls;  // bound of the builds' waits
    uint32_t hooks = 0;                // failure injection of the error-path tests
    uint64_t* states = nullptr;        // [state_words_for(max_n)]
    uint64_t state_words = 0;
    uint32_t* indices = nullptr;       // [max_n] (PSA only)
    float* pweights = nullptr;         // [max_n] weights in partition order (PSA only)
    float* prefix_hi = nullptr;        // [max_n] (PSA debug builds only)
    float* prefix_lo = nullptr;        // [max_n] (PSA debug builds only, modes with hi + lo)
    double* samp_h = nullptr;          // [max_n/kSample + 1] heavy prefix samples (PSA only)
    double* samp_l = nullptr;          // [max_n/kSample + 1] light prefix samples (PSA only)
    wrsk::Coarse* bounds = nullptr;    // [max_n/kMinMergeTile + 2] block-boundary splits (PSA only)
    uint32_t* lightpos = nullptr;      // [max_n/kSample + 1] direct form: element of every
                                       // kSample-th light (PSA only)
    wrsk::SplitRec* splits = nullptr;  // [split_cap] split records of the serial pack form
    uint64_t split_cap = 0;
    // the streamed pack form's heavy keys (16 bytes), element indices, rows and
    // crossing offsets ([max_n] each) and the tiles' alias ranges
    // ([tiles_for(max_n) + 1])
    void* skey = nullptr;
    uint32_t* shidx = nullptr;
    wrsk::AliasEntry* shrow = nullptr;
    uint16_t* scross = nullptr;
    uint32_t* sqa = nullptr;
    uint32_t* sseg = nullptr;
    uint64_t* shsmp = nullptr;

    static Workspace allocate(uint64_t max_n, bool partition_buffers);
    // the full prefix arrays of debug builds and of the sample statistics,
    // allocated on first use
    void ensure_prefix(bool lo);
    // the split records of the serial pack form, allocated on first use
    void ensure_splits(uint64_t K);
    // the streamed pack form's buffers, allocated on first use
    void ensure_stream();
    void release();
};

class Mean {
  public:
    explicit Mean(wrs_precision prec) : m_prec(prec) {}
    // Pass 1: FP64 total/mean into ws.params, also resets ws.states for the
    // look-back kernel that follows.  pivot: NaN -> the mean.
    void run(hipStream_t s, const float* weights, uint64_t N, Workspace& ws,
             double pivot = __builtin_nan("")) const;

  private:
    wrs_precision m_prec;
};

class PrefixPartition {
  public:
    struct Buffers {
        const float* elements;
        uint32_t* partitionIndices;
        float* partitionPrefix;    // hi word
        float* partitionPrefixLo;  // may be null
    };
    explicit PrefixPartition(wrs_precision prec) : m_prec(prec) {}
    // Requires Mean::run on the same stream first (pivot and reset records).
    void run(hipStream_t s, const Buffers& b, uint64_t N, Workspace& ws) const;

  private:
    wrs_precision m_prec;
};

class PrefixSum {
  public:
    struct Buffers {
        const float* elements;
        float* prefix;    // hi word
        float* prefixLo;  // may be null
    };
    explicit PrefixSum(wrs_precision prec, bool normalise = false)
        : m_prec(prec), m_normalise(normalise) {}
    void run(hipStream_t s, const Buffers& b, uint64_t N, Workspace& ws) const;

  private:
    wrs_precision m_prec;
    bool m_normalise;
};

struct PSAConfig {
    uint32_t splitSize = 32;
    wrs_precision precision = WRS_PREC_COMPENSATED;
    wrs_tuning tuning{};  // launch geometry (zero fields: defaults)
};

// Merge shape of a tuning (defaults filled in); throws WRS_EINVAL for a shape
// or partition form that is not supported.
wrsk::MergeShape merge_shape_of(const wrs_tuning& t);
void require_tuning(const wrs_tuning& t);

class PSA {
  public:
    struct Buffers {
        const float* weights;
        wrs_alias_entry* aliasTable;
        Workspace* internals;
    };
    explicit PSA(PSAConfig cfg) : m_cfg(cfg) {}
    uint32_t splitCount(uint64_t N) const {
        const uint64_t K = (N + m_cfg.splitSize - 1) / m_cfg.splitSize;
        return (uint32_t)(K ? K : 1);
    }
    // events (optional, 6 entries) bracket the passes for phase timing:
    // reduce, finalize, partition, coarse splits, merge pack.
    void run(hipStream_t s, const Buffers& b, uint64_t N, const wrs_psa_debug* dbg = nullptr,
             hipEvent_t* events = nullptr) const;
    const PSAConfig& config() const { return m_cfg; }

  private:
    PSAConfig m_cfg;
};

// cooperativeSampleSize 0: independent samples; T > 0: groups of 64
// consecutive counters draw from one window of T buckets (SampleAliasTableConfig,
// src/device/wrs/alias/sampling/AliasTableSampling.hpp:26-33)
struct SampleAliasTableConfig {
    uint32_t cooperativeSampleSize = 0;
};

class SampleAliasTable {
  public:
    struct Buffers {
        const wrs_alias_entry* aliasTable;
        uint32_t* samples;
    };
    explicit SampleAliasTable(SampleAliasTableConfig c = {}) : m_cfg(c) {}
    void run(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed,
             uint64_t counterBase = 0) const;

  private:
    SampleAliasTableConfig m_cfg;
};

// cooperativeSamplingSize 0: independent samples (the exact lower-bound search
// per sample); W > 0: groups of 64 consecutive counters share a window of at
// most W CDF entries, narrowed by 64-part searches (pArraySearch) or by halving
// (InverseTransformSamplingConfig, src/device/wrs/its/sampling/
// InverseTransformSampling.hpp:69-81, whose default is W = 4096 with 64-part
// search)
struct InverseTransformSamplingConfig {
    uint32_t cooperativeSamplingSize = 0;
    bool pArraySearch = true;
};

class InverseTransformSampling {
  public:
    struct Buffers {
        const float* cmf;
        const float* cmfLo;  // may be null
        uint32_t* samples;
    };
    explicit InverseTransformSampling(InverseTransformSamplingConfig c = {}) : m_cfg(c) {}
    void run(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed,
             uint64_t counterBase = 0) const;

  private:
    InverseTransformSamplingConfig m_cfg;
};

// AliasTable = PSA build + alias sampling (src/device/wrs/alias/AliasTable.hpp:75-112)
class AliasTable {
  public:
    struct Buffers {
        const float* weights;
        wrs_alias_entry* aliasTable;
        uint32_t* samples;
        Workspace* internals;
    };
    explicit AliasTable(PSAConfig cfg) : m_psa(cfg) {}
    void build(hipStream_t s, const Buffers& b, uint64_t N) const;
    void sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed) const;

  private:
    PSA m_psa;
    SampleAliasTable m_sampler;
};

// ITS = normalised scan + binary-search sampling (src/device/wrs/its/ITS.hpp:79-135)
class ITS {
  public:
    struct Buffers {
        const float* weights;
        float* cdf;
        float* cdfLo;  // may be null
        uint32_t* samples;
        Workspace* internals;
    };
    explicit ITS(wrs_precision prec) : m_scan(prec, true) {}
    void build(hipStream_t s, const Buffers& b, uint64_t N) const;
    void sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed) const;

  private:
    PrefixSum m_scan;
    InverseTransformSampling m_sampler;
};

// Cutpoint = normalised scan + guide table + guided sampling
// (src/device/wrs/cutpoint/Cutpoint.hpp:22-150); guidingEntrySize weights per
// guide entry, m = ceil(N / guidingEntrySize).
struct CutpointConfig {
    wrs_precision precision = WRS_PREC_COMPENSATED;
    uint32_t guidingEntrySize = 1;
};
class Cutpoint {
  public:
    struct Buffers {
        const float* weights;
        float* cdf;
        float* cdfLo;     // may be null
        uint32_t* guide;  // [guideSize(N) + 1]
        uint32_t* samples;
        Workspace* internals;
    };
    explicit Cutpoint(CutpointConfig cfg) : m_cfg(cfg), m_scan(cfg.precision, true) {}
    uint64_t guideSize(uint64_t N) const {
        const uint64_t e = m_cfg.guidingEntrySize ? m_cfg.guidingEntrySize : 1u;
        const uint64_t m = (N + e - 1) / e;
        return m ? m : 1;
    }
    void build(hipStream_t s, const Buffers& b, uint64_t N) const;
    void sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S, uint64_t seed = 12345u,
                uint64_t counterBase = 0) const;

  private:
    CutpointConfig m_cfg;
    PrefixSum m_scan;
};

// WRS facade: variant dispatch as in src/device/wrs/WRS.hpp:91-172.
using WRSConfig = std::variant<PSAConfig /* alias */, wrs_precision /* ITS */>;
class WRS {
  public:
    struct Buffers {
        const float* weights;
        uint32_t* samples;
        wrs_alias_entry* aliasTable;  // alias method
        float* cdf;                   // ITS method
        float* cdfLo;
        Workspace* internals;
    };
    explicit WRS(WRSConfig cfg);
    void build(hipStream_t s, const Buffers& b, uint64_t N) const;
    void sample(hipStream_t s, const Buffers& b, uint64_t N, uint64_t S,
                uint64_t seed = 12345u) const;

  private:
    std::variant<AliasTable, ITS> m_method;
};

}  // namespace wrs
