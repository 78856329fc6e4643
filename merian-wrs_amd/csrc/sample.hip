// sample.hip -- alias-table sampling, ITS sampling, Philox streams and the
// seeded synthetic weight generators.
//
// Alias sampling replaces src/device/wrs/alias/sampling/shader.comp:80-111
// (non-cooperative path): sample t uses Philox4x32-10 with
// counter = (lo(base+t), hi(base+t), 0, 0) and key = seed; the 64-bit word
// (x1:x0) picks the bucket by a 64x64->128 multiply-high (no modulo bias at
// any N < 2^32), x2 gives the 24-bit coin u2, and the acceptance rule is the
// reference CPU's (src/host/reference/sample_alias_table.hpp:15-24):
// u2 > p -> alias, else the bucket.
//
// ITS sampling replaces src/device/wrs/its/sampling/shader.comp:142-177
// (non-cooperative): u in (0,1] from the 53-bit word (x0:x1), then the
// reference's lower-bound binary search (smallest i with cdf(i) >= u,
// shader.comp:131-140) over cdf(i) = hi[i] (+ lo[i] when given), in FP64.
#include <algorithm>
#include <map>
#include <mutex>

#include "device_common.hpp"
#include "internal.hpp"

namespace wrsk {
using namespace wrsdev;

constexpr int kSampleBlock = 256;
// CDFs longer than this are searched through line indexes (k_its_sample_lines);
// shorter ones stay in the caches anyway
constexpr uint64_t kItsLineIndexMin = 1u << 16;

__global__ __launch_bounds__(kSampleBlock) void k_alias_sample(const uint2* __restrict__ table,
                                                               uint64_t n, uint64_t s,
                                                               uint64_t seed, uint64_t base,
                                                               uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kSampleBlock + threadIdx.x; g < s; g += stride) {
        const U4 r = philox_at(seed, base + g, 0u, 0u);
        const uint64_t x = ((uint64_t)r.y << 32) | r.x;
        const uint64_t ix = __umul64hi(x, n);
        const float u2 = u24(r.z);
        const uint2 e = table[ix];
        out[g] = (u2 > __uint_as_float(e.x)) ? e.y : (uint32_t)ix;
    }
}

template <bool LO>
__global__ __launch_bounds__(kSampleBlock) void k_its_sample(const float* __restrict__ hi,
                                                             const float* __restrict__ lo,
                                                             uint64_t n, uint64_t s,
                                                             uint64_t seed, uint64_t base,
                                                             uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kSampleBlock + threadIdx.x; g < s; g += stride) {
        const U4 r = philox_at(seed, base + g, 0u, 0u);
        const double u = u53_open0(r.x, r.y);
        uint64_t a = 0, b = n - 1;
        while (a < b) {
            const uint64_t mid = (a + b) >> 1;
            double c = (double)hi[mid];
            if (LO) c += (double)lo[mid];
            if (u > c) a = mid + 1;
            else b = mid;
        }
        out[g] = (uint32_t)a;
    }
}

// c(i) = hi[i] (+ lo[i]), in FP64
template <bool LO>
__device__ __forceinline__ double cdf_at(const float* __restrict__ hi, const float* __restrict__ lo,
                                         uint64_t i) {
    double c = (double)hi[i];
    if (LO) c += (double)lo[i];
    return c;
}

// ---- ITS sampling through line indexes -------------------------------------------
//
// The binary search above reads a new 128-byte line of hi (and of lo) at each
// of its lowest ~11 levels: on N = 2^24 that is ~2.8 KB of random lines per
// sample, and the sampler runs at the rate of those reads.  Here the CDF gets
// a hierarchy of line indexes: level 1 holds the last hi of every line of 32
// hi words, level l + 1 the last entry of every line of level l, up to a top
// level of at most kItsTop entries that each block keeps in LDS.  A sample
// searches the top level in LDS, then reads one 128-byte line per level below
// it (as eight 16-byte loads, one round trip) and counts its entries below the
// key -- the position of the first entry at or above it -- down to one line of
// hi; lo is read only for the final exact comparisons.  Since |lo| <= ulp(hi)
// / 2 and hi is non-decreasing, every element whose hi is below u_lo (the
// float two ulps under u, rounded down) has hi + lo < u, so the first element
// with hi >= u_lo bounds the answer from below; the answer -- the smallest i
// with hi[i] + lo[i] >= u, at most n - 1, as in k_its_sample -- follows by
// exact comparisons from there (a short step forward; a binary search if many
// elements share that hi).  On N = 2^24: ~2 lines of HBM traffic per sample.
constexpr uint32_t kLine = 32;      // floats per 128-byte line
constexpr uint32_t kItsLevels = 6;  // index levels at most (n < 2^32)
constexpr uint32_t kItsTop = 4096;  // entries of the top level, kept in LDS
constexpr uint32_t kItsBlocks = 2048;  // sampler blocks (8 per CU of the 256)

struct ItsIndex {
    const float* lvl[kItsLevels + 1];  // lvl[0] = hi, lvl[l] = index level l
    uint64_t len[kItsLevels + 1];      // entries of each level
    uint32_t top;                      // the top level (its length <= kItsTop)
};

// Index levels 1..top from hi: entry k of level 1 is hi of element 32k + 31
// (the last element for the last, partial line); thread k of level 1 also
// writes entry k / 32^(l-1) of level l when it holds that line's last entry.
__global__ __launch_bounds__(256) void k_its_index(const float* __restrict__ hi, uint64_t n,
                                                   float* __restrict__ idx, ItsIndex ix) {
    const uint64_t m1 = ix.len[1];
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x; k < m1; k += stride) {
        const uint64_t i = k * kLine + kLine - 1u;
        const float v = hi[i < n ? i : n - 1u];
        uint64_t span = 1;  // level-1 entries per entry of level l
        for (uint32_t l = 1; l <= ix.top; ++l) {
            if ((k + 1u) % span != 0 && k + 1u != m1) break;
            const_cast<float*>(ix.lvl[l])[k / span] = v;
            span *= kLine;
        }
    }
    (void)idx;
}

// u > hi[i] + lo[i], reading lo only when u lies within an ulp of hi[i] (a
// split double word has |lo| <= ulp(hi) / 2): random samples almost never do,
// so the search touches one line of hi per sample and none of lo.
template <bool LO>
__device__ __forceinline__ bool u_above(double u, const float* __restrict__ hi,
                                        const float* __restrict__ lo, uint64_t i) {
    const float h = hi[i];
    if (!LO) return u > (double)h;
    const double hd = (double)h;
    const double ulp = (double)nextafterf(fabsf(h), __builtin_huge_valf()) - (double)fabsf(h);
    if (u > hd + ulp) return true;
    if (u < hd - ulp) return false;
    return u > hd + (double)lo[i];
}

// One level of the descent for the 64 samples of a wave: lane l holds its
// sample's entry e (a line of the level below) and key; returns the position
// in the level below of the first entry >= key in that line (entries past
// `len` count as >= key).  The lines are read cooperatively: in round r the
// eight lanes of group g read the line of lane 8r + g as eight 16-byte words
// (one 128-byte request for the line) and sum their counts, so a wave issues
// one request per line instead of eight.
template <bool LEAF>
__device__ __forceinline__ uint64_t descend_wave(const float* __restrict__ lv, uint64_t len,
                                                 uint64_t e, float key) {
    const int lane = lane_id();
    const int grp = lane >> 3, sub = lane & 7;
    uint32_t mine = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int src = 8 * r + grp;
        const uint64_t es = shfl(e, src);
        const float ks = shfl(key, src);
        const uint64_t p0 = es * kLine + (uint64_t)sub * 4u;
        uint32_t c;
        if (p0 + 4u <= len) {
            const float4 v = *reinterpret_cast<const float4*>(lv + p0);
            c = (v.x < ks) + (v.y < ks) + (v.z < ks) + (v.w < ks);
        } else {
            c = 0;
            for (uint64_t p = p0; p < len && p < p0 + 4u; ++p) c += lv[p] < ks;
        }
        c += __shfl_xor(c, 1, kWave);
        c += __shfl_xor(c, 2, kWave);
        c += __shfl_xor(c, 4, kWave);
        const uint32_t back = shfl(c, ((lane - 8 * r) & 7) * 8);  // group (lane - 8r)'s sum
        if (grp == r) mine = back;
    }
    return e * kLine + mine;
}

template <bool LO>
__global__ __launch_bounds__(kSampleBlock) void k_its_sample_lines(
    const float* __restrict__ hi, const float* __restrict__ lo, ItsIndex ix, uint64_t n,
    uint64_t s, uint64_t seed, uint64_t base, uint32_t* __restrict__ out) {
    __shared__ float s_top[kItsTop];
    const uint32_t top = ix.top;
    const uint64_t ntop = ix.len[top];
    for (uint32_t k = threadIdx.x; k < ntop; k += kSampleBlock) s_top[k] = ix.lvl[top][k];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    // the descent is wave-cooperative: every lane runs every round of its
    // wave (lanes past s compute a dummy sample and store nothing)
    const uint64_t wave0 = (uint64_t)blockIdx.x * kSampleBlock + (threadIdx.x & ~(kWave - 1));
    for (uint64_t gw = wave0; gw < s; gw += stride) {
        const uint64_t g = gw + (uint64_t)lane_id();
        const U4 r = philox_at(seed, base + (g < s ? g : 0u), 0u, 0u);
        const double u = u53_open0(r.x, r.y);
        const float f = __double2float_rd(u);
        const float key = LO ? nextafterf(nextafterf(f, -__builtin_huge_valf()),
                                         -__builtin_huge_valf())
                             : f;
        // top level: first entry >= key, in LDS
        uint32_t a = 0, b = (uint32_t)ntop;
        while (a < b) {
            const uint32_t mid = (a + b) >> 1;
            if (s_top[mid] < key) a = mid + 1;
            else b = mid;
        }
        // levels below, down to hi: the whole wave descends together (a
        // sample past the top level's end follows line 0 and is then clamped)
        uint64_t e = a < ntop ? a : 0u;  // line of the level below holding the answer
        for (uint32_t l = top; l > 1; --l)
            e = descend_wave<false>(ix.lvl[l - 1], ix.len[l - 1], e, key);
        e = descend_wave<true>(hi, n, e, key);
        uint64_t i = a < ntop ? (e < n ? e : n - 1u) : n - 1u;
        // exact: the smallest i' >= i with cdf(i') >= u, at most n - 1
        int steps = 0;
        while (i < n - 1u && u_above<LO>(u, hi, lo, i) && steps < 8) {
            ++i;
            ++steps;
        }
        if (i < n - 1u && u_above<LO>(u, hi, lo, i)) {
            // a long run of elements whose CDF words share this hi (weights far
            // below an ulp of the CDF): gallop over it, then bisect
            uint64_t p = i + 1u, k = 8;
            while (p + k < n - 1u && u_above<LO>(u, hi, lo, p + k)) {
                p += k + 1u;
                k *= 2u;
            }
            uint64_t q = p + k < n - 1u ? p + k : n - 1u;
            while (p < q) {
                const uint64_t mid = (p + q) >> 1;
                if (u_above<LO>(u, hi, lo, mid)) p = mid + 1;
                else q = mid;
            }
            i = p;
        }
        if (g < s) out[g] = (uint32_t)i;
    }
}

// ---- cutpoint (guide table) acceleration of ITS sampling ---------------------
//
// Chen & Asau's guide table over the normalised CDF c(i) = hi[i] (+ lo[i]):
// guide[g] = smallest i in [0, n) with c(i) >= t_g, t_g = g / m, for g < m
// (n - 1 when there is none), and guide[m] = n - 1.  A sample's u lies in
// [t_g, t_{g+1}) for g = G(u) below, so its lower bound -- the smallest i with
// c(i) >= u -- lies in [guide[g], guide[g+1]]: the search over that range
// returns exactly what the full binary search of k_its_sample returns for the
// same u.  Replaces the reference's Cutpoint method
// (src/device/wrs/cutpoint/guiding_table/shader.comp:37-54,
// src/device/wrs/cutpoint/sampling/shader.comp:118-177), whose sampler picks the
// guide cell with a second uniform instead of from u and so weights every cell
// by 1/m whatever its mass.

__device__ __forceinline__ double cut_t(uint64_t g, uint64_t m) { return (double)g / (double)m; }

// largest g in [-1, m-1] with t_g <= c (-1: none), by the same comparison as
// the table's thresholds
__device__ __forceinline__ int64_t cut_floor(double c, uint64_t m) {
    if (!(c >= 0.0)) return -1;
    double x = floor(c * (double)m);
    int64_t g = x >= (double)m ? (int64_t)m - 1 : (int64_t)x;
    while (g + 1 <= (int64_t)m - 1 && cut_t((uint64_t)(g + 1), m) <= c) ++g;
    while (g >= 0 && cut_t((uint64_t)g, m) > c) --g;
    return g;
}


// lower bound of t in [a, b]: smallest i with c(i) >= t, b if none
template <bool LO>
__device__ __forceinline__ uint64_t cdf_lower_bound(const float* __restrict__ hi,
                                                    const float* __restrict__ lo, uint64_t a,
                                                    uint64_t b, double t) {
    while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if (t > cdf_at<LO>(hi, lo, mid)) a = mid + 1;
        else b = mid;
    }
    return a;
}

// The table is built in two passes, one thread per entry, so that the work
// does not depend on how the mass is spread: the anchors g = 64k (and g = m)
// by a search of the whole CDF, then every other entry by a search between
// its two anchors (the entries are non-decreasing in g).
constexpr uint32_t kCutAnchor = 64;

template <bool LO>
__global__ __launch_bounds__(256) void k_cutpoint_anchors(const float* __restrict__ hi,
                                                          const float* __restrict__ lo,
                                                          uint64_t n, uint32_t* __restrict__ guide,
                                                          uint64_t m) {
    const uint64_t na = (m + kCutAnchor - 1) / kCutAnchor;  // anchors below m
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= na; k += stride) {
        const uint64_t g = k < na ? k * kCutAnchor : m;
        guide[g] = g == m ? (uint32_t)(n - 1)
                          : (uint32_t)cdf_lower_bound<LO>(hi, lo, 0, n - 1, cut_t(g, m));
    }
}

template <bool LO>
__global__ __launch_bounds__(256) void k_cutpoint_fill(const float* __restrict__ hi,
                                                       const float* __restrict__ lo, uint64_t n,
                                                       uint32_t* __restrict__ guide, uint64_t m) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < m; g += stride) {
        if (g % kCutAnchor == 0) continue;
        const uint64_t ga = g - g % kCutAnchor;
        const uint64_t gb = ga + kCutAnchor < m ? ga + kCutAnchor : m;
        guide[g] = (uint32_t)cdf_lower_bound<LO>(hi, lo, guide[ga], guide[gb], cut_t(g, m));
    }
}

template <bool LO>
__global__ __launch_bounds__(kSampleBlock) void k_cutpoint_sample(
    const float* __restrict__ hi, const float* __restrict__ lo, uint64_t n,
    const uint32_t* __restrict__ guide, uint64_t m, uint64_t s, uint64_t seed, uint64_t base,
    uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    for (uint64_t t = (uint64_t)blockIdx.x * kSampleBlock + threadIdx.x; t < s; t += stride) {
        const U4 r = philox_at(seed, base + t, 0u, 0u);
        const double u = u53_open0(r.x, r.y);  // the u of k_its_sample
        int64_t g = cut_floor(u, m);
        g = g < 0 ? 0 : g;
        uint64_t a = guide[g], b = guide[g + 1];
        // a table built from another CDF must not send the search out of bounds
        if (b > n - 1) b = n - 1;
        if (a > b) a = 0;
        out[t] = (uint32_t)cdf_lower_bound<LO>(hi, lo, a, b, u);
    }
}

__global__ void k_philox(uint64_t seed, uint32_t stream_word, uint64_t base, uint64_t n,
                         uint4* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += stride) {
        const U4 r = philox_at(seed, base + g, stream_word, 0u);
        out[g] = make_uint4(r.x, r.y, r.z, r.w);
    }
}

// Seeded bijection of [0,n): 4-round Feistel network over 2*half bits with a
// Philox round function, restricted to [0,n) by cycle walking.
__device__ uint64_t feistel_perm(uint64_t seed, uint64_t n, uint64_t i) {
    unsigned bits = 1;
    while (bits < 64 && (1ull << bits) < n) ++bits;
    const unsigned half = (bits + 1) / 2;
    const uint64_t mask = (1ull << half) - 1ull;
    uint64_t x = i;
    do {
        uint64_t L = x >> half, R = x & mask;
        for (uint32_t round = 0; round < 4; ++round) {
            const U4 r = philox_at(seed, ((uint64_t)round << 32) | (uint32_t)R, 0u, kTagFeistel);
            const uint64_t F = (((uint64_t)r.y << 32) | r.x) & mask;
            const uint64_t nl = R;
            R = L ^ F;
            L = nl;
        }
        x = (L << half) | R;
    } while (x >= n);
    return x;
}

// kind: 0 U(0,1); 1 LogNormal(a, b); 2 Zipf exponent a over a seeded rank
// permutation; 3 spike mixture (1% of positions U(1e5,1e6), rest U(1e-3,1)).
__global__ void k_generate(int kind, double a, double b, uint64_t seed, uint64_t n,
                           float* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float v;
        if (kind == 0) {
            const U4 r = philox_at(seed, i, 0u, kTagGen);
            v = u24(r.x);
        } else if (kind == 1) {
            const U4 r = philox_at(seed, i, 0u, kTagGen);
            const double u1 = u53_open0(r.x, r.y);
            const double u2 = u53(r.z, r.w);
            const double z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
            v = (float)exp(a + b * z);
        } else if (kind == 2) {
            const uint64_t rank = 1u + feistel_perm(seed, n, i);
            v = (float)pow((double)rank, -a);
        } else {
            const U4 r = philox_at(seed, i, 0u, kTagGen);
            const double u = u53(r.x, r.y);
            const bool spike = feistel_perm(seed, n, i) < n / 100u;
            v = spike ? (float)(1e5 + 9e5 * u) : (float)(1e-3 + (1.0 - 1e-3) * u);
        }
        out[i] = v;
    }
}

// Batched alias sampling: sample g belongs to segment g / per_segment and
// returns a segment-local index; counter = counter_base + g.
__global__ __launch_bounds__(kSampleBlock) void k_seg_sample(const uint2* __restrict__ tables,
                                                             const uint64_t* __restrict__ seg_off,
                                                             uint64_t total, uint32_t per_segment,
                                                             uint64_t seed, uint64_t base,
                                                             uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kSampleBlock + threadIdx.x; g < total; g += stride) {
        const uint64_t sg = g / per_segment;
        const uint64_t off = seg_off[sg];
        const uint64_t n = seg_off[sg + 1] - off;
        const U4 r = philox_at(seed, base + g, 0u, 0u);
        const uint64_t x = ((uint64_t)r.y << 32) | r.x;
        const uint64_t ix = __umul64hi(x, n);
        const float u2 = u24(r.z);
        const uint2 e = tables[off + ix];
        out[g] = (u2 > __uint_as_float(e.x)) ? e.y : (uint32_t)ix;
    }
}

// Plain copy of known size, the bandwidth calibration kernel.
__global__ __launch_bounds__(256) void k_copy_f32(float4* __restrict__ dst,
                                                  const float4* __restrict__ src, uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride)
        st_nt4(dst + i, ld_nt4(src + i));
}

static uint32_t grid_for(uint64_t work, uint32_t block, uint32_t cap) {
    uint64_t g = (work + block - 1) / block;
    if (g > cap) g = cap;
    return (uint32_t)(g ? g : 1);
}

hipError_t launch_alias_sample(const AliasEntry* table, uint64_t n, uint64_t s_count,
                               uint64_t seed, uint64_t counter_base, uint32_t* out,
                               hipStream_t s) {
    k_alias_sample<<<grid_for(s_count, kSampleBlock, 1u << 16), kSampleBlock, 0, s>>>(
        reinterpret_cast<const uint2*>(table), n, s_count, seed, counter_base, out);
    return hipGetLastError();
}

// ---- windowed (cooperative) samplers ----------------------------------------
//
// Samples come in groups of kWindowGroup consecutive counters (group
// g = counter / kWindowGroup, with its own Philox stream: counter (g, kTagGroup)).
// All samples of a group draw from one window of the table or of the CDF,
// chosen per group, so that a wave's reads stay in a small region.  Every
// sample's marginal distribution is exact; the samples of a group are not
// independent of each other (they share the window).  Replaces the cooperative
// paths of the reference samplers, with their biases removed:
// alias/sampling/shader.comp:68-90,103-110 centres a window on one lane's
// bucket and clamps it to [0, N-1] (so the edge buckets are over-sampled) and
// that lane reuses the random pair that placed the centre;
// its/sampling/shader.comp:100-129 partitions the range into 64 parts of
// floor(range/64) entries, so a u above the last part's end finds no part.
constexpr uint64_t kWindowGroup = 64;

// Alias: the window holds T' = min(T, n) buckets from c_g = floor(x_g n / 2^64),
// cyclic, and sample t takes bucket (c_g + o_t) mod n with o_t = floor(x_t T' / 2^32):
// c_g is uniform and independent of o_t, so every bucket is equally likely.
// Coin and acceptance as k_alias_sample (x_t, coin: words 0 and 2 of the
// sample's own stream).
__global__ __launch_bounds__(kSampleBlock) void k_alias_sample_window(
    const uint2* __restrict__ table, uint64_t n, uint64_t s, uint64_t seed, uint64_t base,
    uint32_t T, uint32_t* __restrict__ out) {
    const uint64_t Tw = (uint64_t)T < n ? (uint64_t)T : n;
    const uint64_t stride = (uint64_t)gridDim.x * kSampleBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kSampleBlock + threadIdx.x; g < s; g += stride) {
        const uint64_t c = base + g;
        const U4 rg = philox_at(seed, c / kWindowGroup, kTagGroup, 0u);
        const uint64_t start = __umul64hi(((uint64_t)rg.y << 32) | rg.x, n);
        const U4 r = philox_at(seed, c, 0u, 0u);
        uint64_t ix = start + (((uint64_t)r.x * Tw) >> 32);  // < 2n
        ix = ix >= n ? ix - n : ix;
        const float u2 = u24(r.z);
        const uint2 e = table[ix];
        out[g] = (u2 > __uint_as_float(e.x)) ? e.y : (uint32_t)ix;
    }
}

// ITS: one wave per group.  u1 (the group's stream) is narrowed to a window of
// at most W entries holding it -- by 64-part searches (lane l tests the end of
// part l, parts of ceil(range/64) entries, the first part whose CDF end is
// >= u1 wins; none: the last part) or by halving -- so that u1 lies in
// (cdf(x-1), cdf(y)]; each sample then draws u2 uniformly in that interval from
// its own stream (the u of k_its_sample) and returns the lower bound of u2 in
// [x, y].  P(window) is the window's mass and u2 is uniform within it, so the
// marginal is the CDF's distribution.
template <bool LO>
__global__ __launch_bounds__(kSampleBlock) void k_its_sample_window(
    const float* __restrict__ hi, const float* __restrict__ lo, uint64_t n, uint64_t s,
    uint64_t seed, uint64_t base, uint64_t W, int parts, uint32_t* __restrict__ out) {
    const int lane = lane_id();
    const uint64_t g0 = base / kWindowGroup;
    const uint64_t ngroups = (base + s - 1) / kWindowGroup - g0 + 1;
    const uint64_t waves = (uint64_t)gridDim.x * (kSampleBlock / kWave);
    for (uint64_t k = (uint64_t)blockIdx.x * (kSampleBlock / kWave) + threadIdx.x / kWave;
         k < ngroups; k += waves) {  // wave-uniform
        const uint64_t g = g0 + k;
        const U4 rg = philox_at(seed, g, kTagGroup, 0u);
        const double u1 = u53_open0(rg.x, rg.y);
        uint64_t x = 0, y = n - 1;
        while (y - x + 1 > W) {
            if (parts) {
                const uint64_t ps = (y - x + 1 + kWave - 1) / kWave;
                const uint64_t st = x + (uint64_t)lane * ps;
                const uint64_t en = st + ps - 1 < y ? st + ps - 1 : y;
                const bool valid = st <= y;
                const uint64_t hit = __ballot(valid && u1 <= cdf_at<LO>(hi, lo, valid ? en : y));
                const int f = hit ? __builtin_ctzll(hit) : 63 - __builtin_clzll(__ballot(valid));
                x = __shfl(st, f, kWave);
                y = __shfl(en, f, kWave);
            } else {
                const uint64_t mid = (x + y) >> 1;
                if (u1 > cdf_at<LO>(hi, lo, mid)) x = mid + 1;
                else y = mid;
            }
        }
        const uint64_t c = g * kWindowGroup + (uint64_t)lane;
        if (c >= base && c - base < s) {
            const U4 r = philox_at(seed, c, 0u, 0u);
            const double v = u53_open0(r.x, r.y);
            const double cl = x > 0 ? cdf_at<LO>(hi, lo, x - 1) : 0.0;
            const double ch = cdf_at<LO>(hi, lo, y);
            const double u2 = cl + (ch - cl) * v;
            out[c - base] = (uint32_t)cdf_lower_bound<LO>(hi, lo, x, y, u2);
        }
    }
}

hipError_t launch_alias_sample_window(const AliasEntry* table, uint64_t n, uint64_t s_count,
                                      uint64_t seed, uint64_t counter_base, uint32_t window,
                                      uint32_t* out, hipStream_t s) {
    k_alias_sample_window<<<grid_for(s_count, kSampleBlock, 1u << 16), kSampleBlock, 0, s>>>(
        reinterpret_cast<const uint2*>(table), n, s_count, seed, counter_base, window, out);
    return hipGetLastError();
}

hipError_t launch_its_sample_window(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                    uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                    uint64_t window, int parts, uint32_t* out, hipStream_t s) {
    const uint64_t groups = (counter_base + s_count - 1) / kWindowGroup -
                            counter_base / kWindowGroup + 1;
    const uint32_t g = grid_for(groups * kWave, kSampleBlock, 1u << 16);
    if (cdf_lo)
        k_its_sample_window<true><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, s_count, seed,
                                                             counter_base, window, parts, out);
    else
        k_its_sample_window<false><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, s_count, seed,
                                                              counter_base, window, parts, out);
    return hipGetLastError();
}

// Device memory of the ITS line indexes: one buffer per (device, stream),
// kept between calls and grown in stream order (hipFreeAsync of the old one,
// hipMallocAsync of the new one on the caller's stream) when a longer CDF
// comes; a call on a stream whose buffer is large enough allocates nothing.
static hipError_t index_buffer(hipStream_t s, size_t bytes, float** out) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, std::pair<float*, size_t>> bufs;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto& b = bufs[{dev, s}];
    if (b.second < bytes) {
        if (b.first) {
            e = hipFreeAsync(b.first, s);
            if (e != hipSuccess) return e;
            b = {nullptr, 0};
        }
        const size_t want = bytes + bytes / 4;
        e = hipMallocAsync(reinterpret_cast<void**>(&b.first), want, s);
        if (e != hipSuccess) return e;
        b.second = want;
    }
    *out = b.first;
    return hipSuccess;
}

// Layout of the line indexes of an n-entry CDF in one buffer: level k >= 1
// holds ceil(len[k-1] / 32) floats, each level starting on a 16-byte boundary
// for the line loads; returns the floats the buffer needs.
static uint64_t its_index_layout(const float* cdf_hi, uint64_t n, ItsIndex& ix) {
    ix = ItsIndex{};
    ix.lvl[0] = cdf_hi;
    ix.len[0] = n;
    uint64_t words = 0, len = n;
    uint32_t l = 0;
    while (len > kItsTop && l < kItsLevels) {
        len = (len + kLine - 1u) / kLine;
        ++l;
        ix.len[l] = len;
        words += (len + 3u) & ~(uint64_t)3u;
    }
    ix.top = l;
    return words;
}

// The same count at compile time (sizes the static index array below).
constexpr uint64_t its_index_words(uint64_t n) {
    uint64_t words = 0, len = n;
    for (uint32_t l = 0; len > kItsTop && l < kItsLevels; ++l) {
        len = (len + kLine - 1u) / kLine;
        words += (len + 3u) & ~(uint64_t)3u;
    }
    return words;
}

static bool its_uses_index(const float* cdf_hi, uint64_t n) {
    // (a caller's CDF at a pointer that is not 16-byte aligned takes the plain
    // search)
    return n > kItsLineIndexMin && reinterpret_cast<uintptr_t>(cdf_hi) % 16u == 0;
}

// Index build + indexed search, the index in buf (its_index_layout(...) floats).
static hipError_t its_sample_indexed(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                     uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                     uint32_t* out, float* buf, ItsIndex ix, hipStream_t s) {
    uint64_t off = 0;
    for (uint32_t k = 1; k <= ix.top; ++k) {
        ix.lvl[k] = buf + off;
        off += (ix.len[k] + 3u) & ~(uint64_t)3u;
    }
    k_its_index<<<grid_for(ix.len[1], 256, 1u << 14), 256, 0, s>>>(cdf_hi, n, buf, ix);
    // every block stages the top level in LDS first: a grid of about one
    // block per resident slot, each looping over its samples
    const uint32_t gl = std::min(grid_for(s_count, kSampleBlock, 1u << 16), kItsBlocks);
    if (cdf_lo)
        k_its_sample_lines<true><<<gl, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, ix, n, s_count, seed,
                                                             counter_base, out);
    else
        k_its_sample_lines<false><<<gl, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, ix, n, s_count,
                                                              seed, counter_base, out);
    return hipGetLastError();
}

static hipError_t its_sample_plain(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                   uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                   uint32_t* out, hipStream_t s) {
    const uint32_t g = grid_for(s_count, kSampleBlock, 1u << 16);
    if (cdf_lo)
        k_its_sample<true><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, s_count, seed,
                                                      counter_base, out);
    else
        k_its_sample<false><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, s_count, seed,
                                                       counter_base, out);
    return hipGetLastError();
}

hipError_t launch_its_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                             uint64_t s_count, uint64_t seed, uint64_t counter_base,
                             uint32_t* out, hipStream_t s) {
    if (!its_uses_index(cdf_hi, n))
        return its_sample_plain(cdf_hi, cdf_lo, n, s_count, seed, counter_base, out, s);
    // the line indexes, in stream-ordered memory of this stream
    ItsIndex ix;
    const uint64_t words = its_index_layout(cdf_hi, n, ix);
    float* buf = nullptr;
    hipError_t e = index_buffer(s, words * sizeof(float), &buf);
    if (e != hipSuccess) return e;
    return its_sample_indexed(cdf_hi, cdf_lo, n, s_count, seed, counter_base, out, buf, ix, s);
}

// ---- ITS sampling without allocation ----------------------------------------
// The line indexes live in one static device array, `scratch`, sized for CDFs
// of up to kItsNoAllocMaxN entries (its_index_words(2^25) floats, 4.2 MB, held
// by every GPU the library runs on until the process ends).  Host code never
// passes the array's name to a kernel: its address on the current device comes
// from hipGetSymbolAddress, once per device, and is kept per device.  Every
// call on one GPU rewrites that GPU's array before reading it (nothing in it
// must start at zero), so calls that overlap on one GPU -- on two streams, or
// from two threads -- race; calls in stream order do not.
constexpr uint64_t kItsScratchWords = its_index_words(kItsNoAllocMaxN);
static __device__ __align__(256) float scratch[kItsScratchWords];

constexpr int kMaxDevices = 64;

static hipError_t scratch_ptr(float** out) {
    static std::mutex mu;
    static float* ptrs[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptrs[dev] = static_cast<float*>(p);
    }
    *out = ptrs[dev];
    return hipSuccess;
}

hipError_t launch_its_sample_noalloc(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                     uint64_t s_count, uint64_t seed, uint64_t counter_base,
                                     uint32_t* out, hipStream_t s) {
    if (n > kItsNoAllocMaxN) return hipErrorInvalidValue;  // the host checks first
    if (!its_uses_index(cdf_hi, n))
        return its_sample_plain(cdf_hi, cdf_lo, n, s_count, seed, counter_base, out, s);
    ItsIndex ix;
    if (its_index_layout(cdf_hi, n, ix) > kItsScratchWords) return hipErrorInvalidValue;
    float* buf = nullptr;
    hipError_t e = scratch_ptr(&buf);
    if (e != hipSuccess) return e;
    return its_sample_indexed(cdf_hi, cdf_lo, n, s_count, seed, counter_base, out, buf, ix, s);
}

hipError_t launch_cutpoint_build(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                 uint32_t* guide, uint64_t m, hipStream_t s) {
    const uint32_t ga = grid_for(m / kCutAnchor + 2, 256, 1u << 16);
    const uint32_t gf = grid_for(m, 256, 1u << 16);
    if (cdf_lo) {
        k_cutpoint_anchors<true><<<ga, 256, 0, s>>>(cdf_hi, cdf_lo, n, guide, m);
        k_cutpoint_fill<true><<<gf, 256, 0, s>>>(cdf_hi, cdf_lo, n, guide, m);
    } else {
        k_cutpoint_anchors<false><<<ga, 256, 0, s>>>(cdf_hi, cdf_lo, n, guide, m);
        k_cutpoint_fill<false><<<gf, 256, 0, s>>>(cdf_hi, cdf_lo, n, guide, m);
    }
    return hipGetLastError();
}

hipError_t launch_cutpoint_sample(const float* cdf_hi, const float* cdf_lo, uint64_t n,
                                  const uint32_t* guide, uint64_t m, uint64_t s_count,
                                  uint64_t seed, uint64_t counter_base, uint32_t* out,
                                  hipStream_t s) {
    const uint32_t g = grid_for(s_count, kSampleBlock, 1u << 16);
    if (cdf_lo)
        k_cutpoint_sample<true><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, guide, m, s_count,
                                                           seed, counter_base, out);
    else
        k_cutpoint_sample<false><<<g, kSampleBlock, 0, s>>>(cdf_hi, cdf_lo, n, guide, m, s_count,
                                                            seed, counter_base, out);
    return hipGetLastError();
}

hipError_t launch_philox(uint64_t seed, uint32_t stream_word, uint64_t base, uint64_t n,
                         uint32_t* out4, hipStream_t s) {
    k_philox<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(seed, stream_word, base, n,
                                                        reinterpret_cast<uint4*>(out4));
    return hipGetLastError();
}

hipError_t launch_generate(int kind, double a, double b, uint64_t seed, uint64_t n, float* out,
                           hipStream_t s) {
    k_generate<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(kind, a, b, seed, n, out);
    return hipGetLastError();
}

hipError_t launch_seg_sample(const AliasEntry* tables, const uint64_t* seg_off, uint32_t nseg,
                             uint32_t per_segment, uint64_t seed, uint64_t counter_base,
                             uint32_t* out, hipStream_t s) {
    const uint64_t total = (uint64_t)nseg * per_segment;
    k_seg_sample<<<grid_for(total, kSampleBlock, 1u << 16), kSampleBlock, 0, s>>>(
        reinterpret_cast<const uint2*>(tables), seg_off, total, per_segment, seed, counter_base,
        out);
    return hipGetLastError();
}

hipError_t launch_copy_f32(float* dst, const float* src, uint64_t n, hipStream_t s) {
    // callers pass 16-byte aligned buffers with n a multiple of 4 (hipMalloc'd, bench sizes)
    const uint64_t n4 = n / 4;
    k_copy_f32<<<grid_for(n4, 256, 8192), 256, 0, s>>>(reinterpret_cast<float4*>(dst),
                                                       reinterpret_cast<const float4*>(src), n4);
    return hipGetLastError();
}

}  // namespace wrsk
