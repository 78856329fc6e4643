"""CPU tests that pin the oracle (oracle/wrs_oracle.c) before it is trusted as
the checker of the HIP path.  Each check is independent of the restatement it
checks: published known-answer vectors, an independent library, brute force,
closed forms and invariants of the reference's own assertion helpers
(src/host/assert/*.hpp), and planted mistakes that those checks must catch."""
import json
import math
import os

import numpy as np
import pytest

from stats import pooled_chi_square, rmse_ratio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ----------------------------------------------------------------- Philox --
# Random123 (Salmon et al. SC'11) published known-answer vectors for
# philox4x32 with 10 rounds.
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert oracle.philox4x32_10(ctr, key).tolist() == expect


def test_philox_matches_rocrand(oracle):
    vecs = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    assert len(vecs) >= 20
    for v in vecs:
        got = oracle.philox_block(v["seed"], 1, base=v["counter"])[0].tolist()
        assert got == v["out"], v


def test_philox_planted_mistake_is_caught(oracle):
    # a 9-round variant (a plausible off-by-one) must fail the KAT
    ctr, key, expect = KAT[2]
    c = np.array(ctr, np.uint64)
    k = np.array(key, np.uint64)
    M = 0xFFFFFFFF
    for r in range(9):
        if r:
            k = (k + np.array([0x9E3779B9, 0xBB67AE85], np.uint64)) & M
        m0 = 0xD2511F53 * int(c[0])
        m1 = 0xCD9E8D57 * int(c[2])
        c = np.array([(m1 >> 32) ^ int(c[1]) ^ int(k[0]), m1 & M,
                      (m0 >> 32) ^ int(c[3]) ^ int(k[1]), m0 & M], np.uint64)
    assert [int(x) for x in c] != expect


# ------------------------------------------------------------ generators --
def test_generators_ranges(oracle):
    u = oracle.gen_uniform01(0, 1 << 16)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    ln = oracle.gen_lognormal(1, 0.0, 2.0, 1 << 16)
    z = np.log(ln.astype(np.float64))
    assert abs(z.mean()) < 0.05 and abs(z.std() - 2.0) < 0.05
    sp = oracle.gen_spike(4, 10000)
    assert (sp >= 1e5).sum() == 100
    assert ((sp >= 1e-3) & (sp < 1.0)).sum() == 9900


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1000, 4097])
def test_feistel_is_a_permutation(oracle, n):
    perm = [oracle.feistel_perm(2, n, i) for i in range(n)]
    assert sorted(perm) == list(range(n))


def test_zipf_is_a_permuted_power_law(oracle):
    n = 5000
    z = oracle.gen_zipf(2, 1.1, n).astype(np.float64)
    expect = np.sort(np.arange(1, n + 1, dtype=np.float64) ** -1.1).astype(np.float32)
    assert np.array_equal(np.sort(z.astype(np.float32)), expect)


def test_segment_sizes(oracle):
    s = oracle.segment_sizes(3, 4096)
    assert s.min() >= 256 and s.max() <= 1 << 18
    lg = np.log(s.astype(np.float64) / 256.0) / math.log(1024.0)
    assert abs(lg.mean() - 0.5) < 0.02  # log-uniform


# ---------------------------------------------------------- partition --
@pytest.mark.parametrize("n", [1, 2, 7, 100, 4096, 10007])
def test_partition_matches_brute_force(oracle, n):
    w = oracle.gen_uniform01(n, n)
    pivot = float(np.mean(w.astype(np.float64)))
    idx, H = oracle.stable_partition_indices(w, pivot)
    heavy = [i for i in range(n) if float(w[i]) > pivot]
    light = [i for i in range(n) if not float(w[i]) > pivot]
    assert H == len(heavy)
    assert idx.tolist() == heavy + light
    g = oracle.cpu_to_gpu_layout(idx, H)
    assert g[:H].tolist() == heavy and g[H:].tolist() == light[::-1]


def test_partition_f32_pivot_is_strict(oracle):
    w = np.array([1.0, 2.0, 2.0, 3.0], np.float32)
    idx, H = oracle.stable_partition_indices(w, 2.0, f64=False)
    assert H == 1 and idx.tolist() == [3, 0, 1, 2]


# -------------------------------------------------------------- prefix --
def test_prefix_matches_exact_sums(oracle):
    x = oracle.gen_lognormal(5, 0.0, 2.0, 3000).astype(np.float64)
    p = oracle.prefix_sum_f64(x)
    exact = np.array([math.fsum(x[: i + 1]) for i in range(0, 3000, 37)])
    assert np.max(np.abs(p[::37] - exact) / exact) < 1e-15
    ones = oracle.prefix_sum_f32(np.ones(1 << 16, np.float32))
    assert np.array_equal(ones, np.arange(1, (1 << 16) + 1, dtype=np.float32))


# --------------------------------------------------------------- split --
def _split_prefixes(oracle, w, mean):
    idx, H = oracle.stable_partition_indices(w, mean)
    wv = w.astype(np.float64)[idx]
    return oracle.prefix_sum_f64(wv[:H]), oracle.prefix_sum_f64(wv[H:]), H


def test_ceil_mul_div(oracle):
    for a, b, c in [(10, 3, 4), (1 << 28, 12345, 1 << 23), ((1 << 32) - 1, (1 << 20) + 1, 3 << 20)]:
        assert oracle.ceil_mul_div(a, b, c) == -(-a * b // c)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_split_invariant(oracle, seed):
    """assert_is_split (src/host/assert/is_split.hpp:134-300): i + j = n,
    sigma(i,j) <= mean*n < sigma(i,j+1), spill = sigma2 - target."""
    n_w = 5000
    w = oracle.gen_lognormal(seed, 0.0, 1.5, n_w)
    mean = math.fsum(w.astype(np.float64)) / n_w
    hp, lp, H = _split_prefixes(oracle, w, mean)
    K = n_w // 16
    sp = oracle.splitK_f64(hp, lp, mean, n_w, K)
    for k in range(1, K):
        s = sp[k - 1]
        i, j = int(s["i"]), int(s["j"])
        n = oracle.ceil_mul_div(n_w, k, K)
        assert i + j == n
        target = mean * n
        light = 0.0 if i == 0 else lp[i - 1]
        heavy = 0.0 if j == 0 else hp[j - 1]
        assert light + heavy <= target < light + hp[j]
        assert s["spill"] == (light + hp[j]) - target
    assert (int(sp[-1]["i"]), int(sp[-1]["j"]), float(sp[-1]["spill"])) == (n_w - H, H - 1, 0.0)


def test_split_brute_force_tiny(oracle):
    """For every n, the binary search returns one of the j that brute-force
    enumeration finds satisfying the split condition."""
    rng = np.random.default_rng(7)
    for trial in range(40):
        N = int(rng.integers(3, 40))
        w = (rng.random(N) ** 3).astype(np.float32)
        mean = math.fsum(w.astype(np.float64)) / N
        hp, lp, H = _split_prefixes(oracle, w, mean)
        if H == 0 or H == N:
            continue
        L = N - H
        for n in range(1, N):
            i, j, spill = oracle.split_f64(hp, lp, mean, n)
            valid = []
            for jj in range(0, min(n, H - 1) + 1):
                ii = n - jj
                if ii > L or ii < 0:
                    continue
                light = 0.0 if ii == 0 else lp[ii - 1]
                heavy = 0.0 if jj == 0 else hp[jj - 1]
                if light + heavy <= mean * n < light + hp[jj]:
                    valid.append((ii, jj))
            assert (i, j) in valid, (trial, n, (i, j), valid)


@pytest.mark.parametrize("gen", ["uniform", "lognormal", "zipf"])
def test_lower_bound_split_is_the_sweep_state(oracle, gen):
    """The split of the product (lower bound, reference GPU kernels) equals the
    state (i, j) of the split-free sequential sweep at every n: the semantic
    ground truth of a PSA split, computed by running the sweep itself."""
    n_w = 3000
    w = {"uniform": lambda: oracle.gen_uniform01(11, n_w),
         "lognormal": lambda: oracle.gen_lognormal(12, 0.0, 2.0, n_w),
         "zipf": lambda: oracle.gen_zipf(13, 1.1, n_w)}[gen]()
    mean = math.fsum(w.astype(np.float64)) / n_w
    si, sj, H = oracle.sweep_states(w, mean)
    hp, lp, _ = _split_prefixes(oracle, w, mean)
    cpu_rule_off = 0
    for n in range(1, n_w):
        i, j, _ = oracle.split_lb_f64(hp, lp, mean, n)
        assert (i, j) == (si[n], sj[n]), n
        ic, jc, _ = oracle.split_f64(hp, lp, mean, n)
        if (ic, jc) != (i, j):
            # the reference CPU rule's other admissible answers have fewer
            # heavies (and as many more lights)
            assert ic + jc == n and jc < j, n
            cpu_rule_off += 1
    print(f"{gen}: reference CPU split differs from the sweep at {cpu_rule_off} of {n_w - 1} n")


def test_reference_cpu_split_can_break_the_table(oracle):
    """Documents the reference CPU split's failure mode (split size ~1): the
    table it leads to misses w/sum(w) by >10%, the lower-bound split's does not."""
    w = oracle.gen_uniform01(4, 50)
    p, a, _ = oracle.psa_alias_table_f64(w, 49, split_rule=oracle.SPLIT_REFERENCE_CPU)
    assert oracle.distribution_error(oracle.implied_distribution(p, a), w)["max_rel"] > 0.1
    p, a, _ = oracle.psa_alias_table_f64(w, 49, split_rule=oracle.SPLIT_LOWER_BOUND)
    assert oracle.distribution_error(oracle.implied_distribution(p, a), w)["max_rel"] < 1e-12


# ----------------------------------------------------------------- PSA --
def _alias_table_error(w, p, a):
    """max_i |implied mass_i - w_i|, the implied mass p_i + sum_{a_j=i}(1-p_j)
    scaled by total/N (src/host/assert/is_alias_table.hpp:103-195)."""
    n = len(w)
    assert np.all(a < n), "alias out of range"
    sampled = np.zeros(n)
    np.add.at(sampled, np.arange(n), p.astype(np.float64))
    np.add.at(sampled, a.astype(np.int64), 1.0 - p.astype(np.float64))
    total = math.fsum(np.asarray(w, np.float64))
    sampled *= total / n
    return float(np.abs(sampled - np.asarray(w, np.float64)).max())


def _assert_is_alias_table(w, p, a, margin):
    """Restates assert_is_alias_table (src/host/assert/is_alias_table.hpp:103-195):
    implied mass p_i + sum_{a_j=i}(1-p_j), scaled by total/N, within margin of w_i."""
    return _alias_table_error(w, p, a) <= margin


@pytest.mark.parametrize("gen,n", [("uniform", 1 << 14), ("lognormal", 1 << 14), ("zipf", 1 << 14),
                                   ("spike", 20000)])
def test_psa_f64_is_exact_distribution(oracle, gen, n):
    w = {"uniform": lambda: oracle.gen_uniform01(0, n),
         "lognormal": lambda: oracle.gen_lognormal(1, 0.0, 2.0, n),
         "zipf": lambda: oracle.gen_zipf(2, 1.1, n),
         "spike": lambda: oracle.gen_spike(4, n)}[gen]()
    K = n // 32
    p, a, d = oracle.psa_alias_table_f64(w, K)
    q = oracle.implied_distribution(p, a)
    err = oracle.distribution_error(q, w)
    assert err["max_rel"] < 1e-9, err
    assert np.all((p >= 0) & (p <= 1 + 1e-12))


def test_psa_f32_reference_path_invariant(oracle):
    """The float instantiation the reference tests (testAliasTableTest margin
    1e-2, src/host/assert/test.cpp:317-346)."""
    w = oracle.gen_uniform01(0, 1 << 16)
    for scale in (False, True):
        p, a, _ = oracle.psa_alias_table_f32(w, (1 << 16) // 32, reference_scale=scale)
        assert _assert_is_alias_table(w, p, a, 1e-2)


def test_invariant_catches_planted_mistake(oracle):
    w = oracle.gen_uniform01(3, 4096)
    p, a, _ = oracle.psa_alias_table_f64(w, 128)
    assert _assert_is_alias_table(w, p, a, 1e-6)
    bad = a.copy()
    # redirect one light bucket's alias to a different element
    i = int(np.argmin(p))
    bad[i] = (bad[i] + 1) % len(w)
    assert not _assert_is_alias_table(w, p, bad, 1e-6)
    badp = p.copy()
    badp[i] = min(1.0, badp[i] + 0.05)
    assert not _assert_is_alias_table(w, badp, a, 1e-6)


def test_psa_special_cases(oracle):
    # all-equal weights: identity table (sweeping_alias_table.hpp:66-73)
    p, a, d = oracle.psa_alias_table_f64(np.full(100, 0.25, np.float32), 10)
    assert d["H"] == 0 and np.all(p == 1.0) and np.array_equal(a, np.arange(100))
    # N = 1
    p, a, _ = oracle.psa_alias_table_f64(np.array([3.0], np.float32), 1)
    assert p.tolist() == [1.0] and a.tolist() == [0]
    # one dominant weight: every other bucket aliases to it
    w = np.full(64, 1e-3, np.float32)
    w[17] = 100.0
    p, a, _ = oracle.psa_alias_table_f64(w, 8)
    q = oracle.implied_distribution(p, a)
    assert oracle.distribution_error(q, w)["max_rel"] < 1e-12
    assert np.all(a[np.arange(64) != 17] == 17)


def test_golden_psa_fixtures(oracle):
    cases = json.load(open(os.path.join(GOLDEN, "psa_tiny.json")))
    assert len(cases) >= 10
    for c in cases:
        w = np.array(c["weights"], np.float32)
        K = c["K"]
        p, a, d = oracle.psa_alias_table_f64(w, K)
        assert p.tolist() == c["f64"]["p"] and a.tolist() == c["f64"]["a"], c["label"]
        assert d["H"] == c["f64"]["H"] and d["idx_cpu"].tolist() == c["f64"]["idx_cpu"]
        sp = [[int(s["i"]), int(s["j"]), float(s["spill"])] for s in d["splits"]]
        assert sp == c["f64"]["splits"]
        pc, ac, _ = oracle.psa_alias_table_f64(w, K, split_rule=oracle.SPLIT_REFERENCE_CPU)
        assert pc.tolist() == c["f64_cpu_split"]["p"] and ac.tolist() == c["f64_cpu_split"]["a"]
        p32, a32, _ = oracle.psa_alias_table_f32(w, K)
        assert p32.astype(np.float64).tolist() == c["f32"]["p"] and a32.tolist() == c["f32"]["a"]
        p32s, a32s, _ = oracle.psa_alias_table_f32(w, K, reference_scale=True)
        assert p32s.astype(np.float64).tolist() == c["f32_scaled"]["p"]
        assert a32s.tolist() == c["f32_scaled"]["a"]
        # each golden f64 table is an exact alias table of its weights
        q = oracle.implied_distribution(np.array(c["f64"]["p"]), np.array(c["f64"]["a"], np.uint32))
        if w.sum() > 0:
            assert oracle.distribution_error(q, w)["max_rel"] < 1e-12, c["label"]


def test_hand_case_closed_form(oracle):
    # w = [4,1,1,1,1,1,1,2], mean 1.5: lights carry p = 2/3; the implied
    # masses are w/12 exactly.
    w = np.array([4, 1, 1, 1, 1, 1, 1, 2], np.float32)
    p, a, d = oracle.psa_alias_table_f64(w, 4)
    assert d["H"] == 2 and d["mean"] == 1.5
    assert np.allclose(p[1:7], 2.0 / 3.0, rtol=0, atol=1e-15)
    q = oracle.implied_distribution(p, a)
    assert np.allclose(q, w.astype(np.float64) / 12.0, rtol=1e-15, atol=0)


# ------------------------------------------------- SubgroupPack (grouped) --
def _subgroup_shader_sweep(d, w, n, K, f32):
    """pack/subgroup/shader.comp:68-246 read literally, with one lane per
    subproblem (its broadcasts become plain reads): the heavy/light counters
    (:85-112), the start (:114-126), `hasHeavy && w <= mean` (:147), the
    residual step `w += weight - mean` (:240) and the last heavy (:244-246).
    Pure Python floats (double) or numpy float32 scalars.  Returns None where
    the shader would take a light its interval does not hold."""
    num = np.float32 if f32 else float
    mean = num(d["mean"])
    H = d["H"]
    heavy = d["idx_cpu"][:H].tolist()
    light = d["idx_cpu"][H:].tolist()
    wv = [num(x) for x in np.asarray(w, np.float32 if f32 else np.float64)]
    sp = [(int(s["i"]), int(s["j"]), num(s["spill"])) for s in d["splits"]]
    p = [None] * n
    a = [None] * n
    for k in range(K):
        i0, j0, spill = (0, 0, num(0)) if k == 0 else sp[k - 1]
        i1, j1, _ = sp[k]
        j1 += 1
        rem_h, rem_l = j1 - j0, i1 - i0
        nj, ni = j0, i0
        h = heavy[nj]
        r = wv[h] if k == 0 else spill
        nj += 1
        rem_h -= 1
        while rem_l > 0 or rem_h > 0:
            take_heavy = rem_h > 0 and r <= mean
            if not take_heavy and rem_l <= 0:
                return None
            if take_heavy:
                x = heavy[nj]
                nj += 1
                rem_h -= 1
                p[h], a[h] = r / mean, x
                h = x
            else:
                x = light[ni]
                ni += 1
                rem_l -= 1
                p[x], a[x] = wv[x] / mean, h
            r += wv[x] - mean
        if j1 == H:
            p[h], a[h] = num(1), h
    return p, a


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("gen,n,K", [("uniform", 2000, 63), ("zipf", 1500, 47),
                                     ("lognormal", 999, 500), ("spike", 1200, 5)])
def test_grouped_pack_matches_shader_reading(oracle, gen, n, K, f32):
    """The C restatement of SubgroupPack against a literal reading of the
    shader's loop (counters instead of clamped interval ends): same bits."""
    w = {"uniform": lambda: oracle.gen_uniform01(5, n),
         "lognormal": lambda: oracle.gen_lognormal(6, 0.0, 2.0, n),
         "zipf": lambda: oracle.gen_zipf(7, 1.1, n),
         "spike": lambda: oracle.gen_spike(8, n)}[gen]()
    fn = oracle.psa_alias_table_f32 if f32 else oracle.psa_alias_table_f64
    kw = dict(split_rule=oracle.SPLIT_LOWER_BOUND)
    p, a, d = fn(w, K, pack=oracle.PACK_GROUPED, **kw)
    lit = _subgroup_shader_sweep(d, w, n, K, f32)
    assert lit is not None
    assert a.tolist() == lit[1]
    assert p.tolist() == [float(x) for x in lit[0]]


@pytest.mark.parametrize("f32", [False, True])
def test_grouped_pack_equals_pack2_when_exact(oracle, f32):
    """Integer weights over a power-of-two N: every residual is a multiple of
    2^-12 below 2^5, exact in either rounding order, so SubgroupPack's sweep
    and pack2 must give the same table bit for bit (at every split size)."""
    n = 4096
    w = (np.floor(oracle.gen_uniform01(9, n) * 16) + 1).astype(np.float32)
    fn = oracle.psa_alias_table_f32 if f32 else oracle.psa_alias_table_f64
    for K in (1, 7, 128, 2048):
        ps, as_, _ = fn(w, K, split_rule=oracle.SPLIT_LOWER_BOUND)
        pg, ag, _ = fn(w, K, split_rule=oracle.SPLIT_LOWER_BOUND, pack=oracle.PACK_GROUPED)
        assert np.array_equal(as_, ag) and np.array_equal(ps, pg), K


@pytest.mark.parametrize("gen", ["uniform", "lognormal", "zipf"])
def test_grouped_pack_is_an_alias_table(oracle, gen):
    """On inexact weights the two roundings differ (so the grouped sweep is
    not pack2 under another name), the grouped table is still an exact alias
    table of w in FP64, and in FP32 its implied masses are as close to w as
    pack2's (the reference's margin of 1e-2 holds for both on uniform
    weights; lognormal weights up to 4e3 take both just past it)."""
    n = 1 << 14
    w = {"uniform": lambda: oracle.gen_uniform01(10, n),
         "lognormal": lambda: oracle.gen_lognormal(11, 0.0, 2.0, n),
         "zipf": lambda: oracle.gen_zipf(12, 1.1, n)}[gen]()
    ps, as_, _ = oracle.psa_alias_table_f32(w, n // 32, split_rule=oracle.SPLIT_LOWER_BOUND)
    pg, ag, _ = oracle.psa_alias_table_f32(w, n // 32, split_rule=oracle.SPLIT_LOWER_BOUND,
                                           pack=oracle.PACK_GROUPED)
    assert not np.array_equal(ps, pg)
    e_s, e_g = (_alias_table_error(w, p_, a_) for p_, a_ in ((ps, as_), (pg, ag)))
    assert e_g <= 1.5 * e_s + 1e-9, (e_g, e_s)
    if gen == "uniform":
        assert _assert_is_alias_table(w, pg, ag, 1e-2)
    p, a, _ = oracle.psa_alias_table_f64(w, n // 32, pack=oracle.PACK_GROUPED)
    err = oracle.distribution_error(oracle.implied_distribution(p, a), w)
    assert err["max_rel"] < 1e-9, err


# ------------------------------------------------------------ sampling --
def test_alias_sampling_statistics(oracle):
    n, s = 1 << 12, 1 << 20
    w = oracle.gen_uniform01(0, n)
    p, a, _ = oracle.psa_alias_table_f64(w, n // 32)
    smp = oracle.sample_alias(p.astype(np.float32), a, s, seed=12345)
    q = oracle.target_distribution(w)
    chi = pooled_chi_square(smp, q)
    assert chi["p"] > 1e-4, chi
    r = rmse_ratio(smp, q)
    assert 0.9 < r["ratio"] < 1.1, r


def test_alias_sampling_acceptance_rule(oracle):
    # p = 0 always takes the alias, p = 1 never does (u2 < 1 by construction)
    p = np.array([0.0, 1.0], np.float32)
    a = np.array([1, 0], np.uint32)
    smp = oracle.sample_alias(p, a, 10000, seed=1)
    assert np.all(smp == 1)


def test_its_sampling_matches_lower_bound(oracle):
    w = oracle.gen_lognormal(1, 0.0, 2.0, 3000).astype(np.float64)
    cdf = oracle.prefix_sum_f64(w) / math.fsum(w)
    hi = cdf.astype(np.float32)
    lo = (cdf - hi.astype(np.float64)).astype(np.float32)
    s = 5000
    smp = oracle.its_sample(hi, lo, s, seed=9)
    r = oracle.philox_block(9, s)
    u = ((((r[:, 0].astype(np.uint64) << np.uint64(32)) | r[:, 1].astype(np.uint64))
          >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    c = hi.astype(np.float64) + lo.astype(np.float64)
    expect = np.minimum(np.searchsorted(c, u, side="left"), w.size - 1)
    assert np.array_equal(smp, expect.astype(np.uint32))
    chi = pooled_chi_square(smp, w / w.sum(), min_expected=16)
    assert chi["p"] > 1e-4, chi
