"""GPU parity at the five benchmark workloads, at their full sizes, through the
C ABI (one test per workload; the synthetic weights are generated on the
device by wrs_generate, whose bits are checked against the oracle's generators
here by spot checks and at small sizes in test_gpu_sampling.py).

  c1  N = 2^20 U(0,1) seed 0: PSA build + S = 2^20 alias samples
  c2  N = 2^24 LogNormal(0, 2) seed 1: normalised scan (the ITS CDF) and
      S = 2^24 binary-search samples
  c3  N = 2^26 Zipf(1.1) seed 2: partition, table in the production (direct)
      pack form, the four arithmetic modes, S = 2^26 samples
  c4  4096 log-uniform segment sizes in [256, 2^18] (seed 3) over one
      LogNormal(0, 3) stream (seed 3), built and sampled (64 per segment) as
      one batch
  c5  N = 2^28 spike mixture seed 4: direct-form build, S = 2^30 samples
      checked by the on-device statistics

Bars: partition indices, heavy counts, alias indices, split points and samples
bit-exact against the oracle; probabilities within float rounding (3e-7
relative) of the FP64 oracle; implied distribution within 1e-6 relative of
w / sum(w) in FP64; sampling chi-square and RMSE against their expectations.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats as sps

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case(*a):
    print("case:", *a, flush=True)


def device_weights(wrs, kind, n, seed, a=0.0, b=0.0):
    w = wrs.empty(n, np.float32)
    wrs.generate(kind, n, seed, w, a=a, b=b)
    wrs.synchronize()
    return w


def implied_error(table, w):
    """Max abs / rel error of the table's implied distribution vs FP64 w/sum(w)
    (vectorised; sorted summation for the total)."""
    n = w.size
    p = np.clip(table["p"].astype(np.float64), 0.0, 1.0)
    q = p + np.bincount(table["a"].astype(np.int64), weights=1.0 - p, minlength=n)
    q /= n
    w64 = w.astype(np.float64)
    t = w64 / math.fsum(np.sort(w64).tolist()) if n <= (1 << 22) else w64 / np.sum(np.sort(w64))
    d = np.abs(q - t)
    nz = t > 0
    return float(d.max()), float((d[nz] / t[nz]).max()), float(q[~nz].sum())


def psa_debug(wrs, plan, dw, n, with_splits=False):
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    dbg = dict(mean=wrs.empty(1, np.float64), heavy_count=wrs.empty(1, np.uint32),
               indices=wrs.empty(n, np.uint32))
    if with_splits:
        K = plan.split_count(n)
        dbg.update(prefix_hi=wrs.empty(n, np.float32), prefix_lo=wrs.empty(n, np.float32),
                   splits=wrs.empty(K, wrs.SPLIT_DTYPE))
    plan.psa_build_debug(dw, n, table, dbg)
    wrs.synchronize()
    out = {k: v.numpy() for k, v in dbg.items()}
    out["mean"] = float(out["mean"][0])
    out["H"] = int(out["heavy_count"][0])
    out["table"] = table.numpy()
    return out


def sample_check(wrs, plan, dw, table, n, S, seed, expect=None):
    """S alias samples of table: bit-exact against `expect` (a prefix of the
    stream, when given), and the pooled chi-square / RMSE of all S samples by
    the on-device statistics."""
    out = wrs.empty(S, np.uint32)
    wrs.alias_sample(table, n, S, seed, out)
    wrs.synchronize()
    if expect is not None:
        got = out.numpy(expect.size)
        assert np.array_equal(got, expect)
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(out, S, n, hist)
    st = plan.sample_stats(dw, n, hist, S, min_expected=32.0)
    del out, hist
    pval = float(sps.chi2.sf(st["chi2"], st["df"]))
    ratio = st["rmse"] / st["expected_rmse"]
    print(f"  chi2 {st['chi2']:.1f} df {st['df']} p {pval:.3g} rmse ratio {ratio:.5f}")
    assert st["zero_mass_hits"] == 0
    assert pval > 1e-4
    assert abs(ratio - 1.0) < 0.02
    return st


def test_c1_uniform_2e20(wrs, oracle):
    n, S = 1 << 20, 1 << 20
    case("c1 PSA build + samples", n)
    w = oracle.gen_uniform01(0, n)
    dw = device_weights(wrs, wrs.GEN_UNIFORM01, n, 0)
    assert np.array_equal(dw.numpy(), w)
    plan = wrs.Plan(n)
    g = psa_debug(wrs, plan, dw, n, with_splits=True)
    K = plan.split_count(n)
    p, a, d = oracle.psa_alias_table_f64(w, K, mean=g["mean"], split_rule=oracle.SPLIT_LOWER_BOUND)
    assert g["H"] == d["H"]
    assert np.array_equal(g["indices"], oracle.cpu_to_gpu_layout(d["idx_cpu"], d["H"]))
    assert np.array_equal(g["splits"]["i"], d["splits"]["i"])
    assert np.array_equal(g["splits"]["j"], d["splits"]["j"])
    # the production build (the plan's default pack form) gives the debug table
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw, n, t)
    wrs.synchronize()
    table = t.numpy()
    assert np.array_equal(table.view(np.uint64), g["table"].view(np.uint64))
    assert np.array_equal(table["a"], a)
    np.testing.assert_allclose(table["p"].astype(np.float64), p, rtol=3e-7, atol=1e-7)
    mx_abs, mx_rel, zm = implied_error(table, w)
    print(f"  max_abs {mx_abs:.3e} max_rel {mx_rel:.3e}")
    assert mx_rel <= 1e-6 and zm == 0.0
    ref = oracle.sample_alias(table["p"], table["a"], S, 12345)
    sample_check(wrs, plan, dw, t, n, S, 12345, expect=ref)


def test_c2_lognormal_scan_its_2e24(wrs, oracle):
    n, S, seed = 1 << 24, 1 << 24, 7
    case("c2 scan + ITS", n)
    dw = device_weights(wrs, wrs.GEN_LOGNORMAL, n, 1, a=0.0, b=2.0)
    w = dw.numpy()
    assert np.array_equal(w[:4096], oracle.gen_lognormal(1, 0.0, 2.0, 4096))
    plan = wrs.Plan(n)
    hi, lo = wrs.empty(n, np.float32), wrs.empty(n, np.float32)
    plan.its_build(dw, n, hi, lo)
    wrs.synchronize()
    h, l = hi.numpy(), lo.numpy()
    w64 = w.astype(np.float64)
    total = math.fsum(w64.tolist())
    exact = np.cumsum(w64) / total  # FP64 prefix: error ~1e-16 relative per term
    cdf = h.astype(np.float64) + l.astype(np.float64)
    err = float(np.abs(cdf - exact).max())
    print(f"  CDF max |err| vs FP64 prefix {err:.3e}")
    assert err < 1e-12
    assert np.all(np.diff(cdf) >= 0) and abs(cdf[-1] - 1.0) < 1e-15
    out = wrs.empty(S, np.uint32)
    wrs.its_sample(hi, lo, n, S, seed, out)
    wrs.synchronize()
    ref = oracle.its_sample(h, l, S, seed)
    assert np.array_equal(out.numpy(), ref)
    # the samples follow the weights
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(out, S, n, hist)
    st = plan.sample_stats(dw, n, hist, S, min_expected=32.0)
    pval = float(sps.chi2.sf(st["chi2"], st["df"]))
    print(f"  chi2 {st['chi2']:.1f} df {st['df']} p {pval:.3g}")
    assert pval > 1e-4 and st["zero_mass_hits"] == 0


def test_c3_zipf_2e26(wrs, oracle):
    n = 1 << 26
    case("c3 Zipf partition + direct table", n)
    dw = device_weights(wrs, wrs.GEN_ZIPF, n, 2, a=1.1)
    w = dw.numpy()
    rng = np.random.default_rng(0)
    for i in rng.integers(0, n, 64).tolist() + [0, n - 1]:
        rank = 1 + oracle.feistel_perm(2, n, i)
        assert w[i] == np.float32(math.pow(rank, -1.1)), i
    plan = wrs.Plan(n, tune=dict(direct=1, merge_tile=1024, merge_per=4))
    g = psa_debug(wrs, plan, dw, n)
    mean = g["mean"]
    exact_mean = math.fsum(w.astype(np.float64).tolist()) / n
    assert abs(mean - exact_mean) <= 4e-16 * exact_mean
    idx_cpu, H = oracle.stable_partition_indices(w, mean)
    print(f"  H = {H}")
    assert g["H"] == H
    assert np.array_equal(g["indices"], oracle.cpu_to_gpu_layout(idx_cpu, H))
    del idx_cpu, g
    # the production pack form, against the oracle given the same mean
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw, n, t)
    wrs.synchronize()
    table = t.numpy()
    K = plan.split_count(n)
    p, a, d = oracle.psa_alias_table_f64(w, K, mean=mean, split_rule=oracle.SPLIT_LOWER_BOUND)
    assert d["H"] == H
    diff = np.nonzero(table["a"] != a)[0]
    assert diff.size == 0, f"{diff.size} alias entries differ, first at {diff[:5]}"
    np.testing.assert_allclose(table["p"].astype(np.float64), p, rtol=3e-7, atol=1e-7)
    del p, a, d
    mx_abs, mx_rel, zm = implied_error(table, w)
    print(f"  compensated: max_abs {mx_abs:.3e} max_rel {mx_rel:.3e}")
    assert mx_rel <= 1e-6 and zm == 0.0
    # S = 2^26 samples: the first 2^22 bit-exact, all of them by the statistics
    ref = oracle.sample_alias(table["p"], table["a"], 1 << 22, 99)
    sample_check(wrs, plan, dw, t, n, n, 99, expect=ref)


@pytest.mark.parametrize("precision", [0, 1, 3])
def test_c3_zipf_2e26_other_modes(wrs, oracle, precision):
    """The precision study's other modes at full size: every row written, aliases
    in range, p in [0, 1]; FP64 tile sums meet the 1e-6 bar, plain FP32 cannot
    resolve a mean of 1.3e-7 against prefix sums of about 9 and is only valid."""
    n = 1 << 26
    case("c3 Zipf mode", precision, n)
    dw = device_weights(wrs, wrs.GEN_ZIPF, n, 2, a=1.1)
    plan = wrs.Plan(n, precision=precision, tune=dict(direct=1, merge_tile=1024, merge_per=4))
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    t.fill_bytes(0xFF)
    plan.psa_build(dw, n, t)
    wrs.synchronize()
    table = t.numpy()
    assert np.all(table["a"] < n)
    assert np.all((table["p"] >= 0.0) & (table["p"] <= 1.0))
    mx_abs, mx_rel, zm = implied_error(table, dw.numpy())
    print(f"  mode {precision}: max_abs {mx_abs:.3e} max_rel {mx_rel:.3e}")
    assert zm == 0.0
    if precision == 3:
        assert mx_rel <= 1e-6
    elif precision == 1:
        assert mx_abs < 1e-7
    else:
        assert mx_abs < 1e-5


def test_c4_batched_4096(wrs, oracle):
    nseg, per, seed = 4096, 64, 42
    case("c4 batched", nseg)
    sizes = wrs.segment_sizes(3, nseg)
    assert np.array_equal(sizes, oracle.segment_sizes(3, nseg, 256, 1 << 18))
    off = wrs.segment_offsets(sizes)
    tot = int(off[-1])
    dw = device_weights(wrs, wrs.GEN_LOGNORMAL, tot, 3, a=0.0, b=3.0)
    w = dw.numpy()
    assert np.array_equal(w[:4096], oracle.gen_lognormal(3, 0.0, 3.0, 4096))
    doff = wrs.to_device(off)
    bp = wrs.BatchPlan(off, tune=dict(direct=1, merge_tile=1024, merge_per=4))
    tb = wrs.empty(tot, wrs.ALIAS_DTYPE)
    sb = wrs.empty(nseg * per, np.uint32)
    st = wrs.Stream()
    g = wrs.Graph(st)
    with g:
        bp.build(dw, tb, stream=st)
        wrs.alias_sample_batched(tb, doff, nseg, per, seed, sb, stream=st)
    g.launch()
    st.synchronize()
    # one launch sequence per distribution, with the same tuning
    eplan = wrs.Plan(int(sizes.max()), tune=dict(direct=1, merge_tile=1024, merge_per=4))
    te = wrs.empty(tot, wrs.ALIAS_DTYPE)
    se = wrs.empty(nseg * per, np.uint32)
    wrs.build_sample_each(eplan, dw, off, te, per, seed, se)
    wrs.synchronize()
    tables, samples = tb.numpy(), sb.numpy()
    assert np.array_equal(tables.view(np.uint64), te.numpy().view(np.uint64))
    assert np.array_equal(samples, se.numpy())
    worst = 0.0
    for b in range(nseg):
        seg = slice(int(off[b]), int(off[b + 1]))
        n = int(sizes[b])
        t = tables[seg]
        assert np.all(t["a"] < n)
        mx_abs, mx_rel, zm = implied_error(t, w[seg])
        worst = max(worst, mx_rel)
        assert zm == 0.0
        ref = oracle.sample_alias(t["p"], t["a"], per, seed, counter_base=b * per)
        assert np.array_equal(samples[b * per:(b + 1) * per], ref), b
    print(f"  worst segment max_rel {worst:.3e}")
    assert worst <= 1e-6


def test_c5_spike_2e28(wrs, oracle):
    n, S = 1 << 28, 1 << 30
    case("c5 spike build", n)
    dw = device_weights(wrs, wrs.GEN_SPIKE, n, 4)
    w = dw.numpy()
    spikes = w >= 1e5
    assert int(np.count_nonzero(spikes)) == n // 100
    rng = np.random.default_rng(4)
    for i in rng.integers(0, n, 64).tolist():
        assert bool(spikes[i]) == (oracle.feistel_perm(4, n, i) < n // 100), i
    del spikes
    plan = wrs.Plan(n, tune=dict(direct=1, merge_tile=1024, merge_per=4))
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    t.fill_bytes(0xFF)
    plan.psa_build(dw, n, t)
    wrs.synchronize()
    table = t.numpy()
    assert np.all(table["a"] < n)
    assert np.all((table["p"] >= 0.0) & (table["p"] <= 1.0))
    mx_abs, mx_rel, zm = implied_error(table, w)
    print(f"  max_abs {mx_abs:.3e} max_rel {mx_rel:.3e}")
    assert mx_rel <= 1e-6 and zm == 0.0
    del table, w
    case("c5 spike samples", S)
    sample_check(wrs, plan, dw, t, n, S, 2024)


def test_golden_tiny_cases_on_gpu(wrs, oracle):
    """The committed tiny fixtures (hand cases such as a single non-zero weight,
    all-equal weights, zero weights; seeded N <= 64) through the GPU build in
    both pack forms: heavy count, partition, alias indices exactly and p within
    float rounding of the fixture's FP64 lower-bound table; the streamed form
    (exact keys) gives the same aliases on every fixture."""
    cases = json.load(open(os.path.join(GOLDEN, "psa_tiny.json")))
    for c in cases:
        w = np.array(c["weights"], np.float32)
        n = w.size
        case("golden", c["label"], n)
        ref = c["f64"]
        dw = wrs.to_device(w)
        for direct in (2, 1, 6):
            plan = wrs.Plan(n, tune=dict(direct=direct))
            g = psa_debug(wrs, plan, dw, n)
            assert abs(g["mean"] - ref["mean"]) <= 1e-15 * ref["mean"], c["label"]
            assert g["H"] == ref["H"], c["label"]
            idx = oracle.cpu_to_gpu_layout(np.array(ref["idx_cpu"], np.uint32), ref["H"])
            assert np.array_equal(g["indices"], idx), c["label"]
            t = wrs.empty(n, wrs.ALIAS_DTYPE)
            plan.psa_build(dw, n, t)
            wrs.synchronize()
            table = t.numpy()
            assert table["a"].tolist() == ref["a"], (c["label"], direct)
            np.testing.assert_allclose(table["p"].astype(np.float64), ref["p"], rtol=3e-7,
                                       atol=1e-7, err_msg=f"{c['label']} direct={direct}")
