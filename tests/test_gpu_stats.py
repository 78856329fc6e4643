"""GPU tests of the on-device sample statistics: histogram counts equal
numpy's bincount of the same samples, and the chi-square, squared error and
zero-mass counts equal the host statistics of tests/stats.py (per-bin form
to FP64 rounding; pooled form up to pool boundaries that sit within rounding
of a multiple of the pool size)."""
import numpy as np
import pytest

from stats import pooled_chi_square

pytestmark = pytest.mark.gpu


def case(*a):
    print("case:", *a, flush=True)


def host_stats(samples, w, n):
    q = np.asarray(w, np.float64) / np.sum(np.asarray(w, np.float64))
    S = samples.size
    o = np.bincount(samples.astype(np.int64), minlength=n).astype(np.float64)
    e = q * S
    keep = e > 0
    chi2_bin = float(np.sum((o[keep] - e[keep]) ** 2 / e[keep]))
    sse = float(np.sum((o / S - q) ** 2))
    esse = float(np.sum(q * (1 - q)) / S)
    return dict(chi2_bin=chi2_bin, sse=sse, esse=esse, zero=int(o[q == 0].sum()), q=q)


def test_histogram_counts_and_out_of_range(wrs):
    n = 1000
    rng = np.random.default_rng(5)
    smp = rng.integers(0, n + 50, size=200_003).astype(np.uint32)
    ds = wrs.to_device(smp)
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(ds, smp.size, n, hist)
    wrs.histogram(ds, smp.size, n, hist, clear=False)  # accumulates
    wrs.synchronize()
    h = hist.numpy()
    ref = np.bincount(np.minimum(smp, n).astype(np.int64), minlength=n + 1) * 2
    assert np.array_equal(h, ref)


@pytest.mark.parametrize("n", [4097, (1 << 20) + 3])
@pytest.mark.parametrize("kind", ["zipf", "zeros"])
def test_stats_match_host(wrs, oracle, n, kind):
    case("stats", kind, n)
    w = oracle.gen_zipf(2, 1.1, n) if kind == "zipf" else oracle.gen_uniform01(3, n)
    if kind == "zeros":
        w[::5] = 0.0
    plan = wrs.Plan(n)
    dw = wrs.to_device(w)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw, n, table)
    S = 1 << 21
    smp = wrs.empty(S, np.uint32)
    wrs.alias_sample(table, n, S, 9, smp)
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(smp, S, n, hist)
    wrs.synchronize()
    samples = smp.numpy()
    h = hist.numpy()
    assert h[n] == 0
    assert np.array_equal(h[:n], np.bincount(samples.astype(np.int64), minlength=n))
    ref = host_stats(samples, w, n)
    per_bin = plan.sample_stats(dw, n, hist, S, min_expected=0.0)
    print(f"  per-bin chi2 {per_bin['chi2']:.6g} (host {ref['chi2_bin']:.6g}), "
          f"sse {per_bin['sse']:.4g} / expected {per_bin['expected_sse']:.4g}")
    assert per_bin["pools"] == n and per_bin["df"] == n - 1
    assert per_bin["chi2"] == pytest.approx(ref["chi2_bin"], rel=1e-9)
    assert per_bin["sse"] == pytest.approx(ref["sse"], rel=1e-9)
    assert per_bin["expected_sse"] == pytest.approx(ref["esse"], rel=1e-9)
    assert per_bin["zero_mass_hits"] == ref["zero"] == 0
    pooled = plan.sample_stats(dw, n, hist, S, min_expected=32.0)
    host = pooled_chi_square(samples, ref["q"], 32.0)
    print(f"  pooled chi2 {pooled['chi2']:.6g} df {pooled['df']} (host {host['chi2']:.6g} df {host['df']})")
    assert abs(pooled["df"] - host["df"]) <= 1
    assert pooled["chi2"] == pytest.approx(host["chi2"], rel=1e-3, abs=1.0)
    # a correct sampler: chi2 near its degrees of freedom
    assert abs(pooled["chi2"] - pooled["df"]) < 6.0 * np.sqrt(2.0 * pooled["df"])


def test_stats_flag_zero_mass_hits(wrs, oracle):
    """Samples in zero-weight bins are counted, not divided by."""
    n = 10000
    w = oracle.gen_uniform01(4, n)
    w[100:200] = 0.0
    smp = np.arange(n, dtype=np.uint32)  # one sample per bin, 100 of them at zero weight
    plan = wrs.Plan(n)
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(wrs.to_device(smp), n, n, hist)
    st = plan.sample_stats(wrs.to_device(w), n, hist, n, min_expected=0.0)
    assert st["zero_mass_hits"] == 100
    assert np.isfinite(st["chi2"])
