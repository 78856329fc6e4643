"""GPU parity tests of sampling (alias and ITS), the Philox stream, the
weight generators, the scan / ITS build, against the CPU oracle and against
closed-form statistics."""
import math

import numpy as np
import pytest

from stats import pooled_chi_square, rmse_ratio

pytestmark = pytest.mark.gpu


def case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("seed,stream_word,base", [(0, 0, 0), (12345, 0, 1 << 33), (7, 99, 5)])
def test_philox_bit_exact(wrs, oracle, seed, stream_word, base):
    case("philox", seed, stream_word, base)
    n = 100003
    d = wrs.empty(4 * n, np.uint32)
    wrs.philox(seed, n, d, base=base, stream_word=stream_word)
    wrs.synchronize()
    assert np.array_equal(d.numpy().reshape(n, 4), oracle.philox_block(seed, n, base, stream_word))


@pytest.mark.parametrize("kind", ["uniform", "lognormal", "zipf", "spike"])
def test_generators_match_oracle(wrs, oracle, kind):
    case("generator", kind)
    n = 200003
    d = wrs.empty(n, np.float32)
    if kind == "uniform":
        wrs.generate(wrs.GEN_UNIFORM01, n, 0, d)
        ref = oracle.gen_uniform01(0, n)
    elif kind == "lognormal":
        wrs.generate(wrs.GEN_LOGNORMAL, n, 1, d, a=0.0, b=2.0)
        ref = oracle.gen_lognormal(1, 0.0, 2.0, n)
    elif kind == "zipf":
        wrs.generate(wrs.GEN_ZIPF, n, 2, d, a=1.1)
        ref = oracle.gen_zipf(2, 1.1, n)
    else:
        wrs.generate(wrs.GEN_SPIKE, n, 4, d)
        ref = oracle.gen_spike(4, n)
    wrs.synchronize()
    got = d.numpy()
    # integer-derived values are bit-exact; values through exp/log/cos/pow may
    # differ by one float ulp between the device and host math libraries
    ulp = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    print(f"  mismatches {int((ulp > 0).sum())} of {n}, max ulp {int(ulp.max())}")
    assert ulp.max() <= 1
    if kind == "uniform":
        assert ulp.max() == 0
    assert (ulp > 0).mean() < 1e-3


def test_segment_sizes_match_oracle(wrs, oracle):
    assert np.array_equal(wrs.segment_sizes(3, 4096), oracle.segment_sizes(3, 4096))


def _build_table(wrs, w):
    n = w.size
    plan = wrs.Plan(n)
    dw = wrs.to_device(w)
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw, n, t)
    return t


@pytest.mark.parametrize("n,s", [(1, 1000), (1000, 1 << 16), ((1 << 20) + 7, 1 << 20)])
def test_alias_samples_bit_exact(wrs, oracle, n, s):
    case("alias samples", n, s)
    w = oracle.gen_uniform01(0, n)
    t = _build_table(wrs, w)
    out = wrs.empty(s, np.uint32)
    wrs.alias_sample(t, n, s, 12345, out, counter_base=77)
    wrs.synchronize()
    tab = t.numpy()
    ref = oracle.sample_alias(tab["p"], tab["a"], s, 12345, counter_base=77)
    assert np.array_equal(out.numpy(), ref)


def test_alias_sampling_chi_square_and_rmse(wrs, oracle):
    """Uniform(0,1) weights, N = 2^20, S = 2^24: pooled chi-square and the RMSE
    ratio to its closed-form expectation (sampling error only)."""
    n, s = 1 << 20, 1 << 24
    case("alias statistics", n, s)
    w = oracle.gen_uniform01(0, n)
    t = _build_table(wrs, w)
    out = wrs.empty(s, np.uint32)
    wrs.alias_sample(t, n, s, 2024, out)
    wrs.synchronize()
    smp = out.numpy()
    q = oracle.target_distribution(w)
    chi = pooled_chi_square(smp, q)
    r = rmse_ratio(smp, q)
    print(f"  chi2 {chi['chi2']:.1f} df {chi['df']} p {chi['p']:.3g}  rmse ratio {r['ratio']:.4f}")
    assert chi["p"] > 1e-5 and chi["zero_mass_hits"] == 0
    assert 0.97 < r["ratio"] < 1.03


@pytest.mark.parametrize("n", [1, 4097, (1 << 20) + 3])
@pytest.mark.parametrize("with_lo", [True, False])
def test_its_build_and_samples(wrs, oracle, n, with_lo):
    case("its", n, with_lo)
    w = oracle.gen_lognormal(1, 0.0, 2.0, n)
    plan = wrs.Plan(n)
    dw = wrs.to_device(w)
    hi = wrs.empty(n, np.float32)
    lo = wrs.empty(n, np.float32) if with_lo else None
    plan.its_build(dw, n, hi, lo)
    s = 1 << 18
    out = wrs.empty(s, np.uint32)
    wrs.its_sample(hi, lo, n, s, 99, out)
    wrs.synchronize()
    chi_ = hi.numpy()
    clo = lo.numpy() if with_lo else None
    # CDF accuracy against the FP64 prefix
    w64 = w.astype(np.float64)
    exact = oracle.prefix_sum_f64(w64) / math.fsum(w64.tolist())
    got = chi_.astype(np.float64) + (clo.astype(np.float64) if with_lo else 0.0)
    err = np.abs(got - exact).max()
    print(f"  max |cdf - fp64| = {err:.3e}")
    assert err < (1e-14 if with_lo else 1e-6)
    # samples: bit-exact lower bound over the same CDF words
    ref = oracle.its_sample(chi_, clo, s, 99)
    assert np.array_equal(out.numpy(), ref)


def test_its_sampling_statistics(wrs, oracle):
    n, s = 1 << 16, 1 << 22
    case("its statistics", n, s)
    w = oracle.gen_lognormal(1, 0.0, 2.0, n)
    plan = wrs.Plan(n)
    dw = wrs.to_device(w)
    hi = wrs.empty(n, np.float32)
    lo = wrs.empty(n, np.float32)
    plan.its_build(dw, n, hi, lo)
    out = wrs.empty(s, np.uint32)
    wrs.its_sample(hi, lo, n, s, 5, out)
    wrs.synchronize()
    chi = pooled_chi_square(out.numpy(), oracle.target_distribution(w))
    print(f"  chi2 p {chi['p']:.3g}")
    assert chi["p"] > 1e-5 and chi["zero_mass_hits"] == 0


# 64 * 4096 - 5 takes the single-pass look-back scan, the others the
# reduce-then-scan form (more than 64 tiles; 3 * 2^20 + 7 spans several spine
# blocks of 1024 tiles)
@pytest.mark.parametrize("n", [64 * 4096 - 5, (1 << 20) + 9, 3 * (1 << 22) + 7])
@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_scan_inclusive(wrs, oracle, precision, n):
    case("scan", precision, n)
    w = oracle.gen_uniform01(3, n)
    plan = wrs.Plan(n, precision=precision)
    dw = wrs.to_device(w)
    hi = wrs.empty(n, np.float32)
    lo = wrs.empty(n, np.float32)
    plan.scan_inclusive(dw, n, hi, lo if precision >= 2 else None)
    wrs.synchronize()
    got_hi = hi.numpy()
    got_lo = lo.numpy() if precision >= 2 else None
    exact = oracle.prefix_sum_f64(w.astype(np.float64))
    got = got_hi.astype(np.float64)
    if precision >= 2:
        got += got_lo.astype(np.float64)
    rel = np.abs(got - exact).max() / exact[-1]
    print(f"  precision {precision}: max rel error {rel:.3e}")
    bound = {0: 1e-4, 1: 1e-7, 2: 1e-12, 3: 1e-14}[precision]
    assert rel < bound
    if precision >= 1:
        # exact carries: a second run reproduces every bit
        plan.scan_inclusive(dw, n, hi, lo if precision >= 2 else None)
        wrs.synchronize()
        assert np.array_equal(hi.numpy(), got_hi)
        if precision >= 2:
            assert np.array_equal(lo.numpy(), got_lo)


@pytest.mark.parametrize("n", [64 * 4096 - 5, (1 << 20) + 9])
@pytest.mark.parametrize("shift", [1, 3])
def test_scan_unaligned_same_bits(wrs, oracle, n, shift):
    """Input and outputs off 16-byte alignment: rows are loaded around the
    input and stored element-wise; the prefixes are bit-identical."""
    case("scan unaligned", n, shift)
    w = oracle.gen_uniform01(4, n)
    plan = wrs.Plan(n + 4)
    pad = np.full(shift, 5.0, np.float32)
    dw = wrs.to_device(np.concatenate([pad, w, np.full(4, 5.0, np.float32)]))
    hi1 = wrs.empty(n + 4, np.float32)
    lo1 = wrs.empty(n + 4, np.float32)
    plan.scan_inclusive(dw.offset(shift), n, hi1.offset(shift), lo1.offset(shift))
    hi2 = wrs.empty(n, np.float32)
    lo2 = wrs.empty(n, np.float32)
    plan.scan_inclusive(wrs.to_device(w), n, hi2, lo2)
    wrs.synchronize()
    assert np.array_equal(hi1.numpy()[shift:shift + n], hi2.numpy())
    assert np.array_equal(lo1.numpy()[shift:shift + n], lo2.numpy())


def test_scan_all_ones_exact(wrs):
    n = 1 << 24
    plan = wrs.Plan(n)
    dw = wrs.to_device(np.ones(n, np.float32))
    hi = wrs.empty(n, np.float32)
    plan.scan_inclusive(dw, n, hi)
    wrs.synchronize()
    assert np.array_equal(hi.numpy(), np.arange(1, n + 1, dtype=np.float32))
