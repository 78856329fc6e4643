"""GPU tests of the streamed pack form (wrs_tuning.direct = WRS_PACK_STREAM):
exact integer keys, the table written tile by tile in element order.

Bars: the table bit-identical to the oracle's exact-key sweep
(oracle.psa_stream_table, given the build's mean) on every layout of the
heavies and in every arithmetic mode; alias indices identical to the merge
forms' (which the other tests pin to the FP64 sweep) except where the merge
forms' FP64 block offsets decide a near-tie, probabilities within float
rounding of them; a plan reused for a smaller distribution; a table
pointer that is 8 but not 16-byte aligned."""
import numpy as np
import pytest

from test_gpu_configs import implied_error
from test_gpu_direct import KINDS, weights

pytestmark = pytest.mark.gpu

STREAM = 6


def case(*a):
    print("case:", *a, flush=True)


def build_mean(wrs, plan, dw, n, table):
    """A streamed build into `table`, and the build's mean (one debug build of
    the same plan: the same reduce)."""
    mean = wrs.empty(1, np.float64)
    scratch = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build_debug(dw, n, scratch, dict(mean=mean))
    table.fill_bytes(0xFF)
    plan.psa_build(dw, n, table)
    wrs.synchronize()
    return table.numpy(), float(mean.numpy()[0])


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 65553, (1 << 20) + 5])
@pytest.mark.parametrize("kind", KINDS)
def test_stream_equals_oracle(wrs, oracle, kind, n):
    case("stream vs oracle", kind, n)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n, tune=dict(direct=STREAM))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    t, mean = build_mean(wrs, plan, dw, n, table)
    p, a = oracle.psa_stream_table(w, mean)
    assert np.array_equal(t["a"], a)
    assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["zipf", "runs", "ascending", "spike"])
def test_stream_every_mode(wrs, oracle, kind, precision):
    n = 3_000_017
    case("stream modes", kind, precision, n)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n, precision=precision, tune=dict(direct=STREAM))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    t, mean = build_mean(wrs, plan, dw, n, table)
    p, a = oracle.psa_stream_table(w, mean, f32=precision == wrs.PREC_FP32)
    assert np.array_equal(t["a"], a)
    assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))
    # the merge form of the same arithmetic: same aliases and p within rounding
    # in the FP64-carry modes; FP32 mode's merge splits round in FP32, so
    # there the two tables are compared by their accuracy
    ref = wrs.Plan(n, precision=precision, tune=dict(direct=1))
    tm = wrs.empty(n, wrs.ALIAS_DTYPE)
    ref.psa_build(dw, n, tm)
    wrs.synchronize()
    tm = tm.numpy()
    e_s, e_m = (oracle.distribution_error(oracle.implied_distribution(u["p"], u["a"]), w)
                for u in (t, tm))
    assert e_s["max_abs"] <= e_m["max_abs"] * 1.01 + 1e-12
    if precision in (wrs.PREC_FP32, wrs.PREC_FP64_CARRY):
        return  # their merge splits come from FP32 in-tile sums: compared by accuracy
    # the merge forms take the split offsets of their blocks in FP64, which can
    # decide a near-tie of a light and a heavy key the other way
    diff = np.nonzero(t["a"] != tm["a"])[0]
    print(f"  aliases differing from the merge form: {diff.size}")
    assert diff.size <= max(2, n // 1_000_000)
    same_a = t["a"] == tm["a"]
    np.testing.assert_allclose(t["p"][same_a], tm["p"][same_a], rtol=3e-7, atol=1e-7)


def test_stream_reuses_plan_and_unaligned_table(wrs, oracle):
    """A plan built for 2^20 + 5 weights builds a smaller distribution next
    (the heavy keys and ranges of the first build are stale), and a table that
    starts 8 bytes past a 16-byte boundary gets the same rows."""
    n1, n2 = (1 << 20) + 5, 70001
    case("stream reuse", n1, n2)
    plan = wrs.Plan(n1, tune=dict(direct=STREAM))
    for n, seed in ((n1, 2), (n2, 3)):
        w = oracle.gen_zipf(seed, 1.1, n)
        dw = wrs.to_device(w)
        table = wrs.empty(n, wrs.ALIAS_DTYPE)
        t, mean = build_mean(wrs, plan, dw, n, table)
        p, a = oracle.psa_stream_table(w, mean)
        assert np.array_equal(t["a"], a)
        assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))
        big = wrs.empty(n + 1, wrs.ALIAS_DTYPE)
        big.fill_bytes(0xFF)
        plan.psa_build(dw, n, big.offset(1))
        wrs.synchronize()
        assert same(big.numpy()[1:], t)


def test_stream_zipf_2e26(wrs, oracle):
    """The headline workload: N = 2^26 Zipf(1.1) seed 2, compensated mode."""
    n = 1 << 26
    case("stream c3", n)
    w = wrs.empty(n, np.float32)
    wrs.generate(wrs.GEN_ZIPF, n, 2, w, a=1.1)
    wrs.synchronize()
    wh = w.numpy()
    plan = wrs.Plan(n, tune=dict(direct=STREAM))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    t, mean = build_mean(wrs, plan, w, n, table)
    p, a = oracle.psa_stream_table(wh, mean)
    assert np.array_equal(t["a"], a)
    assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))
    del table
    ref = wrs.Plan(n, tune=dict(direct=1))
    tm = wrs.empty(n, wrs.ALIAS_DTYPE)
    ref.psa_build(w, n, tm)
    wrs.synchronize()
    tm = tm.numpy()
    assert np.array_equal(t["a"], tm["a"])
    np.testing.assert_allclose(t["p"], tm["p"], rtol=3e-7, atol=1e-7)


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_stream_unaligned_weights(wrs, oracle, shift):
    """Weights that start 4, 8 or 12 bytes past a 16-byte boundary take the
    per-element loads of the streamed passes: the same table as from an
    aligned copy, and the oracle's."""
    n = 300007
    case("stream unaligned weights", shift, n)
    w = oracle.gen_zipf(4, 1.1, n)
    dw = wrs.to_device(np.concatenate([np.full(shift, 7.0, np.float32), w,
                                       np.full(4, 9.0, np.float32)]))
    plan = wrs.Plan(n, tune=dict(direct=STREAM))
    t1 = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw.offset(shift), n, t1)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    t, mean = build_mean(wrs, plan, wrs.to_device(w), n, table)
    assert same(t1.numpy(), t)
    p, a = oracle.psa_stream_table(w, mean)
    assert np.array_equal(t["a"], a)
    assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))


def test_stream_spike_2e28(wrs, oracle):
    """The 2^28 spike mixture (seed 4) in the streamed form, the tuner's choice
    for it: bit-identical to the oracle's exact-key sweep, and an implied
    distribution within 1e-6 relative of FP64."""
    n = 1 << 28
    case("stream c5", n)
    w = wrs.empty(n, np.float32)
    wrs.generate(wrs.GEN_SPIKE, n, 4, w)
    wrs.synchronize()
    wh = w.numpy()
    plan = wrs.Plan(n, tune=dict(direct=STREAM))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    t, mean = build_mean(wrs, plan, w, n, table)
    del table, plan
    p, a = oracle.psa_stream_table(wh, mean)
    assert np.array_equal(t["a"], a)
    assert np.array_equal(t["p"].view(np.uint32), p.view(np.uint32))
    del p, a
    mx_abs, mx_rel, zm = implied_error(t, wh)
    print(f"  max_abs {mx_abs:.3e} max_rel {mx_rel:.3e}")
    assert mx_rel <= 1e-6 and zm == 0.0
