"""GPU tests of the reference's own split + pack strategies, kept as pack
forms (wrs_tuning.direct): SERIAL (ScalarSplit + ScalarPack: K split records,
then one thread sweeping each split interval) and INLINE (InlineSplitPack: the
split points found inside the sweeping kernel).  Both run the same sweep over
the same split points, so their tables must be bit-identical; each must be the
oracle's restatement of the reference CPU path (pack.hpp:14-83 over the
lower-bound splits of ceil(N/split_size) intervals) -- the same aliases, p
within FP32 rounding -- at every split size the reference benchmarks sweep
(src/bench/psa_split2.cpp:34-74: 2 ... 1024)."""
import numpy as np
import pytest

from test_gpu_direct import KINDS, weights

pytestmark = pytest.mark.gpu

SERIAL, INLINE = 3, 4


def case(*a):
    print("case:", *a, flush=True)


def build(wrs, dw, n, form, split, precision=2):
    plan = wrs.Plan(n, split_size=split, precision=precision, tune=dict(direct=form))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    table.fill_bytes(0xFF)
    plan.psa_build(dw, n, table)
    plan.status()
    return table.numpy()


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def debug_mean(wrs, dw, n):
    plan = wrs.Plan(n)
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    mean = wrs.empty(1, np.float64)
    plan.psa_build_debug(dw, n, t, dict(mean=mean))
    wrs.synchronize()
    return float(mean.numpy()[0])


@pytest.mark.parametrize("split", [2, 32, 1024])
@pytest.mark.parametrize("n", [1, 2, 63, 1000, 65553, (1 << 20) + 5])
@pytest.mark.parametrize("kind", KINDS)
def test_sweep_forms_match_oracle(wrs, oracle, kind, n, split):
    case("sweep forms", kind, n, split)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    ts = build(wrs, dw, n, SERIAL, split)
    ti = build(wrs, dw, n, INLINE, split)
    assert same(ts, ti)
    p, a, _ = oracle.psa_alias_table_f64(w, -(-n // split), mean=debug_mean(wrs, dw, n),
                                         split_rule=oracle.SPLIT_LOWER_BOUND)
    assert np.array_equal(ts["a"], a)
    np.testing.assert_allclose(ts["p"].astype(np.float64), np.clip(p, 0.0, 1.0),
                               rtol=3e-7, atol=1e-7)
    if n > 1 and w.max() > 0:
        err = oracle.distribution_error(oracle.implied_distribution(ts["p"], ts["a"]), w)
        assert err["max_rel"] <= 1e-6


@pytest.mark.parametrize("precision", [0, 1, 3])
@pytest.mark.parametrize("split", [4, 128])
def test_sweep_forms_other_modes(wrs, oracle, precision, split):
    n = 3_000_017
    case("sweep modes", precision, split, n)
    w = weights(oracle, "zipf", n)
    dw = wrs.to_device(w)
    ts = build(wrs, dw, n, SERIAL, split, precision)
    ti = build(wrs, dw, n, INLINE, split, precision)
    assert same(ts, ti)
    tm = build(wrs, dw, n, 2, 32, precision)  # the merge pack of the same mode
    assert np.all(ts["a"] < n) and np.all((ts["p"] >= 0) & (ts["p"] <= 1))
    e_s, e_m = (oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
                for t in (ts, tm))
    if precision <= 1:
        # these modes export the prefix as one float (as the reference does):
        # a split's spill is off by up to an ulp of the total, i.e. 2^-23 in
        # q; FP32 mode adds the FP32 residual over at most `split` steps
        assert e_s["max_abs"] <= 2 * e_m["max_abs"] + 2.0 ** -22 + split * 2.0 ** -22 / n
    else:
        assert e_s["max_rel"] <= 1e-6


def test_sweep_forms_reuse_plan_and_time_phases(wrs, oracle):
    """One SERIAL plan builds a large and then a smaller distribution (its
    split records and prefix arrays are allocated on the first build); the
    phase timer brackets the split and pack kernels."""
    n1, n2 = (1 << 21) + 3, 70001
    case("sweep reuse", n1, n2)
    plan = wrs.Plan(n1, split_size=64, tune=dict(direct=SERIAL))
    plan.set_timing(True)
    for n, seed in ((n1, 2), (n2, 3)):
        w = oracle.gen_zipf(seed, 1.1, n)
        dw = wrs.to_device(w)
        table = wrs.empty(n, wrs.ALIAS_DTYPE)
        plan.psa_build(dw, n, table)
        plan.status()
        ph = dict(zip(wrs.Plan.PHASES, plan.phase_times()))
        assert ph["boundary_splits"] > 0 and ph["merge_pack"] > 0
        ref = build(wrs, dw, n, INLINE, 64)
        assert same(table.numpy(), ref)
    plan.set_timing(False)


def test_batch_plans_reject_sweep_forms(wrs):
    off = np.array([0, 1000, 5000], np.uint64)
    for form in (SERIAL, INLINE):
        with pytest.raises(wrs.WrsError) as e:
            wrs.BatchPlan(off, tune=dict(direct=form))
        assert e.value.status == wrs.WRS_EINVAL
