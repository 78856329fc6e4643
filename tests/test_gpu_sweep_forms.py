"""GPU tests of the reference's own split + pack strategies, kept as pack
forms (wrs_tuning.direct): SERIAL (ScalarSplit + ScalarPack: K split records,
then one thread sweeping each split interval), INLINE (InlineSplitPack: the
split points found inside the sweeping kernel) and SUBGROUP (ScalarSplit +
SubgroupPack: the same records, 16 lanes per interval).  SERIAL and INLINE run
the same sweep over the same split points, so their tables must be
bit-identical, and each must be the oracle's restatement of the reference CPU
path (pack.hpp:14-83 over the lower-bound splits of ceil(N/split_size)
intervals) -- the same aliases, p within FP32 rounding; SUBGROUP must be the
oracle's SubgroupPack sweep (residual r + (w - mean)) over the same splits to
the same bar.  At every split size the reference benchmarks sweep
(src/bench/psa_split2.cpp:34-74: 2 ... 1024)."""
import numpy as np
import pytest

from test_gpu_direct import KINDS, weights

pytestmark = pytest.mark.gpu

SERIAL, INLINE, SUBGROUP = 3, 4, 5


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
    tg = build(wrs, dw, n, SUBGROUP, split)
    pg, ag, _ = oracle.psa_alias_table_f64(w, -(-n // split), mean=debug_mean(wrs, dw, n),
                                           split_rule=oracle.SPLIT_LOWER_BOUND,
                                           pack=oracle.PACK_GROUPED)
    assert np.array_equal(tg["a"], ag)
    np.testing.assert_allclose(tg["p"].astype(np.float64), np.clip(pg, 0.0, 1.0),
                               rtol=3e-7, atol=1e-7)
    if n > 1 and w.max() > 0:
        for t in (ts, tg):
            err = oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
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
    tg = build(wrs, dw, n, SUBGROUP, split, precision)
    tm = build(wrs, dw, n, 2, 32, precision)  # the merge pack of the same mode
    for t in (ts, tg):
        assert np.all(t["a"] < n) and np.all((t["p"] >= 0) & (t["p"] <= 1))
    e_m = oracle.distribution_error(oracle.implied_distribution(tm["p"], tm["a"]), w)
    for t in (ts, tg):
        e_s = oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
        if precision <= 1:
            # these modes export the prefix as one float (as the reference
            # does): a split's spill is off by up to an ulp of the total, i.e.
            # 2^-23 in q; FP32 mode adds the FP32 residual over at most
            # `split` steps
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


def test_subgroup_form_reuses_plan_and_matches_fp32_oracle(wrs, oracle):
    """A SUBGROUP plan builds a large and then a smaller distribution, each
    the oracle's SubgroupPack table over the same splits; in FP32 mode its
    table stays as close to w as the scalar pack's (two roundings of the
    same FP32 sweep)."""
    n1, n2 = (1 << 21) + 3, 70001
    case("subgroup reuse", n1, n2)
    plan = wrs.Plan(n1, split_size=64, tune=dict(direct=SUBGROUP))
    for n, seed in ((n1, 2), (n2, 3)):
        w = oracle.gen_zipf(seed, 1.1, n)
        dw = wrs.to_device(w)
        table = wrs.empty(n, wrs.ALIAS_DTYPE)
        table.fill_bytes(0xFF)
        plan.psa_build(dw, n, table)
        plan.status()
        t = table.numpy()
        p, a, _ = oracle.psa_alias_table_f64(w, -(-n // 64), mean=debug_mean(wrs, dw, n),
                                             split_rule=oracle.SPLIT_LOWER_BOUND,
                                             pack=oracle.PACK_GROUPED)
        assert np.array_equal(t["a"], a)
        np.testing.assert_allclose(t["p"].astype(np.float64), np.clip(p, 0.0, 1.0),
                                   rtol=3e-7, atol=1e-7)
    n = 1 << 20
    w = oracle.gen_uniform01(4, n)
    dw = wrs.to_device(w)
    tg = build(wrs, dw, n, SUBGROUP, 32, precision=0)
    ts = build(wrs, dw, n, SERIAL, 32, precision=0)
    assert not same(tg, ts)  # the two rounding orders differ somewhere
    e_g, e_s = (oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
                for t in (tg, ts))
    assert e_g["max_abs"] <= 1.5 * e_s["max_abs"] + 1e-12


def test_batch_plans_reject_sweep_forms(wrs):
    off = np.array([0, 1000, 5000], np.uint64)
    for form in (SERIAL, INLINE, SUBGROUP):
        with pytest.raises(wrs.WrsError) as e:
            wrs.BatchPlan(off, tune=dict(direct=form))
        assert e.value.status == wrs.WRS_EINVAL
