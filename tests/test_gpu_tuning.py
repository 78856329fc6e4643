"""GPU tests of the launch geometry (wrs_tuning) and the autotuner: every
supported merge shape builds the oracle's table, batched builds with a tuning
equal per-segment builds with the same tuning, and the tuner measures every
candidate once and then answers from its cache."""
import numpy as np
import pytest

from test_gpu_batched import make_batch

pytestmark = pytest.mark.gpu

SHAPES = [(1024, 2), (1024, 4), (1024, 8), (2048, 2), (2048, 4), (2048, 8), (4096, 4), (4096, 8)]


def case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["uniform", "zipf"])
def test_merge_shape_matches_oracle(wrs, oracle, kind, shape):
    n = (1 << 20) + 5
    case("merge shape", kind, shape, n)
    w = oracle.gen_uniform01(0, n) if kind == "uniform" else oracle.gen_zipf(2, 1.1, n)
    plan = wrs.Plan(n, tune=dict(merge_tile=shape[0], merge_per=shape[1]))
    dw = wrs.to_device(w)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    table.fill_bytes(0xFF)
    mean = wrs.empty(1, np.float64)
    plan.psa_build_debug(dw, n, table, dict(mean=mean))
    wrs.synchronize()
    t = table.numpy()
    p, a, d = oracle.psa_alias_table_f64(w, -(-n // 32), mean=float(mean.numpy()[0]),
                                         split_rule=oracle.SPLIT_LOWER_BOUND)
    assert np.array_equal(t["a"], a)
    np.testing.assert_allclose(t["p"].astype(np.float64), p, rtol=3e-7, atol=1e-7)
    err = oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
    assert err["max_rel"] <= 1e-6


@pytest.mark.parametrize("n", [1, 4097, (1 << 20) + 5])
@pytest.mark.parametrize("precision", [1, 2, 3])
def test_lookback_partition_same_build(wrs, oracle, precision, n):
    """The single-pass look-back partition gives the bits of the reduce-then-scan
    form: exact carries, the same tile arithmetic."""
    case("look-back partition", precision, n)
    w = oracle.gen_zipf(2, 1.1, n)
    dw = wrs.to_device(w)
    out = []
    for form in (0, 1):
        plan = wrs.Plan(n, precision=precision, tune=dict(partition=form))
        table = wrs.empty(n, wrs.ALIAS_DTYPE)
        dbg = dict(indices=wrs.empty(n, np.uint32), prefix_hi=wrs.empty(n, np.float32),
                   prefix_lo=wrs.empty(n, np.float32), heavy_count=wrs.empty(1, np.uint32))
        plan.psa_build_debug(dw, n, table, dbg)
        wrs.synchronize()
        out.append((table.numpy(), {k: v.numpy() for k, v in dbg.items()}))
    assert np.array_equal(out[0][0], out[1][0])
    for k in out[0][1]:
        if k == "prefix_lo" and precision < 2:
            continue
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


@pytest.mark.parametrize("tune", [dict(partition=1), dict(merge_tile=1024, merge_per=8),
                                  dict(partition=1, merge_tile=4096, merge_per=4)])
def test_batched_tuning_equals_per_segment(wrs, oracle, tune):
    nseg, per = 150, 64
    case("batched tuning", tune)
    sizes, off, w = make_batch(wrs, oracle, nseg, 1 << 15, seed=9)
    total = int(off[-1])
    dw = wrs.to_device(w)
    doff = wrs.to_device(off)
    bp = wrs.BatchPlan(off, tune=tune)
    tb = wrs.empty(total, wrs.ALIAS_DTYPE)
    sb = wrs.empty(nseg * per, np.uint32)
    bp.build(dw, tb)
    wrs.alias_sample_batched(tb, doff, nseg, per, 42, sb)
    plan = wrs.Plan(int(sizes.max()), tune=tune)
    te = wrs.empty(total, wrs.ALIAS_DTYPE)
    se = wrs.empty(nseg * per, np.uint32)
    wrs.build_sample_each(plan, dw, off, te, per, 42, se)
    wrs.synchronize()
    assert np.array_equal(tb.numpy(), te.numpy())
    assert np.array_equal(sb.numpy(), se.numpy())


def test_autotune_psa_measures_then_caches(wrs, oracle):
    n = 1 << 20
    case("autotune psa", n)
    w = oracle.gen_zipf(2, 1.1, n)
    dw = wrs.to_device(w)
    best, log = wrs.autotune_psa(dw, n, no_cache=True)
    print("  best", best, [(e["partition"], e["direct"], e["merge_tile"], e["merge_per"],
                            round(e["ms"], 4)) for e in log])
    # partition form at the default shape (2), the pack forms (direct, then the
    # serial, inline and subgroup sweeps and the streamed form), then -- when a
    # merge form leads, as it does at 2^20 -- the 7 other shapes, then the two
    # fastest again and the faster of those
    assert {e["partition"] for e in log[:2]} == {0, 1}
    assert [e["direct"] for e in log[:7]] == [2, 2, 1, 3, 4, 5, 6]
    lead = min(log[:7], key=lambda e: e["ms"])
    shapes = len(SHAPES) - 1 if lead["direct"] in (1, 2) else 0
    assert len(log) == 2 + 5 + shapes + 2
    if shapes:
        assert {(e["merge_tile"], e["merge_per"]) for e in log} == set(SHAPES)
    assert all(e["ms"] > 0 for e in log)
    assert best["direct"] in (1, 2, 6)
    assert (best["merge_tile"], best["merge_per"]) in SHAPES
    key = lambda e: (e["partition"], e["direct"], e["merge_tile"], e["merge_per"])
    first = sorted(log[:-2], key=lambda e: e["ms"])[:2]
    assert [key(e) for e in log[-2:]] == [key(e) for e in first]
    assert key(best) == key(min(log[-2:], key=lambda e: e["ms"]))
    best2, log2 = wrs.autotune_psa(dw, n)
    assert best2 == best and log2 == []
    # the tuned plan builds the oracle's table
    plan = wrs.Plan(n, tune=best)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw, n, table)
    wrs.synchronize()
    t = table.numpy()
    err = oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
    assert err["max_rel"] <= 1e-6


def test_autotune_batched_measures_then_caches(wrs, oracle):
    nseg = 200
    case("autotune batched", nseg)
    sizes, off, w = make_batch(wrs, oracle, nseg, 1 << 15, seed=10)
    dw = wrs.to_device(w)
    best, log = wrs.autotune_batched(off, dw, no_cache=True)
    print("  best", best, [(e["partition"], e["direct"], e["merge_tile"], e["merge_per"],
                            round(e["ms"], 4)) for e in log])
    # partition form at the default shape (2), the pack form (1: direct, with the
    # reduce-then-scan partition a batch builds it with), then the 7 other
    # shapes, then the two fastest again
    assert len(log) == 2 + 1 + len(SHAPES) - 1 + 2
    assert {e["partition"] for e in log[:2]} == {0, 1}
    assert [e["direct"] for e in log[:3]] == [2, 2, 1] and log[2]["partition"] == 0
    assert (best["merge_tile"], best["merge_per"]) in SHAPES
    best2, log2 = wrs.autotune_batched(off, dw)
    assert best2 == best and log2 == []


def test_set_tuning_on_an_existing_plan(wrs, oracle):
    """A plan switched to another geometry builds what a plan created with it
    builds (bit for bit), and rejects an unsupported one."""
    n = (1 << 20) + 5
    case("set tuning", n)
    w = oracle.gen_zipf(2, 1.1, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n)
    for tune in (dict(merge_tile=2048, merge_per=4, direct=1), dict(direct=2, partition=1),
                 dict(direct=4), dict(direct=5), dict()):
        plan.set_tuning(tune)
        t0 = wrs.empty(n, wrs.ALIAS_DTYPE)
        plan.psa_build(dw, n, t0)
        fresh = wrs.Plan(n, tune=tune)
        t1 = wrs.empty(n, wrs.ALIAS_DTYPE)
        fresh.psa_build(dw, n, t1)
        wrs.synchronize()
        assert np.array_equal(t0.numpy().view(np.uint64), t1.numpy().view(np.uint64)), tune
    with pytest.raises(wrs.WrsError) as e:
        plan.set_tuning(dict(merge_tile=3000))
    assert e.value.status == wrs.WRS_EINVAL
