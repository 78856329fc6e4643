"""GPU parity tests of the PSA build (weights -> alias table) against the CPU
oracle, through the C ABI.  Partition indices, heavy counts, split points and
alias indices are compared bit for bit; probabilities within float rounding of
the FP64 oracle; the implied distribution against w/sum(w) in FP64."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 100, 4095, 4096, 4097, 65536 + 3, 1 << 20]


def gen_weights(oracle, kind, n, seed=0):
    if kind == "uniform":
        return oracle.gen_uniform01(seed, n)
    if kind == "lognormal":
        return oracle.gen_lognormal(seed + 1, 0.0, 2.0, n)
    if kind == "zipf":
        return oracle.gen_zipf(seed + 2, 1.1, n)
    if kind == "spike":
        return oracle.gen_spike(seed + 4, n)
    raise ValueError(kind)


def gpu_build(wrs, w, split_size=32, precision=2):
    n = w.size
    plan = wrs.Plan(n, split_size, precision)
    K = plan.split_count(n)
    dw = wrs.to_device(w)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    dbg = dict(mean=wrs.empty(1, np.float64), heavy_count=wrs.empty(1, np.uint32),
               indices=wrs.empty(n, np.uint32), prefix_hi=wrs.empty(n, np.float32),
               prefix_lo=wrs.empty(n, np.float32), splits=wrs.empty(K, wrs.SPLIT_DTYPE))
    plan.psa_build_debug(dw, n, table, dbg)
    wrs.synchronize()
    out = {k: v.numpy() for k, v in dbg.items()}
    out["mean"] = float(out["mean"][0])
    out["H"] = int(out["heavy_count"][0])
    out["table"] = table.numpy()
    out["K"] = K
    return out


def print_case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", ["uniform", "zipf"])
def test_partition_bit_exact(wrs, oracle, kind, n):
    print_case("partition", kind, n)
    w = gen_weights(oracle, kind, n)
    g = gpu_build(wrs, w)
    # the FP64 mean is a deterministic tree sum: compare with the correctly rounded sum
    exact_mean = math.fsum(w.astype(np.float64).tolist()) / n
    assert abs(g["mean"] - exact_mean) <= 4e-16 * exact_mean
    idx_cpu, H = oracle.stable_partition_indices(w, g["mean"])
    assert g["H"] == H
    assert np.array_equal(g["indices"], oracle.cpu_to_gpu_layout(idx_cpu, H))


@pytest.mark.parametrize("n", [4097, 1 << 20])
def test_partition_given_pivot_fp32(wrs, oracle, n):
    """FP32 mode: heavy <=> w > mean_f32, compared with the float instantiation."""
    print_case("partition fp32", n)
    w = gen_weights(oracle, "uniform", n, seed=3)
    g = gpu_build(wrs, w, precision=wrs.PREC_FP32)
    idx_cpu, H = oracle.stable_partition_indices(w, np.float32(g["mean"]), f64=False)
    assert g["H"] == H
    assert np.array_equal(g["indices"], oracle.cpu_to_gpu_layout(idx_cpu, H))


@pytest.mark.parametrize("kind", ["uniform", "lognormal", "zipf", "spike"])
def test_prefix_accuracy(wrs, oracle, kind):
    n = (1 << 20) + 17
    print_case("prefix", kind, n)
    w = gen_weights(oracle, kind, n)
    g = gpu_build(wrs, w)
    H = g["H"]
    pos = g["indices"]
    vals = w.astype(np.float64)[pos]
    # heavies forward, lights reversed: exact-ish FP64 prefixes of each stream
    hp = oracle.prefix_sum_f64(vals[:H])
    lp = oracle.prefix_sum_f64(vals[H:][::-1])[::-1]
    ref = np.concatenate([hp, lp])
    got = g["prefix_hi"].astype(np.float64) + g["prefix_lo"].astype(np.float64)
    total = math.fsum(w.astype(np.float64).tolist())
    err = np.abs(got - ref).max() / total
    print(f"  max |prefix - fp64| / total = {err:.3e}")
    assert err < 1e-13


@pytest.mark.parametrize("n,split_size", [(1000, 1), (4097, 4), (65536 + 3, 32), (1 << 20, 32),
                                          ((1 << 20) + 5, 256)])
@pytest.mark.parametrize("kind", ["uniform", "lognormal", "zipf"])
def test_splits_and_table_match_oracle(wrs, oracle, kind, n, split_size):
    print_case("table", kind, n, split_size)
    w = gen_weights(oracle, kind, n)
    g = gpu_build(wrs, w, split_size=split_size)
    K = g["K"]
    p, a, d = oracle.psa_alias_table_f64(w, K, mean=g["mean"],
                                         split_rule=oracle.SPLIT_LOWER_BOUND)
    assert d["H"] == g["H"]
    si = d["splits"]["i"]
    sj = d["splits"]["j"]
    gi = g["splits"]["i"]
    gj = g["splits"]["j"]
    flips = np.nonzero((si != gi) | (sj != gj))[0]
    assert flips.size == 0, f"{flips.size} split points differ, first at {flips[:5]}"
    total = g["mean"] * n
    np.testing.assert_allclose(g["splits"]["spill"], d["splits"]["spill"], rtol=0,
                               atol=1e-13 * total)
    t = g["table"]
    assert np.array_equal(t["a"], a)
    np.testing.assert_allclose(t["p"].astype(np.float64), p, rtol=3e-7, atol=1e-7)


@pytest.mark.parametrize("kind", ["uniform", "lognormal", "zipf", "spike"])
def test_implied_distribution_error_compensated(wrs, oracle, kind):
    """The accuracy bar of the compensated build: max per-element relative error <= 1e-6."""
    n = 1 << 20
    print_case("distribution error", kind, n)
    w = gen_weights(oracle, kind, n)
    g = gpu_build(wrs, w)
    t = g["table"]
    assert np.all(t["a"] < n)
    q = oracle.implied_distribution(t["p"], t["a"])
    err = oracle.distribution_error(q, w)
    print(f"  {kind}: max_abs {err['max_abs']:.3e} max_rel {err['max_rel']:.3e}")
    assert err["max_rel"] <= 1e-6
    assert err["zero_weight_mass"] == 0.0


@pytest.mark.parametrize("precision", [0, 1, 3])
def test_other_precisions_build_valid_tables(wrs, oracle, precision):
    n = 1 << 18
    print_case("precision", precision)
    w = gen_weights(oracle, "zipf", n)
    g = gpu_build(wrs, w, precision=precision)
    t = g["table"]
    assert np.all(t["a"] < n)
    assert np.all(np.isfinite(t["p"]))
    q = oracle.implied_distribution(t["p"], t["a"])
    err = oracle.distribution_error(q, w)
    print(f"  precision {precision}: max_rel {err['max_rel']:.3e}")
    assert err["max_abs"] < 1e-3
    if precision == 3:  # FP64 tile sums meet the same bar as compensated FP32
        assert err["max_rel"] <= 1e-6


def test_special_cases(wrs, oracle):
    print_case("special cases")
    # all-equal weights: identity table
    g = gpu_build(wrs, np.full(10000, 0.25, np.float32))
    assert g["H"] == 0
    assert np.all(g["table"]["p"] == 1.0) and np.array_equal(g["table"]["a"], np.arange(10000))
    # N = 1
    g = gpu_build(wrs, np.array([3.0], np.float32))
    assert g["table"]["p"].tolist() == [1.0] and g["table"]["a"].tolist() == [0]
    # one dominant weight, tail of the last tile
    w = np.full(5000, 1e-3, np.float32)
    w[4321] = 1e4
    g = gpu_build(wrs, w)
    q = oracle.implied_distribution(g["table"]["p"], g["table"]["a"])
    assert oracle.distribution_error(q, w)["max_rel"] < 1e-6
    # zero weights are never reachable
    w = oracle.gen_uniform01(9, 70001)
    w[::7] = 0.0
    g = gpu_build(wrs, w)
    q = oracle.implied_distribution(g["table"]["p"], g["table"]["a"])
    assert oracle.distribution_error(q, w)["zero_weight_mass"] == 0.0


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_unaligned_input_gives_same_table(wrs, oracle, shift):
    """A weights pointer that is not 16-byte aligned is loaded as the aligned
    16-byte rows around it; the table must be bit-identical."""
    print_case("unaligned", shift)
    n = 100003
    w = gen_weights(oracle, "lognormal", n)
    plan = wrs.Plan(n + 3)
    dw = wrs.to_device(np.concatenate([np.full(shift, 7.0, np.float32), w,
                                       np.full(4, 9.0, np.float32)]))
    t1 = wrs.empty(n, wrs.ALIAS_DTYPE)
    t2 = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(dw.offset(shift), n, t1)
    dw2 = wrs.to_device(w)
    plan.psa_build(dw2, n, t2)
    wrs.synchronize()
    assert np.array_equal(t1.numpy(), t2.numpy())


def test_determinism_repeated_builds(wrs, oracle):
    print_case("determinism")
    n = (1 << 20) + 123
    w = gen_weights(oracle, "lognormal", n)
    plan = wrs.Plan(n)
    dw = wrs.to_device(w)
    tabs = []
    for _ in range(3):
        t = wrs.empty(n, wrs.ALIAS_DTYPE)
        plan.psa_build(dw, n, t)
        wrs.synchronize()
        tabs.append(t.numpy())
    assert np.array_equal(tabs[0], tabs[1]) and np.array_equal(tabs[0], tabs[2])


def test_plan_rejects_oversize(wrs):
    plan = wrs.Plan(100)
    dw = wrs.empty(200, np.float32)
    t = wrs.empty(200, wrs.ALIAS_DTYPE)
    with pytest.raises(wrs.WrsError) as e:
        plan.psa_build(dw, 200, t)
    assert e.value.status == wrs.WRS_EINVAL


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["zipf", "uniform"])
def test_table_valid_in_every_mode(wrs, oracle, kind, precision):
    """Every row written, every alias in range, every p in [0, 1] -- in every
    mode, including plain FP32, whose prefix sums at this size are too coarse
    to keep the merge keys ordered (the boundary repair and the block key
    envelope keep the table valid).  The FP64-carry modes also meet the 1e-6
    relative bar."""
    n = 1 << 22
    print_case("valid table", kind, precision, n)
    w = gen_weights(oracle, kind, n)
    plan = wrs.Plan(n, 32, precision)
    dw = wrs.to_device(w)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    table.fill_bytes(0xFF)  # a = 0xFFFFFFFF and p = NaN mark an unwritten row
    plan.psa_build(dw, n, table)
    wrs.synchronize()
    t = table.numpy()
    assert np.all(t["a"] < n)
    assert np.all((t["p"] >= 0.0) & (t["p"] <= 1.0))
    q = oracle.implied_distribution(t["p"], t["a"])
    assert abs(q.sum() - 1.0) < 1e-9
    err = oracle.distribution_error(q, w)
    print(f"  max_abs {err['max_abs']:.3e} max_rel {err['max_rel']:.3e}")
    if precision >= 2:
        assert err["max_rel"] <= 1e-6


@pytest.mark.parametrize("direct", [2, 1, 6])  # classic, direct and streamed forms
def test_single_build_graph_replays_and_plan_reuse(wrs, oracle, direct):
    """A captured single build replays to the eager table, with eager builds of
    other sizes on the same plan between the replays: the reduce restores its
    own state (the block-total words it polls) at the end of every launch."""
    n1, n2 = (1 << 21) + 3, 70001
    print("case: single-build graph", direct, flush=True)
    w1 = oracle.gen_zipf(81, 1.1, n1)
    w2 = oracle.gen_lognormal(82, 0.0, 2.0, n2)
    d1, d2 = wrs.to_device(w1), wrs.to_device(w2)
    plan = wrs.Plan(n1, tune=dict(direct=direct))
    st = wrs.Stream()
    t1 = wrs.empty(n1, wrs.ALIAS_DTYPE)
    t2 = wrs.empty(n2, wrs.ALIAS_DTYPE)
    plan.psa_build(d1, n1, t1, stream=st)
    st.synchronize()
    eager1 = t1.numpy()
    plan.psa_build(d2, n2, t2, stream=st)
    st.synchronize()
    eager2 = t2.numpy()
    tg = wrs.empty(n1, wrs.ALIAS_DTYPE)
    g = wrs.Graph(st)
    with g:
        plan.psa_build(d1, n1, tg, stream=st)
    for rep in range(4):
        tg.fill_bytes(0xAB)
        wrs.synchronize()
        g.launch()
        st.synchronize()
        assert np.array_equal(tg.numpy(), eager1), rep
        t2.fill_bytes(0xCD)
        wrs.synchronize()
        plan.psa_build(d2, n2, t2, stream=st)  # another size between the replays
        st.synchronize()
        assert np.array_equal(t2.numpy(), eager2), rep
    assert np.all(eager1["a"] < n1) and np.all(eager2["a"] < n2)
