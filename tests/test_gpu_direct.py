"""GPU tests of the direct pack form (wrs_tuning.direct = 1) of a single build:
the partition writes the heavies and the position of every 64th light only,
and the split search and the merge pack read the lights from the weights.  Its
tables must be bit-identical to the classic form's (which the other tests pin
to the oracle) on every layout of the heavies: scattered, all before the
lights, all after them, in runs longer than a merge block between lights (the
pack then locates each light by bisection over the heavy positions), none, or
a single one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1024, 2), (1024, 4), (1024, 8), (2048, 2), (2048, 4), (2048, 8), (4096, 4), (4096, 8)]


def case(*a):
    print("case:", *a, flush=True)


def weights(oracle, kind, n):
    rng = np.random.default_rng(11)
    if kind == "zipf":
        return oracle.gen_zipf(2, 1.1, n)
    if kind == "uniform":
        return oracle.gen_uniform01(3, n)
    if kind == "lognormal":
        return oracle.gen_lognormal(6, 0.0, 3.0, n)
    if kind == "spike":
        return oracle.gen_spike(5, n)
    if kind == "ascending":  # every heavy after every light
        return np.sort(oracle.gen_uniform01(4, n))
    if kind == "descending":  # every heavy before every light
        return np.sort(oracle.gen_uniform01(4, n))[::-1].copy()
    if kind == "runs":  # runs of 10000 heavies (> 2 merge blocks) every 25000 elements
        w = rng.uniform(0.05, 0.5, n).astype(np.float32)
        for s in range(7000, n, 25000):
            w[s:s + 10000] = rng.uniform(4.0, 6.0, min(10000, n - s)).astype(np.float32)
        return w
    if kind == "alternate":  # heavy, light, heavy, ...: ranges twice the lights
        w = np.where(np.arange(n) % 2 == 0, 1.5, 0.5).astype(np.float32)
        return w * rng.uniform(0.9, 1.1, n).astype(np.float32)
    if kind == "one_heavy":
        w = np.ones(n, np.float32)
        w[n // 3] = float(n)
        return w
    if kind == "equal":  # no heavy at all: identity table
        return np.full(n, 0.25, np.float32)
    if kind == "zeros":  # 90% zero weights
        w = rng.uniform(0.0, 1.0, n).astype(np.float32)
        w[rng.random(n) < 0.9] = 0.0
        return w
    raise ValueError(kind)


def build(wrs, dw, n, precision, tune):
    plan = wrs.Plan(n, precision=precision, tune=tune)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    table.fill_bytes(0xFF)
    plan.psa_build(dw, n, table)
    wrs.synchronize()
    return table.numpy()


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


KINDS = ["zipf", "uniform", "lognormal", "spike", "ascending", "descending", "runs",
         "alternate", "one_heavy", "equal", "zeros"]


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4097, 65553, (1 << 20) + 5])
@pytest.mark.parametrize("kind", KINDS)
def test_direct_equals_classic(wrs, oracle, kind, n):
    case("direct vs classic", kind, n)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    t0 = build(wrs, dw, n, wrs.PREC_COMPENSATED, dict(direct=2))
    t1 = build(wrs, dw, n, wrs.PREC_COMPENSATED, dict(direct=1))
    assert same(t0, t1)


@pytest.mark.parametrize("partition", [0, 1])
@pytest.mark.parametrize("precision", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["zipf", "runs", "ascending"])
def test_direct_every_mode(wrs, oracle, kind, precision, partition):
    n = 3_000_017
    case("direct modes", kind, precision, partition, n)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    t0 = build(wrs, dw, n, precision, dict(partition=partition, direct=2))
    t1 = build(wrs, dw, n, precision, dict(partition=partition, direct=1))
    if precision == wrs.PREC_FP32 and partition == 1:
        # FP32 mode's look-back carries are float sums in look-back order, so
        # two builds of that form round differently: compare their accuracy
        e0, e1 = (oracle.distribution_error(oracle.implied_distribution(t["p"], t["a"]), w)
                  for t in (t0, t1))
        assert np.all(t1["a"] < n) and np.all((t1["p"] >= 0) & (t1["p"] <= 1))
        assert e1["max_abs"] <= 2 * e0["max_abs"] + 1e-9
        return
    assert same(t0, t1)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["zipf", "runs", "alternate"])
def test_direct_every_shape(wrs, oracle, kind, shape):
    n = (1 << 20) + 5
    case("direct shapes", kind, shape, n)
    w = weights(oracle, kind, n)
    dw = wrs.to_device(w)
    tune = dict(merge_tile=shape[0], merge_per=shape[1])
    t0 = build(wrs, dw, n, wrs.PREC_COMPENSATED, dict(tune, direct=2))
    t1 = build(wrs, dw, n, wrs.PREC_COMPENSATED, dict(tune, direct=1))
    assert same(t0, t1)


def test_direct_matches_oracle_and_reuses_plan(wrs, oracle):
    """A direct plan builds the oracle's table, then a smaller distribution
    with the same plan (the light positions of the first build are stale)."""
    n1, n2 = (1 << 20) + 5, 70001
    case("direct oracle", n1, n2)
    plan = wrs.Plan(n1, tune=dict(direct=1))
    ref = wrs.Plan(n1)
    for n, seed in ((n1, 2), (n2, 3)):
        w = oracle.gen_zipf(seed, 1.1, n)
        dw = wrs.to_device(w)
        table = wrs.empty(n, wrs.ALIAS_DTYPE)
        table.fill_bytes(0xFF)
        plan.psa_build(dw, n, table)
        classic = wrs.empty(n, wrs.ALIAS_DTYPE)
        mean = wrs.empty(1, np.float64)
        ref.psa_build_debug(dw, n, classic, dict(mean=mean))
        wrs.synchronize()
        t = table.numpy()
        assert same(t, classic.numpy())
        p, a, _ = oracle.psa_alias_table_f64(w, -(-n // 32), mean=float(mean.numpy()[0]),
                                             split_rule=oracle.SPLIT_LOWER_BOUND)
        assert np.array_equal(t["a"], a)
        np.testing.assert_allclose(t["p"].astype(np.float64), p, rtol=3e-7, atol=1e-7)


def test_direct_tuning_in_debug_and_batched_builds(wrs, oracle):
    """Debug builds (which read the full partition arrays) build in the
    classic form whatever the tuning says, and batch plans in either form:
    same tables."""
    n = 200003
    case("direct ignored", n)
    w = oracle.gen_zipf(4, 1.1, n)
    dw = wrs.to_device(w)
    ref = build(wrs, dw, n, wrs.PREC_COMPENSATED, dict(direct=2))
    plan = wrs.Plan(n, tune=dict(direct=1))
    t = wrs.empty(n, wrs.ALIAS_DTYPE)
    idx = wrs.empty(n, np.uint32)
    plan.psa_build_debug(dw, n, t, dict(indices=idx))
    wrs.synchronize()
    assert same(t.numpy(), ref)
    off = np.array([0, 1000, 70000, n], np.uint64)
    tables = {}
    for d in (2, 1):
        bp = wrs.BatchPlan(off, tune=dict(direct=d))
        tb = wrs.empty(n, wrs.ALIAS_DTYPE)
        bp.build(dw, tb)
        wrs.synchronize()
        tables[d] = tb.numpy()
    assert same(tables[2], tables[1])
