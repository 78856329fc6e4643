"""GPU tests of the segmented (batched) build + sample: bit-identical to
building and sampling each segment with its own launches, correct against the
oracle, and replayable from a captured graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(wrs, oracle, nseg, hi, seed=3):
    sizes = oracle.segment_sizes(seed, nseg, 256, hi)
    off = wrs.segment_offsets(sizes)
    total = int(off[-1])
    w = np.empty(total, np.float32)
    for b in range(nseg):
        w[off[b]:off[b + 1]] = oracle.gen_lognormal(1000 + b, 0.0, 3.0, int(sizes[b]))
    return sizes, off, w


@pytest.mark.parametrize("precision", [2, 3, 1, 0])
def test_batched_equals_per_segment(wrs, oracle, precision):
    """FP64 modes: tables and samples bit-identical to per-segment builds (the
    carry is exact, so its bits do not depend on look-back timing).  FP32 mode
    sums its carry in float in look-back order, as the reference GPU does, so
    only the partition is order-independent there: compare validity instead."""
    nseg, per = 200, 64
    print("case: batched vs per-segment", nseg, precision, flush=True)
    sizes, off, w = make_batch(wrs, oracle, nseg, 1 << 15)
    total = int(off[-1])
    dw = wrs.to_device(w)
    doff = wrs.to_device(off)
    bp = wrs.BatchPlan(off, precision=precision)
    tb = wrs.empty(total, wrs.ALIAS_DTYPE)
    sb = wrs.empty(nseg * per, np.uint32)
    bp.build(dw, tb)
    wrs.alias_sample_batched(tb, doff, nseg, per, 42, sb)
    plan = wrs.Plan(int(sizes.max()), precision=precision)
    te = wrs.empty(total, wrs.ALIAS_DTYPE)
    se = wrs.empty(nseg * per, np.uint32)
    wrs.build_sample_each(plan, dw, off, te, per, 42, se)
    wrs.synchronize()
    if precision != 0:
        assert np.array_equal(tb.numpy(), te.numpy())
        assert np.array_equal(sb.numpy(), se.numpy())
        return
    a, b = tb.numpy(), te.numpy()
    for g in range(nseg):
        seg = slice(int(off[g]), int(off[g + 1]))
        n = int(off[g + 1] - off[g])
        for t in (a[seg], b[seg]):
            assert np.all(t["a"] < n)
            q = oracle.implied_distribution(t["p"], t["a"])
            # plain FP32 keeps only ~1e-7 relative on the prefix sums: the
            # implied probabilities stay within 1e-5 absolute
            assert oracle.distribution_error(q, w[seg])["max_abs"] < 1e-5


def test_batched_against_oracle(wrs, oracle):
    nseg, per = 64, 64
    print("case: batched vs oracle", flush=True)
    sizes, off, w = make_batch(wrs, oracle, nseg, 1 << 14, seed=5)
    total = int(off[-1])
    dw = wrs.to_device(w)
    doff = wrs.to_device(off)
    bp = wrs.BatchPlan(off)
    tb = wrs.empty(total, wrs.ALIAS_DTYPE)
    sb = wrs.empty(nseg * per, np.uint32)
    bp.build(dw, tb)
    wrs.alias_sample_batched(tb, doff, nseg, per, 7, sb)
    wrs.synchronize()
    tab = tb.numpy()
    smp = sb.numpy()
    for b in range(nseg):
        seg = slice(int(off[b]), int(off[b + 1]))
        t = tab[seg]
        ws = w[seg]
        n = ws.size
        assert np.all(t["a"] < n)
        q = oracle.implied_distribution(t["p"], t["a"])
        assert oracle.distribution_error(q, ws)["max_rel"] <= 1e-6, b
        ref = oracle.sample_alias(t["p"], t["a"], per, 7, counter_base=b * per)
        assert np.array_equal(smp[b * per:(b + 1) * per], ref), b


def test_graph_replays_match_eager(wrs, oracle):
    nseg, per = 128, 64
    print("case: graph replay", flush=True)
    sizes, off, w = make_batch(wrs, oracle, nseg, 1 << 14, seed=8)
    total = int(off[-1])
    dw = wrs.to_device(w)
    doff = wrs.to_device(off)
    bp = wrs.BatchPlan(off)
    st = wrs.Stream()
    te = wrs.empty(total, wrs.ALIAS_DTYPE)
    se = wrs.empty(nseg * per, np.uint32)
    bp.build(dw, te, stream=st)
    wrs.alias_sample_batched(te, doff, nseg, per, 11, se, stream=st)
    st.synchronize()
    eager_t, eager_s = te.numpy(), se.numpy()
    tg = wrs.empty(total, wrs.ALIAS_DTYPE)
    sg = wrs.empty(nseg * per, np.uint32)
    g = wrs.Graph(st)
    with g:
        bp.build(dw, tg, stream=st)
        wrs.alias_sample_batched(tg, doff, nseg, per, 11, sg, stream=st)
    for rep in range(3):
        tg.fill_bytes(0xAB)
        wrs.synchronize()
        g.launch()
        st.synchronize()
        assert np.array_equal(tg.numpy(), eager_t), rep
        assert np.array_equal(sg.numpy(), eager_s), rep
        # eager work between replays
        bp.build(dw, te, stream=st)
        st.synchronize()
