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


def edge_batch(oracle):
    """Segments of 1 to 2^18 weights, including layouts of heavies that stress
    the direct form: none (equal weights), one huge weight, heavies clustered
    at the end (ascending) or the start (descending), and Zipf."""
    segs = []
    sizes = [1, 2, 3, 63, 64, 65, 100, 4095, 4096, 4097, 8191, 70001, 1 << 18]
    for i, n in enumerate(sizes):
        segs.append(oracle.gen_zipf(50 + i, 1.1, n))
    segs.append(np.ones(5000, np.float32))
    one = oracle.gen_uniform01(71, 5000)
    one[1234] = np.float32(1e6)
    segs.append(one)
    segs.append(np.sort(oracle.gen_lognormal(72, 0.0, 2.0, 3000)))
    segs.append(np.sort(oracle.gen_lognormal(73, 0.0, 2.0, 4099))[::-1].copy())
    sizes = np.array([s.size for s in segs], np.uint64)
    return sizes, np.concatenate(segs).astype(np.float32)


@pytest.mark.parametrize("tile,per", [(1024, 4), (2048, 4), (4096, 8), (1024, 2)])
@pytest.mark.parametrize("precision", [2, 1])
def test_batched_direct_equals_classic(wrs, oracle, tile, per, precision):
    """The direct pack form of a batch (each segment's light positions at its
    own offset of one array): tables bit-identical to the classic form's and to
    building every segment alone."""
    print("case: batched direct", tile, per, precision, flush=True)
    sizes, w = edge_batch(oracle)
    off = wrs.segment_offsets(sizes)
    total = int(off[-1])
    dw = wrs.to_device(w)
    tables = {}
    for d in (2, 1):
        bp = wrs.BatchPlan(off, precision=precision,
                           tune=dict(merge_tile=tile, merge_per=per, partition=0, direct=d))
        tb = wrs.empty(total, wrs.ALIAS_DTYPE)
        tb.fill_bytes(0xFF)
        bp.build(dw, tb)
        wrs.synchronize()
        tables[d] = tb.numpy()
    assert np.array_equal(tables[2], tables[1])
    plan = wrs.Plan(int(sizes.max()), precision=precision,
                    tune=dict(merge_tile=tile, merge_per=per))
    te = wrs.empty(total, wrs.ALIAS_DTYPE)
    se = wrs.empty(len(sizes), np.uint32)
    wrs.build_sample_each(plan, dw, off, te, 1, 42, se)
    wrs.synchronize()
    assert np.array_equal(tables[1], te.numpy())
