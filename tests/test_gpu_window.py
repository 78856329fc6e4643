"""GPU tests of the windowed ("cooperative") samplers: bit-identical to a
numpy restatement of the window rules on the same Philox words (group stream
(g, 0x47525031), sample stream (t, 0)), marginals equal to the target
distribution (no edge bias: the reference's clamped window over-samples the
first and last buckets), and the ITS window spanning the whole CDF returns the
exact sampler's samples."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAG_GROUP = 0x47525031
GROUP = 64


def case(*a):
    print("case:", *a, flush=True)


def table_of(wrs, w):
    n = w.size
    plan = wrs.Plan(n)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    plan.psa_build(wrs.to_device(w), n, table)
    wrs.synchronize()
    return table


def its_cdf(wrs, w, with_lo):
    n = w.size
    plan = wrs.Plan(n)
    hi = wrs.empty(n, np.float32)
    lo = wrs.empty(n, np.float32) if with_lo else None
    plan.its_build(wrs.to_device(w), n, hi, lo)
    wrs.synchronize()
    return hi, lo


def words(oracle, seed, base, count, stream):
    return oracle.philox_block(seed, count, base=base, stream=stream).astype(np.uint64)


def u53_open0(hi, lo):
    """((hi:lo) >> 11) + 1) * 2^-53, the device's u53_open0(word 0, word 1)."""
    x = (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)
    return ((x >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53


def alias_window_ref(oracle, p, a, n, s, seed, base, T):
    c = np.arange(base, base + s, dtype=np.uint64)
    g = c // np.uint64(GROUP)
    g0 = int(g[0])
    rg = words(oracle, seed, g0, int(g[-1]) - g0 + 1, TAG_GROUP)[(g - np.uint64(g0)).astype(np.int64)]
    xg = [(int(h) << 32) | int(l) for h, l in zip(rg[:, 1], rg[:, 0])]
    start = np.array([(x * n) >> 64 for x in xg], dtype=np.uint64)
    r = words(oracle, seed, base, s, 0)
    tw = np.uint64(min(T, n))
    ix = start + ((r[:, 0] * tw) >> np.uint64(32))
    ix = np.where(ix >= np.uint64(n), ix - np.uint64(n), ix).astype(np.int64)
    u2 = (r[:, 2] >> np.uint64(8)).astype(np.float32) * np.float32(2.0 ** -24)
    return np.where(u2 > p[ix], a[ix], ix).astype(np.uint32)


def its_window_ref(oracle, cdf, n, s, seed, base, W, parts):
    out = np.empty(s, np.uint32)
    r = words(oracle, seed, base, s, 0)
    v = u53_open0(r[:, 0], r[:, 1])
    g_first, g_last = base // GROUP, (base + s - 1) // GROUP
    rg = words(oracle, seed, g_first, g_last - g_first + 1, TAG_GROUP)
    for k, g in enumerate(range(g_first, g_last + 1)):
        u1 = u53_open0(rg[k:k + 1, 0], rg[k:k + 1, 1])[0]
        x, y = 0, n - 1
        while y - x + 1 > W:
            if parts:
                ps = (y - x + 1 + 63) // 64
                st = np.array([x + l * ps for l in range(64)], dtype=np.int64)
                en = np.minimum(st + ps - 1, y)
                valid = st <= y
                hit = valid & (u1 <= cdf[np.where(valid, en, y)])
                f = int(np.argmax(hit)) if hit.any() else int(np.nonzero(valid)[0][-1])
                x, y = int(st[f]), int(en[f])
            else:
                mid = (x + y) >> 1
                if u1 > cdf[mid]:
                    x = mid + 1
                else:
                    y = mid
        lo_t = max(g * GROUP, base)
        hi_t = min(g * GROUP + GROUP, base + s)
        cl = cdf[x - 1] if x > 0 else 0.0
        ch = cdf[y]
        u2 = cl + (ch - cl) * v[lo_t - base:hi_t - base]
        # lower bound of u2 in [x, y] (y when u2 rounds above cdf(y))
        out[lo_t - base:hi_t - base] = np.minimum(
            x + np.searchsorted(cdf[x:y + 1], u2, side="left"), y)
    return out


@pytest.mark.parametrize("n,T", [(1, 128), (7, 128), (1000, 1), (1000, 64), (1000, 128),
                                 (100003, 128), (100003, 4096), (300, 5000)])
def test_alias_window_matches_restatement(wrs, oracle, n, T):
    case("alias window", n, T)
    w = oracle.gen_lognormal(21, 0.0, 1.5, n)
    table = table_of(wrs, w)
    s, base, seed = 20011, 37, 99
    out = wrs.empty(s, np.uint32)
    wrs.alias_sample_window(table, n, s, seed, T, out, counter_base=base)
    wrs.synchronize()
    t = table.numpy()
    ref = alias_window_ref(oracle, t["p"], t["a"], n, s, seed, base, T)
    assert np.array_equal(out.numpy(), ref)


@pytest.mark.parametrize("n,W,parts", [(1, 64, True), (5000, 1, True), (5000, 64, True),
                                       (5000, 64, False), (100003, 128, True),
                                       (100003, 4096, True), (100003, 4096, False)])
@pytest.mark.parametrize("with_lo", [True, False])
def test_its_window_matches_restatement(wrs, oracle, n, W, parts, with_lo):
    case("its window", n, W, parts, with_lo)
    w = oracle.gen_lognormal(23, 0.0, 2.0, n)
    hi, lo = its_cdf(wrs, w, with_lo)
    s, base, seed = 6007, 100, 5
    out = wrs.empty(s, np.uint32)
    wrs.its_sample_window(hi, lo, n, s, seed, W, out, parts=parts, counter_base=base)
    wrs.synchronize()
    cdf = hi.numpy().astype(np.float64)
    if lo is not None:
        cdf = cdf + lo.numpy().astype(np.float64)
    ref = its_window_ref(oracle, cdf, n, s, seed, base, W, parts)
    assert np.array_equal(out.numpy(), ref)


def test_its_window_over_whole_cdf_is_exact_sampler(wrs, oracle):
    """A window of n entries never narrows; with cdf(n-1) == 1 the sample's
    u2 is its own uniform, so the samples are wrs_its_sample's."""
    n = 4096
    case("its window = n", n)
    w = np.ones(n, np.float32)  # exact CDF: i / n, ending at 1
    hi, lo = its_cdf(wrs, w, True)
    assert hi.numpy()[-1] + lo.numpy()[-1] == 1.0
    s = 50000
    a = wrs.empty(s, np.uint32)
    b = wrs.empty(s, np.uint32)
    wrs.its_sample_window(hi, lo, n, s, 8, n, a)
    wrs.its_sample(hi, lo, n, s, 8, b)
    wrs.synchronize()
    assert np.array_equal(a.numpy(), b.numpy())


def marginal_z(wrs, samples, s, w):
    n = w.size
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(samples, s, n, hist)
    wrs.synchronize()
    h = hist.numpy().astype(np.float64)
    assert h[n] == 0
    q = w.astype(np.float64) / w.astype(np.float64).sum()
    e = q * s
    # the samples of a group of 64 share one window: allow 64x the variance
    return np.abs(h[:n] - e) / np.sqrt(64.0 * np.maximum(e, 1.0))


@pytest.mark.parametrize("kind", ["alias", "its"])
def test_window_marginals_have_no_edge_bias(wrs, oracle, kind):
    n, s = 1000, 1 << 22
    case("window marginals", kind, n, s)
    w = oracle.gen_uniform01(31, n) + np.float32(0.5)
    out = wrs.empty(s, np.uint32)
    if kind == "alias":
        wrs.alias_sample_window(table_of(wrs, w), n, s, 3, 128, out)
    else:
        hi, lo = its_cdf(wrs, w, True)
        wrs.its_sample_window(hi, lo, n, s, 3, 128, out)
    z = marginal_z(wrs, out, s, w)
    print(f"  max z {z.max():.2f}, edge z {z[0]:.2f} {z[-1]:.2f}")
    assert z.max() < 6.0


def test_window_rejects_zero(wrs):
    out = wrs.empty(4, np.uint32)
    with pytest.raises(wrs.WrsError):
        wrs.alias_sample_window(out, 4, 4, 1, 0, out)
