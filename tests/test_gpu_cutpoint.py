"""GPU tests of the cutpoint method (guide table over the ITS CDF): the table
equals numpy's lower-bound search of the same CDF words at every threshold
g/m, and guided samples are bit-identical to the ITS sampler's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(*a):
    print("case:", *a, flush=True)


def its_cdf(wrs, oracle, w, with_lo):
    n = w.size
    plan = wrs.Plan(n)
    hi = wrs.empty(n, np.float32)
    lo = wrs.empty(n, np.float32) if with_lo else None
    plan.its_build(wrs.to_device(w), n, hi, lo)
    wrs.synchronize()
    return hi, lo


def guide_reference(c, m):
    """smallest i with c[i] >= g/m (n-1 if none), g < m; guide[m] = n-1."""
    t = np.arange(m, dtype=np.float64) / np.float64(m)
    g = np.searchsorted(c, t, side="left")
    g = np.minimum(g, c.size - 1)
    return np.concatenate([g, [c.size - 1]]).astype(np.uint32)


@pytest.mark.parametrize("n,m", [(1, 1), (1, 7), (5000, 64), (5000, 5000), ((1 << 20) + 3, 1 << 18),
                                 ((1 << 20) + 3, (1 << 20) + 3), (1 << 16, (1 << 18) + 5)])
@pytest.mark.parametrize("with_lo", [True, False])
def test_guide_table_and_samples(wrs, oracle, n, m, with_lo):
    case("cutpoint", n, m, with_lo)
    w = oracle.gen_lognormal(11, 0.0, 2.0, n)
    hi, lo = its_cdf(wrs, oracle, w, with_lo)
    guide = wrs.empty(m + 1, np.uint32)
    wrs.cutpoint_build(hi, lo, n, guide, m)
    s = 1 << 18
    out_c = wrs.empty(s, np.uint32)
    out_i = wrs.empty(s, np.uint32)
    wrs.cutpoint_sample(hi, lo, n, guide, m, s, 77, out_c, counter_base=5)
    wrs.its_sample(hi, lo, n, s, 77, out_i, counter_base=5)
    wrs.synchronize()
    c = hi.numpy().astype(np.float64)
    if with_lo:
        c += lo.numpy().astype(np.float64)
    assert np.all(np.diff(c) >= 0)
    assert np.array_equal(guide.numpy(), guide_reference(c, m))
    assert np.array_equal(out_c.numpy(), out_i.numpy())


def test_guide_with_dominant_weights(wrs, oracle):
    """A few weights holding most of the mass: long runs of equal guide entries."""
    n, m = (1 << 18) + 1, 1 << 18
    case("cutpoint dominant", n, m)
    w = oracle.gen_uniform01(12, n)
    w[[7, 1000, n - 1]] = [3e5, 1e5, 2e5]
    hi, lo = its_cdf(wrs, oracle, w, True)
    guide = wrs.empty(m + 1, np.uint32)
    wrs.cutpoint_build(hi, lo, n, guide, m)
    s = 1 << 18
    out_c = wrs.empty(s, np.uint32)
    out_i = wrs.empty(s, np.uint32)
    wrs.cutpoint_sample(hi, lo, n, guide, m, s, 3, out_c)
    wrs.its_sample(hi, lo, n, s, 3, out_i)
    wrs.synchronize()
    c = hi.numpy().astype(np.float64) + lo.numpy().astype(np.float64)
    g = guide.numpy()
    assert np.array_equal(g, guide_reference(c, m))
    assert np.count_nonzero(g == 7) > m // 4
    assert np.array_equal(out_c.numpy(), out_i.numpy())
