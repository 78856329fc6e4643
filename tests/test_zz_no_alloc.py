"""ITS sampling without allocation (wrs_its_sample_noalloc) against
wrs_its_sample, through the C ABI: the samples must be bit-identical, on two
calls in a row with the second on a smaller CDF (the static index then holds
the first call's levels beyond the second's), with and without the low words,
at the size limit, below the indexed range, and on an unaligned CDF."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(*a):
    print("case:", *a, flush=True)


def cdf_of(wrs, oracle, seed, n):
    w = oracle.gen_lognormal(seed, 0.0, 2.0, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n)
    hi, lo = wrs.empty(n, np.float32), wrs.empty(n, np.float32)
    plan.its_build(dw, n, hi, lo)
    wrs.synchronize()
    return hi, lo


def both(wrs, hi, lo, n, s, seed, base):
    ref, got = wrs.empty(s, np.uint32), wrs.empty(s, np.uint32)
    ref.fill_bytes(0xFF)
    got.fill_bytes(0xEE)
    wrs.its_sample(hi, lo, n, s, seed, ref, counter_base=base)
    wrs.its_sample_noalloc(hi, lo, n, s, seed, got, counter_base=base)
    wrs.synchronize()
    return ref.numpy(), got.numpy()


@pytest.mark.parametrize("with_lo", [True, False])
def test_noalloc_equals_allocating_sampler_twice(wrs, oracle, with_lo):
    sizes = [(wrs.ITS_NOALLOC_MAX_N, 1 << 22), (3_000_017, 1_000_003)]
    for k, (n, s) in enumerate(sizes):
        case("noalloc its", with_lo, "call", k, n, s)
        hi, lo = cdf_of(wrs, oracle, 7 + k, n)
        ref, got = both(wrs, hi, lo if with_lo else None, n, s, 11 + k, 5 * k)
        assert np.all(ref < n)
        assert np.array_equal(ref, got)


@pytest.mark.parametrize("n", [1, 5000, 65536, 65537, 300_001])
def test_noalloc_small_and_unaligned(wrs, oracle, n):
    case("noalloc small", n)
    hi, lo = cdf_of(wrs, oracle, 3, n + 1)
    s = 200_003
    ref, got = both(wrs, hi, lo, n, s, 5, 0)
    assert np.array_equal(ref, got)
    # a CDF one float into the buffer (not 16-byte aligned): the plain search
    if n > 1:
        hs, ls = hi.offset(1), lo.offset(1)
        ref, got = both(wrs, hs, ls, n, s, 6, 0)
        assert np.array_equal(ref, got)


def test_noalloc_rejects_beyond_limit(wrs):
    case("noalloc limit")
    n = wrs.ITS_NOALLOC_MAX_N + 1
    out = wrs.empty(16, np.uint32)
    hi = wrs.empty(16, np.float32)  # never read: the size check comes first
    with pytest.raises(wrs.WrsError) as e:
        wrs.its_sample_noalloc(hi, None, n, 16, 1, out)
    assert e.value.status == wrs.WRS_EINVAL
