"""GPU test at the largest distribution a plan takes, N = 2^32 - 1 (about 115 GB
of device memory for the weights, the table and the plan): the 32-bit index
arithmetic at the last tile, the last block boundary (b * merge_tile >= 2^32)
and the last merge block.  The weights have a closed-form table: N - 1 equal
weights c and one weight 2c at the end.  The mean c (1 + 2^-32) rounds down
to the pivot c, so the last element is the only heavy; every light bucket gets
p = (float)(c / mean) = 1 and alias N - 1, and the last heavy gets (1, itself):
every row is (1.0, N - 1).  Checked on slices at both ends and in the middle,
in the classic, direct and streamed pack forms (the streamed one with about
120 GB more for its keys and rows, one heavy and no crossing tile), and by
alias samples (p = 1 keeps every bucket); the ITS
CDF of the same weights is (i + 1) / 2^32.  Before these tests the tile count
in the emit and scan kernels ((n + 4095) / 4096 in 32 bits), the last block
boundary (b * merge_tile) and the last merge block's end wrapped at
n > 2^32 - 4096."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 32) - 1


def case(*a):
    print("case:", *a, flush=True)


def slice_rows(wrs, table, start, count):
    out = np.empty(count, wrs.ALIAS_DTYPE)
    wrs.check(wrs.lib().wrs_rt_memcpy_dtoh(out.ctypes.data, table.offset(start),
                                           count * wrs.ALIAS_DTYPE.itemsize))
    return out


@pytest.mark.parametrize("direct", [2, 1, 6])  # classic, direct, streamed (30 bytes a weight more)
def test_largest_n_closed_form_table(wrs, direct):
    case("max N", N, "direct", direct)
    w = wrs.empty(N, np.float32)
    w.fill_bytes(0x3F)  # c = 0x3F3F3F3F, about 0.747
    c = np.frombuffer(np.uint32(0x3F3F3F3F).tobytes(), np.float32)[0]
    last = np.array([2.0 * c], np.float32)
    wrs.check(wrs.lib().wrs_rt_memcpy_htod(w.offset(N - 1), last.ctypes.data, 4))
    table = wrs.empty(N, wrs.ALIAS_DTYPE)
    table.fill_bytes(0xFF)
    plan = wrs.Plan(N, tune=dict(direct=direct))
    plan.psa_build(w, N, table)
    wrs.synchronize()
    for start in (0, (1 << 31) - 777, N - (1 << 20)):
        rows = slice_rows(wrs, table, start, 1 << 20)
        assert np.all(rows["p"] == np.float32(1.0)), start
        assert np.all(rows["a"] == np.uint32(N - 1)), start
    s = 1 << 22
    out = wrs.empty(s, np.uint32)
    wrs.alias_sample(table, N, s, 7, out)
    wrs.synchronize()
    smp = out.numpy().astype(np.uint64)
    assert smp.max() < N
    # p = 1: every sample is its uniform bucket, so the top 1/1000 of the range is hit
    assert np.count_nonzero(smp >= N - N // 1000) > 0
    plan.close()
    del table, w, out


def test_largest_n_its_cdf(wrs):
    """The ITS build (normalised scan) at N = 2^32 - 1 on the same weights:
    cdf(i) = (i + 1) c / (2^32 c) = (i + 1) / 2^32 for i < N - 1, cdf(N - 1) = 1."""
    case("max N its", N)
    w = wrs.empty(N, np.float32)
    w.fill_bytes(0x3F)
    c = np.frombuffer(np.uint32(0x3F3F3F3F).tobytes(), np.float32)[0]
    last = np.array([2.0 * c], np.float32)
    wrs.check(wrs.lib().wrs_rt_memcpy_htod(w.offset(N - 1), last.ctypes.data, 4))
    hi = wrs.empty(N, np.float32)
    lo = wrs.empty(N, np.float32)
    plan = wrs.Plan(N)
    plan.its_build(w, N, hi, lo)
    wrs.synchronize()
    k = 1 << 16
    for start in (0, (1 << 31) + 12345, N - k):
        h = np.empty(k, np.float32)
        l = np.empty(k, np.float32)
        wrs.check(wrs.lib().wrs_rt_memcpy_dtoh(h.ctypes.data, hi.offset(start), 4 * k))
        wrs.check(wrs.lib().wrs_rt_memcpy_dtoh(l.ctypes.data, lo.offset(start), 4 * k))
        got = h.astype(np.float64) + l.astype(np.float64)
        i = np.arange(start, start + k, dtype=np.float64)
        ref = np.where(i < N - 1, (i + 1.0) / 2.0 ** 32, 1.0)
        assert np.abs(got - ref).max() < 1e-12, start
    plan.close()
    del w, hi, lo
