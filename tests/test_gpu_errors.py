"""The error path of the device-side waits, forced from the host.

A build waits on the device in two places: the last block of the total reads
the other blocks' totals, and the look-back scans read their predecessors'
records.  Each wait is bounded; the plan's test hooks withhold one publication
(or publish a block total under the previous build's tag, as a late store of
that build would land: the reader must not sum it) and shrink the bound so
that a wait must give up.  Then the plan reports
WRS_EWAIT, no table row is written (the kernels after the failing wait return
at once), every kernel finishes (the stream synchronises), and the next
build on the same plan is correct again.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HOOK_SKIP_PARTIAL, HOOK_SKIP_SPINE, HOOK_STALE_TOTAL = 1, 2, 4


def case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("hook,n", [(HOOK_SKIP_PARTIAL, (1 << 20) + 7),
                                    (HOOK_SKIP_SPINE, (1 << 23) + 5),
                                    (HOOK_STALE_TOTAL, (1 << 22) + 9)])
@pytest.mark.parametrize("direct", [1, 2, 6])  # direct, classic and streamed forms
def test_forced_wait_failure_is_reported_and_writes_nothing(wrs, oracle, hook, n, direct):
    case("forced wait failure", hook, n, direct)
    w = oracle.gen_lognormal(5, 0.0, 2.0, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n, tune=dict(direct=direct))
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    # a clean build first: the plan is then reused after the failure
    plan.psa_build(dw, n, table)
    plan.status()
    good = table.numpy()
    table.fill_bytes(0x5A)
    plan.set_test_hooks(wait_polls=1000, hooks=hook)
    plan.psa_build(dw, n, table)
    with pytest.raises(wrs.WrsError) as e:
        plan.status()
    assert e.value.status == wrs.WRS_EWAIT
    t = table.numpy()
    assert np.all(t.view(np.uint8) == 0x5A), "a failed build wrote table rows"
    # the next build starts clean and gives the table of the first
    plan.set_test_hooks(0, 0)
    plan.psa_build(dw, n, table)
    plan.status()
    assert np.array_equal(table.numpy().view(np.uint64), good.view(np.uint64))


def test_status_of_a_clean_scan_and_batch(wrs, oracle):
    case("status ok")
    n = (1 << 22) + 3
    w = oracle.gen_uniform01(6, n)
    dw = wrs.to_device(w)
    plan = wrs.Plan(n)
    hi, lo = wrs.empty(n, np.float32), wrs.empty(n, np.float32)
    plan.its_build(dw, n, hi, lo)
    plan.status()
    off = np.array([0, 1000, 70000, n], np.uint64)
    bp = wrs.BatchPlan(off, tune=dict(partition=1))
    tb = wrs.empty(n, wrs.ALIAS_DTYPE)
    bp.build(dw, tb)
    bp.status()
