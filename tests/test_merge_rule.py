"""CPU check of the rule the merge-pack kernel uses in place of walking the
merged order of a block (csrc/mergepack.hip, merge_core): with the block's
light keys A (non-decreasing, A[nl] = +inf for the walk) and heavy keys B
(non-decreasing), the walk takes heavy q before light r iff B[q] <= A[r]
(heavy first on ties), and each item's row depends only on how many items of
the other kind precede it:
    light r:  q(r) = #{q < nh : B[q] <= A[r]}
    heavy q:  r(q) = #{r < nl : A[r] <  B[q]}
The kernel finds q(r) for the first light of each thread by binary search and
continues from it for the thread's next lights (one comparison, a search only
past it), and r(q) by binary search per heavy.  Checked against the walk on
random key sequences with ties and long runs of either kind."""
import random

import pytest


def walk(A, B, nl, nh):
    q = r = 0
    got = {}
    for _ in range(nl + nh):
        if q < nh and (r >= nl or B[q] <= A[r]):
            got[("h", q)] = r
            q += 1
        else:
            got[("l", r)] = q
            r += 1
    return got


def counts(A, B, nl, nh, per):
    inf = float("inf")

    def q_above(lo, a):  # first q in [lo, nh] with B[q] > a
        hi = nh
        while lo < hi:
            mid = (lo + hi) >> 1
            if B[mid] <= a:
                lo = mid + 1
            else:
                hi = mid
        return lo

    got = {}
    for e0 in range(0, nl, per):  # one thread's lights
        q = q_above(0, A[e0])
        bq = B[q] if q < nh else inf
        for k in range(per):
            r = e0 + k
            if r >= nl:
                break
            if k > 0 and bq <= A[r]:
                q = q_above(q + 1, A[r])
                bq = B[q] if q < nh else inf
            got[("l", r)] = q
    for q in range(nh):  # one heavy per thread
        lo, hi = 0, nl
        while lo < hi:
            mid = (lo + hi) >> 1
            if A[mid] < B[q]:
                lo = mid + 1
            else:
                hi = mid
        got[("h", q)] = lo
    return got


@pytest.mark.parametrize("per", [2, 4, 8])
def test_merge_counts_equal_the_merged_walk(per):
    rnd = random.Random(per)
    for _ in range(3000):
        nl, nh = rnd.randint(0, 40), rnd.randint(0, 40)
        spread = rnd.choice([1, 3, 10, 1000])  # 1: many ties
        A = [0]
        for _ in range(nl):
            A.append(A[-1] + rnd.randint(0, spread))
        B, x = [], rnd.randint(-50, 50)
        for _ in range(nh):
            x += rnd.randint(0, spread)
            B.append(x)
        assert counts(A, B, nl, nh, per) == walk(A, B, nl, nh), (nl, nh, A, B)
