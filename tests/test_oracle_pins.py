"""Independent pins of the oracle's reductions and of its scaled FP32 PSA path.

The reference's own tests of its reductions compare them against exact or
library sums (src/host/assert/test.cpp:166-250, restating
src/host/reference/reduce.hpp:81-92).  Here the checks are exact rational sums
(math.fsum) on inputs where an uncompensated or sign-flipped Kahan loop gives a
visibly different answer, and closed forms for the 1e-4 weight scaling of
psa_alias_table.hpp:25-39.  Each check is also run against a planted mistake
(an uncompensated sum, a flipped compensation sign, a dropped scale) to show
that it rejects it.
"""
import math

import numpy as np
import pytest


def _kahan_variant(x, dtype, flip=False, compensate=True):
    """Python restatement with a planted mistake (flip: c = (t - sum) + y;
    compensate=False: plain running sum), evaluated in `dtype` scalars."""
    s = dtype(0)
    c = dtype(0)
    with np.errstate(over="ignore", invalid="ignore"):
        for v in x:
            s, c = _kahan_step(s, c, dtype(v), dtype, flip, compensate)
    return s


def _kahan_step(s, c, v, dtype, flip, compensate):
    if not compensate:
        return dtype(s + v), c
    y = dtype(v - c)
    t = dtype(s + y)
    return t, (dtype(dtype(t - s) + y) if flip else dtype(dtype(t - s) - y))


def _adversarial(dtype, tiny_exp):
    # 1.0 then 4096 terms of 2^tiny_exp: each term alone is absorbed by 1.0
    # (below half an ulp of 1), their sum 2^(tiny_exp + 12) is not.
    x = np.array([1.0] + [2.0 ** tiny_exp] * 4096, dtype)
    exact = math.fsum(x.astype(np.float64).tolist())
    assert exact == 1.0 + 2.0 ** (tiny_exp + 12)
    return x, exact


@pytest.mark.parametrize("dtype,tiny_exp", [(np.float32, -25), (np.float64, -54)])
def test_kahan_recovers_absorbed_terms(oracle, dtype, tiny_exp):
    x, exact = _adversarial(dtype, tiny_exp)
    got = oracle.kahan_reduce_f32(x) if dtype is np.float32 else oracle.kahan_reduce_f64(x)
    assert got == exact  # exactly representable, and Kahan's error here is 0
    # the same check rejects the planted mistakes
    assert _kahan_variant(x, dtype, compensate=False) != exact
    assert _kahan_variant(x, dtype, flip=True) != exact
    assert _kahan_variant(x, dtype) == exact  # the restatement itself


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kahan_within_bound_of_exact_sum(oracle, seed):
    # Kahan's bound |E| <= 2 eps |S| + O(n eps^2) sum|x| (Knuth, TAOCP 4.2.2):
    # within 2 ulps of the exact sum for these positive inputs, where a plain
    # float32 running sum of 2e5 terms drifts by about ten
    rng = np.random.default_rng(seed)
    x = rng.random(200_000).astype(np.float32)
    exact = math.fsum(x.astype(np.float64).tolist())
    got = float(oracle.kahan_reduce_f32(x))
    ulp = float(np.spacing(np.float32(exact)))
    assert abs(got - exact) <= 2 * ulp
    naive = float(np.cumsum(x, dtype=np.float32)[-1])  # sequential, uncompensated
    assert abs(naive - exact) > 4 * ulp
    x64 = rng.random(200_000) * 10.0 ** rng.integers(-8, 8, 200_000)
    exact64 = math.fsum(x64.tolist())
    assert abs(oracle.kahan_reduce_f64(x64) - exact64) <= 2 * np.spacing(exact64)


def test_reference_scale_closed_form(oracle):
    # w = 10^4 k: (float)((double)w * 1e-4) == k exactly (the double product is
    # within 5e-17 relative of k), so the scaled path sees the integers k, its
    # Kahan mean is exactly sum(k) / N and its table implies k / sum(k)
    k = np.arange(1, 9, dtype=np.float64)
    w = (1e4 * k).astype(np.float32)
    assert np.array_equal((w.astype(np.float64) * 0.0001).astype(np.float32), k.astype(np.float32))
    p, a, d = oracle.psa_alias_table_f32(w, 4, reference_scale=True)
    assert d["mean"] == 4.5 and d["H"] == 4
    q = oracle.implied_distribution(p, a)
    assert np.allclose(q, k / k.sum(), rtol=1e-6, atol=0)
    # a dropped scale shows up in the mean; the table alone cannot show it
    _, _, d_unscaled = oracle.psa_alias_table_f32(w, 4, reference_scale=False)
    assert d_unscaled["mean"] == 45000.0 and d_unscaled["mean"] != d["mean"]


def test_reference_scale_mean_is_kahan_of_scaled_weights(oracle):
    # 1e4 followed by 4096 weights that scale to about 2^-25: after the scaling
    # a plain float32 sum absorbs every small term, the compensated one keeps them
    w = np.array([1e4] + [2.0 ** -25 * 1e4] * 4096, np.float32)
    scaled = (w.astype(np.float64) * 0.0001).astype(np.float32)
    n = np.float32(w.size)
    exact = math.fsum(scaled.astype(np.float64).tolist())
    want = np.float32(exact) / n  # mean of the correctly rounded total
    p, a, d = oracle.psa_alias_table_f32(w, 64, reference_scale=True)
    ulp = float(np.spacing(want))
    assert abs(d["mean"] - float(want)) <= ulp

    def rejected(mean):
        return abs(mean - float(want)) > ulp

    # planted: no compensation, and no scale
    assert rejected(float(np.cumsum(scaled, dtype=np.float32)[-1] / n))
    assert rejected(float(np.float32(oracle.kahan_reduce_f32(w)) / n))
    # the table of the scaled path is a valid alias table of the scaled weights
    q = oracle.implied_distribution(p, a)
    err = oracle.distribution_error(q, scaled)
    assert err["max_rel"] < 1e-3
