"""CPU pins of the oracle's exact-key sweep (oracle.psa_stream_table, the
restatement of the streamed pack form).

It is checked against what does not share its arithmetic:
  * the FP64 restatement of the reference's sweep (pack.hpp:14-83 through
    oro_psa_alias_table_f64 with the lower-bound split): identical aliases,
    probabilities within float rounding, on several weight families;
  * the FP64 implied distribution of the table (inverse_alias_table.hpp:11-27)
    against w / sum(w);
  * hand cases whose exact answers are worked out below, one of them an exact
    tie of a light and a heavy key (the reference packs the heavy first:
    residual <= mean);
  * plausible mistakes -- light-first ties, the residual taken before the
    crossing light instead of after it, the heavy compared by its own key
    Y(q) instead of the next heavy's Y(q+1) -- planted in a pure-Python exact
    sweep, each of which these checks reject.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

def gen(oracle, kind, n):
    if kind == "uniform":
        return oracle.gen_uniform01(3, n)
    if kind == "lognormal":
        return oracle.gen_lognormal(6, 0.0, 3.0, n)
    if kind == "zipf":
        return oracle.gen_zipf(2, 1.1, n)
    if kind == "spike":
        return oracle.gen_spike(5, n)
    raise ValueError(kind)


def fp64_mean(w):
    return math.fsum(w.astype(np.float64).tolist()) / w.size


@pytest.mark.parametrize("kind,n", [("uniform", 20011), ("lognormal", 30011), ("zipf", 50021),
                                    ("spike", 40003), ("zipf", 7), ("uniform", 2)])
def test_stream_table_matches_fp64_sweep(oracle, kind, n):
    w = gen(oracle, kind, n)
    mean = fp64_mean(w)
    p, a, _ = oracle.psa_alias_table_f64(w, max(1, -(-n // 32)), mean=mean,
                                         split_rule=oracle.SPLIT_LOWER_BOUND)
    ps, as_ = oracle.psa_stream_table(w, mean)
    assert np.array_equal(as_, a)
    np.testing.assert_allclose(ps.astype(np.float64), p, rtol=3e-7, atol=1e-7)
    err = oracle.distribution_error(oracle.implied_distribution(ps, as_), w)
    assert err["max_rel"] <= 1e-6


def test_stream_table_tie_hand_case(oracle):
    """w = [0.5, 0.5, 1.5, 1.5], mean 1: lights 0, 1 (keys X = 0, 0.5, 1),
    heavies 2, 3 (Y(1) = 0.5, Y(2) = 1).  Light 0 goes first (0.5 > 0), then
    heavy 2 ties with light 1 (0.5 <= 0.5) and is packed first, its residual
    Y(1) - X(1) + mean = 1 -> (1, heavy 3); light 1 then aliases heavy 3; the
    last heavy keeps itself."""
    w = np.array([0.5, 0.5, 1.5, 1.5], np.float32)
    p, a = oracle.psa_stream_table(w, 1.0)
    assert a.tolist() == [2, 3, 3, 3]
    assert p.tolist() == [0.5, 0.5, 1.0, 1.0]
    # the FP64 sweep of the reference's pack agrees (all values are exact)
    p64, a64, _ = oracle.psa_alias_table_f64(w, 1, mean=1.0, split_rule=oracle.SPLIT_LOWER_BOUND)
    assert a64.tolist() == [2, 3, 3, 3] and p64.tolist() == [0.5, 0.5, 1.0, 1.0]


def test_stream_table_special_cases(oracle):
    # all equal: no heavy, identity
    p, a = oracle.psa_stream_table(np.full(5, 0.25, np.float32), 0.25)
    assert a.tolist() == list(range(5)) and p.tolist() == [1.0] * 5
    # a single non-zero weight: every bucket aliases it
    w = np.zeros(8, np.float32)
    w[3] = 1.0
    p, a = oracle.psa_stream_table(w, 1.0 / 8)
    assert a.tolist() == [3] * 8 and p[3] == 1.0 and np.all(p[np.arange(8) != 3] == 0.0)
    # one weight: itself
    p, a = oracle.psa_stream_table(np.array([2.5], np.float32), 2.5)
    assert a.tolist() == [0] and p.tolist() == [1.0]


def python_sweep(w, mean, mistake=None):
    """The sweep with exact rational keys (Fractions), for the planted
    mistakes; `mistake` None gives the correct table."""
    n = len(w)
    wf = [Fraction(float(x)) for x in w]
    m = Fraction(mean)
    heavy = [e for e in range(n) if wf[e] > m]
    light = [e for e in range(n) if not wf[e] > m]
    p = np.ones(n, np.float64)
    a = np.arange(n, dtype=np.uint32)
    H, L = len(heavy), len(light)
    if H == 0 or L == 0:
        return p, a
    X = [Fraction(0)]
    for e in light:
        X.append(X[-1] + m - wf[e])
    Y = [Fraction(0)]
    for e in heavy:
        Y.append(Y[-1] + wf[e] - m)
    i = q = 0
    while True:
        y = Y[q] if mistake == "own_key" else Y[q + 1]
        if q + 1 < H and (i == L or (y < X[i] if mistake == "light_first" else y <= X[i])):
            xi = X[i - 1] if (mistake == "before" and i > 0) else X[i]
            p[heavy[q]] = min(max(float((Y[q + 1] - xi + m) / m), 0.0), 1.0)
            a[heavy[q]] = heavy[q + 1]
            q += 1
        elif i < L:
            p[light[i]] = min(float(wf[light[i]] / m), 1.0)
            a[light[i]] = heavy[q]
            i += 1
        else:
            break
    p[heavy[H - 1]] = 1.0
    a[heavy[H - 1]] = heavy[H - 1]
    return p, a


@pytest.mark.parametrize("kind", ["uniform", "lognormal", "zipf"])
def test_python_sweep_agrees_with_oracle(oracle, kind):
    w = gen(oracle, kind, 600)
    mean = fp64_mean(w)
    ps, as_ = oracle.psa_stream_table(w, mean)
    p, a = python_sweep(w, mean)
    assert np.array_equal(as_, a)
    np.testing.assert_allclose(ps.astype(np.float64), p, rtol=3e-7, atol=1e-9)


@pytest.mark.parametrize("mistake", ["light_first", "before", "own_key"])
def test_planted_mistakes_are_caught(oracle, mistake):
    """Each mistake fails the checks the oracle passes: the hand case, the
    FP64 sweep's table, or the implied distribution."""
    caught = []
    w = np.array([0.5, 0.5, 1.5, 1.5], np.float32)
    p, a = python_sweep(w, 1.0, mistake)
    caught.append(a.tolist() != [2, 3, 3, 3] or p.tolist() != [0.5, 0.5, 1.0, 1.0])
    w = gen(oracle, "uniform", 600)
    mean = fp64_mean(w)
    p64, a64, _ = oracle.psa_alias_table_f64(w, 19, mean=mean,
                                             split_rule=oracle.SPLIT_LOWER_BOUND)
    p, a = python_sweep(w, mean, mistake)
    caught.append(not np.array_equal(a, a64) or
                  not np.allclose(p, p64, rtol=3e-7, atol=1e-7))
    err = oracle.distribution_error(oracle.implied_distribution(p, a), w)
    caught.append(err["max_rel"] > 1e-6)
    assert any(caught), mistake


def test_stream_table_matches_golden_fixtures(oracle):
    """The committed tiny fixtures (hand cases and seeded N <= 64, FP64 tables
    of the lower-bound sweep, tests/golden/psa_tiny.json): the exact-key sweep
    gives the same aliases and p within float rounding."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psa_tiny.json")
    cases = json.load(open(path))
    assert len(cases) >= 10
    for c in cases:
        w = np.array(c["weights"], np.float32)
        ref = c["f64"]
        p, a = oracle.psa_stream_table(w, ref["mean"])
        assert a.tolist() == ref["a"], c["label"]
        np.testing.assert_allclose(p.astype(np.float64), ref["p"], rtol=3e-7, atol=1e-7,
                                   err_msg=c["label"])
