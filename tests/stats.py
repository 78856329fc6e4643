"""Sample statistics used by the tests (written for this project):
chi-square over pooled bins with scipy's survival function for the p-value,
and the histogram RMSE against its closed-form expectation."""
from __future__ import annotations

import numpy as np
from scipy import stats as sps


def pooled_chi_square(samples: np.ndarray, q: np.ndarray, min_expected: float = 32.0) -> dict:
    """Chi-square of `samples` (indices) against probabilities q, pooling
    consecutive bins until each pool expects at least `min_expected` samples
    (the reference's chi_square, src/host/statistics/chi_square.hpp:17-37,
    uses one bin per index; with S/N ~ 1 pooling keeps the statistic valid)."""
    q = np.asarray(q, np.float64)
    S = samples.size
    counts = np.bincount(samples.astype(np.int64), minlength=q.size).astype(np.float64)
    if counts.size > q.size:
        raise AssertionError("sample index out of range")
    exp = q * S
    # pool boundaries: cumulative expected mass crossing multiples of min_expected
    cum = np.cumsum(exp)
    pool = np.floor(cum / min_expected).astype(np.int64)
    pool = np.minimum(pool, max(pool[-1] - 1, 0))  # merge the last partial pool
    e = np.bincount(pool, weights=exp)
    o = np.bincount(pool, weights=counts)
    keep = e > 0
    chi2 = float(np.sum((o[keep] - e[keep]) ** 2 / e[keep]))
    df = int(keep.sum()) - 1
    p = float(sps.chi2.sf(chi2, df)) if df > 0 else 1.0
    zero_mass_hits = int(counts[q == 0].sum())
    return dict(chi2=chi2, df=df, p=p, z=(chi2 - df) / np.sqrt(2.0 * max(df, 1)),
                zero_mass_hits=zero_mass_hits)


def rmse_ratio(samples: np.ndarray, q: np.ndarray) -> dict:
    """RMSE of the empirical distribution vs q, and its ratio to the expected
    RMSE of a perfect sampler, E[RMSE^2] = sum q(1-q) / (N S)."""
    q = np.asarray(q, np.float64)
    S = samples.size
    h = np.bincount(samples.astype(np.int64), minlength=q.size).astype(np.float64) / S
    rmse = float(np.sqrt(np.mean((h - q) ** 2)))
    expected = float(np.sqrt(np.sum(q * (1 - q)) / (q.size * S)))
    return dict(rmse=rmse, expected=expected, ratio=rmse / expected if expected > 0 else 0.0)
