"""Shared helpers for the parity tests."""
from __future__ import annotations

import math

import numpy as np


def z_scores(mean_a, var_a, n_a, mean_b, var_b, n_b):
    se2 = var_a / np.maximum(n_a, 1) + var_b / np.maximum(n_b, 1)
    ok = se2 > 0
    z = np.zeros_like(mean_a)
    z[ok] = (mean_a[ok] - mean_b[ok]) / np.sqrt(se2[ok])
    return z, ok


def check_four_se(z, ok, label=""):
    """Multiplicity-aware 4-SE rule: the count of |z| > 4 must stay within a binomial
    bound (p = 6.3e-5 per point, 6 sigma of the binomial + 3), and the z-scores of
    independent estimates must look standard normal (mean ~ 0, std ~ 1)."""
    zz = z[ok]
    n = len(zz)
    p = 6.334e-5
    exceed = int((np.abs(zz) > 4.0).sum())
    bound = n * p + 6.0 * math.sqrt(n * p) + 3
    assert exceed <= bound, f"{label}: {exceed} of {n} points beyond 4 SE (bound {bound:.1f})"
    assert abs(zz.mean()) < 5.0 / math.sqrt(n) + 0.02, f"{label}: z mean {zz.mean():.4f}"
    assert 0.85 < zz.std() < 1.15, f"{label}: z std {zz.std():.4f}"
    return exceed


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))
