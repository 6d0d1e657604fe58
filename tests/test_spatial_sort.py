"""Spatially sorted point indices (getSpatiallySortedPointIndices, utils/nearest_neighbor_finder.h
:112-128): the kd-tree leaf order of the points (leaves of at most 10, nanoflann's build; nanoflann
is not vendored, so the exact in-leaf order is parity unpinned).  Pinned here by what a kd order
fixes: a permutation; every range of more than 10 points splits at an axis-aligned plane into two
ranges that are kd orders themselves; leaves of at most 10 points; a hand-worked case of the
midpoint rule; and the locality the sort exists for.  Host-only: runs without a GPU."""
import time

import numpy as np
import pytest

from zombie_amd import engine as E


def _is_kd_order(p, lo, hi, depth=0):
    """True when p[lo:hi] splits recursively at axis-aligned planes down to ranges of <= 10 points
    (for each level the split candidates are the positions m where the ranges' coordinates along
    some axis do not interleave; points of continuous random inputs make them unique)."""
    n = hi - lo
    if n <= 10:
        return True
    q = p[lo:hi]
    for a in range(p.shape[1]):
        pre = np.maximum.accumulate(q[:, a])[:-1]
        suf = np.minimum.accumulate(q[::-1, a])[::-1][1:]
        for m in np.nonzero(pre <= suf)[0] + 1:
            if _is_kd_order(p, lo, lo + m, depth + 1) and _is_kd_order(p, lo + m, hi, depth + 1):
                return True
    return False


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [0, 1, 10, 11, 97, 5000])
def test_permutation_and_kd_structure(dim, n):
    rng = np.random.default_rng(n + dim)
    pts = rng.random((n, dim)).astype(np.float32)
    idx = E.spatially_sorted_point_indices(pts)
    assert idx.dtype == np.int32 and np.array_equal(np.sort(idx), np.arange(n))
    assert _is_kd_order(pts[idx], 0, n)
    # deterministic
    assert np.array_equal(idx, E.spatially_sorted_point_indices(pts))


def test_midpoint_rule_on_a_line():
    """21 points x = 0..20: root box [0, 20] cuts at 10 into {0..9} (a leaf) and {10..20}; that
    box [10, 20] cuts at 15 into {10..14} and {15..20}, keeping the point equal to the cut on the
    larger side as the count/2 rule places it."""
    x = np.arange(21, dtype=np.float32)
    rng = np.random.default_rng(1)
    perm = rng.permutation(21)
    pts = np.stack([x[perm], np.zeros(21, np.float32)], 1)
    order = x[perm][E.spatially_sorted_point_indices(pts)]
    assert set(order[:10]) == set(range(10))
    assert set(order[10:15]) == set(range(10, 15))
    assert set(order[15:]) == set(range(15, 21))


def test_identical_points_and_ties():
    """ties: all-equal coordinates split in the middle (no empty side); points of an integer
    grid (equal coordinates everywhere) still give a kd order"""
    pts = np.zeros((100, 3), np.float32)
    idx = E.spatially_sorted_point_indices(pts)
    assert np.array_equal(np.sort(idx), np.arange(100))
    g = np.arange(6, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.random.default_rng(2).permutation(len(pts))]
    idx = E.spatially_sorted_point_indices(pts)
    assert np.array_equal(np.sort(idx), np.arange(len(pts)))
    assert _is_kd_order(pts[idx], 0, len(pts))


def test_locality_and_rejections():
    rng = np.random.default_rng(3)
    pts = rng.random((200_000, 3)).astype(np.float32)
    t = time.perf_counter()
    idx = E.spatially_sorted_point_indices(pts)
    dt = time.perf_counter() - t
    step = lambda q: float(np.mean(np.linalg.norm(np.diff(q, axis=0), axis=1)))
    sorted_step, random_step = step(pts[idx]), step(pts)
    print(f"kd order of 200k points: {dt * 1e3:.0f} ms, mean neighbour step {sorted_step:.4f} vs {random_step:.4f}")
    assert sorted_step < 0.05 * random_step
    with pytest.raises(ValueError):
        E.spatially_sorted_point_indices(np.zeros((4, 4), np.float32))
    bad = pts[:10].copy()
    bad[3, 1] = np.nan
    with pytest.raises(Exception):
        E.spatially_sorted_point_indices(bad)
