"""Independence of the per-walk random number streams (oracle, no GPU).

Walk w of point p draws from PCG32 stream p started at mix(seed, p, w), a chain of SplitMix64
finalisers (the GPU kernels use the same function).  Pinned here by:
- SplitMix64's published outputs (the finaliser) and bijectivity of the walk seeding;
- a draw-index scan: at every draw index of the first 1024, the uniforms of N walk pairs
  (adjacent walks, walks further apart, the same walk of neighbouring points) correlate by no
  more than 5/sqrt(N);
- the estimator: the across-seed variance of per-point means at 4096 walks per point equals the
  mean squared standard error the walks report (ratio 1 +- 0.05), which correlated walks break.
Both checks are shown to catch the flawed scheme of one shared initstate on per-walk
increments (oracle rng_mode 2), whose walks move in lockstep.
"""
import json
import math
import os

import numpy as np
import pytest

from zombie_amd import problems as PB

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = 2 ** 64 - 1


def test_splitmix64_known_answers(oracle_lib):
    with open(os.path.join(GOLDEN, "splitmix64_seed0.json")) as f:
        kat = json.load(f)
    inc = int(kat["increment"], 16)
    want = [int(v, 16) for v in kat["u64"]]
    got = [oracle_lib.splitmix64_mix((k + 1) * inc & M64) for k in range(len(want))]
    assert got == want


def test_walk_seeding_is_bijective_in_the_walk_index(oracle_lib):
    O = oracle_lib
    for seed, p in ((0x5EED, 0), (0x5EED, 123457), (~0x5EED, 9)):
        init = {O.walk_init_state(seed, p, w) for w in range(20000)}
        assert len(init) == 20000
    # walks of different points and seeds start elsewhere too
    a = {O.walk_init_state(7, p, 0) for p in range(5000)}
    assert len(a) == 5000
    assert O.walk_init_state(7, 0, 0) != O.walk_init_state(8, 0, 0)
    # the generator of walk w is PCG32 stream p from that state
    f = O.walk_floats(0x5EED, [11], [4], 16)[0]
    np.testing.assert_array_equal(f, O.pcg32_f32(O.walk_init_state(0x5EED, 11, 4), 11, 16))


def _max_abs_corr(a, b):
    a = a.astype(np.float64) - a.mean(0)
    b = b.astype(np.float64) - b.mean(0)
    c = (a * b).sum(0) / np.sqrt((a * a).sum(0) * (b * b).sum(0))
    k = int(np.argmax(np.abs(c)))
    return float(abs(c[k])), k


N_PAIRS = 10000
N_DRAWS = 1024


def _pairs(lag_walk, lag_point):
    rng = np.random.default_rng(17)
    p = rng.permutation(1 << 20)[:N_PAIRS].astype(np.uint32)
    w = rng.integers(0, 1 << 16, N_PAIRS).astype(np.uint32)
    return p, w, p + np.uint32(lag_point), w + np.uint32(lag_walk)


@pytest.mark.parametrize("lag_walk,lag_point", [(1, 0), (2, 0), (17, 0), (256, 0), (0, 1)])
def test_draw_index_scan_shows_no_correlation(oracle_lib, lag_walk, lag_point):
    O = oracle_lib
    p0, w0, p1, w1 = _pairs(lag_walk, lag_point)
    a = O.walk_floats(0x5EED, p0, w0, N_DRAWS)
    b = O.walk_floats(0x5EED, p1, w1, N_DRAWS)
    c, k = _max_abs_corr(a, b)
    assert c < 5.0 / math.sqrt(N_PAIRS), f"walk lag {lag_walk}, point lag {lag_point}: |corr| {c:.4f} at draw {k}"


def test_draw_index_scan_catches_lockstep_walks(oracle_lib):
    """The flawed scheme (shared initstate 0x5EED, increment 2((p << 32) | w) + 1): adjacent
    walks correlate far beyond the bound at some draw index."""
    O = oracle_lib
    p0, w0, p1, w1 = _pairs(1, 0)
    n = 2000
    a = np.stack([O.pcg32_f32(0x5EED, (int(p) << 32) | int(w), N_DRAWS) for p, w in zip(p0[:n], w0[:n])])
    b = np.stack([O.pcg32_f32(0x5EED, (int(p) << 32) | int(w), N_DRAWS) for p, w in zip(p1[:n], w1[:n])])
    c, _ = _max_abs_corr(a, b)
    assert c > 5.0 / math.sqrt(n)


def _variance_ratio(O, mode, n_points, n_walks, n_seeds):
    """Across-seed variance of per-point means / mean reported SE^2 (3D Poisson, sin 2x sin 2y
    sin 2z on a 320-triangle harmonic-displaced icosphere, shell 1e-2)."""
    prob = PB.sphere3d(level=2, eps=1e-2)
    sc = O.Scene(prob.positions, prob.indices, None, prob.domain_min, prob.domain_max)
    rng = np.random.default_rng(1)
    cand = (rng.random((20000, 3)) * 1.6 - 0.8).astype(np.float32)
    pts = cand[sc.inside_domain(cand)]
    d = sc.closest_point(pts)[0]
    keep = d > 0.02   # points in the shell take one walk
    pts, d = pts[keep][:n_points], d[keep][:n_points]
    assert len(pts) == n_points
    means, se2 = [], []
    for s in range(n_seeds):
        st = O.Stats(n_points)
        O.solve(sc, prob.pde, prob.settings, pts, d, n_walks, st, seed=1000 + s, rng_mode=mode)
        assert (st.n == n_walks).all()
        means.append(st.mean.copy())
        se2.append(st.variance() / st.n)
    return float(np.var(np.array(means), axis=0, ddof=1).mean() / np.array(se2).mean())


def test_walk_estimates_are_independent(oracle_lib):
    """600 points x 4096 walks x 8 seeds (4200 degrees of freedom: the ratio's own standard
    error is about 0.022)."""
    r = _variance_ratio(oracle_lib, oracle_lib.RNG_PER_WALK, 600, 4096, 8)
    assert abs(r - 1.0) < 0.05, r


def test_variance_ratio_catches_lockstep_walks(oracle_lib):
    r = _variance_ratio(oracle_lib, oracle_lib.RNG_LOCKSTEP, 600, 2048, 4)
    assert r > 1.1, r
