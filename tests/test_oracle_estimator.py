"""The oracle estimator against manufactured solutions (FP64 analytic ground truth), and its
bookkeeping rules (no GPU).  These pin the whole restated walk (walk_on_spheres.h:228-414,
walk_on_stars.h:345-662) against the mathematics rather than against itself."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB

from helpers import rmse


def _z_vs_exact(st, u):
    se = np.sqrt(st.variance() / np.maximum(st.n, 1))
    ok = se > 0
    return (st.mean[ok] - u[ok]) / se[ok]


def _pooled_z(st, u):
    """z-score of the average error over all points.  Unlike the mean of per-point
    t-statistics it is not pulled negative by the correlation between a skewed sample's
    mean and its variance estimate (walk estimates near the boundary are very skewed)."""
    se2 = st.variance() / np.maximum(st.n, 1)
    return float(np.sum(st.mean - u) / np.sqrt(np.sum(se2)))


def test_laplace_2d_star_matches_exact(oracle_lib):
    O = oracle_lib
    p = PB.star2d()
    pts = p.meta["points"][::40]
    sc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = sc.closest_point(pts)
    st = O.Stats(len(pts))
    rep = O.solve(sc, p.pde, p.settings, pts, d, 512, st, rng_mode=O.RNG_PER_POINT)
    u = p.exact(pts)
    z = _z_vs_exact(st, u)
    assert 0.85 < z.std() < 1.25
    # the epsilon shell biases each estimate by O(eps |grad u|) ~ 1e-3 at eps = 1e-3
    assert abs(_pooled_z(st, u)) < 6.0
    assert rep["codes"][0] == st.n.sum()
    ws = E.walk_settings_host(1e-5, 1e-5, 1024)
    st5 = O.Stats(len(pts))
    O.solve(sc, p.pde, ws, pts, d, 512, st5, rng_mode=O.RNG_PER_POINT)
    assert abs(_pooled_z(st5, u)) < 4.0


def test_poisson_3d_source_term_matches_exact(oracle_lib):
    O = oracle_lib
    V, F = PB.icosphere(3)
    V = PB.sh_displace(V, 11).astype(np.float32)
    p = PB.sphere3d()
    sc = O.Scene(V, F.astype(np.int32), None, V.min(0) - 1e-3, V.max(0) + 1e-3)
    cand = PB.sample_box(V.min(0), V.max(0), 3000, 1)
    pts = cand[sc.inside_domain(cand)][:600]
    d, *_ = sc.closest_point(pts)
    st = O.Stats(len(pts))
    O.solve(sc, p.pde, p.settings, pts, d, 256, st, rng_mode=O.RNG_PER_POINT)
    z = _z_vs_exact(st, PB.sin_product(pts))
    assert abs(_pooled_z(st, PB.sin_product(pts))) < 4.0 and 0.8 < z.std() < 1.25
    # dropping the source term must be detected: u != E[g(X_T)] for f = 12 u != 0
    ws = E.walk_settings_host(1e-4, 1e-4, 1024, ignore_source_contribution=1)
    st2 = O.Stats(len(pts))
    O.solve(sc, p.pde, ws, pts, d, 256, st2, rng_mode=O.RNG_PER_POINT)
    z2 = _z_vs_exact(st2, PB.sin_product(pts))
    assert np.mean(z2 ** 2) > 4.0


def test_screened_poisson_matches_exact_and_rr_unbiased(oracle_lib):
    O = oracle_lib
    V, F = PB.icosphere(3)
    V = V.astype(np.float32)
    sigma = 6.0
    a = np.sqrt(sigma) * np.array([0.0, 0.6, 0.8])
    u = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a))
    pde = E.make_pde(source=E.zero(), dirichlet=u, absorption=sigma, is_source_constant=True)
    sc = O.Scene(V, F.astype(np.int32), None, V.min(0) - 1e-3, V.max(0) + 1e-3)
    rng = np.random.default_rng(2)
    pts = (rng.normal(size=(400, 3)) * 0.3).astype(np.float32)
    pts = pts[np.linalg.norm(pts, axis=1) < 0.9]
    d, *_ = sc.closest_point(pts)
    for rr in (0.0, 0.3):
        ws = E.walk_settings_host(1e-4, 1e-4, 1024, russian_roulette_threshold=rr)
        st = O.Stats(len(pts))
        rep = O.solve(sc, pde, ws, pts, d, 256, st, rng_mode=O.RNG_PER_POINT)
        z = _z_vs_exact(st, PB.exp_linear(pts, a))
        pz = _pooled_z(st, PB.exp_linear(pts, a))
        assert abs(pz) < 4.0 and 0.8 < z.std() < 1.25, (rr, pz, z.std())
        if rr > 0:
            assert rep["codes"][1] > 0
    # without the Yukawa throughput the estimate is biased high
    pde0 = E.make_pde(source=E.zero(), dirichlet=u, absorption=0.0, is_source_constant=True)
    st = O.Stats(len(pts))
    O.solve(sc, pde0, E.walk_settings_host(1e-4, 1e-4, 1024), pts, d, 256, st, rng_mode=O.RNG_PER_POINT)
    assert _pooled_z(st, PB.exp_linear(pts, a)) > 10.0


def test_mixed_2d_walk_on_stars_matches_exact(oracle_lib):
    O = oracle_lib
    p = PB.mixed2d(n_segments=1024, n_slits=8)
    sc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    V, I = p.meta["outer"]
    cand = PB.sample_box(p.domain_min, p.domain_max, 1500, 9)
    pts = cand[PB.points_in_polygon(cand, V, I)][:300]
    d, *_ = sc.closest_point(pts, boundary=0)
    st = O.Stats(len(pts))
    rep = O.solve(sc, p.pde, p.settings, pts, d, 256, st, wost=True, rng_mode=O.RNG_PER_POINT)
    z = _z_vs_exact(st, p.exact(pts))
    pz = _pooled_z(st, p.exact(pts))
    assert abs(pz) < 4.0 and 0.8 < z.std() < 1.3, (pz, z.std())
    assert rep["codes"][0] > 0


def test_progressive_solves_merge_exactly(oracle_lib):
    """Chan's merge: two calls of 32 walks (stream offsets 0, 32) equal one call of 64."""
    O = oracle_lib
    p = PB.star2d()
    pts = p.meta["points"][::300]
    sc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = sc.closest_point(pts)
    a = O.Stats(len(pts))
    O.solve(sc, p.pde, p.settings, pts, d, 32, a, seed=5, stream_offset=0)
    O.solve(sc, p.pde, p.settings, pts, d, 32, a, seed=5, stream_offset=32)
    b = O.Stats(len(pts))
    O.solve(sc, p.pde, p.settings, pts, d, 64, b, seed=5, stream_offset=0)
    m = d > p.settings.epsilon_shell_for_absorbing_boundary
    np.testing.assert_array_equal(a.n[m], b.n[m])
    np.testing.assert_allclose(a.mean[m], b.mean[m], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(a.m2[m], b.m2[m], rtol=1e-9, atol=1e-13)


def test_sample_types_and_shell_rules(oracle_lib):
    O = oracle_lib
    p = PB.star2d()
    pts = p.meta["points"][::500]
    sc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = sc.closest_point(pts)
    types = np.zeros(len(pts), np.uint8)
    types[0] = A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY
    dd = d.copy()
    dd[1] = 0.5 * p.settings.epsilon_shell_for_absorbing_boundary
    st = O.Stats(len(pts))
    O.solve(sc, p.pde, p.settings, pts, dd, 16, st, types=types)
    assert st.n[0] == 1 and st.mean[0] == pytest.approx(float(p.exact(pts[:1])[0]), rel=1e-5)
    assert st.n[1] == 1      # inside the shell: one walk
    assert (st.n[2:] == 16).all()
    O.solve(sc, p.pde, p.settings, pts, dd, 16, st, types=types)
    assert st.n[0] == 1      # the boundary value is recorded once


def test_robin_walk_on_stars_matches_exact(oracle_lib):
    """Robin conditions du/dn + kappa u = h on the reflecting segments (walk_on_stars.h:277-307
    reflectance, :372-376 epsilon radius on the boundary, Robin star-radius bound): the estimate
    matches u = x^3 - 3xy^2; treating the same data as Neumann (reflectance 1) is detected."""
    O = oracle_lib
    p = PB.mixed2d_robin(kappa=2.0, n_segments=1024, n_slits=8)
    sc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    sc.set_robin(*p.meta["robin"])
    V, I = p.meta["outer"]
    cand = PB.sample_box(p.domain_min, p.domain_max, 1500, 21)
    pts = cand[PB.points_in_polygon(cand, V, I)][:200]
    d, *_ = sc.closest_point(pts, boundary=0)
    st = O.Stats(len(pts))
    rep = O.solve(sc, p.pde, p.settings, pts, d, 256, st, wost=True)
    u = p.exact(pts)
    z = _z_vs_exact(st, u)
    pz = _pooled_z(st, u)
    assert abs(pz) < 4.0 and 0.8 < z.std() < 1.3, (pz, z.std())
    pde_n = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=p.pde.robin, is_source_constant=True)
    st2 = O.Stats(len(pts))
    O.solve(sc, pde_n, p.settings, pts, d, 256, st2, wost=True)
    assert abs(_pooled_z(st2, u)) > 8.0
