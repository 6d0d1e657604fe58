"""The oracle's SDF-grid Dirichlet geometry (utils/sdf_grid_geometric_queries.h:87-205, 247-272)
against the exact sphere: distances, projections, normals and sphere-traced ray hits of an
interpolated signed-distance grid, and walk on spheres over it against a harmonic closed form."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB


@pytest.mark.parametrize("dim", [2, 3])
def test_sdf_queries_match_the_sphere(oracle_lib, dim):
    O = oracle_lib
    vals, lo, hi = PB.sphere_sdf_grid(dim, 96 if dim == 2 else 48)
    g = O.SdfGrid(vals, lo, hi)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(500, dim)); x = x / np.linalg.norm(x, axis=1, keepdims=True) * rng.uniform(0.2, 1.15, (500, 1))
    d, proj, nrm = g.query(x)
    r = np.linalg.norm(x, axis=1)
    h = (hi[0] - lo[0]) / vals.shape[0]
    assert np.max(np.abs(d - np.abs(r - 1))) < 0.2 * h           # trilinear interpolation of |x| - 1
    assert np.max(np.abs(np.linalg.norm(proj, axis=1) - 1)) < 1e-3
    u = x / r[:, None]
    assert np.min(np.sum(nrm * u, 1)) > 0.995
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-4)


@pytest.mark.parametrize("dim", [2, 3])
def test_sdf_walk_on_spheres_matches_exact(oracle_lib, dim):
    O = oracle_lib
    vals, lo, hi = PB.sphere_sdf_grid(dim, 96 if dim == 2 else 48)
    g = O.SdfGrid(vals, lo, hi)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    ws = E.walk_settings_host(1e-3, 1e-3, 1024)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(300, dim)); x = x / np.linalg.norm(x, axis=1, keepdims=True) * rng.random((300, 1)) * 0.8
    d, _, _ = g.query(x)
    st = O.Stats(len(x))
    rep = g.solve(pde, ws, x, d, 256, st)
    u = np.exp(x[:, 0].astype(np.float64)) * np.cos(x[:, 1].astype(np.float64))
    se2 = st.variance() / st.n
    pooled = np.sum(st.mean - u) / np.sqrt(np.sum(se2))
    z = (st.mean - u) / np.sqrt(se2)
    assert abs(pooled) < 4.0 and 0.8 < z.std() < 1.25, (pooled, z.std())
    assert rep["codes"][0] == 256 * len(x)
    # the walk length of walk on spheres in a ball grows like log(1 / eps)
    assert 5 < rep["total_steps"] / st.n.sum() < 60


@pytest.mark.parametrize("dim,res,tol", [(2, 96, 5e-4), (3, 48, 3e-3)])
def test_sphere_tracing_matches_the_exact_sphere(oracle_lib, dim, res, tol):
    """intersectZeroLevelSet (sdf_grid_geometric_queries.h:147-205) on the sampled unit sphere:
    rays from inside hit at t = -o.d + sqrt((o.d)^2 - |o|^2 + 1) with the outward normal; rays
    from outside aimed inward hit at distance 0.15; rays aimed away and rays stopped at half the
    hit distance miss.  Errors stay within the grid's interpolation error."""
    O = oracle_lib
    vals, lo, hi = PB.sphere_sdf_grid(dim, res)
    g = O.SdfGrid(vals, lo, hi)
    rng = np.random.default_rng(1)
    o = rng.normal(size=(1500, dim))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * rng.random((1500, 1)) * 0.9
    d = rng.normal(size=(1500, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hit, t, pt, nr = g.ray(o.astype(np.float32), d.astype(np.float32), 10.0)
    od = (o * d).sum(1)
    t_exact = -od + np.sqrt(od ** 2 - (o ** 2).sum(1) + 1)
    assert hit.all()
    assert np.abs(t - t_exact).max() < tol
    assert np.abs(np.linalg.norm(pt, axis=1) - 1).max() < tol
    assert np.abs((nr * pt).sum(1) - 1).max() < tol
    outside = (d * 1.15).astype(np.float32)
    hit2, t2, _, nr2 = g.ray(outside, (-d).astype(np.float32), 10.0)
    assert hit2.all() and np.abs(t2 - 0.15).max() < tol and np.all((nr2 * d).sum(1) > 0.99)
    assert not g.ray(outside, d.astype(np.float32), 10.0)[0].any()
    assert not g.ray(o.astype(np.float32), d.astype(np.float32), (0.5 * t_exact).astype(np.float32))[0].any()


@pytest.mark.parametrize("dim", [2, 3])
def test_sdf_gradient_walks_match_analytic(oracle_lib, dim):
    """estimateSolutionAndGradient continued over the SDF geometry (walk_on_spheres.h:416-593 with
    the SdfGrid queries): gradients and solutions of u = e^x cos y on the sampled unit disk / ball
    within 4 pooled standard errors; dropping the first step's boundary term is detected."""
    O = oracle_lib
    vals, lo, hi = PB.sphere_sdf_grid(dim, 96 if dim == 2 else 48)
    g = O.SdfGrid(vals, lo, hi)
    rng = np.random.default_rng(31)
    x = rng.normal(size=(300, dim))
    x = (x / np.linalg.norm(x, axis=1, keepdims=True) * rng.random((300, 1)) ** (1 / dim) * 0.7).astype(np.float32)
    d, _, _ = g.query(x)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)),
                     is_source_constant=True)
    ws = E.walk_settings_host(1e-3, 1e-3, 1024)
    st = O.GradStats(len(x), dim)
    g.solve_grad(pde, ws, x, d, 256, st, seed=3)
    assert (st.n_grad == 256).all() and (st.n == 256).all()
    xd = x.astype(np.float64)
    e = np.exp(xd[:, 0])
    gr = np.zeros_like(xd)
    gr[:, 0] = e * np.cos(xd[:, 1])
    gr[:, 1] = -e * np.sin(xd[:, 1])
    pooled = np.sum(st.grad_mean - gr, 0) / np.sqrt(np.sum(st.grad_variance() / st.n_grad[:, None], 0))
    assert np.all(np.abs(pooled) < 4.0), pooled
    u = e * np.cos(xd[:, 1])
    assert abs(np.sum(st.mean - u) / np.sqrt(np.sum(st.variance() / st.n))) < 4.0
    # the gradient is not the trivial one: its x component is far from zero in every point's units
    assert np.abs(np.sum(st.grad_mean[:, 0]) / np.sqrt(np.sum(st.grad_variance()[:, 0] / st.n_grad))) > 10.0


def test_sdf_walk_splitting_matches_exact(oracle_lib):
    """applyWeightWindow's splitting (walk_on_spheres.h:192-226) over the SDF walks: a walk whose
    throughput exceeds the threshold continues at the threshold and queues parts; the estimate
    is the sum over the parts.  Screened u = exp(a . x) on the sampled unit ball; a walk that
    dropped its parts would lose weight and be detected."""
    O = oracle_lib
    vals, lo, hi = PB.sphere_sdf_grid(3, 48)
    g = O.SdfGrid(vals, lo, hi)
    sigma = 6.0
    a = np.sqrt(sigma) * np.array([0.0, 0.6, 0.8])
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a)),
                     absorption=sigma, is_source_constant=True)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(300, 3))
    x = (x / np.linalg.norm(x, axis=1, keepdims=True) * rng.random((300, 1)) ** (1 / 3) * 0.8).astype(np.float32)
    d, _, _ = g.query(x)
    u = np.exp(x.astype(np.float64) @ a)
    for thr in (0.4, 3.4e38):   # split, unsplit
        ws = E.walk_settings_host(1e-3, 1e-3, 1024, splitting_threshold=thr, russian_roulette_threshold=0.05)
        st = O.Stats(len(x))
        g.solve(pde, ws, x, d, 256, st, seed=9)
        pooled = np.sum(st.mean - u) / np.sqrt(np.sum(st.variance() / st.n))
        assert abs(pooled) < 4.0, (thr, pooled)
