"""Boundary value caching on the GPU: the splat kernel against the oracle's restatement on the
same inputs (every boundary form, both orientations, source samples, cutoffs, regularization,
repeated calls), and the whole solver (boundary_value_caching.h:178-946) on manufactured
Dirichlet problems in 2D and 3D and a mixed Dirichlet / Neumann problem in 3D.  The oracle's splat itself is pinned against Green's
representation formula in tests/test_oracle_bvc.py."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import bvc as BV
from zombie_amd import engine as E
from zombie_amd import problems as PB

pytestmark = pytest.mark.gpu


def _t(torch, a, dt=np.float32):
    return torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()


def _random_cache(rng, dim, n):
    pt = rng.normal(size=(n, dim)) * 0.6
    nr = rng.normal(size=(n, dim)); nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    ty = rng.choice([A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY, A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY, A.ZG_SAMPLE_IN_DOMAIN], n)
    kappa = np.where(ty == A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY, rng.choice([0.0, 0.3, 3.0], n), 0.0)
    return dict(pt=pt, normal=nr, type=ty.astype(np.uint8), normal_aligned=(rng.random(n) < 0.3).astype(np.uint8),
                pdf=rng.uniform(0.5, 2.0, n), solution=rng.normal(size=n), normal_derivative=rng.normal(size=n),
                contribution=rng.normal(size=n), robin_coeff=kappa)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kreg", [0.0, 0.05])
def test_splat_matches_oracle(gpu, oracle_lib, dim, kreg):
    torch = gpu
    O = oracle_lib
    rng = np.random.default_rng(dim * 10 + int(kreg * 100))
    S = _random_cache(rng, dim, 3000)
    m = 1500
    x = rng.normal(size=(m, dim)) * 0.5
    ety = rng.choice([A.ZG_SAMPLE_IN_DOMAIN, A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY, A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY],
                     m, p=[0.8, 0.1, 0.1]).astype(np.uint8)
    dA = rng.uniform(0.0, 1.0, m); dR = rng.uniform(0.0, 1.0, m)
    ost = O.BvcStats(m, dim)
    O.bvc_splat(dim, S, dict(pt=x, type=ety, dist_abs=dA, dist_refl=dR), ost, radius_clamp=1e-3,
                kernel_regularization=kreg, robin_cutoff=1.0, cutoff_abs=0.05, cutoff_refl=0.02)
    samples = E.BvcSamples(pt=_t(torch, S["pt"]), normal=_t(torch, S["normal"]), type=_t(torch, S["type"], np.uint8),
                           pdf=_t(torch, S["pdf"]), solution=_t(torch, S["solution"]),
                           normal_derivative=_t(torch, S["normal_derivative"]), contribution=_t(torch, S["contribution"]),
                           robin_coeff=_t(torch, S["robin_coeff"]), normal_aligned=_t(torch, S["normal_aligned"], np.uint8))
    ev = E.BvcEvalPoints(pt=_t(torch, x), dist_to_absorbing=_t(torch, dA), dist_to_reflecting=_t(torch, dR),
                         type=_t(torch, ety, np.uint8))
    st = E.BvcStats.zeros(m, dim)
    E.bvc_splat(dim, samples, ev, st, radius_clamp=1e-3, kernel_regularization=kreg, robin_coeff_cutoff=1.0,
                cutoff_dist_to_absorbing=0.05, cutoff_dist_to_reflecting=0.02)
    n = st.n.cpu().numpy()
    assert np.array_equal(n, ost.n)
    assert (n.sum(1) == 0).sum() == ((dA < 0.05) | (dR < 0.02)).sum()
    scale = np.abs(ost.mean).max() + 1
    np.testing.assert_allclose(st.mean.cpu().numpy(), ost.mean, rtol=2e-4, atol=2e-5 * scale)
    gscale = np.abs(ost.grad_mean).max() + 1
    np.testing.assert_allclose(st.grad_mean.cpu().numpy(), ost.grad_mean, rtol=2e-4, atol=2e-5 * gscale)
    np.testing.assert_allclose(st.m2.cpu().numpy(), ost.m2, rtol=2e-3, atol=1e-4 * scale ** 2 * n.max())
    # a second call Chan-merges like the oracle
    E.bvc_splat(dim, samples, ev, st, radius_clamp=1e-3, kernel_regularization=kreg, robin_coeff_cutoff=1.0,
                cutoff_dist_to_absorbing=0.05, cutoff_dist_to_reflecting=0.02)
    assert np.array_equal(st.n.cpu().numpy(), 2 * ost.n)
    np.testing.assert_allclose(st.mean.cpu().numpy(), ost.mean, rtol=2e-4, atol=2e-5 * scale)


def test_splat_argument_errors(gpu):
    torch = gpu
    z = torch.zeros(4, 3, device="cuda")
    s = E.BvcSamples(pt=z, normal=z, type=torch.zeros(4, dtype=torch.uint8, device="cuda"),
                     pdf=torch.ones(4, device="cuda"), solution=torch.zeros(4, device="cuda"),
                     normal_derivative=torch.zeros(4, device="cuda"))
    ev = E.BvcEvalPoints(pt=z, dist_to_absorbing=torch.ones(4, device="cuda"))
    s2 = E.BvcSamples(pt=z[:, :2].contiguous(), normal=z[:, :2].contiguous(), type=s.type, pdf=s.pdf,
                      solution=s.solution, normal_derivative=s.normal_derivative)
    ev2 = E.BvcEvalPoints(pt=z[:, :2].contiguous(), dist_to_absorbing=ev.dist_to_absorbing)
    with pytest.raises(A.ZombieError):            # screened kernels are 3D only
        E.bvc_splat(2, s2, ev2, E.BvcStats.zeros(4, 2), absorption=1.0)
    with pytest.raises(A.ZombieError):
        E.bvc_splat(3, s, ev, E.BvcStats.zeros(4, 3), absorption=-1.0)
    with pytest.raises(A.ZombieError):
        E.bvc_splat(4, s, ev, E.BvcStats.zeros(4, 3))


def _splat_se(st):
    """standard error of each point's estimate: the five category means are independent sums"""
    n = st.n.cpu().numpy().astype(np.float64)
    m2 = st.m2.cpu().numpy()
    var = np.where(n > 1, m2 / np.maximum(n - 1, 1) / np.maximum(n, 1), 0.0)
    return np.sqrt(var.sum(1))


def _exp_cos(x):
    x = x.astype(np.float64)
    return np.exp(x[:, 0]) * np.cos(x[:, 1])


def test_bvc_solver_3d_dirichlet(gpu, oracle_lib):
    """Laplace on the unit icosphere with u = e^x cos y: cached absorbing samples (solution and
    normal derivative from GPU walks) splatted into interior points reproduce u."""
    torch = gpu
    V, F = PB.icosphere(4)
    V = V.astype(np.float32)
    sc = E.Scene(V, F.astype(np.int32), None, V.min(0) - 1e-3, V.max(0) + 1e-3)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    ws = E.walk_settings_host(1e-4, 1e-4, 1024)
    solver = BV.BoundaryValueCachingSolver(sc, BV.UniformBoundarySampler(V, F), None, None, seed=3)
    solver.generate_samples(40000, 0, 0, 2e-3, 0.0, False)
    assert sum(len(c) for c in solver.caches) == 40000
    solver.compute_sample_estimates(pde, ws, 64, 64)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2000, 3)); x = x / np.linalg.norm(x, axis=1, keepdims=True) * rng.random((2000, 1)) ** (1 / 3) * 0.7
    xt = torch.from_numpy(x.astype(np.float32)).cuda()
    d = sc.closest_point(xt, want_point=False)[0]
    ev = E.BvcEvalPoints(pt=xt, dist_to_absorbing=d.contiguous())
    st = E.BvcStats.zeros(len(x), 3)
    solver.splat(ev, st, radius_clamp=1e-3)
    u = st.solution().cpu().numpy()
    err = u - _exp_cos(x)
    rmse = float(np.sqrt(np.mean(err ** 2)))
    se = _splat_se(st)
    print(f"bvc 3d: rmse {rmse:.4g}, mean error {err.mean():.3g}, median se {np.median(se):.3g}")
    # every point shares the one sample set, so errors are correlated: the mean error is bounded
    # by the standard error of a single point, not by its average over points, and what varies
    # from point to point is small
    assert abs(err.mean()) < 3 * np.median(se) + 0.005
    assert float(np.sqrt(np.mean((err - err.mean()) ** 2))) < 0.03
    # near-boundary points: a direct walk estimate replaces the splat
    xn = (x / np.linalg.norm(x, axis=1, keepdims=True) * 0.97)[:50].astype(np.float32)
    xnt = torch.from_numpy(xn).cuda()
    dn = sc.closest_point(xnt, want_point=False)[0].contiguous()
    evn = E.BvcEvalPoints(pt=xnt, dist_to_absorbing=dn)
    stn = E.BvcStats.zeros(50, 3)
    solver.splat(evn, stn, radius_clamp=1e-3, cutoff_dist_to_absorbing=0.05)
    assert stn.n.sum().item() == 0
    solver.estimate_solution_near_boundary(pde, ws, evn, stn, 0.05, 0.0, 256)
    un = stn.solution().cpu().numpy()
    assert np.sqrt(np.mean((un - _exp_cos(xn)) ** 2)) < 0.05


def test_bvc_solver_2d_dirichlet_with_source(gpu, oracle_lib):
    """2D star domain, Poisson with source: boundary caches plus a domain cache of source samples"""
    torch = gpu
    p = PB.star2d()
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    u_fn = E.function(A.ZG_FN_SIN_PRODUCT, A.ZG_EVAL_VALUE, (1.0, 2.0))
    f_fn = E.function(A.ZG_FN_SIN_PRODUCT, A.ZG_EVAL_SCREENED_SOURCE, (1.0, 2.0))
    pde = E.make_pde(source=f_fn, dirichlet=u_fn)
    ws = E.walk_settings_host(1e-4, 1e-4, 1024)
    lo, hi = p.domain_min, p.domain_max
    # area of the star polygon (shoelace) for the domain density
    Vp = p.positions.astype(np.float64)
    area = 0.5 * abs(np.sum(Vp[p.indices[:, 0], 0] * Vp[p.indices[:, 1], 1] - Vp[p.indices[:, 1], 0] * Vp[p.indices[:, 0], 1]))
    solver = BV.BoundaryValueCachingSolver(sc, BV.UniformBoundarySampler(p.positions, p.indices), None,
                                           BV.UniformDomainSampler(sc, lo, hi, area), seed=5)
    solver.generate_samples(8000, 0, 60000, 1e-3, 0.0, False)
    solver.compute_sample_estimates(pde, ws, 128, 128)
    x = p.meta["points"][::97].astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    d = sc.closest_point(xt, want_point=False)[0].contiguous()
    keep = (d > 0.05).cpu().numpy()
    ev = E.BvcEvalPoints(pt=xt[keep].contiguous(), dist_to_absorbing=d[keep].contiguous())
    st = E.BvcStats.zeros(int(keep.sum()), 2)
    solver.splat(ev, st, radius_clamp=1e-3)
    xs = x[keep].astype(np.float64)
    exact = np.sin(2 * xs[:, 0]) * np.sin(2 * xs[:, 1])
    err = st.solution().cpu().numpy() - exact
    rmse = float(np.sqrt(np.mean(err ** 2)))
    se = _splat_se(st)
    print(f"bvc 2d: {keep.sum()} points, rmse {rmse:.4g}, mean error {err.mean():.3g}, median se {np.median(se):.3g}")
    assert rmse < 4 * np.sqrt(np.mean(se ** 2)) + 0.01
    assert st.n[:, 4].min().item() > 0          # the source cache contributed everywhere


def test_bvc_solver_3d_mixed_neumann(gpu, oracle_lib):
    """L-shaped prism with Dirichlet outer walls and Neumann re-entrant walls and top cap:
    absorbing samples take solution and normal derivative from 3D walk-on-stars gradient walks,
    reflecting samples the solution from 3D walk-on-stars walks and the Neumann data h; the
    splat reproduces u = e^x cos y within its standard errors."""
    torch = gpu
    p = PB.lprism3d(plate=False)
    refl = p.is_reflecting.astype(bool)
    sc = E.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max)
    solver = BV.BoundaryValueCachingSolver(sc, BV.UniformBoundarySampler(p.positions, p.indices[~refl]),
                                           BV.UniformBoundarySampler(p.positions, p.indices[refl]), None, seed=11)
    solver.generate_samples(30000, 20000, 0, 2e-3, 0.0, False)
    assert sum(len(c) for c in solver.caches) >= 50000
    solver.compute_sample_estimates(p.pde, p.settings, 64, 64)
    rng = np.random.default_rng(12)
    x = rng.uniform(-0.85, 0.85, (6000, 3)).astype(np.float32)
    x = x[p.meta["inside"](x, 0.15)][:1500]
    xt = torch.from_numpy(x).cuda()
    dA = sc.closest_point(xt, boundary=0, want_point=False)[0].contiguous()
    dR = sc.closest_point(xt, boundary=1, want_point=False)[0].contiguous()
    ev = E.BvcEvalPoints(pt=xt, dist_to_absorbing=dA, dist_to_reflecting=dR)
    st = E.BvcStats.zeros(len(x), 3)
    solver.splat(ev, st, radius_clamp=1e-3)
    n = st.n.cpu().numpy()
    assert (n[:, 0] > 0).all() and (n[:, 2] > 0).all()     # absorbing and reflecting caches contributed
    err = st.solution().cpu().numpy() - _exp_cos(x)
    rmse = float(np.sqrt(np.mean(err ** 2)))
    se = _splat_se(st)
    print(f"bvc 3d mixed: rmse {rmse:.4g}, mean error {err.mean():.3g}, median se {np.median(se):.3g}")
    assert rmse < 4 * np.sqrt(np.mean(se ** 2)) + 0.01
    assert abs(err.mean()) < 3 * np.median(se) + 0.005


@pytest.mark.parametrize("kreg", [0.0, 0.05])
def test_screened_splat_matches_oracle(gpu, oracle_lib, kreg):
    """3D Yukawa kernels (absorption > 0): the GPU splat equals the oracle's, whose kernels are
    pinned by finite differences and the screened Green's representation."""
    torch = gpu
    O = oracle_lib
    dim, lam = 3, 3.0
    rng = np.random.default_rng(77 + int(kreg * 100))
    S = _random_cache(rng, dim, 3000)
    m = 1200
    x = rng.normal(size=(m, dim)) * 0.5
    dA = rng.uniform(0.0, 1.0, m)
    ost = O.BvcStats(m, dim)
    O.bvc_splat(dim, S, dict(pt=x, dist_abs=dA), ost, radius_clamp=1e-3, kernel_regularization=kreg,
                robin_cutoff=1.0, cutoff_abs=0.05, absorption=lam)
    samples = E.BvcSamples(pt=_t(torch, S["pt"]), normal=_t(torch, S["normal"]), type=_t(torch, S["type"], np.uint8),
                           pdf=_t(torch, S["pdf"]), solution=_t(torch, S["solution"]),
                           normal_derivative=_t(torch, S["normal_derivative"]), contribution=_t(torch, S["contribution"]),
                           robin_coeff=_t(torch, S["robin_coeff"]), normal_aligned=_t(torch, S["normal_aligned"], np.uint8))
    ev = E.BvcEvalPoints(pt=_t(torch, x), dist_to_absorbing=_t(torch, dA))
    st = E.BvcStats.zeros(m, dim)
    E.bvc_splat(dim, samples, ev, st, radius_clamp=1e-3, kernel_regularization=kreg, robin_coeff_cutoff=1.0,
                cutoff_dist_to_absorbing=0.05, absorption=lam)
    assert np.array_equal(st.n.cpu().numpy(), ost.n)
    scale = np.abs(ost.mean).max() + 1
    np.testing.assert_allclose(st.mean.cpu().numpy(), ost.mean, rtol=2e-4, atol=2e-5 * scale)
    gscale = np.abs(ost.grad_mean).max() + 1
    np.testing.assert_allclose(st.grad_mean.cpu().numpy(), ost.grad_mean, rtol=2e-4, atol=2e-5 * gscale)
    # the screened kernels differ from the harmonic ones
    st0 = E.BvcStats.zeros(m, dim)
    E.bvc_splat(dim, samples, ev, st0, radius_clamp=1e-3, kernel_regularization=kreg, robin_coeff_cutoff=1.0,
                cutoff_dist_to_absorbing=0.05)
    assert np.abs(st0.mean.cpu().numpy() - ost.mean).max() > 1e-3 * scale


def test_bvc_solver_3d_screened_dirichlet(gpu):
    """Screened Poisson -Lap u + lambda u = 0 on the unit icosphere with u = exp(a.x), |a|^2 =
    lambda: cached boundary estimates (GPU walk on spheres with the Yukawa ball) splatted with the
    Yukawa free-space kernels (boundary_value_caching.h:386-387,422-423) reproduce u."""
    torch = gpu
    lam = 4.0
    a = np.array([0.8, -1.1, 1.2]); a = a / np.linalg.norm(a) * np.sqrt(lam)
    V, F = PB.icosphere(4)
    V = V.astype(np.float32)
    sc = E.Scene(V, F.astype(np.int32), None, V.min(0) - 1e-3, V.max(0) + 1e-3)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a)),
                     absorption=lam, is_source_constant=True)
    ws = E.walk_settings_host(1e-4, 1e-4, 1024)
    solver = BV.BoundaryValueCachingSolver(sc, BV.UniformBoundarySampler(V, F), None, None, seed=5)
    solver.generate_samples(40000, 0, 0, 2e-3, 0.0, False)
    print("bvc screened: boundary estimates", flush=True)
    solver.compute_sample_estimates(pde, ws, 64, 64)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(2000, 3)); x = x / np.linalg.norm(x, axis=1, keepdims=True) * rng.random((2000, 1)) ** (1 / 3) * 0.7
    xt = torch.from_numpy(x.astype(np.float32)).cuda()
    d = sc.closest_point(xt, want_point=False)[0]
    ev = E.BvcEvalPoints(pt=xt, dist_to_absorbing=d.contiguous())
    st = E.BvcStats.zeros(len(x), 3)
    solver.splat(ev, st, radius_clamp=1e-3)
    u = st.solution().cpu().numpy()
    exact = np.exp(x.astype(np.float64) @ a)
    err = u - exact
    rmse = float(np.sqrt(np.mean(err ** 2)))
    se = _splat_se(st)
    spread = float(np.sqrt(np.mean((err - err.mean()) ** 2)))
    print(f"bvc 3d screened: rmse {rmse:.4g}, mean error {err.mean():.3g}, spread {spread:.3g}, "
          f"median se {np.median(se):.3g}", flush=True)
    # the sample set is shared by all points: a common offset within the standard error, a small spread
    assert abs(err.mean()) < 3 * np.median(se) + 0.005
    assert spread < 0.04
    # the harmonic kernels would not reproduce the screened solution
    solver.absorption = 0.0
    st0 = E.BvcStats.zeros(len(x), 3)
    solver.splat(ev, st0, radius_clamp=1e-3)
    rmse0 = float(np.sqrt(np.mean((st0.solution().cpu().numpy() - exact) ** 2)))
    print(f"bvc 3d screened with harmonic kernels: rmse {rmse0:.4g}", flush=True)
    assert rmse0 > 2 * rmse
