"""Gradient estimation on the GPU (estimateSolutionAndGradient, walk_on_spheres.h:416-593):
the stratified tables bit for bit against the oracle, gradient means against the oracle
(4 standard errors) and against analytic gradients of manufactured solutions, and the
bookkeeping of estimation quantities."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB

pytestmark = pytest.mark.gpu


def _settings(**kw):
    ws = E.walk_settings_host(1e-4, 1e-4, 1024)
    for k, v in kw.items():
        setattr(ws, k, v)
    return ws


def _ball_points(n, rmax, seed, dim=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(4 * n, dim))
    p = p / np.linalg.norm(p, axis=1, keepdims=True) * (rng.random((4 * n, 1)) ** (1 / dim)) * rmax
    return p[:n].astype(np.float32)


def _exp_cos_grad(x, a=1.0, amp=1.0):
    x = x.astype(np.float64)
    e = amp * np.exp(a * x[:, 0])
    g = np.zeros_like(x)
    g[:, 0] = a * e * np.cos(a * x[:, 1])
    g[:, 1] = -a * e * np.sin(a * x[:, 1])
    return g


def _points(torch, pts, dist, n_walks, quantity=1, direction=None):
    n = len(pts)
    q = np.broadcast_to(np.asarray(quantity, np.uint8), (n,)).copy()
    return E.DevicePoints(
        pt=torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda(),
        dist_to_absorbing=torch.from_numpy(np.ascontiguousarray(dist, np.float32)).cuda(),
        n_walks=torch.full((n,), n_walks, dtype=torch.int32, device="cuda"), n_walks_max=n_walks,
        first_sphere_radius=torch.zeros(n, dtype=torch.float32, device="cuda"),
        estimation_quantity=torch.from_numpy(q).cuda(),
        direction_for_derivative=None if direction is None else torch.from_numpy(
            np.ascontiguousarray(np.broadcast_to(np.asarray(direction, np.float32), pts.shape))).cuda())


class _Host:
    def __init__(self, st):
        self.n = st.n.cpu().numpy()
        self.mean = st.mean.cpu().numpy()
        self.var = st.variance().cpu().numpy()
        self.n_grad = st.n_grad.cpu().numpy()
        self.grad_mean = st.grad_mean.cpu().numpy()
        self.grad_var = st.grad_variance().cpu().numpy()
        self.first_source = st.first_source.cpu().numpy()
        self.derivative = st.derivative.cpu().numpy()


def _gpu_grad(torch, scene, pde, ws, pts, dist, n_walks, seed=1, stream_offset=0, stats=None, quantity=1,
              direction=None, opts=None):
    dp = _points(torch, pts, dist, n_walks, quantity, direction)
    st = stats or E.DeviceStats.zeros(len(pts), dim=pts.shape[1])
    o = opts or E.SolveOptions(seed=seed, stream_offset=stream_offset)
    rep = E.wos_solve(scene, pde, ws, dp, st, o)
    return st, rep, dp


def _pooled(mean, var, n, exact):
    se2 = var / np.maximum(n, 1)[:, None]
    return np.sum(mean - exact, axis=0) / np.sqrt(np.sum(se2, axis=0))


def _vs_oracle(h, ost, label):
    """per-point, per-component 4-SE comparison of GPU and oracle gradient means"""
    zs, oks = [], []
    for k in range(h.grad_mean.shape[1]):
        se2 = h.grad_var[:, k] / h.n_grad + ost.grad_variance()[:, k] / ost.n_grad
        ok = se2 > 0
        z = np.zeros(len(se2))
        z[ok] = (h.grad_mean[ok, k] - ost.grad_mean[ok, k]) / np.sqrt(se2[ok])
        zs.append(z); oks.append(ok)
    z = np.concatenate(zs); ok = np.concatenate(oks)
    zz = z[ok]
    assert int((np.abs(zz) > 4.0).sum()) <= 3, f"{label}: {(np.abs(zz) > 4).sum()} beyond 4 SE"
    assert abs(zz.mean()) < 5.0 / np.sqrt(len(zz)) + 0.05, f"{label}: z mean {zz.mean():.3f}"
    return zz


@pytest.fixture(scope="module")
def sphere(oracle_lib):
    V, F = PB.icosphere(3)
    V = V.astype(np.float32)
    F = F.astype(np.int32)
    lo, hi = V.min(0) - 1e-3, V.max(0) + 1e-3
    osc = oracle_lib.Scene(V, F, None, lo, hi)
    return V, F, lo, hi, osc


def _gpu_scene(V, F, lo, hi, width=None):
    return E.Scene(V, F, None, lo, hi) if width is None else E.Scene(V, F, None, lo, hi, bvh_width=width)


# ------------------------------------------------------------------------------------
def test_stratified_tables_bit_exact(gpu, oracle_lib):
    torch = gpu
    O = oracle_lib
    for dims, ns, nt in ((1, 2, 5), (1, 1000, 40), (2, 1, 3), (2, 513, 64), (2, 2048, 7)):
        for seed, off in ((1, 0), (0xFEEDFACECAFEBEEF, 123456)):
            g = E.stratified_samples(dims, nt, ns, seed, off).cpu().numpy()
            o = O.stratified_samples(dims, nt, ns, seed, off)
            assert np.array_equal(g, o), (dims, ns, nt, seed, off)
    with pytest.raises(A.ZombieError):
        E.stratified_samples(3, 1, 4, 1)


@pytest.mark.parametrize("variant", ["default", "no_cv", "no_antithetic", "cosine"])
def test_harmonic_3d_gradient_vs_oracle_and_exact(gpu, oracle_lib, sphere, variant):
    torch = gpu
    O = oracle_lib
    V, F, lo, hi, osc = sphere
    kw = dict(default={}, no_cv=dict(use_gradient_control_variates=0),
              no_antithetic=dict(use_gradient_antithetic_variates=0),
              cosine=dict(use_cosine_sampling_for_derivatives=1))[variant]
    ws = _settings(**kw)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    pts = _ball_points(400, 0.75, 21)
    d, *_ = osc.closest_point(pts)
    direction = np.array([0.6, -0.8, 0.0], np.float32)
    st, rep, dp = _gpu_grad(torch, _gpu_scene(V, F, lo, hi), pde, ws, pts, d, 512, seed=3, direction=direction)
    h = _Host(st)
    assert np.all(h.n_grad == h.n) and np.all(h.n == 512)
    assert rep.total_walks == 512 * len(pts)
    assert np.allclose(dp.first_sphere_radius.cpu().numpy(), np.float32(0.99) * d)
    g = _exp_cos_grad(pts)
    pooled = _pooled(h.grad_mean, h.grad_var, h.n_grad, g)
    assert np.all(np.abs(pooled) < 4.0), pooled
    ost = O.GradStats(len(pts), 3)
    O.solve_grad(osc, pde, ws, pts, d, 512, ost, direction=direction, seed=77, cv_mode=O.CV_STAGED)
    _vs_oracle(h, ost, variant)
    # the staged control variates reduce the GPU's variance as much as the oracle's (not with
    # cosine sampling: its throughput P / pdf ~ 1 / |cos| has infinite variance, so sample
    # variances do not settle)
    if variant != "cosine":
        ratio = np.mean(h.grad_var) / np.mean(ost.grad_variance())
        assert 0.75 < ratio < 1.33, ratio
    # directional derivative along `direction`
    dd = h.derivative / h.n
    se = np.sqrt(np.sum(h.grad_var * direction.astype(np.float64) ** 2, axis=1) / h.n)
    zd = np.sum(dd - g @ direction.astype(np.float64)) / np.sqrt(np.sum(se ** 2))
    assert abs(zd) < 4.5, zd
    # solution statistics of the same walks
    u = np.exp(pts[:, 0].astype(np.float64)) * np.cos(pts[:, 1].astype(np.float64))
    zs = np.sum(h.mean - u) / np.sqrt(np.sum(h.var / h.n))
    assert abs(zs) < 4.0, zs


def test_poisson_source_gradient(gpu, oracle_lib, sphere):
    torch = gpu
    O = oracle_lib
    V, F, lo, hi, osc = sphere
    p = PB.sphere3d()
    pts = _ball_points(400, 0.7, 22)
    d, *_ = osc.closest_point(pts)
    st, _, _ = _gpu_grad(torch, _gpu_scene(V, F, lo, hi), p.pde, _settings(), pts, d, 512, seed=5)
    h = _Host(st)
    k = 2.0
    x = pts.astype(np.float64)
    s, c = np.sin(k * x), np.cos(k * x)
    g = k * np.stack([c[:, 0] * s[:, 1] * s[:, 2], s[:, 0] * c[:, 1] * s[:, 2], s[:, 0] * s[:, 1] * c[:, 2]], 1)
    assert np.all(np.abs(_pooled(h.grad_mean, h.grad_var, h.n_grad, g)) < 4.0)
    ost = O.GradStats(len(pts), 3)
    O.solve_grad(osc, p.pde, _settings(), pts, d, 512, ost, seed=9, cv_mode=O.CV_STAGED)
    _vs_oracle(h, ost, "poisson")
    assert np.all(h.first_source != 0.0)


def test_screened_gradient(gpu, oracle_lib, sphere):
    torch = gpu
    V, F, lo, hi, osc = sphere
    sigma = 6.0
    a = np.sqrt(sigma) * np.array([0.0, 0.6, 0.8])
    u = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a))
    pde = E.make_pde(source=E.zero(), dirichlet=u, absorption=sigma, is_source_constant=True)
    pts = _ball_points(400, 0.7, 23)
    d, *_ = osc.closest_point(pts)
    for kw in ({}, dict(use_cosine_sampling_for_derivatives=1)):
        st, _, _ = _gpu_grad(torch, _gpu_scene(V, F, lo, hi), pde, _settings(**kw), pts, d, 512, seed=7)
        h = _Host(st)
        g = np.exp(pts.astype(np.float64) @ a)[:, None] * a[None, :]
        assert np.all(np.abs(_pooled(h.grad_mean, h.grad_var, h.n_grad, g)) < 4.0), kw


def test_harmonic_2d_gradient(gpu, oracle_lib):
    torch = gpu
    O = oracle_lib
    p = PB.star2d()
    pts = p.meta["points"][::30]
    osc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = osc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    for kw in ({}, dict(use_cosine_sampling_for_derivatives=1)):
        ws = _settings(**kw)
        st, _, _ = _gpu_grad(torch, sc, pde, ws, pts, d, 512, seed=11)
        h = _Host(st)
        assert np.all(np.abs(_pooled(h.grad_mean, h.grad_var, h.n_grad, _exp_cos_grad(pts))) < 4.0), kw
        ost = O.GradStats(len(pts), 2)
        O.solve_grad(osc, pde, ws, pts, d, 512, ost, seed=13, cv_mode=O.CV_STAGED)
        _vs_oracle(h, ost, f"2d {kw}")


def test_gradient_with_splitting_and_bvh_width(gpu, oracle_lib, sphere):
    torch = gpu
    V, F, lo, hi, osc = sphere
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    pts = _ball_points(300, 0.7, 24)
    d, *_ = osc.closest_point(pts)
    g = _exp_cos_grad(pts)
    ws = _settings(splitting_threshold=0.6, russian_roulette_threshold=0.3)
    for width in (4, 8):
        st, _, _ = _gpu_grad(torch, _gpu_scene(V, F, lo, hi, width), pde, ws, pts, d, 256, seed=15)
        h = _Host(st)
        assert st.splits.sum().item() > 0
        assert np.all(np.abs(_pooled(h.grad_mean, h.grad_var, h.n_grad, g)) < 4.0), width


def test_quantities_and_progressive_calls(gpu, oracle_lib, sphere):
    torch = gpu
    V, F, lo, hi, osc = sphere
    sc = _gpu_scene(V, F, lo, hi)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    ws = _settings()
    pts = _ball_points(300, 0.7, 25)
    d, *_ = osc.closest_point(pts)
    q = np.array([0, 1, 2] * 100, np.uint8)
    st, rep, _ = _gpu_grad(torch, sc, pde, ws, pts, d, 64, seed=17, quantity=q)
    h = _Host(st)
    assert np.all(h.n[q == 2] == 0) and np.all(h.n_grad[q != 1] == 0)
    assert np.all(h.n[q == 0] == 64) and np.all(h.n_grad[q == 1] == 64)
    assert rep.total_walks == 64 * 200
    # solution points: the plain solve's walks (same seed and streams)
    plain = E.DeviceStats.zeros(len(pts))
    dp = _points(torch, pts, d, 64, quantity=0)
    dp.estimation_quantity = None
    E.wos_solve(sc, pde, ws, dp, plain, E.SolveOptions(seed=17))
    pm = plain.mean.cpu().numpy()
    assert np.allclose(h.mean[q == 0], pm[q == 0], rtol=1e-9, atol=1e-12)
    # a second call accumulates: counts double, first-step totals add up
    st2, _, _ = _gpu_grad(torch, sc, pde, ws, pts, d, 64, seed=17, stream_offset=64, quantity=q, stats=st)
    h2 = _Host(st2)
    assert np.all(h2.n_grad[q == 1] == 128) and np.all(h2.n[q == 0] == 128)
    g = _exp_cos_grad(pts)[q == 1]
    assert np.all(np.abs(_pooled(h2.grad_mean[q == 1], h2.grad_var[q == 1], h2.n_grad[q == 1], g)) < 4.0)


def test_walk_on_stars_gradient(gpu, oracle_lib):
    """walk_on_stars.h:664-848 on the mixed Dirichlet/Neumann problem with interior slits:
    first ball 0.99 min(d_absorbing, d_reflecting), then walk-on-stars steps"""
    torch = gpu
    O = oracle_lib
    p = PB.mixed2d(n_segments=1024, n_slits=8)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    V, I = p.meta["outer"]
    cand = PB.sample_box(p.domain_min, p.domain_max, 3000, 29)
    pts = cand[PB.points_in_polygon(cand, V, I)][:400]
    d, *_ = osc.closest_point(pts, boundary=0)
    dr, *_ = osc.closest_point(pts, boundary=1)
    x = pts.astype(np.float64)
    g = np.stack([3 * x[:, 0] ** 2 - 3 * x[:, 1] ** 2, -6 * x[:, 0] * x[:, 1]], 1)
    sc = E.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    ws = p.settings
    for give_dr in (True, False):
        dp = _points(torch, pts, d, 256)
        if give_dr:
            dp.dist_to_reflecting = torch.from_numpy(dr.astype(np.float32)).cuda()
        st = E.DeviceStats.zeros(len(pts), dim=2)
        rep = E.wost_solve(sc, p.pde, ws, dp, st, E.SolveOptions(seed=31 + give_dr))
        h = _Host(st)
        # rare escapes through the slit ends (numerical leaks) are dropped, as in the reference
        assert np.array_equal(h.n_grad, h.n) and h.n.min() >= 250
        assert rep.completion_codes[3] <= 1e-4 * rep.total_walks + 2
        assert np.allclose(dp.first_sphere_radius.cpu().numpy(), np.float32(0.99) * np.minimum(d, dr), rtol=1e-5)
        assert np.all(np.abs(_pooled(h.grad_mean, h.grad_var, h.n_grad, g)) < 4.0), give_dr
    ost = O.GradStats(len(pts), 2)
    O.solve_grad(osc, p.pde, ws, pts, d, 256, ost, wost=True, seed=41, cv_mode=O.CV_STAGED, dist_reflecting=dr)
    _vs_oracle(h, ost, "wost")


def test_gradient_argument_errors(gpu, sphere):
    torch = gpu
    V, F, lo, hi, osc = sphere
    sc = _gpu_scene(V, F, lo, hi)
    pde = E.make_pde(source=E.zero(), dirichlet=E.constant(1.0))
    pts = _ball_points(10, 0.5, 26)
    d, *_ = osc.closest_point(pts)
    # gradient points need the gradient statistics (refused on the host, before the C-ABI)
    with pytest.raises((A.ZombieError, ValueError)):
        _gpu_grad(torch, sc, pde, _settings(), pts, d, 8, stats=E.DeviceStats.zeros(len(pts)))
    # unknown quantity value
    with pytest.raises(A.ZombieError):
        _gpu_grad(torch, sc, pde, _settings(), pts, d, 8, quantity=7)
    # all skipped: nothing changes
    st, rep, _ = _gpu_grad(torch, sc, pde, _settings(), pts, d, 8, quantity=2)
    assert st.n.sum().item() == 0 and rep.total_walks == 0
