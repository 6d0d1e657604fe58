"""Pins of the oracle's gradient path (estimateSolutionAndGradient, walk_on_spheres.h:416-593)
against mathematics: the bounded PCG32 draw, the Latin hypercube table, the boundary
direction samplers, the ball Green's function gradients (through PDE identities with closed
forms) and the whole estimator against analytic gradients of manufactured solutions."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB


# ------------------------------------------------------------------------------------
# bounded draws and Latin hypercube tables
# ------------------------------------------------------------------------------------
def test_bounded_draw_power_of_two_is_low_bits(oracle_lib):
    O = oracle_lib
    raw = O.pcg32_u32(17, 5, 4000)
    for bound in (1, 2, 8, 1 << 20):
        b = O.pcg32_bounded(17, 5, bound, 4000)
        assert np.array_equal(b, raw & np.uint32(bound - 1))   # 2^32 mod bound = 0: no rejection


def test_bounded_draw_is_unbiased(oracle_lib):
    O = oracle_lib
    # bound = 3 * 2^30: plain `raw % bound` lands below 2^30 half the time; with the rejection of
    # raw < 2^32 mod bound (= 2^30) every residue is equally likely, so a third of the time
    bound = 3 << 30
    b = O.pcg32_bounded(99, 3, bound, 60000).astype(np.int64)
    assert b.max() < bound
    frac = float(np.mean(b < (1 << 30)))
    assert abs(frac - 1.0 / 3.0) < 0.012, frac
    for bound in (3, 7, 1000):
        b = O.pcg32_bounded(5, 9, bound, 70000)
        counts = np.bincount(b, minlength=bound)
        assert len(counts) == bound
        e = 70000 / bound
        chi2 = float(np.sum((counts - e) ** 2 / e))
        assert chi2 < bound - 1 + 6 * np.sqrt(2 * (bound - 1)) + 10, (bound, chi2)


def test_stratified_table_is_a_latin_hypercube(oracle_lib):
    O = oracle_lib
    for dims, n in ((1, 1), (1, 64), (2, 2), (2, 513)):
        t = O.stratified_samples(dims, 3, n, seed=0x1234, stream_offset=7)
        assert t.shape == (3, n, dims)
        assert t.min() >= 0.0 and t.max() <= 1.0 - np.finfo(np.float32).eps
        for k in range(3):
            for j in range(dims):
                strata = np.floor(t[k, :, j].astype(np.float64) * n).astype(np.int64)
                assert np.array_equal(np.sort(strata), np.arange(n)), (dims, n, k, j)
        # deterministic per (seed, offset, table), different across tables / offsets
        assert np.array_equal(t, O.stratified_samples(dims, 3, n, seed=0x1234, stream_offset=7))
        if n > 1:
            assert not np.array_equal(t[0], t[1])
            assert not np.array_equal(t, O.stratified_samples(dims, 3, n, seed=0x1234, stream_offset=8))


def test_stratified_table_known_answer_two_rows(oracle_lib):
    """Two rows, one coordinate, worked out by hand from the raw stream: the diagonal draws
    (0 + f0)/2 and (1 + f1)/2, then the shuffle swaps row 0 with row 0 + draw(2) (the low bit
    of the third raw value, since 2 divides 2^32) and row 1 with itself."""
    O = oracle_lib
    seed, off = 0xBEEF, 3
    for t in range(6):
        init, seq = O.walk_init_state(~seed, t, off), t
        raw = O.pcg32_u32(init, seq, 3)
        f = ((raw[:2] >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
        rows = [np.float32(np.float32(0.0) + f[0]) * np.float32(0.5), np.float32(np.float32(1.0) + f[1]) * np.float32(0.5)]
        if raw[2] & 1:
            rows = rows[::-1]
        got = O.stratified_samples(1, t + 1, 2, seed=seed, stream_offset=off)[t, :, 0]
        assert np.array_equal(got, np.array(rows, np.float32)), (t, got, rows)


# ------------------------------------------------------------------------------------
# boundary directions of the first step
# ------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_direction_map(oracle_lib, dim):
    O = oracle_lib
    rng = np.random.default_rng(1)
    for _ in range(50):
        u = rng.random(dim - 1).astype(np.float32)
        d, pdf = O.gradient_direction(dim, False, u)
        assert abs(np.linalg.norm(d) - 1.0) < 1e-6
        if dim == 2:
            assert np.allclose(d, [np.cos(2 * np.pi * u[0]), np.sin(2 * np.pi * u[0])], atol=2e-6)
            assert abs(pdf - 1 / (2 * np.pi)) < 1e-7
        else:
            assert abs(d[2] - (1 - 2 * u[0])) < 1e-6
            assert abs(np.arctan2(d[1], d[0]) % (2 * np.pi) - (2 * np.pi * u[1]) % (2 * np.pi)) < 1e-4 or np.hypot(d[0], d[1]) < 1e-3
            assert abs(pdf - 1 / (4 * np.pi)) < 1e-7


@pytest.mark.parametrize("dim", [2, 3])
def test_cosine_direction_density(oracle_lib, dim):
    """Cosine-weighted directions about +-n: unit length, symmetric in the sign of d.n, pdf equal
    to half the hemisphere density of |d.n|, and the moment E|d.n| of the cosine law
    (pi/4 in 2D, 2/3 in 3D) -- a wrong frame or a dropped factor fails one of them."""
    O = oracle_lib
    rng = np.random.default_rng(2)
    cos_abs, cos_signed = [], []
    for i in range(4000):
        n = rng.normal(size=dim).astype(np.float32)
        n /= np.linalg.norm(n)
        u = rng.random(dim).astype(np.float32)        # dim-1 direction uniforms + the sign coin
        d, pdf = O.gradient_direction(dim, True, u, n)
        c = float(np.dot(d, n))
        assert abs(np.linalg.norm(d) - 1.0) < 2e-5
        dens = abs(c) / 2.0 if dim == 2 else abs(c) / np.pi
        assert abs(pdf - 0.5 * dens) < 2e-5
        cos_abs.append(abs(c)); cos_signed.append(c)
    m = np.mean(cos_abs)
    expect = np.pi / 4 if dim == 2 else 2.0 / 3.0
    se = np.std(cos_abs) / np.sqrt(len(cos_abs))
    assert abs(m - expect) < 4 * se + 1e-4, (m, expect)
    assert abs(np.mean(cos_signed)) < 4 * np.std(cos_signed) / np.sqrt(len(cos_signed))


# ------------------------------------------------------------------------------------
# Green's function gradients through PDE identities
# ------------------------------------------------------------------------------------
def _i1(z):
    """modified spherical Bessel function i_1(z) = (z cosh z - sinh z) / z^2"""
    return (z * np.cosh(z) - np.sinh(z)) / (z * z)


@pytest.mark.parametrize("dim,lam", [(2, 0.0), (3, 0.0), (3, 4.0), (3, 30.0)])
def test_greens_gradient_reproduces_linear_source_solution(oracle_lib, dim, lam):
    """grad u(c) = int_B f(y) (y - c) gradientNorm(|y - c|) dy for -Lap u + lam u = f in the ball,
    u = 0 on the sphere.  f = y_0 has the closed-form solutions y_0 (R^2 - r^2) / (2 dim + 4)
    (harmonic) and y_0 / lam - R y_0 i_1(mu r) / (lam r i_1(mu R)) (Yukawa), so
    du/dy_0(c) = R^2/8 (2D), R^2/10 (3D), 1/lam - R mu / (3 lam i_1(mu R))."""
    O = oracle_lib
    R = 0.7
    x, w = np.polynomial.legendre.leggauss(80)
    r = 0.5 * R * (x + 1)
    w = 0.5 * R * w
    gn = np.array([O.greens_ball_gradient(dim, R, lam, float(ri))["gradient_norm"] for ri in r])
    if dim == 2:
        integral = np.pi * np.sum(w * r ** 3 * gn)
        exact = R * R / 8
    else:
        integral = 4 * np.pi / 3 * np.sum(w * r ** 4 * gn)
        if lam == 0.0:
            exact = R * R / 10
        else:
            mu = np.sqrt(lam)
            exact = 1 / lam - R * mu / (3 * lam * _i1(mu * R))
    assert abs(integral - exact) < 2e-4 * abs(exact), (integral, exact)


@pytest.mark.parametrize("dim,lam", [(2, 0.0), (3, 0.0), (3, 4.0), (3, 30.0)])
def test_poisson_kernel_gradient_reproduces_linear_boundary_data(oracle_lib, dim, lam):
    """grad u(c) = oint u(y) poissonKernelGradient(y) d(omega) (solid angle) for -Lap u + lam u = 0
    with u = y_0 on the sphere: 1 (harmonic) and R mu / (3 i_1(mu R)) (Yukawa)."""
    O = oracle_lib
    R = 0.45
    k = O.greens_ball_gradient(dim, R, lam, 0.1)["poisson_kernel_gradient"] / R   # pkg(y) = k (y - c)
    integral = k * R * R * (np.pi if dim == 2 else 4 * np.pi / 3)
    exact = 1.0 if lam == 0.0 else R * np.sqrt(lam) / (3 * _i1(np.sqrt(lam) * R))
    assert abs(integral - exact) < 1e-5 * max(1.0, abs(exact)), (integral, exact)


# ------------------------------------------------------------------------------------
# the estimator against analytic gradients
# ------------------------------------------------------------------------------------
def _gz(st, g_exact):
    """per-component z-scores of the gradient means and the pooled z over all points"""
    se2 = st.grad_variance() / np.maximum(st.n_grad, 1)[:, None]
    ok = se2 > 0
    z = np.where(ok, (st.grad_mean - g_exact) / np.sqrt(np.where(ok, se2, 1)), 0.0)
    pooled = np.sum(st.grad_mean - g_exact, axis=0) / np.sqrt(np.sum(se2, axis=0))
    return z[ok], pooled


def _sphere_scene(O, level=3):
    V, F = PB.icosphere(level)
    V = V.astype(np.float32)
    return V, F.astype(np.int32), O.Scene(V, F.astype(np.int32), None, V.min(0) - 1e-3, V.max(0) + 1e-3)


def _ball_points(n, rmax, seed, dim=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(4 * n, dim))
    p = p / np.linalg.norm(p, axis=1, keepdims=True) * (rng.random((4 * n, 1)) ** (1 / dim)) * rmax
    return p[:n].astype(np.float32)


def _exp_cos_grad(x, a=1.0):
    e = np.exp(a * x[:, 0])
    g = np.zeros_like(x, dtype=np.float64)
    g[:, 0] = a * e * np.cos(a * x[:, 1])
    g[:, 1] = -a * e * np.sin(a * x[:, 1])
    return g


def _settings(**kw):
    ws = E.walk_settings_host(1e-4, 1e-4, 1024)
    for k, v in kw.items():
        setattr(ws, k, v)
    return ws


VARIANTS = [dict(), dict(use_gradient_control_variates=0), dict(use_gradient_antithetic_variates=0),
            dict(use_cosine_sampling_for_derivatives=1)]


@pytest.mark.parametrize("variant", range(len(VARIANTS)))
def test_harmonic_3d_gradient_matches_exact(oracle_lib, variant):
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(160, 0.75, 3)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    ws = _settings(**VARIANTS[variant])
    st = O.GradStats(len(pts), 3)
    direction = np.array([0.6, -0.8, 0.0], np.float32)
    fsr = np.zeros(len(pts), np.float32)
    O.solve_grad(sc, pde, ws, pts, d, 512, st, direction=direction, first_sphere_radius=fsr)
    g = _exp_cos_grad(pts)
    z, pooled = _gz(st, g)
    assert np.all(np.abs(pooled) < 4.0), pooled
    # the halves of an antithetic pair are negatively correlated, so the per-estimate variance
    # overstates the error of the mean (z narrower than N(0,1)); independent walks give std ~ 1
    lo = 0.8 if not ws.use_gradient_antithetic_variates else 0.2
    assert lo < z.std() < 1.2, z.std()
    assert np.array_equal(st.n_grad, st.n) and st.n.min() > 0
    assert np.allclose(fsr, np.float32(0.99) * d)
    assert np.all(st.first_source == 0.0)
    # directional derivative: total / n estimates the derivative along `direction`
    dd = st.derivative_mean()
    se = np.sqrt(np.sum(st.grad_variance() * direction.astype(np.float64) ** 2, axis=1) / st.n)
    zd = np.sum(dd - g @ direction.astype(np.float64)) / np.sqrt(np.sum(se ** 2))
    assert abs(zd) < 4.5, zd
    # the solution estimate of the same walks stays unbiased
    u = np.exp(pts[:, 0].astype(np.float64)) * np.cos(pts[:, 1].astype(np.float64))
    zs = np.sum(st.mean - u) / np.sqrt(np.sum(st.variance() / st.n))
    assert abs(zs) < 4.0


def test_gradient_control_variates_and_antithetic_pairs_reduce_variance(oracle_lib):
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(120, 0.6, 4)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (2.0, 1.0)))
    var = {}
    for name, kw in (("plain", dict(use_gradient_control_variates=0, use_gradient_antithetic_variates=0)),
                     ("cv", dict(use_gradient_antithetic_variates=0)), ("both", dict())):
        st = O.GradStats(len(pts), 3)
        O.solve_grad(sc, pde, _settings(**kw), pts, d, 512, st)
        var[name] = float(np.mean(st.grad_variance()))
    # u ~ 2 e^x is large against |grad u| R: subtracting the running mean removes most variance
    assert var["cv"] < 0.3 * var["plain"], var
    assert var["both"] < var["cv"] * 1.5, var


def test_staged_control_variates_match_per_pair(oracle_lib):
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(150, 0.7, 5)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (2.0, 1.0)))
    g = 2.0 * _exp_cos_grad(pts)
    out = {}
    for mode in (O.CV_PER_PAIR, O.CV_STAGED):
        st = O.GradStats(len(pts), 3)
        O.solve_grad(sc, pde, _settings(), pts, d, 512, st, cv_mode=mode)
        z, pooled = _gz(st, g)
        assert np.all(np.abs(pooled) < 4.0), (mode, pooled)
        out[mode] = float(np.mean(st.grad_variance()))
    # stages refresh the control variates at walks 1, 2, 4, 8, ...: nearly the same variance
    assert out[O.CV_STAGED] < 1.3 * out[O.CV_PER_PAIR], out


def test_reference_rng_order_matches_exact(oracle_lib):
    """rng mode 0 follows the reference's generator order (table, then walks; the antithetic
    half replays the generator saved after the first step)"""
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(150, 0.75, 6)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    st = O.GradStats(len(pts), 3)
    O.solve_grad(sc, pde, _settings(use_cosine_sampling_for_derivatives=1), pts, d, 512, st,
                 rng_mode=O.RNG_PER_POINT)
    z, pooled = _gz(st, _exp_cos_grad(pts))
    assert np.all(np.abs(pooled) < 4.0) and 0.2 < z.std() < 1.2


def test_poisson_source_gradient_matches_exact(oracle_lib):
    """-Lap u = 3 k^2 u with u = prod sin(k x_i): the source term's gradient direction
    (gradient(r, y) / (pdf norm)) carries part of grad u"""
    O = oracle_lib
    p = PB.sphere3d()
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(150, 0.7, 7)
    d, *_ = sc.closest_point(pts)
    st = O.GradStats(len(pts), 3)
    O.solve_grad(sc, p.pde, _settings(), pts, d, 512, st)
    k = 2.0
    x = pts.astype(np.float64)
    s, c = np.sin(k * x), np.cos(k * x)
    g = k * np.stack([c[:, 0] * s[:, 1] * s[:, 2], s[:, 0] * c[:, 1] * s[:, 2], s[:, 0] * s[:, 1] * c[:, 2]], 1)
    z, pooled = _gz(st, g)
    assert np.all(np.abs(pooled) < 4.0) and 0.2 < z.std() < 1.2, (pooled, z.std())
    assert np.all(st.first_source != 0.0)
    # without the source term's gradient the estimate is biased: detect it
    st2 = O.GradStats(len(pts), 3)
    O.solve_grad(sc, p.pde, _settings(ignore_source_contribution=1), pts, d, 512, st2)
    _, pooled2 = _gz(st2, g)
    assert np.max(np.abs(pooled2)) > 6.0


def test_screened_gradient_matches_exact(oracle_lib):
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    sigma = 6.0
    a = np.sqrt(sigma) * np.array([0.0, 0.6, 0.8])
    u = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a))
    pde = E.make_pde(source=E.zero(), dirichlet=u, absorption=sigma, is_source_constant=True)
    pts = _ball_points(150, 0.7, 8)
    d, *_ = sc.closest_point(pts)
    st = O.GradStats(len(pts), 3)
    O.solve_grad(sc, pde, _settings(), pts, d, 512, st)
    uu = np.exp(pts.astype(np.float64) @ a)
    g = uu[:, None] * a[None, :]
    z, pooled = _gz(st, g)
    assert np.all(np.abs(pooled) < 4.0) and 0.2 < z.std() < 1.2, (pooled, z.std())


def test_constant_source_path_agrees_with_sampled_source(oracle_lib):
    """is_source_constant evaluates norm * f(x0) with no source gradient; sampling the same
    constant source gives the same expectation"""
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(100, 0.6, 9)
    d, *_ = sc.closest_point(pts)
    res = []
    for const in (True, False):
        pde = E.make_pde(source=E.constant(3.0), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)),
                         is_source_constant=const)
        st = O.GradStats(len(pts), 3)
        O.solve_grad(sc, pde, _settings(), pts, d, 512, st)
        res.append(st)
    a, b = res
    se2 = a.grad_variance() / a.n_grad[:, None] + b.grad_variance() / b.n_grad[:, None]
    pooled = np.sum(a.grad_mean - b.grad_mean, axis=0) / np.sqrt(np.sum(se2, axis=0))
    assert np.all(np.abs(pooled) < 4.0), pooled
    # first-step source term: norm * f = R0^2 / 6 * 3 exactly on the constant path
    R0 = (np.float32(0.99) * d).astype(np.float64)
    assert np.allclose(a.first_source / a.n, R0 ** 2 / 6 * 3, rtol=1e-5)


def test_harmonic_2d_gradient_matches_exact(oracle_lib):
    O = oracle_lib
    p = PB.star2d()
    pts = p.meta["points"][::80]
    sc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    for kw in (dict(), dict(use_cosine_sampling_for_derivatives=1)):
        st = O.GradStats(len(pts), 2)
        O.solve_grad(sc, pde, _settings(**kw), pts, d, 512, st)
        z, pooled = _gz(st, _exp_cos_grad(pts))
        assert np.all(np.abs(pooled) < 4.0) and 0.2 < z.std() < 1.2, (kw, pooled, z.std())


def test_quantities_skip_and_solution(oracle_lib):
    """Skip leaves the statistics alone; Solution points are the plain solve, bit for bit"""
    O = oracle_lib
    _, _, sc = _sphere_scene(O)
    pts = _ball_points(60, 0.7, 10)
    d, *_ = sc.closest_point(pts)
    pde = E.make_pde(source=E.zero(), dirichlet=E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0)))
    ws = _settings()
    q = np.array([0, 1, 2] * 20, np.uint8)
    st = O.GradStats(len(pts), 3)
    O.solve_grad(sc, pde, ws, pts, d, 64, st, quantity=q)
    ref = O.Stats(len(pts))
    O.solve(sc, pde, ws, pts, d, 64, ref)
    s0 = q == 0
    assert np.array_equal(st.mean[s0], ref.mean[s0]) and np.array_equal(st.n[s0], ref.n[s0])
    assert np.all(st.n[q == 2] == 0) and np.all(st.n_grad[q != 1] == 0)
    assert np.all(st.n_grad[q == 1] == 64)


def _cubic_grad(x):
    x = x.astype(np.float64)
    return np.stack([3 * x[:, 0] ** 2 - 3 * x[:, 1] ** 2, -6 * x[:, 0] * x[:, 1]], 1)


def test_walk_on_stars_gradient_matches_exact(oracle_lib):
    """walk_on_stars.h:664-848: the first ball avoids both boundaries (radius
    0.99 min(d_absorbing, d_reflecting)), then a walk-on-stars walk with Neumann reflections"""
    O = oracle_lib
    p = PB.mixed2d(n_segments=1024, n_slits=8)
    sc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    V, I = p.meta["outer"]
    cand = PB.sample_box(p.domain_min, p.domain_max, 1500, 19)
    pts = cand[PB.points_in_polygon(cand, V, I)][:160]
    d, *_ = sc.closest_point(pts, boundary=0)
    dr, *_ = sc.closest_point(pts, boundary=1)
    g = _cubic_grad(pts)
    for kw, rng_mode in ((dict(), O.RNG_PER_WALK), (dict(use_gradient_control_variates=0), O.RNG_PER_POINT)):
        ws = E.walk_settings_host(1e-3, 1e-3, 16384, True)
        for k, v in kw.items():
            setattr(ws, k, v)
        st = O.GradStats(len(pts), 2)
        fsr = np.zeros(len(pts), np.float32)
        O.solve_grad(sc, p.pde, ws, pts, d, 256, st, wost=True, rng_mode=rng_mode, first_sphere_radius=fsr,
                     dist_reflecting=dr if rng_mode == O.RNG_PER_WALK else None)
        z, pooled = _gz(st, g)
        assert np.all(np.abs(pooled) < 4.0) and 0.2 < z.std() < 1.2, (kw, pooled, z.std())
        assert np.allclose(fsr, np.float32(0.99) * np.minimum(d, dr), rtol=1e-6)
        assert np.array_equal(st.n_grad, st.n) and np.all(st.n == 256)
