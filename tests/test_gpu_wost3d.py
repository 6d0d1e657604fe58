"""Walk on stars on the GPU in 3D (Neumann and Robin reflecting triangle meshes, harmonic and
screened): the
silhouette-edge star radius (with the triangles' Robin bound), line of sight and
reflecting-boundary sampler against the oracle and closed forms, and the estimator (solution
and gradient) against the oracle (4 standard errors) and the manufactured solution
u = e^x cos y on an L-shaped prism with reflecting re-entrant walls, top cap and an interior
double-sided plate."""
import math

import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB

from helpers import check_four_se, z_scores

pytestmark = pytest.mark.gpu


def _t(torch, a, dtype=np.float32):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


def _interior(p, n, seed, margin=0.02):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.95, 0.95, (4 * n, 3)).astype(np.float32)
    return x[p.meta["inside"](x, margin)][:n]


@pytest.fixture(scope="module", params=[False, True], ids=["closed", "plate"])
def prism(request, oracle_lib):
    p = PB.lprism3d(plate=request.param)
    osc = oracle_lib.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max,
                           solve_double_sided=p.solve_double_sided)
    pts = _interior(p, 600, 7)
    d = osc.closest_point(pts, boundary=0)[0]
    return p, osc, pts, d


def _scene(p, width):
    return E.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, bvh_width=width,
                   leaf_size=4, solve_double_sided=p.solve_double_sided)


def _seg_dist(x, a, b):
    u = b - a
    t = np.clip(((x - a) * u).sum(-1) / (u * u).sum(-1), 0.0, 1.0)
    return np.linalg.norm(x - (a + u * t[..., None]), axis=-1)


@pytest.mark.parametrize("width", [4, 8])
def test_star_radius_matches_oracle(gpu, prism, width):
    torch = gpu
    p, osc, pts, d = prism
    sc = _scene(p, width)
    assert sc.info.n_silhouette_vertices == osc.stats()["silhouette_candidates"]
    eps = p.settings.epsilon_shell_for_reflecting_boundary
    mn = np.full(len(pts), eps, np.float32)
    g = sc.star_radius(_t(torch, pts), _t(torch, mn), _t(torch, d)).cpu().numpy()
    o = osc.star_radius(pts, mn, d)
    assert np.isclose(g, o, rtol=1e-5, atol=1e-6).mean() > 0.995
    assert (g > o * (1 + 1e-4) + 1e-5).sum() == 0
    assert np.mean(g < d - 1e-6) > 0.1
    fl = torch.ones(len(pts), dtype=torch.uint8, device="cuda")
    gf = sc.star_radius(_t(torch, pts), _t(torch, mn), _t(torch, d), flip=fl).cpu().numpy()
    of = osc.star_radius(pts, mn, d, flip=np.ones(len(pts), np.uint8))
    assert np.isclose(gf, of, rtol=1e-5, atol=1e-6).mean() > 0.995
    # no silhouette edge lies strictly inside a radius above the clamp
    for i in range(0, len(pts), 8):
        e, s = osc.silhouette_candidates(pts[i])
        if s.any():
            ds = _seg_dist(pts[i].astype(np.float64), e[s, 0].astype(np.float64), e[s, 1].astype(np.float64))
            assert ds.min() >= g[i] * (1 - 1e-5) - 1e-6 or g[i] <= eps * (1 + 1e-6)


@pytest.mark.parametrize("width", [4, 8])
def test_line_of_sight_matches_oracle(gpu, prism, width):
    torch = gpu
    p, osc, _, _ = prism
    sc = _scene(p, width)
    rng = np.random.default_rng(12)
    a = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    b = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    g = sc.line_of_sight(_t(torch, a), _t(torch, b)).cpu().numpy().astype(bool)
    o = np.array([osc.line_of_sight(u, v) for u, v in zip(a, b)])
    assert (g != o).sum() <= 2 and (~o).sum() > 100


def _tri_support_area(V, T, x, R):
    A_, B_, C_ = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    lo = np.minimum(np.minimum(A_, B_), C_); hi = np.maximum(np.maximum(A_, B_), C_)
    dd = np.maximum(np.maximum(lo - x, x - hi), 0)
    area = 0.5 * np.linalg.norm(np.cross(B_ - A_, C_ - A_), axis=1)
    return area[(dd * dd).sum(1) <= R * R].sum()


@pytest.mark.parametrize("width", [4, 8])
def test_sample_reflecting_is_unbiased(gpu, width):
    """E[1/pdf] is the area of the reflecting triangles whose boxes meet the ball (the support
    of the sampler); E[1{|y - x| < R}/pdf] above the plate is the disc area pi (R^2 - h^2)."""
    torch = gpu
    p = PB.lprism3d(plate=True)
    sc = _scene(p, width)
    V = p.positions.astype(np.float64)
    T = p.indices[p.is_reflecting.astype(bool)]
    rng = np.random.default_rng(13)
    ns = 200_000
    for x, R in [((-0.55, -0.55, -0.2), 0.2), ((-0.3, 0.5, 0.8), 0.5), ((0.5, -0.3, 0.2), 0.6), ((-0.2, 0.2, 0.0), 0.9)]:
        x = np.array(x)
        xs = np.broadcast_to(np.float32(x), (ns, 3))
        u = rng.random((ns, 3), dtype=np.float32)
        ok, y, nr, pdf = sc.sample_reflecting(_t(torch, xs), _t(torch, np.full(ns, R)), _t(torch, u))
        ok = ok.cpu().numpy().astype(bool); y = y.cpu().numpy().astype(np.float64); pdf = pdf.cpu().numpy()
        support = _tri_support_area(V, T, x, R)
        assert support > 0 and ok.all() and (pdf > 0).all()
        w = 1.0 / pdf.astype(np.float64)
        assert abs(w.mean() - support) < 4 * w.std() / math.sqrt(ns) + 1e-6 * support, (x, w.mean(), support)
        if R == 0.2:
            wi = np.where(np.linalg.norm(y - x, axis=1) < R, w, 0.0)
            disc = math.pi * (R * R - 0.1 ** 2)
            assert abs(wi.mean() - disc) < 4 * wi.std() / math.sqrt(ns), (wi.mean(), disc)


def _host(st):
    return dict(n=st.n.cpu().numpy(), mean=st.mean.cpu().numpy(), var=st.variance().cpu().numpy(),
                len=st.walk_length.cpu().numpy())


@pytest.mark.parametrize("width", [4, 8])
def test_wost3d_parity_with_oracle_and_closed_form(gpu, oracle_lib, prism, width):
    torch = gpu
    O = oracle_lib
    p, osc, pts, d = prism
    pts, d = pts[:400], d[:400]
    n = len(pts)
    sc = _scene(p, width)
    st = E.DeviceStats.zeros(n)
    fsr = torch.zeros(n, dtype=torch.float32, device="cuda")
    dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                        n_walks=torch.full((n,), 256, dtype=torch.int32, device="cuda"), n_walks_max=256,
                        first_sphere_radius=fsr)
    rep = E.wost_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=21))
    g = _host(st)
    ost = O.Stats(n)
    ofsr = np.zeros(n, np.float32)
    orep = O.solve(osc, p.pde, p.settings, pts, d, 256, ost, seed=22, wost=True, first_sphere_radius=ofsr)
    walks = 256 * n
    assert rep.total_walks == walks
    assert rep.completion_codes[3] <= 1e-4 * walks and orep["codes"][3] <= 1e-4 * walks
    z, ok = z_scores(g["mean"], g["var"], g["n"], ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, f"wost 3d gpu vs oracle (width {width})")
    exact = p.exact(pts)
    pooled = (g["mean"] - exact).sum() / np.sqrt((g["var"] / g["n"]).sum())
    assert abs(pooled) < 4.0, pooled
    np.testing.assert_allclose(fsr.cpu().numpy(), ofsr, rtol=1e-5, atol=1e-6)
    ml_g, ml_o = g["len"].sum() / g["n"].sum(), ost.walk_length.sum() / ost.n.sum()
    assert abs(ml_g / ml_o - 1) < 0.05, (ml_g, ml_o)


@pytest.mark.parametrize("case", ["neumann", "robin", "screened"])
def test_wost3d_gradient(gpu, oracle_lib, case):
    """walk_on_stars.h:664-848 in 3D: first ball 0.99 min(d_absorbing, d_reflecting), then
    walk-on-stars steps (Neumann, Robin and screened); gradients against the manufactured
    gradient and the oracle"""
    torch = gpu
    O = oracle_lib
    if case == "neumann":
        p = PB.lprism3d(plate=True)
    elif case == "robin":
        p = PB.lprism3d_robin(kappa=2.0, plate=True)
    else:
        p = PB.lprism3d_screened(sigma=4.0, plate=True)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    sc = _scene(p, 8)
    if "robin" in p.meta:
        osc.set_robin(*p.meta["robin"])
        sc.set_robin_coefficients(*p.meta["robin"])
    pts = _interior(p, 300, 17, margin=0.05)
    n = len(pts)
    d = osc.closest_point(pts, boundary=0)[0]
    x = pts.astype(np.float64)
    if case == "screened":
        a = np.array(p.meta["a"])
        g = np.exp(x @ a)[:, None] * a[None, :]
    else:
        e = np.exp(x[:, 0])
        g = np.stack([e * np.cos(x[:, 1]), -e * np.sin(x[:, 1]), np.zeros(n)], 1)
    dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                        n_walks=torch.full((n,), 256, dtype=torch.int32, device="cuda"), n_walks_max=256,
                        first_sphere_radius=torch.zeros(n, dtype=torch.float32, device="cuda"),
                        estimation_quantity=torch.full((n,), A.ZG_ESTIMATE_SOLUTION_AND_GRADIENT, dtype=torch.uint8,
                                                       device="cuda"))
    st = E.DeviceStats.zeros(n, dim=3)
    rep = E.wost_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=33))
    gm = st.grad_mean.cpu().numpy(); gv = st.grad_variance().cpu().numpy(); ng = st.n_grad.cpu().numpy()
    assert rep.completion_codes[3] <= 1e-4 * rep.total_walks + 2 and ng.min() >= 250
    pooled = np.sum(gm - g, axis=0) / np.sqrt(np.sum(gv / ng[:, None], axis=0))
    assert np.all(np.abs(pooled) < 4.0), (case, pooled)
    ost = O.GradStats(n, 3)
    O.solve_grad(osc, p.pde, p.settings, pts, d, 256, ost, wost=True, seed=34, cv_mode=O.CV_STAGED)
    zs = []
    for k in range(3):
        se2 = gv[:, k] / ng + ost.grad_variance()[:, k] / ost.n_grad
        m = se2 > 0
        zs.append((gm[m, k] - ost.grad_mean[m, k]) / np.sqrt(se2[m]))
    z = np.concatenate(zs)
    assert int((np.abs(z) > 4.0).sum()) <= 3 and abs(z.mean()) < 5.0 / np.sqrt(len(z)) + 0.05, (case, z.mean())


@pytest.mark.parametrize("width", [4, 8])
def test_robin_star_radius_matches_oracle(gpu, oracle_lib, width):
    """with per-triangle Robin bounds the star radius is min(silhouette radius, Robin bound)"""
    torch = gpu
    O = oracle_lib
    p = PB.lprism3d_robin(kappa=3.0, plate=True)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    osc.set_robin(*p.meta["robin"])
    sc = _scene(p, width)
    sc.set_robin_coefficients(*p.meta["robin"])
    pts = _interior(p, 600, 19)
    d = osc.closest_point(pts, boundary=0)[0]
    eps = p.settings.epsilon_shell_for_reflecting_boundary
    mn = np.full(len(pts), eps, np.float32)
    g = sc.star_radius(_t(torch, pts), _t(torch, mn), _t(torch, d)).cpu().numpy()
    o = np.maximum(np.minimum(osc.star_radius(pts, mn, d), osc.robin_radius(pts, d)), eps)
    assert np.isclose(g, o, rtol=3e-4, atol=1e-6).mean() > 0.99
    assert np.mean(osc.robin_radius(pts, d) < osc.star_radius(pts, mn, d) - 1e-6) > 0.02   # the bound binds
    with pytest.raises(ValueError):
        sc.set_robin_coefficients(np.zeros(3, np.float32), np.ones(3, np.float32))


@pytest.mark.parametrize("kappa", [0.5, 10.0])
def test_robin_wost3d_parity_with_oracle_and_closed_form(gpu, oracle_lib, kappa):
    torch = gpu
    O = oracle_lib
    p = PB.lprism3d_robin(kappa=kappa, plate=True)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    osc.set_robin(*p.meta["robin"])
    sc = _scene(p, 8)
    sc.set_robin_coefficients(*p.meta["robin"])
    pts = _interior(p, 400, 23)
    d = osc.closest_point(pts, boundary=0)[0]
    n = len(pts)
    st = E.DeviceStats.zeros(n)
    dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                        n_walks=torch.full((n,), 256, dtype=torch.int32, device="cuda"), n_walks_max=256)
    rep = E.wost_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=41))
    g = _host(st)
    ost = O.Stats(n)
    O.solve(osc, p.pde, p.settings, pts, d, 256, ost, seed=42, wost=True)
    assert rep.completion_codes[3] <= 1e-4 * rep.total_walks
    z, ok = z_scores(g["mean"], g["var"], g["n"], ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, f"robin wost 3d gpu vs oracle (kappa {kappa})")
    pooled = (g["mean"] - p.exact(pts)).sum() / np.sqrt((g["var"] / g["n"]).sum())
    assert abs(pooled) < 4.0, pooled
    ml_g, ml_o = g["len"].sum() / g["n"].sum(), ost.walk_length.sum() / ost.n.sum()
    assert abs(ml_g / ml_o - 1) < 0.05, (ml_g, ml_o)


@pytest.mark.parametrize("case", [(4.0, None, True), (9.0, 2.0, False)], ids=["neumann_plate", "robin"])
def test_screened_wost3d_parity_with_oracle_and_closed_form(gpu, oracle_lib, case):
    """the Yukawa ball in 3D walk on stars: u = exp(a . x), -Lap u + sigma u = 0"""
    torch = gpu
    O = oracle_lib
    sigma, kappa, plate = case
    p = PB.lprism3d_screened(sigma=sigma, kappa=kappa, plate=plate)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max,
                  solve_double_sided=p.solve_double_sided)
    sc = _scene(p, 8)
    if kappa is not None:
        osc.set_robin(*p.meta["robin"])
        sc.set_robin_coefficients(*p.meta["robin"])
    pts = _interior(p, 400, 29)
    d = osc.closest_point(pts, boundary=0)[0]
    n = len(pts)
    st = E.DeviceStats.zeros(n)
    dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                        n_walks=torch.full((n,), 512, dtype=torch.int32, device="cuda"), n_walks_max=512)
    rep = E.wost_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=51))
    g = _host(st)
    ost = O.Stats(n)
    O.solve(osc, p.pde, p.settings, pts, d, 512, ost, seed=52, wost=True)
    assert rep.completion_codes[3] <= 1e-4 * rep.total_walks
    z, ok = z_scores(g["mean"], g["var"], g["n"], ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, f"screened wost 3d gpu vs oracle {case}")
    pooled = (g["mean"] - p.exact(pts)).sum() / np.sqrt((g["var"] / g["n"]).sum())
    assert abs(pooled) < 4.0, pooled
    ml_g, ml_o = g["len"].sum() / g["n"].sum(), ost.walk_length.sum() / ost.n.sum()
    assert abs(ml_g / ml_o - 1) < 0.05, (ml_g, ml_o)


def test_wost3d_rejections(gpu):
    torch = gpu
    p = PB.lprism3d(plate=False)
    sc = _scene(p, 8)
    x = _t(torch, np.zeros((4, 3)) - 0.5)
    d = sc.closest_point(x, want_point=False)[0].contiguous()
    dp = E.DevicePoints(pt=x, dist_to_absorbing=d, n_walks=torch.full((4,), 8, dtype=torch.int32, device="cuda"),
                        n_walks_max=8)
    # splitting with gradient points runs in walk on spheres only
    dq = E.DevicePoints(pt=x, dist_to_absorbing=d, n_walks=dp.n_walks, n_walks_max=8,
                        estimation_quantity=torch.full((4,), A.ZG_ESTIMATE_SOLUTION_AND_GRADIENT, dtype=torch.uint8,
                                                       device="cuda"))
    with pytest.raises(A.ZombieError, match="splitting"):
        E.wost_solve(sc, p.pde, E.walk_settings_host(1e-3, 1e-3, 64, splitting_threshold=0.5), dq,
                     E.DeviceStats.zeros(4, dim=3))
    with pytest.raises(ValueError):       # 2D points for a 3D scene
        E.wost_solve(sc, p.pde, p.settings, E.DevicePoints(pt=x[:, :2].contiguous(), dist_to_absorbing=d,
                                                           n_walks=dp.n_walks, n_walks_max=8), E.DeviceStats.zeros(4))


@pytest.mark.parametrize("case", ["closed", "plate", "robin", "screened"])
def test_interleaved_kernel_walks_equal_run_to_completion(gpu, oracle_lib, case):
    """3D walk on stars: the interleaved kernel runs the same walks as the run-to-completion
    kernel (same steps, completion codes and estimates; FP64 sums up to their accumulation order)."""
    torch = gpu
    O = oracle_lib
    if case in ("closed", "plate"):
        p = PB.lprism3d(plate=case == "plate")
    elif case == "robin":
        p = PB.lprism3d_robin(kappa=2.0, plate=True)
    else:
        p = PB.lprism3d_screened(sigma=4.0, plate=True)
    osc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max,
                  solve_double_sided=p.solve_double_sided)
    sc = _scene(p, 4)
    if "robin" in p.meta:
        sc.set_robin_coefficients(*p.meta["robin"])
    pts = _interior(p, 200, 29)
    d = osc.closest_point(pts, boundary=0)[0]
    n = len(pts)
    res = []
    for inter in (True, False):
        st = E.DeviceStats.zeros(n)
        dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                            n_walks=torch.full((n,), 128, dtype=torch.int32, device="cuda"), n_walks_max=128)
        rep = E.wost_solve(sc, p.pde, p.settings, dp, st,
                           E.SolveOptions(seed=71, distance_tolerance=0.2, step_interval=8,
                                          wost_interleave_queries=inter))
        res.append((rep, _host(st)))
    (ri, gi), (rr, gr) = res
    print(f"interleaved 3d {case}: steps {ri.total_steps} vs {rr.total_steps}, codes {ri.completion_codes} vs "
          f"{rr.completion_codes}, kernel {ri.kernel_ms:.1f} vs {rr.kernel_ms:.1f} ms", flush=True)
    assert ri.total_walks == rr.total_walks == 128 * n
    assert ri.completion_codes == rr.completion_codes
    assert ri.total_steps == rr.total_steps
    np.testing.assert_array_equal(gi["n"], gr["n"])
    np.testing.assert_array_equal(gi["len"], gr["len"])
    np.testing.assert_allclose(gi["mean"], gr["mean"], rtol=1e-9, atol=1e-9)
