"""Pins of the oracle's 3D walk-on-stars geometry and estimator (walk_on_stars.h:189-662 over
triangle meshes; FCPW's closest-silhouette-edge, stochastic sphere sampling and line-of-sight
queries behind fcpw_geometric_queries.h:1218-1270, with the dihedral-angle rule of
:955-967 and :1029-1038).  FCPW is absent, so the silhouette queries are pinned by what the
mathematics fixes: the candidate set of a known polyhedron, the star-shapedness the radius
exists for (no ray from x crosses the reflecting surface twice inside it, which the plain
absorbing distance violates on the same inputs), an open rim's closed-form distance, and the
BVH forms against brute force; the sampler by the surface-area identity E[1{|y - x| < R} / pdf]
= area inside the ball; the triangles' Robin bound by the non-negative reflectance it exists for
(and its tightness); the estimator by a manufactured harmonic solution with Neumann and with
Robin data."""
import math

import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB


def _interior(p, n, seed, margin=0.02):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.95, 0.95, (4 * n, 3)).astype(np.float32)
    return x[p.meta["inside"](x, margin)][:n]


def _scene(O, p):
    return O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max,
                   solve_double_sided=p.solve_double_sided)


def _edges(I):
    e = np.sort(np.concatenate([I[:, [0, 1]], I[:, [1, 2]], I[:, [2, 0]]]), axis=1)
    u, c = np.unique(e, axis=0, return_counts=True)
    return u, c


def test_silhouette_candidates_of_the_l_prism(oracle_lib):
    """One-sided closed L-prism: of the reflecting triangles' edges, flat and convex ones are
    dropped; the concave re-entrant edge (n_z pieces along (0, 0, z)) and the rim between the
    reflecting and absorbing parts (edges with one reflecting triangle) remain.  Double-sided:
    every edge is a candidate."""
    O = oracle_lib
    n_sub, n_z = 4, 4
    p = PB.lprism3d(n_sub=n_sub, n_z=n_z, plate=False)
    sc = _scene(O, p)
    R = p.indices[p.is_reflecting.astype(bool)]
    u, c = _edges(R)
    rim = int((c == 1).sum())
    V, _ = sc.silhouette_candidates(np.zeros(3, np.float32))
    assert len(V) == rim + n_z
    reentrant = np.all(np.abs(V[:, :, :2]) < 1e-6, axis=(1, 2))
    assert reentrant.sum() == n_z
    sc2 = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    assert len(sc2.silhouette_candidates(np.zeros(3, np.float32))[0]) == len(u)


@pytest.mark.parametrize("plate", [False, True])
def test_star_radius_is_star_shaped_and_bvh_equals_brute_force(oracle_lib, plate):
    O = oracle_lib
    p = PB.lprism3d(plate=plate)
    sc = _scene(O, p)
    x = _interior(p, 600, 1)
    dA = sc.closest_point(x, boundary=0)[0]
    r = sc.star_radius(x, 1e-3, dA)
    assert np.array_equal(r, sc.star_radius(x, 1e-3, dA, bvh=True))
    assert np.all(r <= dA) and np.mean(r < dA - 1e-6) > 0.1

    rng = np.random.default_rng(2)

    def max_crossings(radius, idx):
        worst = 0
        for i in idx:
            d = rng.normal(size=(48, 3)).astype(np.float32)
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            for k in range(48):
                t0, cnt = 0.0, 0
                while t0 < radius[i]:
                    o = (x[i] + d[k] * t0).astype(np.float32)
                    hit, t, _ = sc.intersect_ray(o[None], d[k][None], np.array([radius[i] - t0], np.float32), boundary=1)
                    if not hit[0]:
                        break
                    cnt += 1
                    t0 += float(t[0]) + 1e-5
                worst = max(worst, cnt)
        return worst

    idx = np.arange(0, len(x), 6)
    assert max_crossings(r, idx) <= 1
    # the absorbing distance alone is not star-shaped on the same inputs
    assert max_crossings(dA, idx) >= 2


def test_open_rim_radius_closed_form(oracle_lib):
    """A flat open square [-a, a]^2 x {0} seen from (0, 0, h): flat interior edges are never
    silhouettes, every rim edge is, and the nearest rim point is an edge midpoint at
    sqrt(h^2 + a^2)."""
    O = oracle_lib
    m, a = 6, 0.5
    g = np.linspace(-a, a, m + 1)
    V = np.array([(x, y, 0.0) for y in g for x in g], np.float32)
    T = []
    for j in range(m):
        for i in range(m):
            v = j * (m + 1) + i
            T += [(v, v + 1, v + m + 2), (v, v + m + 2, v + m + 1)]
    I = np.array(T, np.int32)
    sc = O.Scene(V, I, np.ones(len(I), np.uint8), np.full(3, -2.0), np.full(3, 2.0), solve_double_sided=True)
    for h in (0.1, 0.3, -0.2):
        x = np.array([[0.0, 0.0, h]], np.float32)
        r = sc.star_radius(x, 1e-4, np.array([10.0], np.float32))[0]
        assert r == pytest.approx(math.hypot(h, a), rel=1e-5)
    # off-centre: the nearest rim edge (x = a)
    x = np.array([[0.3, -0.1, 0.2]], np.float32)
    assert sc.star_radius(x, 1e-4, np.array([10.0], np.float32))[0] == pytest.approx(math.hypot(0.2, 0.2), rel=1e-5)


@pytest.mark.parametrize("bvh", [False, True])
def test_sampler_area_identity(oracle_lib, bvh):
    """E[1{|y - x| < R} / pdf] over the sampler's draws equals the area of the reflecting
    surface inside the ball: a disc of radius sqrt(R^2 - h^2) on a large flat square."""
    O = oracle_lib
    m = 16
    g = np.linspace(-1.0, 1.0, m + 1)
    V = np.array([(x, y, 0.0) for y in g for x in g], np.float32)
    T = []
    for j in range(m):
        for i in range(m):
            v = j * (m + 1) + i
            T += [(v, v + 1, v + m + 2), (v, v + m + 2, v + m + 1)]
    I = np.array(T, np.int32)
    sc = O.Scene(V, I, np.ones(len(I), np.uint8), np.full(3, -2.0), np.full(3, 2.0), solve_double_sided=True)
    x = np.array([0.1, -0.2, 0.25], np.float32)
    R = 0.6
    rng = np.random.default_rng(4)
    vals = []
    for _ in range(20000):
        ok, y, nr, pdf = sc.sample_reflecting(x, R, rng.random(3), bvh=bvh)
        assert ok and abs(y[2]) < 1e-6 and abs(abs(nr[2]) - 1) < 1e-6
        vals.append((np.linalg.norm(y - x) < R) / pdf)
    vals = np.array(vals)
    exact = math.pi * (R * R - 0.25 ** 2)
    assert abs(vals.mean() - exact) < 4 * vals.std() / math.sqrt(len(vals)), (vals.mean(), exact)


def test_line_of_sight_bvh_equals_brute_force_and_plate_blocks(oracle_lib):
    O = oracle_lib
    p = PB.lprism3d(plate=True)
    sc = _scene(O, p)
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    b = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    assert [sc.line_of_sight(u, v) for u, v in zip(a, b)] == [sc.line_of_sight(u, v, bvh=True) for u, v in zip(a, b)]
    # through the plate (z = -0.3, x, y in [-0.8, -0.3], away from its triangles' shared edges)
    # and beside it
    assert not sc.line_of_sight([-0.52, -0.45, -0.1], [-0.52, -0.45, -0.5])
    assert sc.line_of_sight([-0.52, -0.45, -0.1], [-0.52, -0.45, -0.2])
    assert sc.line_of_sight([-0.9, -0.9, -0.1], [-0.9, -0.9, -0.5])


@pytest.mark.parametrize("plate", [False, True])
def test_walk_on_stars_3d_matches_exact(oracle_lib, plate):
    """u = e^x cos y on the L-prism with Neumann walls, cap (and plate): 4-SE agreement; the
    same walks with the Neumann data dropped are detected."""
    O = oracle_lib
    p = PB.lprism3d(plate=plate)
    sc = _scene(O, p)
    x = _interior(p, 200, 3)
    d = sc.closest_point(x, boundary=0)[0]
    st = O.Stats(len(x))
    rep = O.solve(sc, p.pde, p.settings, x, d, 256, st, wost=True)
    u = p.exact(x)
    se = np.sqrt(st.variance() / st.n)
    z = (st.mean - u) / se
    pooled = np.sum(st.mean - u) / np.sqrt(np.sum(se ** 2))
    assert abs(pooled) < 4.0 and 0.8 < z.std() < 1.3, (pooled, z.std())
    assert rep["codes"][0] == 256 * len(x)
    bad = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=E.zero(), is_source_constant=True)
    st0 = O.Stats(len(x))
    O.solve(sc, bad, p.settings, x, d, 256, st0, wost=True)
    se0 = np.sqrt(st0.variance() / st0.n)
    assert abs(np.sum(st0.mean - u) / np.sqrt(np.sum(se0 ** 2))) > 10.0


def _robin_scene(O, kappa, plate=True):
    p = PB.lprism3d_robin(kappa=kappa, plate=plate)
    sc = _scene(O, p)
    sc.set_robin(*p.meta["robin"])
    return p, sc


@pytest.mark.parametrize("kappa", [0.5, 3.0, 20.0])
def test_robin_triangle_bound_keeps_reflectance_nonnegative(oracle_lib, kappa):
    """RobinTriangleBound (robin_bounds.h:52-80): inside the bound R every point y of a Robin
    triangle at distance r < R has |cos(n, y - x)| >= kappa r (1 - r / R), i.e. the 3D
    reflectance 1 - kappa G / P (G ~ 1/r - 1/R, P ~ cos / r^2) of a step from x to y is >= 0 --
    the property the bound exists for -- and the bound is tight: the minimising triangle
    touches the constraint.  BVH pruning equals the scan of every triangle."""
    O = oracle_lib
    p, sc = _robin_scene(O, kappa)
    V = p.positions.astype(np.float64)
    T = p.indices[p.is_reflecting.astype(bool)]
    x = _interior(p, 80, 5)
    R = sc.robin_radius(x, 10.0).astype(np.float64)
    np.testing.assert_array_equal(R, sc.robin_radius(x, 10.0, brute=True))
    g = np.linspace(0, 1, 9)
    bary = np.array([(a, b) for a in g for b in g if a + b <= 1 + 1e-12])
    A_, B_, C_ = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    Y = A_[:, None] + bary[None, :, :1] * (B_ - A_)[:, None] + bary[None, :, 1:] * (C_ - A_)[:, None]
    N = np.cross(B_ - A_, C_ - A_)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    slack = []
    for xi, Ri in zip(x.astype(np.float64), R):
        d = Y - xi
        r = np.linalg.norm(d, axis=2)
        m = (r < Ri) & (r > 1e-9)
        if not m.any():
            continue
        cos = np.abs(np.einsum("tkd,td->tk", d, N)) / np.maximum(r, 1e-30)
        need = kappa * r * (1 - r / Ri)
        sl = (cos - need)[m]
        assert sl.min() >= -1e-4, (xi, Ri, sl.min())
        slack.append(sl.min())
    assert len(slack) >= 40 and np.median(slack) < 0.02, np.median(slack)
    sc.set_robin(None, None)
    np.testing.assert_array_equal(sc.robin_radius(x, 10.0), np.full(len(x), 10.0, np.float32))


@pytest.mark.parametrize("kappa", [0.5, 10.0])
def test_robin_walk_on_stars_3d_matches_exact(oracle_lib, kappa):
    """Robin data du/dn + kappa u = h on the reflecting triangles (reflectance throughput,
    epsilon radius on the Robin boundary, triangle bound of the star radius): 4-SE agreement with
    u = e^x cos y; the same data treated as Neumann is detected."""
    O = oracle_lib
    p, sc = _robin_scene(O, kappa, plate=False)
    x = _interior(p, 200, 3)
    d = sc.closest_point(x, boundary=0)[0]
    st = O.Stats(len(x))
    O.solve(sc, p.pde, p.settings, x, d, 256, st, wost=True)
    u = p.exact(x)
    se = np.sqrt(st.variance() / st.n)
    pooled = np.sum(st.mean - u) / np.sqrt(np.sum(se ** 2))
    z = (st.mean - u) / se
    assert abs(pooled) < 4.0 and 0.8 < z.std() < 1.3, (pooled, z.std())
    neumann = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=p.pde.robin, is_source_constant=True)
    st0 = O.Stats(len(x))
    O.solve(sc, neumann, p.settings, x, d, 256, st0, wost=True)
    se0 = np.sqrt(st0.variance() / st0.n)
    assert abs(np.sum(st0.mean - u) / np.sqrt(np.sum(se0 ** 2))) > 10.0


@pytest.mark.parametrize("case", [(4.0, None, True), (9.0, 2.0, False)], ids=["neumann_plate", "robin"])
def test_screened_walk_on_stars_3d_matches_exact(oracle_lib, case):
    """-Lap u + sigma u = 0, u = exp(a . x): the Yukawa ball's dampening on every step that does
    not hit the reflecting boundary and the Yukawa reflectance Q - kappa G / P on those that do
    (distributions.h:965-985, walk_on_stars.h:277-307); the same data walked with the harmonic
    ball (absorption 0) is detected."""
    O = oracle_lib
    sigma, kappa, plate = case
    p = PB.lprism3d_screened(sigma=sigma, kappa=kappa, plate=plate)
    sc = _scene(O, p)
    if kappa is not None:
        sc.set_robin(*p.meta["robin"])
    x = _interior(p, 200, 3)
    d = sc.closest_point(x, boundary=0)[0]
    st = O.Stats(len(x))
    O.solve(sc, p.pde, p.settings, x, d, 512, st, wost=True)
    u = p.exact(x)
    se = np.sqrt(st.variance() / st.n)
    assert abs(np.sum(st.mean - u) / np.sqrt(np.sum(se ** 2))) < 4.0
    harmonic = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=p.pde.robin, is_source_constant=True,
                          robin_coeff=None if kappa is None else E.constant(kappa))
    st0 = O.Stats(len(x))
    O.solve(sc, harmonic, p.settings, x, d, 256, st0, wost=True)
    se0 = np.sqrt(st0.variance() / st0.n)
    assert abs(np.sum(st0.mean - u) / np.sqrt(np.sum(se0 ** 2))) > 10.0
