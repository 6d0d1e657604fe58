"""Oracle geometry pinned by brute force and independent numpy closed forms (no GPU)."""
import numpy as np
import pytest

from zombie_amd import problems as PB


def _np_closest_segment(Q, A, B):
    """Independent vectorised point-segment distance (projection clamped to [0, 1])."""
    ab = B - A
    t = np.clip(np.einsum("qij,ij->qi", Q[:, None, :] - A[None], ab) / np.einsum("ij,ij->i", ab, ab)[None], 0, 1)
    P = A[None] + t[..., None] * ab[None]
    d = np.linalg.norm(Q[:, None, :] - P, axis=2)
    return d.min(1)


def test_closest_point_2d_bvh_equals_numpy(oracle_lib):
    V, I = PB.star_curve(256)
    sc = oracle_lib.Scene(V, I)
    rng = np.random.default_rng(1)
    Q = rng.uniform(-1.5, 1.5, size=(3000, 2)).astype(np.float32)
    d, cp, nr, pr, _ = sc.closest_point(Q)
    ref = _np_closest_segment(Q.astype(np.float64), V[I[:, 0]].astype(np.float64), V[I[:, 1]].astype(np.float64))
    np.testing.assert_allclose(d, ref, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(Q - cp, axis=1), d, rtol=1e-5, atol=1e-6)
    # returned normals are unit and perpendicular to the returned segment
    seg = V[I[pr, 1]] - V[I[pr, 0]]
    assert np.abs(np.einsum("ij,ij->i", nr, seg)).max() < 1e-4
    np.testing.assert_allclose(np.linalg.norm(nr, axis=1), 1.0, atol=1e-5)


def test_closest_point_3d_bvh_equals_brute_force_and_dense_sampling(oracle_lib):
    V, F = PB.icosphere(2)
    V = PB.sh_displace(V, 3).astype(np.float32)
    F = F.astype(np.int32)
    sc = oracle_lib.Scene(V, F)
    rng = np.random.default_rng(2)
    Q = rng.uniform(-1.4, 1.4, size=(2000, 3)).astype(np.float32)
    d, cp, _, _, _ = sc.closest_point(Q)
    db, *_ = sc.closest_point(Q, brute=True)
    np.testing.assert_array_equal(d, db)
    # an independent upper bound: distance to a dense barycentric sampling of every triangle
    k = 24
    a, b = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    m = a + b <= k
    w = np.stack([a[m], b[m], k - a[m] - b[m]], 1) / k
    S = np.einsum("sw,fwd->fsd", w, V[F]).reshape(-1, 3)
    for q0 in range(0, 2000, 250):
        q = Q[q0:q0 + 250]
        dense = np.sqrt(((q[:, None, :] - S[None]) ** 2).sum(2)).min(1)
        assert (d[q0:q0 + 250] <= dense + 1e-5).all()
        # sampling resolution: the closest point is within half a sample spacing of a sample
        assert (dense - d[q0:q0 + 250]).max() < 0.05


def test_ray_first_hit_matches_circle(oracle_lib):
    n = 2048
    t = 2 * np.pi * np.arange(n) / n
    V = np.stack([np.cos(t), np.sin(t)], 1).astype(np.float32)
    I = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    sc = oracle_lib.Scene(V, I)
    rng = np.random.default_rng(3)
    ang = rng.uniform(0, 2 * np.pi, 500)
    d = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    o = np.zeros((500, 2), np.float32)
    hit, th, _ = sc.intersect_ray(o, d, np.full(500, 10.0, np.float32))
    assert hit.all()
    # polygon inscribed in the unit circle: cos(pi/n) <= t <= 1
    assert (th >= np.cos(np.pi / n) - 1e-6).all() and (th <= 1 + 1e-6).all()
    hb, tb, _ = sc.intersect_ray(o, d, np.full(500, 10.0, np.float32), brute=True)
    np.testing.assert_allclose(th, tb, rtol=1e-6)
    miss, _, _ = sc.intersect_ray(o, d, np.full(500, 0.5, np.float32))
    assert not miss.any()


def test_inside_domain_2d_and_3d(oracle_lib):
    V, I = PB.star_curve(512)
    sc = oracle_lib.Scene(V, I)
    rng = np.random.default_rng(4)
    Q = rng.uniform(-1.4, 1.4, size=(4000, 2)).astype(np.float32)
    assert (sc.inside_domain(Q) == PB.points_in_polygon(Q, V, I)).mean() > 0.999
    Vs, F = PB.icosphere(3)
    sc3 = oracle_lib.Scene(Vs.astype(np.float32), F.astype(np.int32))
    Q3 = rng.uniform(-1.2, 1.2, size=(4000, 3)).astype(np.float32)
    r = np.linalg.norm(Q3, axis=1)
    ins = sc3.inside_domain(Q3)
    clear = (r < 0.97) | (r > 1.01)  # the inscribed polyhedron lies between radius ~0.98 and 1
    assert (ins[clear] == (r[clear] < 0.97)).all()


def test_silhouette_star_radius_invariants(oracle_lib):
    """No silhouette point lies strictly inside the returned star radius; the radius is
    clamped to [min_radius, max_radius]."""
    prob = PB.mixed2d(n_segments=512, n_slits=8)
    sc = oracle_lib.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                          solve_double_sided=True)
    rng = np.random.default_rng(5)
    Q = rng.uniform(-0.8, 0.8, size=(200, 2)).astype(np.float32)
    rad = sc.star_radius(Q, 1e-3, 10.0)
    for q, r in zip(Q, rad):
        verts, sil = sc.silhouette_candidates(q)
        if sil.any():
            dmin = np.linalg.norm(verts[sil] - q, axis=1).min()
            assert r <= dmin + 1e-6
            assert abs(r - max(dmin, 1e-3)) < 1e-5
        else:
            assert r == pytest.approx(10.0)
    # open slit ends are always silhouettes
    verts, sil = sc.silhouette_candidates(np.array([0.0, 0.0], np.float32))
    ends = np.array([p for s in prob.meta["slits"] for p in s], np.float32)
    for e in ends:
        k = np.argmin(np.linalg.norm(verts - e, axis=1))
        assert np.linalg.norm(verts[k] - e) < 1e-6 and sil[k]


def _outer_reflecting_chain(prob):
    """Vertices of the reflecting half of the star curve with both neighbours reflecting:
    (vertex, previous vertex, next vertex) in the curve's counter-clockwise order."""
    V, I = prob.meta["outer"]
    refl = prob.is_reflecting[: len(I)].astype(bool)
    out = []
    for k in range(len(I)):
        kn = (k + 1) % len(I)
        if refl[k] and refl[kn]:
            out.append((V[I[k, 1]], V[I[k, 0]], V[I[kn, 1]]))
    return out


def test_silhouette_rule_matches_visibility_geometry(oracle_lib):
    """Generic views: a vertex is a silhouette exactly when its two neighbours lie on the same
    side of the line through the viewer and the vertex.  Views along an adjacent segment
    (the walker on the boundary): the vertex is a silhouette exactly when the corner is
    concave (a right turn of the counter-clockwise curve), with normals pointing away from
    the walker; flipping the orientation flips the answer."""
    prob = PB.mixed2d(n_segments=256, n_slits=1)
    sc = oracle_lib.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                          solve_double_sided=True)
    chain = _outer_reflecting_chain(prob)
    assert len(chain) > 100

    def cross(a, b):
        return a[0] * b[1] - a[1] * b[0]

    rng = np.random.default_rng(12)
    Q = rng.uniform(-0.5, 0.5, size=(40, 2)).astype(np.float32)
    checked = 0
    for q in Q:
        verts, sil = sc.silhouette_candidates(q)
        for v, u0, u1 in chain[::5]:
            k = np.argmin(np.linalg.norm(verts - v, axis=1))
            assert np.linalg.norm(verts[k] - v) < 1e-6
            d = (q - v).astype(np.float64)
            s0 = cross(d, (u0 - v).astype(np.float64)) / (np.linalg.norm(d) * np.linalg.norm(u0 - v))
            s1 = cross(d, (u1 - v).astype(np.float64)) / (np.linalg.norm(d) * np.linalg.norm(u1 - v))
            if min(abs(s0), abs(s1)) < 1e-2:
                continue                               # near-grazing views: covered below
            assert sil[k] == (s0 * s1 > 0), (q, v)
            checked += 1
    assert checked > 500
    # the walker on a segment, looking at the segment's end vertex
    tangent = 0
    for v, u0, u1 in chain:
        x = (0.5 * (u0 + v)).astype(np.float32)        # on the segment u0 -> v
        turn = cross((v - u0).astype(np.float64), (u1 - v).astype(np.float64))
        if abs(turn) < 1e-4 * np.linalg.norm(v - u0) * np.linalg.norm(u1 - v):
            continue
        for flip in (False, True):
            verts, sil = sc.silhouette_candidates(x, flip=flip)
            k = np.argmin(np.linalg.norm(verts - v, axis=1))
            concave = turn < 0
            assert sil[k] == (concave != flip), (v, turn, flip)
        tangent += 1
    assert tangent > 50


def test_reflecting_sampler_is_unbiased_for_boundary_length(oracle_lib):
    """E[1/pdf * 1{sample in ball}] equals the length of reflecting boundary inside the ball."""
    prob = PB.mixed2d(n_segments=512, n_slits=8)
    sc = oracle_lib.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                          solve_double_sided=True)
    x = np.array([0.1, 0.8], np.float32)
    R = 0.5
    rng = np.random.default_rng(6)
    acc = []
    for _ in range(20000):
        ok, pt, nr, pdf = sc.sample_reflecting(x, R, rng.random(2).astype(np.float32))
        acc.append(1.0 / pdf if ok and np.linalg.norm(pt - x) < R else 0.0)
    est = np.mean(acc)
    se = np.std(acc) / np.sqrt(len(acc))
    # exact clipped length by fine subdivision of the reflecting segments
    V, I = prob.positions, prob.indices[prob.is_reflecting.astype(bool)]
    s = np.linspace(0, 1, 2001)
    length = 0.0
    for a, b in zip(V[I[:, 0]], V[I[:, 1]]):
        P = a[None] + s[:, None] * (b - a)[None]
        inside = np.linalg.norm(P - x, axis=1) < R
        length += np.linalg.norm(b - a) * inside[:-1].mean()
    assert abs(est - length) < 4 * se + 1e-3


def test_bvh_reflecting_queries_equal_brute_force(oracle_lib):
    """The estimator's BVH forms of the star radius and line of sight return exactly what the
    brute-force scans return; the BVH sampler's density is exact: E[1/pdf] is the length of
    the segments whose box meets the ball and E[1/pdf 1{in ball}] the length inside it."""
    prob = PB.mixed2d(n_segments=512, n_slits=8)
    sc = oracle_lib.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                          solve_double_sided=True)
    rng = np.random.default_rng(21)
    Q = rng.uniform(-1.0, 1.0, size=(600, 2)).astype(np.float32)
    mx = rng.uniform(0.01, 1.5, size=600).astype(np.float32)
    for flip in (False, True):
        a = sc.star_radius(Q, 1e-3, mx, flip=flip)
        b = sc.star_radius(Q, 1e-3, mx, flip=flip, bvh=True)
        np.testing.assert_array_equal(a, b)
    P2 = Q + rng.normal(scale=0.4, size=Q.shape).astype(np.float32)
    los = [sc.line_of_sight(p, q) for p, q in zip(Q, P2)]
    los_b = [sc.line_of_sight(p, q, bvh=True) for p, q in zip(Q, P2)]
    assert los == los_b and 0 < sum(los) < len(los)
    V, I = prob.positions, prob.indices[prob.is_reflecting.astype(bool)]
    A, B = V[I[:, 0]].astype(np.float64), V[I[:, 1]].astype(np.float64)
    for x, R in (((0.1, 0.8), 0.5), ((-0.3, 0.1), 0.3)):
        x = np.array(x, np.float32)
        lo, hi = np.minimum(A, B), np.maximum(A, B)
        dd = np.maximum(np.maximum(lo - x, x - hi), 0)
        support = np.linalg.norm(B - A, axis=1)[(dd * dd).sum(1) <= R * R].sum()
        s = np.linspace(0, 1, 2001)
        clipped = 0.0
        for a, b in zip(A, B):
            pts = a[None] + s[:, None] * (b - a)[None]
            clipped += np.linalg.norm(b - a) * (np.linalg.norm(pts - x, axis=1) < R)[:-1].mean()
        w, wi = [], []
        for _ in range(20000):
            ok, pt, nr, pdf = sc.sample_reflecting(x, R, rng.random(2).astype(np.float32), bvh=True)
            w.append(1.0 / pdf if ok else 0.0)
            wi.append(1.0 / pdf if ok and np.linalg.norm(pt - x) < R else 0.0)
        w, wi = np.array(w), np.array(wi)
        assert abs(w.mean() - support) < 4 * w.std() / np.sqrt(len(w)) + 1e-6 * support
        assert abs(wi.mean() - clipped) < 4 * wi.std() / np.sqrt(len(wi)) + 1e-3


def test_line_of_sight_brute_force(oracle_lib):
    prob = PB.mixed2d(n_segments=256, n_slits=8)
    sc = oracle_lib.Scene(prob.positions, prob.indices, prob.is_reflecting, prob.domain_min, prob.domain_max,
                          solve_double_sided=True)
    V, I = prob.positions, prob.indices[prob.is_reflecting.astype(bool)]
    rng = np.random.default_rng(7)

    def crosses(p, q, a, b):
        def orient(u, v, w):
            return (v[0] - u[0]) * (w[1] - u[1]) - (v[1] - u[1]) * (w[0] - u[0])
        return orient(p, q, a) * orient(p, q, b) < 0 and orient(a, b, p) * orient(a, b, q) < 0
    for _ in range(300):
        p, q = rng.uniform(-0.9, 0.9, size=(2, 2)).astype(np.float32)
        blocked = any(crosses(p, q, V[i], V[j]) for i, j in I)
        assert sc.line_of_sight(p, q) == (not blocked)


# ------------------------------------------------------------------------------------
# Robin star-radius bound (reflectance_boundary_bvh/geometry.h:86-140, robin_bounds.h:12-35)
# ------------------------------------------------------------------------------------
def _robin_scene(O, kappa):
    p = PB.mixed2d_robin(kappa=kappa, n_segments=512, n_slits=8)
    sc = O.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, solve_double_sided=True)
    sc.set_robin(*p.meta["robin"])
    return p, sc


def test_robin_bound_bvh_equals_brute_force(oracle_lib):
    O = oracle_lib
    p, sc = _robin_scene(O, 3.0)
    pts = PB.sample_box(p.domain_min, p.domain_max, 3000, 12)
    a = sc.robin_radius(pts, 10.0)
    b = sc.robin_radius(pts, 10.0, brute=True)
    np.testing.assert_array_equal(a, b)
    assert np.all(a <= 10.0) and np.all(a > 0)
    sc.set_robin(None, None)
    np.testing.assert_array_equal(sc.robin_radius(pts, 10.0), np.full(len(pts), 10.0, np.float32))


@pytest.mark.parametrize("kappa", [0.5, 3.0, 20.0])
def test_robin_bound_keeps_reflectance_nonnegative(oracle_lib, kappa):
    """Inside the bound R every point y of a Robin segment at distance r < R has
    |cos(n, y - x)| >= kappa r log(R / r), i.e. the reflectance 1 - kappa G / P of a step from x
    to y is >= 0 -- the property the bound exists for -- and the bound is tight: some segment
    point of the minimising segment comes within a few percent of equality."""
    O = oracle_lib
    p, sc = _robin_scene(O, kappa)
    V, I = p.positions.astype(np.float64), p.indices
    refl = p.is_reflecting.astype(bool)
    segs = [(V[I[k, 0]], V[I[k, 1]]) for k in np.nonzero(refl)[0]]
    rng = np.random.default_rng(5)
    Vo, Io = p.meta["outer"]
    cand = PB.sample_box(p.domain_min, p.domain_max, 600, 13)
    pts = cand[PB.points_in_polygon(cand, Vo, Io)][:120]
    R = sc.robin_radius(pts, 10.0).astype(np.float64)
    t = np.linspace(0.0, 1.0, 65)
    worst_slack = []
    for x, Rx in zip(pts.astype(np.float64), R):
        slack_min = np.inf
        for a, b in segs:
            y = a[None, :] + t[:, None] * (b - a)[None, :]
            d = y - x[None, :]
            r = np.linalg.norm(d, axis=1)
            m = (r < Rx) & (r > 1e-9)
            if not m.any():
                continue
            n = np.array([(b - a)[1], -(b - a)[0]]) / np.linalg.norm(b - a)
            cos = np.abs(d[m] @ n) / r[m]
            need = kappa * r[m] * np.log(Rx / r[m])
            assert np.all(cos >= need - 1e-4 * (1 + need)), (x, Rx, float(np.min(cos - need)))
            slack_min = min(slack_min, float(np.min(cos - need)))
        worst_slack.append(slack_min)
    # tightness: for most points the minimising segment nearly touches the constraint
    ws = np.array([s for s in worst_slack if np.isfinite(s)])
    assert np.median(ws) < 0.05, np.median(ws)
