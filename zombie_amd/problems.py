"""Manufactured benchmark problems (host side, FP64 ground truth).

Each builder returns a Problem with the boundary mesh, the bounding domain, the PDE
descriptor (zg_pde), walk settings and the FP64 analytic solution.  Sizes follow the five
benchmark configurations:

  star2d     2D Laplace on r(t) = 1 + 0.3 cos 5t (1024 segments), g = e^x cos y, 256^2 grid
  sphere3d   3D Poisson on a spherical-harmonic displaced icosphere (81,920 triangles),
             u = sin 2x sin 2y sin 2z, f = 12 u, 1M interior points
  mixed2d    2D Laplace, 4096-segment star curve + 32 interior slits, Neumann on the slits
             and on the arc t in [0, pi], u = x^3 - 3 x y^2
  screened   screened Poisson instances u = exp(a.x), |a|^2 = sigma, on geodesic spheres
  lprism3d   3D Laplace on an L-shaped prism with Neumann re-entrant walls, top cap and an
             interior double-sided plate, u = e^x cos y (walk on stars in 3D)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from . import engine as E


@dataclass
class Problem:
    name: str
    dim: int
    positions: np.ndarray          # [nV, dim] float32
    indices: np.ndarray            # [nP, dim] int32
    is_reflecting: np.ndarray | None
    domain_min: np.ndarray
    domain_max: np.ndarray
    pde: A.zg_pde
    settings: A.zg_walk_settings
    n_walks: int
    exact: callable                # FP64 analytic u(x), x [n, dim]
    solve_double_sided: bool = False
    meta: dict = field(default_factory=dict)


# ------------------------------------------------------------------------------------
# geometry helpers
# ------------------------------------------------------------------------------------
def star_curve(n_segments: int, a: float = 0.3, k: int = 5):
    """Closed CCW polyline r(t) = 1 + a cos(k t); outward normals (s.y, -s.x)."""
    t = 2.0 * np.pi * np.arange(n_segments) / n_segments
    r = 1.0 + a * np.cos(k * t)
    V = np.stack([r * np.cos(t), r * np.sin(t)], 1)
    I = np.stack([np.arange(n_segments), (np.arange(n_segments) + 1) % n_segments], 1)
    return V.astype(np.float32), I.astype(np.int32)


def square_bounding_box(V: np.ndarray):
    """computeBoundingBox(positions, makeSquare=true, scale=1) (fcpw_geometric_queries.h:291-317)."""
    lo, hi = V.min(0).astype(np.float32), V.max(0).astype(np.float32)
    c = 0.5 * (lo + hi)
    half = 0.5 * float((hi - lo).max())
    return (c - half).astype(np.float32), (c + half).astype(np.float32)


def points_in_polygon(P: np.ndarray, V: np.ndarray, I: np.ndarray) -> np.ndarray:
    """Even-odd rule against the segments (V[I[:,0]], V[I[:,1]]), vectorised."""
    P = np.asarray(P, np.float64)
    a = V[I[:, 0]].astype(np.float64)
    b = V[I[:, 1]].astype(np.float64)
    inside = np.zeros(len(P), bool)
    chunk = 4096
    for s in range(0, len(P), chunk):
        x = P[s:s + chunk, 0:1]
        y = P[s:s + chunk, 1:2]
        ay, by = a[None, :, 1], b[None, :, 1]
        ax, bx = a[None, :, 0], b[None, :, 0]
        cond = (ay > y) != (by > y)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = ax + (y - ay) * (bx - ax) / (by - ay)
        cross = cond & (x < xi)
        inside[s:s + chunk] = (cross.sum(1) % 2) == 1
    return inside


def icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return V, F


def icosphere(level: int):
    """Unit icosphere with 20*4^level outward-oriented triangles."""
    V, F = icosahedron()
    verts = [v for v in V]
    for _ in range(level):
        cache = {}
        newF = []

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            m = cache.get(key)
            if m is None:
                p = verts[i] + verts[j]
                p = p / np.linalg.norm(p)
                verts.append(p)
                m = len(verts) - 1
                cache[key] = m
            return m

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            newF += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = np.array(newF, np.int64)
    return np.array(verts), F


def geodesic_sphere(freq: int):
    """Unit geodesic sphere of frequency `freq` (20*freq^2 outward triangles), welded vertices."""
    V0, F0 = icosahedron()
    n = freq
    ii, jj = [a.ravel() for a in np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")]
    keep = ii + jj <= n
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((n + 1, n + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    # local triangles of one face: up (i,j),(i+1,j),(i,j+1); down (i+1,j),(i+1,j+1),(i,j+1)
    ui, uj = [a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), indexing="ij")]
    up = ui + uj < n
    down = ui + uj + 1 < n
    tri_up = np.stack([lid[ui[up], uj[up]], lid[ui[up] + 1, uj[up]], lid[ui[up], uj[up] + 1]], 1)
    tri_dn = np.stack([lid[ui[down] + 1, uj[down]], lid[ui[down] + 1, uj[down] + 1], lid[ui[down], uj[down] + 1]], 1)
    tri_local = np.concatenate([tri_up, tri_dn])
    pts, tris = [], []
    nloc = len(ii)
    for f, (a, b, c) in enumerate(F0):
        A0, B0, C0 = V0[a], V0[b], V0[c]
        P = A0[None] + (B0 - A0)[None] * (ii / n)[:, None] + (C0 - A0)[None] * (jj / n)[:, None]
        pts.append(P / np.linalg.norm(P, axis=1, keepdims=True))
        tris.append(tri_local + f * nloc)
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    key = np.round(P * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return P[first], inv.reshape(-1)[T]


def real_sph_harm_basis(dirs: np.ndarray, lmax: int) -> np.ndarray:
    """Real orthonormal spherical harmonics Y_lm(dirs) for l <= lmax, columns ordered (l, m)."""
    from scipy.special import lpmv
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    ct = np.cos(theta)
    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            P = lpmv(am, l, ct)
            if m == 0:
                cols.append(norm * P)
            elif m > 0:
                cols.append(math.sqrt(2.0) * norm * P * np.cos(am * phi))
            else:
                cols.append(math.sqrt(2.0) * norm * P * np.sin(am * phi))
    return np.stack(cols, 1)


def sh_displace(V: np.ndarray, seed: int, amplitude: float = 0.15, lmax: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    coeffs = []
    for l in range(lmax + 1):
        coeffs += list(rng.standard_normal(2 * l + 1) / (l + 1))
    Y = real_sph_harm_basis(V, lmax)
    r = 1.0 + amplitude * (Y @ np.array(coeffs))
    if r.min() <= 0.05:
        raise ValueError("displacement makes the surface non star-shaped")
    return V * r[:, None]


# ------------------------------------------------------------------------------------
# analytic fields (FP64)
# ------------------------------------------------------------------------------------
def exp_cos(x):
    return np.exp(x[:, 0]) * np.cos(x[:, 1])


def sin_product(x, k=2.0):
    return np.prod(np.sin(k * np.asarray(x, np.float64)), axis=1)


def cubic_harmonic(x):
    x = np.asarray(x, np.float64)
    return x[:, 0] ** 3 - 3.0 * x[:, 0] * x[:, 1] ** 2


def exp_linear(x, a):
    return np.exp(np.asarray(x, np.float64) @ np.asarray(a, np.float64))


# ------------------------------------------------------------------------------------
# problems
# ------------------------------------------------------------------------------------
def star2d(n_segments=1024, grid_res=256, n_walks=512, eps=1e-3, max_len=1024) -> Problem:
    V, I = star_curve(n_segments)
    lo, hi = square_bounding_box(V)
    g = E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0))
    pde = E.make_pde(source=E.zero(), dirichlet=g, is_source_constant=True)
    ws = E.walk_settings_host(eps, eps, max_len, False)
    prob = Problem("star2d", 2, V, I, None, lo, hi, pde, ws, n_walks, exact_fn(exp_cos))
    ext = hi - lo
    ii, jj = np.meshgrid(np.arange(grid_res), np.arange(grid_res), indexing="ij")
    gx = (ii.ravel() / np.float32(grid_res)).astype(np.float32) * ext[0] + lo[0]
    gy = (jj.ravel() / np.float32(grid_res)).astype(np.float32) * ext[1] + lo[1]
    G = np.stack([gx, gy], 1).astype(np.float32)
    inside = points_in_polygon(G, V, I)
    prob.meta["grid"] = G
    prob.meta["points"] = G[inside]
    prob.meta["grid_inside"] = inside
    return prob


def exact_fn(f, **kw):
    def u(x):
        return f(np.asarray(x, np.float64), **kw)
    return u


def sphere3d(level=6, seed=0x5EED, n_walks=256, eps=1e-4, max_len=1024, k=2.0) -> Problem:
    V, F = icosphere(level)
    V = sh_displace(V, seed)
    lo, hi = V.min(0) - 1e-3, V.max(0) + 1e-3
    u = E.function(A.ZG_FN_SIN_PRODUCT, A.ZG_EVAL_VALUE, (1.0, k))
    f = E.function(A.ZG_FN_SIN_PRODUCT, A.ZG_EVAL_SCREENED_SOURCE, (1.0, k))
    pde = E.make_pde(source=f, dirichlet=u, is_source_constant=False)
    ws = E.walk_settings_host(eps, eps, max_len, False)
    return Problem("sphere3d", 3, V.astype(np.float32), F.astype(np.int32), None,
                   lo.astype(np.float32), hi.astype(np.float32), pde, ws, n_walks, exact_fn(sin_product, k=k))


def mixed2d(n_segments=4096, n_slits=32, seed=0x5EED, n_walks=1024, eps=1e-3, max_len=16384) -> Problem:
    """Star curve (Dirichlet on its lower half, Neumann on the upper half) with n_slits interior
    Neumann slits (double-sided), u = x^3 - 3xy^2.  Walks bounce between the slits (about 1,000
    steps on average with 32 slits); a walk that stops at max_len loses its terminal
    contribution, so the default max_len leaves no walk truncated."""
    V, I = star_curve(n_segments)
    rng = np.random.default_rng(seed)
    slits = []
    tries = 0
    while len(slits) < n_slits and tries < 100000:
        tries += 1
        c = rng.uniform(-0.6, 0.6, 2)
        L = rng.uniform(0.05, 0.2)
        ang = rng.uniform(0, np.pi)
        d = 0.5 * L * np.array([np.cos(ang), np.sin(ang)])
        p0, p1 = c - d, c + d
        ok = True
        for p in (p0, p1):
            t = math.atan2(p[1], p[0])
            if np.hypot(*p) > 0.9 * (1.0 + 0.3 * math.cos(5 * t)) - 0.05:
                ok = False
        for q0, q1 in slits:
            if _segments_close(p0, p1, q0, q1, 0.02):
                ok = False
        if ok:
            slits.append((p0, p1))
    nv = len(V)
    SV = np.array([p for s in slits for p in s], np.float32)
    SI = np.arange(2 * len(slits), dtype=np.int32).reshape(-1, 2) + nv
    mid = 0.5 * (V[I[:, 0]] + V[I[:, 1]])
    theta = np.mod(np.arctan2(mid[:, 1], mid[:, 0]), 2 * np.pi)
    refl_outer = theta <= np.pi
    Vall = np.concatenate([V, SV]).astype(np.float32)
    Iall = np.concatenate([I, SI]).astype(np.int32)
    refl = np.concatenate([refl_outer, np.ones(len(SI), bool)]).astype(np.uint8)
    lo, hi = square_bounding_box(V)
    u = E.function(A.ZG_FN_CUBIC_HARMONIC, A.ZG_EVAL_VALUE, (1.0,))
    h = E.function(A.ZG_FN_CUBIC_HARMONIC, A.ZG_EVAL_NORMAL_DERIVATIVE, (1.0,))
    pde = E.make_pde(source=E.zero(), dirichlet=u, robin=h, is_source_constant=True)
    ws = E.walk_settings_host(eps, eps, max_len, True)
    prob = Problem("mixed2d", 2, Vall, Iall, refl, lo, hi, pde, ws, n_walks, exact_fn(cubic_harmonic),
                   solve_double_sided=True)
    prob.meta["outer"] = (V, I)
    prob.meta["slits"] = slits
    return prob


def mixed2d_robin(kappa=2.0, n_segments=1024, n_slits=8, seed=0x5EED, n_walks=256, eps=1e-3,
                  max_len=16384) -> Problem:
    """mixed2d with Robin conditions du/dn + kappa u = h on every reflecting segment (the upper
    half of the star curve and both sides of the slits), u = x^3 - 3xy^2.  meta['robin'] holds
    the per-primitive coefficient bounds (kappa on reflecting primitives)."""
    p = mixed2d(n_segments=n_segments, n_slits=n_slits, seed=seed, n_walks=n_walks, eps=eps, max_len=max_len)
    h = E.function(A.ZG_FN_CUBIC_HARMONIC, A.ZG_EVAL_ROBIN, (1.0, 0.0, 0.0, 0.0, 0.0, kappa))
    p.pde = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=h, is_source_constant=True,
                       robin_coeff=E.constant(kappa))
    k = np.where(p.is_reflecting.astype(bool), np.float32(kappa), np.float32(0.0)).astype(np.float32)
    p.meta["robin"] = (k, k)
    p.meta["kappa"] = kappa
    p.name = "mixed2d_robin"
    return p


def lprism3d(n_sub=4, n_z=4, plate=True, n_walks=256, eps=1e-3, max_len=16384) -> Problem:
    """L-shaped prism ([-1, 1]^2 minus the quadrant (0, 1]^2) x [-1, 1] with outward normals,
    u = e^x cos y (harmonic in 3D).  Neumann data on the two re-entrant walls (x = 0 and y = 0
    inside the notch) and on the top cap z = 1, Dirichlet elsewhere; with `plate` a
    double-sided Neumann square z = -0.3, x, y in [-0.8, -0.3] inside the domain (an open
    surface: its rim edges are silhouettes from everywhere).  The re-entrant edge along
    (0, 0, z) is the concave silhouette candidate of the closed part.  Polygon edges are split
    into n_sub pieces and the height into n_z layers.  meta['inside'] tests interior points."""
    poly = np.array([(-1, -1), (1, -1), (1, 0), (0, 0), (0, 1), (-1, 1)], np.float64)
    ring, edge_of = [], []
    for e in range(6):
        a, b = poly[e], poly[(e + 1) % 6]
        for k in range(n_sub):
            ring.append(a + (b - a) * (k / n_sub))
            edge_of.append(e)
    ring = np.array(ring)
    nb = len(ring)
    zs = np.linspace(-1.0, 1.0, n_z + 1)
    V = np.concatenate([np.column_stack([ring, np.full(nb, z)]) for z in zs])
    tris, refl = [], []
    notch = {2, 3}                     # polygon edges (1,0)->(0,0) and (0,0)->(0,1)
    for k in range(n_z):
        for i in range(nb):
            j = (i + 1) % nb
            a0, b0, a1, b1 = k * nb + i, k * nb + j, (k + 1) * nb + i, (k + 1) * nb + j
            r = edge_of[i] in notch
            tris += [(a0, b0, b1), (a0, b1, a1)]
            refl += [r, r]
    c = 3 * n_sub                      # ring index of the re-entrant vertex (0, 0)
    for i in range(nb):
        j = (i + 1) % nb
        if edge_of[i] in notch:        # segments on the lines through c: degenerate fan triangles
            continue
        top, bot = n_z * nb, 0
        tris.append((top + c, top + i, top + j)); refl.append(True)
        tris.append((bot + c, bot + j, bot + i)); refl.append(False)
    if plate:
        m = 2 * n_sub
        g = np.linspace(-0.8, -0.3, m + 1)
        base = len(V)
        PV = np.array([(x, y, -0.3) for y in g for x in g])
        V = np.concatenate([V, PV])
        for yi in range(m):
            for xi in range(m):
                v00 = base + yi * (m + 1) + xi
                v10, v01, v11 = v00 + 1, v00 + m + 1, v00 + m + 2
                tris += [(v00, v10, v11), (v00, v11, v01)]
                refl += [True, True]
    V = V.astype(np.float32)
    I = np.array(tris, np.int32)
    refl = np.array(refl, np.uint8)
    lo = np.full(3, -1.0 - 1e-3, np.float32)
    hi = np.full(3, 1.0 + 1e-3, np.float32)
    u = E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_VALUE, (1.0, 1.0))
    h = E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_NORMAL_DERIVATIVE, (1.0, 1.0))
    pde = E.make_pde(source=E.zero(), dirichlet=u, robin=h, is_source_constant=True)
    ws = E.walk_settings_host(eps, eps, max_len, plate)
    prob = Problem("lprism3d", 3, V, I, refl, lo, hi, pde, ws, n_walks, exact_fn(exp_cos), solve_double_sided=plate)

    def inside(x, margin=0.0):
        x = np.asarray(x)
        box = np.all(np.abs(x) < 1.0 - margin, axis=1)
        return box & ~((x[:, 0] > -margin) & (x[:, 1] > -margin))
    prob.meta["inside"] = inside
    return prob


def lprism3d_robin(kappa=2.0, plate=False, n_walks=256, eps=1e-3, max_len=16384) -> Problem:
    """lprism3d with Robin conditions du/dn + kappa u = h on every reflecting triangle,
    u = e^x cos y.  meta['robin'] holds the per-primitive coefficient bounds (kappa on
    reflecting primitives)."""
    p = lprism3d(plate=plate, n_walks=n_walks, eps=eps, max_len=max_len)
    h = E.function(A.ZG_FN_EXP_COS, A.ZG_EVAL_ROBIN, (1.0, 1.0, 0.0, 0.0, 0.0, kappa))
    p.pde = E.make_pde(source=E.zero(), dirichlet=p.pde.dirichlet, robin=h, is_source_constant=True,
                       robin_coeff=E.constant(kappa))
    k = np.where(p.is_reflecting.astype(bool), np.float32(kappa), np.float32(0.0)).astype(np.float32)
    p.meta["robin"] = (k, k)
    p.meta["kappa"] = kappa
    p.name = "lprism3d_robin"
    return p


def lprism3d_screened(sigma=4.0, kappa=None, plate=False, n_walks=256, eps=1e-3, max_len=16384) -> Problem:
    """lprism3d for the screened equation -Lap u + sigma u = 0 with u = exp(a . x),
    a = sqrt(sigma) (0.48, 0.6, 0.64): Neumann data du/dn on the reflecting triangles, or with
    kappa Robin data du/dn + kappa u (meta['robin'] then holds the coefficient bounds)."""
    p = lprism3d(plate=plate, n_walks=n_walks, eps=eps, max_len=max_len)
    a = tuple(float(v) for v in math.sqrt(sigma) * np.array([0.48, 0.6, 0.64]))
    u = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, *a))
    if kappa is None:
        h = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_NORMAL_DERIVATIVE, (1.0, *a))
        p.pde = E.make_pde(source=E.zero(), dirichlet=u, robin=h, absorption=sigma, is_source_constant=True)
    else:
        h = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_ROBIN, (1.0, *a, 0.0, kappa))
        p.pde = E.make_pde(source=E.zero(), dirichlet=u, robin=h, absorption=sigma, is_source_constant=True,
                           robin_coeff=E.constant(kappa))
        k = np.where(p.is_reflecting.astype(bool), np.float32(kappa), np.float32(0.0)).astype(np.float32)
        p.meta["robin"] = (k, k)
    p.exact = exact_fn(exp_linear, a=np.array(a))
    p.meta["a"] = a
    p.name = "lprism3d_screened"
    return p


def sphere_sdf_grid(dim=3, res=64, radius=1.0, margin=0.2):
    """Signed distance |x| - radius of a ball sampled at the cell centres of a res^dim grid over
    [-(radius + margin), radius + margin]^dim (DenseGrid layout, last axis fastest)."""
    h = radius + margin
    c = (np.arange(res) + 0.5) / res * 2 * h - h
    mesh = np.meshgrid(*([c] * dim), indexing="ij")
    r = np.sqrt(sum(m.astype(np.float64) ** 2 for m in mesh))
    return (r - radius).astype(np.float32), np.full(dim, -h, np.float32), np.full(dim, h, np.float32)


def _segments_close(p0, p1, q0, q1, tol):
    def seg_dist(a, b, c):
        ab = b - a
        t = np.clip(np.dot(c - a, ab) / max(np.dot(ab, ab), 1e-30), 0, 1)
        return np.linalg.norm(a + t * ab - c)
    d = min(seg_dist(p0, p1, q0), seg_dist(p0, p1, q1), seg_dist(q0, q1, p0), seg_dist(q0, q1, p1))
    # proper intersection test
    def orient(a, b, c):
        return np.sign((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))
    inter = orient(p0, p1, q0) != orient(p0, p1, q1) and orient(q0, q1, p0) != orient(q0, q1, p1)
    return inter or d < tol


def screened_instance(n_tris: int, sigma: float, direction, seed: int, n_walks=128, eps=1e-4,
                      max_len=1024) -> Problem:
    freq = max(1, int(round(math.sqrt(n_tris / 20.0))))
    V, F = geodesic_sphere(freq)
    V = sh_displace(V, seed)
    lo, hi = V.min(0) - 1e-3, V.max(0) + 1e-3
    a = math.sqrt(sigma) * np.asarray(direction, np.float64) / np.linalg.norm(direction)
    u = E.function(A.ZG_FN_EXP_LINEAR, A.ZG_EVAL_VALUE, (1.0, a[0], a[1], a[2]))
    pde = E.make_pde(source=E.zero(), dirichlet=u, absorption=sigma, is_source_constant=True)
    ws = E.walk_settings_host(eps, eps, max_len, False)
    prob = Problem("screened", 3, V.astype(np.float32), F.astype(np.int32), None,
                   lo.astype(np.float32), hi.astype(np.float32), pde, ws, n_walks, exact_fn(exp_linear, a=a))
    prob.meta["sigma"] = sigma
    prob.meta["a"] = a
    return prob


def screened_batch_spec(n_instances: int = 256, seed: int = 0x5EED, tri_range=(2000, 200000),
                        point_range=(1000, 100000), sigma_max: float = 10.0):
    """Instance parameters of the batched screened-Poisson workload: triangle and point
    counts log-uniform over the ranges, sigma ~ U[0, sigma_max], a uniform direction for
    a = sqrt(sigma) * direction.  Returns dicts (n_tris, n_points, sigma, direction, seed)."""
    rng = np.random.default_rng(seed)
    lt = rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]), n_instances)
    lp = rng.uniform(math.log(point_range[0]), math.log(point_range[1]), n_instances)
    sig = rng.uniform(0.0, sigma_max, n_instances)
    dirs = rng.normal(size=(n_instances, 3))
    return [dict(n_tris=int(round(math.exp(lt[i]))), n_points=int(round(math.exp(lp[i]))), sigma=float(sig[i]),
                 direction=dirs[i], seed=int(seed + 1 + i)) for i in range(n_instances)]


def engine_wost(fixture, device="cuda") -> Problem:
    """The reference's engine demo problem (demo/model_problems/engine/wost.json), set up as its
    demo does (demo/model_problem.h:129-328, demo/demo.cpp:94-146, demo/grid.h:23-40) from the
    data in `fixture` (an .npz written by tests/golden/make_engine_fixture.py):

    - the OBJ polyline normalised (centroid at the origin, largest vertex distance 1,
      fcpw_geometric_queries.h:255-273) with every segment's orientation flipped (:284-289);
      the domain is its square bounding box (:291-317);
    - a segment is reflecting (Neumann) when the reflecting-boundary image is > 0 at its
      midpoint (partitionBoundaryMesh, :459-509); the image lookup is Image::get(uv) with
      uv = (x - bMin) / maxLength: the cell containing uv, rows flipped (demo/image.h:103-110);
    - Dirichlet data: the absorbing-value image under the same lookup, as a nearest-cell dense
      grid over [bMin, bMin + maxLength]^2; pure Neumann reflecting conditions (robinCoeff 0);
      source and reflecting data ignored (ignoreSourceContribution, ignoreReflecting...
      Contribution in wost.json);
    - walk settings of wost.json: 128 walks, epsilon 1e-3 on both boundaries, silhouette
      precision 1e-3, Russian roulette 0.99, splitting 1.5, max length 1024.

    meta: 'grid' [res*res, 2] points (index i*res + j at x index i, y index j), 'solution'
    [res, res] the reference's result indexed [x index, y index], 'mask_distance' (pixels
    closer than it to the boundary are written as 0), 'dirichlet_grid' (keeps the device
    values alive), 'reference_walks' (128)."""
    z = np.load(fixture, allow_pickle=False)
    cfg = dict(zip([str(n) for n in z["setting_names"]], z["settings"].tolist()))
    V = z["positions"].astype(np.float64)
    V = V - V.mean(0)
    V = (V / np.linalg.norm(V, axis=1).max()).astype(np.float32)
    I = z["segments"][:, ::-1].astype(np.int32).copy()          # flipOrientation
    lo, hi = square_bounding_box(V)
    max_len = np.float32((hi - lo).max())

    def lookup_index(img_shape, x):
        h, w = img_shape
        uv = (x.astype(np.float32) - lo) / max_len
        i = np.clip((uv[:, 1] * np.float32(h)).astype(np.int64), 0, h - 1)
        j = np.clip((uv[:, 0] * np.float32(w)).astype(np.int64), 0, w - 1)
        return h - i - 1, j

    mask_shape = tuple(int(v) for v in z["reflecting_mask_shape"])
    mask = np.unpackbits(z["reflecting_mask_bits"])[: mask_shape[0] * mask_shape[1]].reshape(mask_shape).astype(bool)
    mid = 0.5 * (V[I[:, 0]] + V[I[:, 1]])
    r, c = lookup_index(mask_shape, mid)
    refl = mask[r, c].astype(np.uint8)

    g = z["absorbing_value"]
    # grid[x cell, y cell] = image buffer[h - 1 - y cell, x cell]
    grid_vals = np.ascontiguousarray(g[::-1, :].T)
    dirichlet = E.GridFunction(grid_vals, lo, lo + max_len, interpolate=False, device=device)
    pde = E.make_pde(source=E.zero(), dirichlet=dirichlet.fn, robin=E.zero(), is_source_constant=False)
    ws = E.walk_settings_host(cfg["epsilonShellForAbsorbingBoundary"], cfg["epsilonShellForReflectingBoundary"],
                              int(cfg["maxWalkLength"]), bool(cfg["solveDoubleSided"]),
                              russian_roulette_threshold=cfg["russianRouletteThreshold"],
                              splitting_threshold=cfg["splittingThreshold"],
                              ignore_absorbing_boundary_contribution=int(cfg["ignoreAbsorbingBoundaryContribution"]),
                              ignore_reflecting_boundary_contribution=int(cfg["ignoreReflectingBoundaryContribution"]),
                              ignore_source_contribution=int(cfg["ignoreSourceContribution"]))
    res = int(cfg["gridRes"])
    ext = hi - lo
    ii, jj = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    grid = np.stack([(ii / np.float32(res)).astype(np.float32) * ext[0] + lo[0],
                     (jj / np.float32(res)).astype(np.float32) * ext[1] + lo[1]], -1).reshape(-1, 2).astype(np.float32)
    sol = z["solution"]   # file rows: row h-1-j holds y index j, column i holds x index i
    prob = Problem("engine_wost", 2, V, I, refl, lo, hi, pde, ws, int(cfg["nWalks"]), None)
    prob.meta.update(grid=grid, solution=np.ascontiguousarray(sol[::-1, :].T), dirichlet_grid=dirichlet,
                     mask_distance=float(cfg["boundaryDistanceMask"]), reference_walks=int(cfg["nWalks"]),
                     watertight=bool(cfg["domainIsWatertight"]))
    return prob


def sample_box(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, len(lo))).astype(np.float32)
