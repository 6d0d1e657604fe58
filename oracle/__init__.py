"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/_build/libzombie_oracle.so, the C++ restatement of the reference
estimator (see zombie_oracle.cpp for the file:line map).  Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from zombie_amd import _abi as A

LIB = Path(__file__).resolve().parent / "_build" / "libzombie_oracle.so"
_lib = None

RNG_PER_POINT = 0   # the reference's scheme: one generator per point, consumed across walks
RNG_PER_WALK = 1    # the GPU's scheme: one generator per walk
RNG_LOCKSTEP = 2    # counter-example only: per-walk increments on one initstate (correlated walks)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        L.orc_pcg32_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, P]
        L.orc_pcg32_f32.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, P]
        L.orc_splitmix64_mix.argtypes = [C.c_uint64]
        L.orc_splitmix64_mix.restype = C.c_uint64
        L.orc_walk_init_state.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_walk_init_state.restype = C.c_uint64
        L.orc_walk_floats.argtypes = [C.c_uint64, P, P, C.c_int64, C.c_int32, P]
        L.orc_pcg32_default_u32.argtypes = [C.c_int64, P]
        L.orc_sample_unit_sphere.argtypes = [C.c_int, C.c_uint64, C.c_int64, P]
        L.orc_sample_ball_radius.argtypes = [C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_int64, P, P]
        L.orc_greens_ball.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, P]
        L.orc_offset_point.argtypes = [C.c_int, P, P, P]
        L.orc_eval_function.argtypes = [C.c_int, C.POINTER(A.zg_function), P, P, C.c_int, C.c_float]
        L.orc_eval_function.restype = C.c_float
        L.orc_scene_create.argtypes = [C.c_int, P, C.c_int64, P, C.c_int64, P, P, P, C.c_int, C.c_int, C.c_int, C.c_float]
        L.orc_scene_create.restype = P
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_set_robin.argtypes = [P, P, P]
        L.orc_robin_radius.argtypes = [P, C.c_int64, P, P, C.c_int, P]
        L.orc_scene_stats.argtypes = [P, P]
        L.orc_closest_point.argtypes = [P, C.c_int, C.c_int64, P, P, P, P, P, C.c_int, P, C.c_int]
        L.orc_intersect_ray.argtypes = [P, C.c_int, C.c_int64, P, P, P, P, P, P, C.c_int]
        L.orc_inside_domain.argtypes = [P, C.c_int64, P, P, C.c_int]
        L.orc_star_radius.argtypes = [P, C.c_int64, P, P, P, C.c_float, P, P]
        L.orc_star_radius_bvh.argtypes = [P, C.c_int64, P, P, P, C.c_float, P, P]
        L.orc_sample_reflecting_bvh.argtypes = [P, P, C.c_float, P, P, P, P]
        L.orc_line_of_sight_bvh.argtypes = [P, P, P]
        L.orc_silhouette_candidates.argtypes = [P, P, C.c_float, C.c_int, P, P]
        L.orc_silhouette_candidates.restype = C.c_int64
        L.orc_sample_reflecting.argtypes = [P, P, C.c_float, P, P, P, P]
        L.orc_sample_reflecting.restype = C.c_int
        L.orc_line_of_sight.argtypes = [P, P, P]
        L.orc_line_of_sight.restype = C.c_int
        L.orc_solve.argtypes = [C.c_int, P, C.POINTER(A.zg_pde), C.POINTER(A.zg_walk_settings), C.c_int64, P, P, P, P,
                                P, P, C.c_uint64, C.c_uint32, C.c_int, C.c_int, P, P, P, P, P, P, P]
        L.orc_solve.restype = C.c_int
        L.orc_solve_grad.argtypes = [C.c_int, P, C.POINTER(A.zg_pde), C.POINTER(A.zg_walk_settings), C.c_int64, P, P,
                                     P, P, P, P, P, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, P, P, P, P, P,
                                     P, P, P, P, P, P, P]
        L.orc_solve_grad.restype = C.c_int
        L.orc_sdf_solve.argtypes = [C.c_int, P, P, P, P, P, P, C.POINTER(A.zg_pde), C.POINTER(A.zg_walk_settings),
                                    C.c_int64, P, P, P, P, C.c_uint64, C.c_uint32, C.c_int, C.c_int, P, P, P, P, P]
        L.orc_sdf_solve.restype = C.c_int
        L.orc_sdf_query.argtypes = [C.c_int, P, P, P, P, C.c_int64, P, P, P, P]
        L.orc_sdf_ray.argtypes = [C.c_int, P, P, P, P, C.c_int64, P, P, P, P, P, P, P]
        L.orc_sdf_solve_grad.argtypes = [C.c_int, P, P, P, P, P, P, C.POINTER(A.zg_pde), C.POINTER(A.zg_walk_settings),
                                         C.c_int64, P, P, P, P, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                         P, P, P, P, P, P, P, P, P]
        L.orc_sdf_solve_grad.restype = C.c_int
        L.orc_bvc_splat.argtypes = [C.c_int, C.c_float, C.POINTER(A.zg_bvc_samples), C.POINTER(A.zg_bvc_eval_points),
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, P, P, P, P, P, C.c_int]
        L.orc_free_space_kernels.argtypes = [C.c_int, P, P, P, C.c_float, C.c_float, C.c_float, P]
        L.orc_yukawa_ball.argtypes = [C.c_float, C.c_float, C.c_float, P]
        L.orc_stratified_samples.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_uint64, C.c_uint32, P]
        L.orc_pcg32_bounded.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64, P]
        L.orc_greens_ball_gradient.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, P]
        L.orc_gradient_direction.argtypes = [C.c_int, C.c_int, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def pcg32_u32(initstate, initseq, n):
    out = np.empty(n, np.uint32)
    lib().orc_pcg32_u32(initstate, initseq, n, _p(out))
    return out


def pcg32_f32(initstate, initseq, n):
    out = np.empty(n, np.float32)
    lib().orc_pcg32_f32(initstate, initseq, n, _p(out))
    return out


def splitmix64_mix(z):
    return int(lib().orc_splitmix64_mix(z & (2 ** 64 - 1)))


def walk_init_state(seed, p, w):
    """initstate of the generator of walk w (stream offset included) of point p; its stream is p"""
    return int(lib().orc_walk_init_state(seed & (2 ** 64 - 1), p, w & (2 ** 32 - 1)))


def walk_floats(seed, p, w, n_draws):
    """[len(p), n_draws] first nextFloat() values of the generators of walks (p[i], w[i])"""
    p = np.ascontiguousarray(p, np.uint32)
    w = np.ascontiguousarray(w, np.uint32)
    out = np.empty((len(p), n_draws), np.float32)
    lib().orc_walk_floats(seed & (2 ** 64 - 1), _p(p), _p(w), len(p), n_draws, _p(out))
    return out


def pcg32_default_u32(n):
    out = np.empty(n, np.uint32)
    lib().orc_pcg32_default_u32(n, _p(out))
    return out


def pcg32_bounded(initstate, initseq, bound, n):
    out = np.empty(n, np.uint32)
    lib().orc_pcg32_bounded(initstate, initseq, bound, n, _p(out))
    return out


def stratified_samples(dims, n_tables, n_samples, seed, stream_offset=0):
    out = np.empty((n_tables, n_samples, dims), np.float32)
    lib().orc_stratified_samples(dims, n_tables, n_samples, seed, stream_offset, _p(out))
    return out


def greens_ball_gradient(dim, R, lam, r):
    out = np.empty(2, np.float32)
    lib().orc_greens_ball_gradient(dim, R, lam, r, _p(out))
    return dict(gradient_norm=float(out[0]), poisson_kernel_gradient=float(out[1]))


def gradient_direction(dim, cosine, u, n=None):
    """Boundary direction of the gradient path from uniforms u (dim-1 values, plus the sign
    coin when cosine); returns (direction, pdf)."""
    u = np.ascontiguousarray(u, np.float32)
    nn = np.zeros(dim, np.float32) if n is None else np.ascontiguousarray(n, np.float32)
    out = np.empty(dim + 1, np.float32)
    lib().orc_gradient_direction(dim, 1 if cosine else 0, _p(u), _p(nn), _p(out))
    return out[:dim].copy(), float(out[dim])


class SdfGrid:
    """SDF-grid Dirichlet geometry (values [shape...] cell-centred over [lo, hi], last axis fastest)."""

    def __init__(self, values, lo, hi, domain_min=None, domain_max=None):
        self.values = np.ascontiguousarray(values, np.float32)
        self.dim = self.values.ndim
        self.shape = np.ascontiguousarray(self.values.shape, np.int32)
        self.lo = np.ascontiguousarray(lo, np.float32)
        self.hi = np.ascontiguousarray(hi, np.float32)
        self.dmin = np.ascontiguousarray(self.lo if domain_min is None else domain_min, np.float32)
        self.dmax = np.ascontiguousarray(self.hi if domain_max is None else domain_max, np.float32)

    def query(self, pts):
        pts = np.ascontiguousarray(pts, np.float32)
        n = len(pts)
        d = np.empty(n, np.float32); pr = np.empty_like(pts); nr = np.empty_like(pts)
        lib().orc_sdf_query(self.dim, _p(self.values), _p(self.shape), _p(self.lo), _p(self.hi), n, _p(pts), _p(d),
                            _p(pr), _p(nr))
        return d, pr, nr


    def ray(self, o, d, tmax):
        """intersectZeroLevelSet: hit, t, point, normal per ray"""
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = len(o)
        tm = np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float32), n), np.float32)
        hit = np.zeros(n, np.uint8)
        t = np.zeros(n, np.float32)
        pt = np.zeros((n, self.dim), np.float32)
        nr = np.zeros((n, self.dim), np.float32)
        lib().orc_sdf_ray(self.dim, _p(self.values), _p(self.shape), _p(self.lo), _p(self.hi), n, _p(o), _p(d), _p(tm),
                          _p(hit), _p(t), _p(pt), _p(nr))
        return hit.astype(bool), t, pt, nr
    def solve_grad(self, pde, ws, pts, dist, n_walks, stats: "GradStats", seed=0x5EED, stream_offset=0,
                   rng_mode=RNG_PER_WALK, cv_mode=None, direction=None, threads=None):
        """estimateSolutionAndGradient at every point, walks continued over the SDF geometry
        (cv_mode: CV_STAGED, the GPU's control-variate stages, by default)"""
        cv_mode = CV_STAGED if cv_mode is None else cv_mode
        pts = np.ascontiguousarray(pts, np.float32)
        n = len(pts)
        dist = np.ascontiguousarray(dist, np.float32)
        nw = np.ascontiguousarray(np.broadcast_to(np.asarray(n_walks, np.int32), n), np.int32)
        d = None if direction is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(direction, np.float32), (n, self.dim)), np.float32)
        status = lib().orc_sdf_solve_grad(self.dim, _p(self.values), _p(self.shape), _p(self.lo), _p(self.hi), _p(self.dmin),
                                 _p(self.dmax), C.byref(pde), C.byref(ws), n, _p(pts), _p(dist), _p(nw), _p(d), seed,
                                 stream_offset, rng_mode, cv_mode, threads or default_threads(), _p(stats.n),
                                 _p(stats.mean), _p(stats.m2), _p(stats.walk_length), _p(stats.n_grad),
                                 _p(stats.grad_mean), _p(stats.grad_m2), _p(stats.first_source), _p(stats.derivative))
        if status != 0:
            raise RuntimeError(f"oracle SDF gradient solve failed with status {status}")

    def solve(self, pde, ws, pts, dist, n_walks, stats: "Stats", seed=0x5EED, stream_offset=0, rng_mode=RNG_PER_WALK,
              threads=None, types=None):
        pts = np.ascontiguousarray(pts, np.float32)
        n = len(pts)
        dist = np.ascontiguousarray(dist, np.float32)
        nw = np.ascontiguousarray(np.broadcast_to(np.asarray(n_walks, np.int32), n), np.int32)
        types = None if types is None else np.ascontiguousarray(types, np.uint8)
        rep = np.zeros(5, np.int64)
        status = lib().orc_sdf_solve(self.dim, _p(self.values), _p(self.shape), _p(self.lo), _p(self.hi), _p(self.dmin),
                            _p(self.dmax), C.byref(pde), C.byref(ws), n, _p(pts), _p(types), _p(dist), _p(nw), seed,
                            stream_offset, rng_mode, threads or default_threads(), _p(stats.n), _p(stats.mean),
                            _p(stats.m2), _p(stats.walk_length), _p(rep))
        if status != 0:
            raise RuntimeError(f"oracle SDF solve failed with status {status}")
        return dict(codes=rep[:4].copy(), total_steps=int(rep[4]))


def free_space_kernels(dim, x, y, n, r, kreg=0.0, absorption=0.0):
    """[G, P, dG, dP, G_reg, P_reg] of the free-space kernels (pole x, sample y, normal n): harmonic,
    or in 3D with absorption > 0 the Yukawa kernels."""
    out = np.zeros(2 + 2 * dim + 2, np.float32)
    f = lambda a: np.ascontiguousarray(a, np.float32)
    xx, yy, nn = f(x), f(y), f(n)
    lib().orc_free_space_kernels(dim, _p(xx), _p(yy), _p(nn), r, kreg, absorption, _p(out))
    return dict(G=float(out[0]), P=float(out[1]), dG=out[2:2 + dim].copy(), dP=out[2 + dim:2 + 2 * dim].copy(),
                G_reg=float(out[2 + 2 * dim]), P_reg=float(out[3 + 2 * dim]))


def yukawa_ball(R, lam, r):
    """The oracle's 3D Yukawa ball terms: norm, dampening(r), gradientNorm(r), Poisson-kernel
    gradient scale, G(r), P (centred)."""
    out = np.zeros(6, np.float32)
    lib().orc_yukawa_ball(R, lam, r, _p(out))
    return dict(norm=float(out[0]), dampening=float(out[1]), gradient_norm=float(out[2]),
                pk_gradient_scale=float(out[3]), G=float(out[4]), P=float(out[5]))


class BvcStats:
    def __init__(self, n, dim):
        self.dim = dim
        self.n = np.zeros((n, 5), np.int64)
        self.mean = np.zeros((n, 5), np.float64)
        self.m2 = np.zeros((n, 5), np.float64)
        self.grad_mean = np.zeros((n, 5, dim), np.float64)
        self.grad_m2 = np.zeros((n, 5, dim), np.float64)

    def solution(self):
        return self.mean.sum(1)

    def gradient(self):
        return self.grad_mean.sum(1)


def bvc_splat(dim, samples: dict, evals: dict, stats: BvcStats, radius_clamp=1e-3, kernel_regularization=0.0,
              robin_cutoff=np.inf, cutoff_abs=0.0, cutoff_refl=0.0, threads=None, absorption=0.0):
    """samples: pt, normal, type, normal_aligned, pdf, solution, normal_derivative, contribution, robin_coeff
    (numpy; the last two and normal_aligned optional); evals: pt, type, dist_abs, dist_refl."""
    keep = []
    def arr(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data
    S = A.zg_bvc_samples()
    S.n = len(samples["pt"])
    S.pt = arr(samples["pt"], np.float32); S.normal = arr(samples["normal"], np.float32)
    S.type = arr(samples["type"], np.uint8); S.normal_aligned = arr(samples.get("normal_aligned"), np.uint8)
    S.pdf = arr(samples["pdf"], np.float32); S.solution = arr(samples["solution"], np.float32)
    S.normal_derivative = arr(samples["normal_derivative"], np.float32)
    S.contribution = arr(samples.get("contribution"), np.float32)
    S.robin_coeff = arr(samples.get("robin_coeff"), np.float32)
    Ev = A.zg_bvc_eval_points()
    Ev.n = len(evals["pt"])
    Ev.pt = arr(evals["pt"], np.float32); Ev.type = arr(evals.get("type"), np.uint8)
    Ev.dist_to_absorbing_boundary = arr(evals["dist_abs"], np.float32)
    Ev.dist_to_reflecting_boundary = arr(evals.get("dist_refl"), np.float32)
    st = lib().orc_bvc_splat(dim, absorption, C.byref(S), C.byref(Ev), radius_clamp, kernel_regularization,
                             float(min(robin_cutoff, 3.0e38)), cutoff_abs, cutoff_refl, _p(stats.n), _p(stats.mean),
                             _p(stats.m2), _p(stats.grad_mean), _p(stats.grad_m2), threads or default_threads())
    if st != 0:
        raise RuntimeError(f"oracle splat refused its arguments (status {st})")
    return stats


def sample_unit_sphere(dim, seed, n):
    out = np.empty((n, dim), np.float32)
    lib().orc_sample_unit_sphere(dim, seed, n, _p(out))
    return out


def sample_ball_radius(dim, R, lam, seed, n):
    r = np.empty(n, np.float32)
    pdf = np.empty(n, np.float32)
    lib().orc_sample_ball_radius(dim, R, lam, seed, n, _p(r), _p(pdf))
    return r, pdf


def greens_ball(dim, R, lam, r):
    out = np.empty(4, np.float32)
    lib().orc_greens_ball(dim, R, lam, r, _p(out))
    return dict(norm=float(out[0]), poisson_kernel=float(out[1]), G=float(out[2]), dampening=float(out[3]))


def offset_point(p, n):
    p = np.ascontiguousarray(p, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.empty_like(p)
    lib().orc_offset_point(len(p), _p(p), _p(n), _p(out))
    return out


def eval_function(dim, f, x, n=None, aligned=False, absorption=0.0):
    x = np.ascontiguousarray(x, np.float32)
    nn = None if n is None else np.ascontiguousarray(n, np.float32)
    return lib().orc_eval_function(dim, C.byref(f), _p(x), _p(nn), 1 if aligned else 0, absorption)


def default_threads():
    return max(1, min(os.cpu_count() or 1, 16))


class Scene:
    def __init__(self, positions, indices, is_reflecting=None, domain_min=None, domain_max=None,
                 leaf_size=4, buckets=8, solve_double_sided=False, silhouette_precision=1e-3):
        self.V = np.ascontiguousarray(positions, np.float32)
        self.I = np.ascontiguousarray(indices, np.int32)
        self.dim = self.V.shape[1]
        self.refl = None if is_reflecting is None else np.ascontiguousarray(is_reflecting, np.uint8)
        self.dmin = np.ascontiguousarray(domain_min if domain_min is not None else self.V.min(0) - 1, np.float32)
        self.dmax = np.ascontiguousarray(domain_max if domain_max is not None else self.V.max(0) + 1, np.float32)
        self.h = lib().orc_scene_create(self.dim, _p(self.V), len(self.V), _p(self.I), len(self.I), _p(self.refl),
                                        _p(self.dmin), _p(self.dmax), leaf_size, buckets,
                                        1 if solve_double_sided else 0, silhouette_precision)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def set_robin(self, min_coeff, max_coeff):
        """Per-primitive Robin coefficient bounds (arrays over all primitives; None clears)."""
        if min_coeff is None:
            lib().orc_scene_set_robin(self.h, None, None)
            return
        self._rmin = np.ascontiguousarray(min_coeff, np.float32)
        self._rmax = np.ascontiguousarray(max_coeff, np.float32)
        lib().orc_scene_set_robin(self.h, _p(self._rmin), _p(self._rmax))

    def robin_radius(self, pts, max_r, brute=False):
        pts = np.ascontiguousarray(pts, np.float32)
        mr = np.ascontiguousarray(np.broadcast_to(np.asarray(max_r, np.float32), len(pts)), np.float32)
        out = np.empty(len(pts), np.float32)
        lib().orc_robin_radius(self.h, len(pts), _p(pts), _p(mr), 1 if brute else 0, _p(out))
        return out

    def stats(self):
        out = np.zeros(5, np.int64)
        lib().orc_scene_stats(self.h, _p(out))
        return dict(abs_nodes=int(out[0]), abs_depth=int(out[1]), refl_nodes=int(out[2]), refl_depth=int(out[3]),
                    silhouette_candidates=int(out[4]))

    def closest_point(self, pts, boundary=0, brute=False, threads=None):
        pts = np.ascontiguousarray(pts, np.float32)
        n = len(pts)
        d = np.empty(n, np.float32)
        cp = np.empty((n, self.dim), np.float32)
        nr = np.empty((n, self.dim), np.float32)
        pr = np.empty(n, np.int32)
        visits = np.zeros(2, np.int64)
        lib().orc_closest_point(self.h, boundary, n, _p(pts), _p(d), _p(cp), _p(nr), _p(pr), 1 if brute else 0,
                                _p(visits), threads or default_threads())
        return d, cp, nr, pr, visits

    def intersect_ray(self, o, d, tmax, boundary=0, brute=False):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        tmax = np.ascontiguousarray(tmax, np.float32)
        n = len(o)
        hit = np.empty(n, np.uint8)
        t = np.empty(n, np.float32)
        pr = np.empty(n, np.int32)
        lib().orc_intersect_ray(self.h, boundary, n, _p(o), _p(d), _p(tmax), _p(hit), _p(t), _p(pr), 1 if brute else 0)
        return hit, t, pr

    def inside_domain(self, pts, threads=None):
        pts = np.ascontiguousarray(pts, np.float32)
        out = np.empty(len(pts), np.uint8)
        lib().orc_inside_domain(self.h, len(pts), _p(pts), _p(out), threads or default_threads())
        return out.astype(bool)

    def star_radius(self, pts, min_r, max_r, precision=1e-3, flip=None, bvh=False):
        pts = np.ascontiguousarray(pts, np.float32)
        n = len(pts)
        mn = np.ascontiguousarray(np.broadcast_to(min_r, n), np.float32)
        mx = np.ascontiguousarray(np.broadcast_to(max_r, n), np.float32)
        fl = None if flip is None else np.ascontiguousarray(np.broadcast_to(flip, n), np.uint8)
        out = np.empty(n, np.float32)
        fn = lib().orc_star_radius_bvh if bvh else lib().orc_star_radius
        fn(self.h, n, _p(pts), _p(mn), _p(mx), precision, _p(fl), _p(out))
        return out

    def silhouette_candidates(self, x, precision=1e-3, flip=False):
        x = np.ascontiguousarray(x, np.float32)
        n = lib().orc_silhouette_candidates(self.h, _p(x), precision, int(flip), None, None)
        # 2D: the candidate vertices [n, 2]; 3D: the candidate edges' end points [n, 2, 3]
        v = np.empty((n, 2) if self.dim == 2 else (n, 2, 3), np.float32)
        s = np.empty(n, np.uint8)
        lib().orc_silhouette_candidates(self.h, _p(x), precision, int(flip), _p(v), _p(s))
        return v, s.astype(bool)

    def sample_reflecting(self, x, radius, u, bvh=False):
        x = np.ascontiguousarray(x, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        pt = np.empty(self.dim, np.float32)
        nr = np.empty(self.dim, np.float32)
        pdf = np.empty(1, np.float32)
        fn = lib().orc_sample_reflecting_bvh if bvh else lib().orc_sample_reflecting
        ok = fn(self.h, _p(x), radius, _p(u), _p(pt), _p(nr), _p(pdf))
        return bool(ok), pt, nr, float(pdf[0])

    def line_of_sight(self, p1, p2, bvh=False):
        fn = lib().orc_line_of_sight_bvh if bvh else lib().orc_line_of_sight
        a = np.ascontiguousarray(p1, np.float32)   # held until the call returns (_p takes a raw address)
        b = np.ascontiguousarray(p2, np.float32)
        return bool(fn(self.h, _p(a), _p(b)))


class Stats:
    def __init__(self, n):
        self.n = np.zeros(n, np.int64)
        self.mean = np.zeros(n, np.float64)
        self.m2 = np.zeros(n, np.float64)
        self.walk_length = np.zeros(n, np.int64)
        self.splits = np.zeros(n, np.int64)

    def variance(self):
        return self.m2 / np.maximum(self.n - 1, 1)


class GradStats(Stats):
    """Stats plus the gradient members of SampleStatistics."""

    def __init__(self, n, dim):
        super().__init__(n)
        self.dim = dim
        self.n_grad = np.zeros(n, np.int64)
        self.grad_mean = np.zeros((n, dim), np.float64)
        self.grad_m2 = np.zeros((n, dim), np.float64)
        self.first_source = np.zeros(n, np.float64)
        self.derivative = np.zeros(n, np.float64)

    def grad_variance(self):
        return self.grad_m2 / np.maximum(self.n_grad - 1, 1)[:, None]

    def derivative_mean(self):
        return self.derivative / np.maximum(self.n, 1)


CV_PER_PAIR = 0   # the reference: control variates refreshed before every antithetic pair
CV_STAGED = 1     # the GPU: refreshed when a stage [0,1), [1,2), [2,4), ... starts


def solve_grad(scene: Scene, pde, ws, pts, dist, n_walks, stats: GradStats, quantity=None, direction=None,
               seed=0x5EED, stream_offset=0, rng_mode=RNG_PER_WALK, cv_mode=CV_PER_PAIR, threads=None,
               types=None, first_sphere_radius=None, wost=False, dist_reflecting=None):
    """Walk on spheres (or stars) with per-point estimation quantities (gradient points included).
    dist_reflecting: distances to the reflecting boundary (computed when None)."""
    pts = np.ascontiguousarray(pts, np.float32)
    n = len(pts)
    dist = np.ascontiguousarray(dist, np.float32)
    nw = np.ascontiguousarray(np.broadcast_to(np.asarray(n_walks, np.int32), n), np.int32)
    q = np.ascontiguousarray(np.broadcast_to(np.asarray(1 if quantity is None else quantity, np.uint8), n), np.uint8)
    d = None if direction is None else np.ascontiguousarray(np.broadcast_to(np.asarray(direction, np.float32),
                                                                            (n, scene.dim)), np.float32)
    types = None if types is None else np.ascontiguousarray(types, np.uint8)
    rep = np.zeros(8, np.int64)
    dr = None if dist_reflecting is None else np.ascontiguousarray(dist_reflecting, np.float32)
    st = lib().orc_solve_grad(1 if wost else 0, scene.h, C.byref(pde), C.byref(ws), n, _p(pts), _p(types), _p(dist),
                              _p(dr), _p(nw), _p(q),
                              _p(d), seed, stream_offset, rng_mode, cv_mode, threads or default_threads(),
                              _p(stats.n), _p(stats.mean), _p(stats.m2), _p(stats.walk_length), _p(stats.splits),
                              _p(stats.n_grad), _p(stats.grad_mean), _p(stats.grad_m2), _p(stats.first_source),
                              _p(stats.derivative), _p(first_sphere_radius), _p(rep))
    if st != 0:
        raise RuntimeError(f"oracle gradient solve failed with status {st}")
    return dict(codes=rep[:4].copy(), total_steps=int(rep[4]))


def solve(scene: Scene, pde, ws, pts, dist, n_walks, stats: Stats, seed=0x5EED, stream_offset=0,
          rng_mode=RNG_PER_WALK, threads=None, wost=False, types=None, normals=None, aligned=None,
          first_sphere_radius=None):
    """Runs the oracle estimator; returns the report dict."""
    pts = np.ascontiguousarray(pts, np.float32)
    n = len(pts)
    dist = np.ascontiguousarray(dist, np.float32)
    nw = np.ascontiguousarray(np.broadcast_to(np.asarray(n_walks, np.int32), n), np.int32)
    rep = np.zeros(8, np.int64)
    types = None if types is None else np.ascontiguousarray(types, np.uint8)
    normals = None if normals is None else np.ascontiguousarray(normals, np.float32)
    aligned = None if aligned is None else np.ascontiguousarray(aligned, np.uint8)
    st = lib().orc_solve(1 if wost else 0, scene.h, C.byref(pde), C.byref(ws), n, _p(pts), _p(normals), _p(types),
                         _p(aligned), _p(dist), _p(nw), seed, stream_offset, rng_mode, threads or default_threads(),
                         _p(stats.n), _p(stats.mean), _p(stats.m2), _p(stats.walk_length), _p(stats.splits),
                         _p(first_sphere_radius), _p(rep))
    if st != 0:
        raise RuntimeError(f"oracle solve failed with status {st}")
    return dict(codes=rep[:4].copy(), total_steps=int(rep[4]), node_visits=int(rep[5]), prim_visits=int(rep[6]))
