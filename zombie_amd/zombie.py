"""The reference's Python surface over the GPU engine.

The reference binds its C++ solvers with nanobind and exposes them as `zombie.Core`,
`zombie.Solvers` and `zombie.Utils`, with factories that take `dim=` and `channels=`
(python/zombie/__init__.py:28-40,181-196; python/zombie_bindings.h).  This module keeps those
names, argument orders and defaults, so the reference's driver code (demo/demo.py) runs with
`from zombie_amd import zombie` in place of `import zombie`; the solves run on the GPU through
the C-ABI (zg_wos_solve, zg_wost_solve).

Differences a caller can see:
- PDE data must be device-evaluable: constants (Core.get_constant_*_callback), dense grids
  (Utils.get_dense_grid_*_callback) or manufactured fields (zombie_amd.engine.function).  A
  Python callable is accepted where only the host evaluates it (the indicator given to
  Utils.partition_boundary_mesh) and refused by the solvers with a TypeError.
- channels > 1: the GPU walks carry one value; a solve with C channels runs C scalar solves
  with the same seed and stream offsets, hence the same walks, which is what one walk carrying
  C values computes.  Statistics hold numpy arrays of C values.
- Solvers take an explicit seed (the reference seeds from the clock, core/sampling.h:36-41):
  WalkOnSpheres(..., seed=) and the solver's `seed` attribute; successive solve() calls draw
  fresh walk streams.
- Walk-state and terminal-contribution callbacks (host std::functions) are not available on
  the GPU; run_single_threaded and report_progress are accepted and ignored.
- Double-sided dense-grid data (two grids per callback) is not available.
"""
from __future__ import annotations

import enum
import math
import types
from typing import Callable

import numpy as np

from . import _abi as A
from . import engine as E

_VALID_DIMS = {2, 3}
_VALID_CHANNELS = {1, 4, 16, 64, 256}


def _validate_dim(dim):
    if dim not in _VALID_DIMS:
        raise ValueError(f"dim must be one of {sorted(_VALID_DIMS)}, got {dim}")


def _validate_channels(channels):
    if channels not in _VALID_CHANNELS:
        raise ValueError(f"channels must be one of {sorted(_VALID_CHANNELS)}, got {channels}")


# ------------------------------------------------------------------------------------
# list types (zombie_bindings.cpp binds std::vector of these as *List)
# ------------------------------------------------------------------------------------
class BoolList(list): pass
class IntList(list): pass
class UintList(list): pass
class FloatList(list): pass
class Float2List(list): pass
class Float3List(list): pass
class Int2List(list): pass
class Int3List(list): pass


def FloatNList(*args, dim):
    _validate_dim(dim)
    return (Float2List if dim == 2 else Float3List)(*args)


def IntNList(*args, dim):
    _validate_dim(dim)
    return (Int2List if dim == 2 else Int3List)(*args)


def convert_list_to_numpy_array(values):
    return np.asarray(values)


# ------------------------------------------------------------------------------------
# device-evaluable PDE data
# ------------------------------------------------------------------------------------
class Callback:
    """A PDE datum the GPU can evaluate: one zg_function per channel, plus the host-side
    evaluation the set-up code needs (Utils.partition_boundary_mesh)."""

    def __init__(self, fns, host: Callable, keep=()):
        self.fns = list(fns)
        self.host = host
        self._keep = list(keep)

    def __call__(self, x, *args):
        return self.host(np.asarray(x, np.float32), *args)


def _constant(value, channels=1) -> Callback:
    vals = np.broadcast_to(np.asarray(value, np.float32), (channels,))
    return Callback([E.constant(float(v)) for v in vals], lambda x, *a: vals[0] if channels == 1 else vals.copy())


class DenseGrid:
    """DenseGrid<float, CHANNELS, DIM> (utils/dense_grid.h:30-356): cell values on a regular grid
    over [grid_min, grid_max], flat data in row order with the last axis fastest; lookups return
    the value of the containing cell, or bi/trilinear interpolation between cell centres clamped
    to the edge values."""

    def __init__(self, *args, dim, channels=1, enable_interpolation=False):
        _validate_dim(dim)
        self.dim, self.channels = dim, channels
        if len(args) >= 4:
            data, shape, gmin, gmax = args[:4]
            if len(args) > 4:
                enable_interpolation = args[4]
            self.set(data, shape)
        else:
            gmin, gmax = args[:2]
            if len(args) > 2:
                enable_interpolation = args[2]
            self.data = np.zeros((0, channels), np.float32)
            self.shape = np.zeros(dim, np.int32)
        self.origin = np.asarray(gmin, np.float32)
        self.extent = np.asarray(gmax, np.float32) - self.origin
        self.enable_interpolation = bool(enable_interpolation)

    def set(self, data, shape=None):
        d = np.asarray(data, np.float32).reshape(-1, self.channels)
        if shape is not None:
            self.shape = np.asarray(shape, np.int32)
        if d.shape[0] != int(np.prod(self.shape)):
            raise ValueError(f"{d.shape[0]} values for a grid of shape {tuple(self.shape)}")
        self.data = d

    def __call__(self, x):
        x = np.atleast_2d(np.asarray(x, np.float32))
        l = (x - self.origin) / self.extent
        sh = self.shape
        if not self.enable_interpolation:
            idx = np.zeros(len(x), np.int64)
            for a in range(self.dim):
                idx = idx * sh[a] + np.clip(np.floor(l[:, a] * sh[a]).astype(np.int64), 0, sh[a] - 1)
            return self.data[idx]
        raise NotImplementedError("host interpolation of dense grids is not needed by the set-up code")

    def function(self, device="cuda"):
        vals = self.data.reshape(*[int(s) for s in self.shape], self.channels)
        hi = self.origin + self.extent
        grids = [E.GridFunction(vals[..., c], self.origin, hi, interpolate=self.enable_interpolation, device=device)
                 for c in range(self.channels)]
        return grids


def _grid_callback(args, dim, channels, what) -> Callback:
    if len(args) == 1 and isinstance(args[0], DenseGrid):
        grid = args[0]
    elif len(args) == 2 and all(isinstance(a, DenseGrid) for a in args):
        raise NotImplementedError(f"{what}: double-sided dense-grid data (two grids) is not available on the GPU")
    else:
        grid = DenseGrid(*args, dim=dim, channels=channels)
    grids = grid.function()
    host = lambda x, *a: grid(x)[:, 0] if channels == 1 else grid(x)
    return Callback([g.fn for g in grids], host, keep=[grid, *grids])


# ------------------------------------------------------------------------------------
# Core
# ------------------------------------------------------------------------------------
class PDE:
    """PDE<T, DIM> (core/pde.h:19-79) with device-evaluable callbacks."""

    def __init__(self, *, dim, channels=1):
        _validate_dim(dim)
        _validate_channels(channels)
        self.dim, self.channels = dim, channels
        self.source = _constant(0.0, channels)
        self.dirichlet = _constant(0.0, channels)
        self.robin = _constant(0.0, channels)
        self.robin_coeff = _constant(0.0)
        self.has_reflecting_boundary_conditions = None
        self.has_non_zero_robin_coeff = None
        self.is_source_constant = True
        self.are_robin_conditions_pure_neumann = True
        self.are_robin_coeffs_nonnegative = True
        self.absorption_coeff = 0.0

    def _fn(self, name, c):
        v = getattr(self, name)
        if isinstance(v, Callback):
            return v.fns[c if len(v.fns) > 1 else 0]
        if isinstance(v, A.zg_function):
            return v
        if isinstance(v, (int, float)):
            return E.constant(float(v))
        raise TypeError(f"PDE.{name} is a {type(v).__name__}; the GPU solvers need a device-evaluable function "
                        "(Core.get_constant_*_callback, Utils.get_dense_grid_*_callback or engine.function)")

    def channel(self, c: int) -> A.zg_pde:
        """The scalar zg_pde of channel c."""
        robin_coeff = None if self.are_robin_conditions_pure_neumann else self._fn("robin_coeff", 0)
        p = E.make_pde(source=self._fn("source", c), dirichlet=self._fn("dirichlet", c), robin=self._fn("robin", c),
                       absorption=float(self.absorption_coeff), is_source_constant=bool(self.is_source_constant),
                       robin_coeff=robin_coeff)
        p.are_robin_coeffs_nonnegative = 1 if self.are_robin_coeffs_nonnegative else 0
        return p


def get_constant_source_callback(value, *, dim, channels=1):
    _validate_dim(dim)
    return _constant(value, channels)


def get_constant_dirichlet_callback(value, value_boundary_normal_aligned=None, *, dim, channels=1):
    _validate_dim(dim)
    if value_boundary_normal_aligned is not None and value_boundary_normal_aligned != value:
        raise NotImplementedError("different normal-aligned Dirichlet constants are not available on the GPU")
    return _constant(value, channels)


def get_constant_robin_callback(value, value_boundary_normal_aligned=None, *, dim, channels=1):
    return get_constant_dirichlet_callback(value, value_boundary_normal_aligned, dim=dim, channels=channels)


def get_constant_indicator_callback(value, *, dim):
    _validate_dim(dim)
    return lambda x: bool(value)


def get_constant_robin_coefficient_callback(value, value_boundary_normal_aligned=None, *, dim):
    _validate_dim(dim)
    return _constant(value)


class GeometricQueries:
    """GeometricQueries<DIM> (core/geometric_queries.h:59-125): the domain box and the absorbing
    and reflecting boundaries attached by Utils.populate_geometric_queries_for_*; the device scene
    is built on first use and rebuilt when a boundary changes."""

    def __init__(self, domain_is_watertight=True, domain_min=None, domain_max=None, *, dim):
        _validate_dim(dim)
        self.dim = dim
        self.domain_is_watertight = bool(domain_is_watertight)
        self.domain_min = None if domain_min is None else np.asarray(domain_min, np.float32)
        self.domain_max = None if domain_max is None else np.asarray(domain_max, np.float32)
        self._absorbing = None
        self._reflecting = None
        self._scene = None
        self._key = None

    def scene(self, solve_double_sided=False, silhouette_precision=1e-3) -> E.Scene:
        parts = [p for p in (self._absorbing, self._reflecting) if p is not None]
        if not parts:
            raise ValueError("no boundary attached: call Utils.populate_geometric_queries_for_*_boundary")
        key = (id(self._absorbing), id(self._reflecting), bool(solve_double_sided), float(silhouette_precision))
        if self._scene is None or self._key != key:
            V, I, R = [], [], []
            base = 0
            for part, refl in ((self._absorbing, 0), (self._reflecting, 1)):
                if part is None:
                    continue
                V.append(part.positions)
                I.append(part.indices + base)
                R.append(np.full(len(part.indices), refl, np.uint8))
                base += len(part.positions)
            V = np.concatenate(V).astype(np.float32)
            lo = self.domain_min if self.domain_min is not None else V.min(0) - 1.0
            hi = self.domain_max if self.domain_max is not None else V.max(0) + 1.0
            self._scene = E.Scene(V, np.concatenate(I).astype(np.int32), np.concatenate(R), lo, hi,
                                  bvh_width=4, leaf_size=4, solve_double_sided=solve_double_sided,
                                  silhouette_precision=silhouette_precision)
            rb = self._reflecting
            if rb is not None and getattr(rb, "max_coeff", None) is not None:
                n_abs = 0 if self._absorbing is None else len(self._absorbing.indices)
                mn = np.concatenate([np.zeros(n_abs, np.float32), rb.min_coeff]).astype(np.float32)
                mx = np.concatenate([np.zeros(n_abs, np.float32), rb.max_coeff]).astype(np.float32)
                self._scene.set_robin_coefficients(mn, mx)
            self._key = key
        return self._scene

    @property
    def has_non_empty_absorbing_boundary(self):
        return self._absorbing is not None and len(self._absorbing.indices) > 0

    @property
    def has_non_empty_reflecting_boundary(self):
        return self._reflecting is not None and len(self._reflecting.indices) > 0


Core = types.SimpleNamespace(
    PDE=PDE, GeometricQueries=GeometricQueries,
    get_constant_source_callback=get_constant_source_callback,
    get_constant_dirichlet_callback=get_constant_dirichlet_callback,
    get_constant_robin_callback=get_constant_robin_callback,
    get_constant_indicator_callback=get_constant_indicator_callback,
    get_constant_robin_coefficient_callback=get_constant_robin_coefficient_callback,
)


# ------------------------------------------------------------------------------------
# Solvers
# ------------------------------------------------------------------------------------
class SampleType(enum.IntEnum):
    InDomain = A.ZG_SAMPLE_IN_DOMAIN
    OnAbsorbingBoundary = A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY
    OnReflectingBoundary = A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY


class EstimationQuantity(enum.IntEnum):
    Solution = A.ZG_ESTIMATE_SOLUTION
    SolutionAndGradient = A.ZG_ESTIMATE_SOLUTION_AND_GRADIENT
    Skip = A.ZG_ESTIMATE_SKIP


class WalkCompletionCode(enum.IntEnum):
    ReachedAbsorbingBoundary = 0
    TerminatedWithRussianRoulette = 1
    ExceededMaxWalkLength = 2
    EscapedDomain = 3


_WS_FIELDS = ("epsilon_shell_for_absorbing_boundary", "epsilon_shell_for_reflecting_boundary", "silhouette_precision",
              "russian_roulette_threshold", "splitting_threshold", "max_walk_length", "steps_before_applying_tikhonov",
              "steps_before_using_maximal_spheres", "solve_double_sided", "use_gradient_control_variates",
              "use_gradient_antithetic_variates", "use_cosine_sampling_for_derivatives",
              "ignore_absorbing_boundary_contribution", "ignore_reflecting_boundary_contribution",
              "ignore_source_contribution", "print_logs")


class WalkSettings:
    """WalkSettings (core/common.h:19-90) with the bindings' two constructors
    (zombie_bindings.h:397-425): (epsilon_shell_for_absorbing_boundary,
    epsilon_shell_for_reflecting_boundary, max_walk_length, solve_double_sided) or all 16 fields."""

    def __init__(self, *args, **kw):
        vals = dict(zip(_WS_FIELDS, [1e-3, 1e-3, 1e-3, 0.0, math.inf, 1024, 0, None, False, True, True, False,
                                      False, False, False, False]))
        if len(args) + len(kw) == 4:
            names = ("epsilon_shell_for_absorbing_boundary", "epsilon_shell_for_reflecting_boundary",
                     "max_walk_length", "solve_double_sided")
        elif len(args) + len(kw) == 16:
            names = _WS_FIELDS
        else:
            raise TypeError("WalkSettings takes 4 or 16 arguments")
        given = dict(zip(names, args))
        for k, v in kw.items():
            if k not in names or k in given:
                raise TypeError(f"unexpected or repeated argument {k}")
            given[k] = v
        if len(given) != len(names):
            raise TypeError(f"WalkSettings needs {[n for n in names if n not in given]}")
        vals.update(given)
        if vals["steps_before_using_maximal_spheres"] is None:
            vals["steps_before_using_maximal_spheres"] = vals["max_walk_length"]
        for k, v in vals.items():
            setattr(self, k, v)

    def to_c(self) -> A.zg_walk_settings:
        ws = A.zg_walk_settings()
        for k in _WS_FIELDS:
            v = getattr(self, k)
            if k == "splitting_threshold":
                v = min(float(v), A.FLT_MAX)
            elif k == "russian_roulette_threshold":
                v = float(v)
            setattr(ws, k, int(v) if isinstance(getattr(ws, k), int) else float(v))
        return ws


class SamplePoint:
    """SamplePoint<T, DIM> (core/common.h:145-195; zombie_bindings.h:1194-1210)."""

    def __init__(self, pt, normal, type, estimation_quantity, pdf, dist_to_absorbing_boundary,
                 dist_to_reflecting_boundary, *, dim=None, channels=1):
        self.pt = np.asarray(pt, np.float32)
        self.normal = np.asarray(normal, np.float32)
        self.direction_for_derivative = np.zeros_like(self.pt)
        self.direction_for_derivative[0] = 1.0
        self.type = SampleType(int(type))
        self.estimation_quantity = EstimationQuantity(int(estimation_quantity))
        self.pdf = float(pdf)
        self.dist_to_absorbing_boundary = float(dist_to_absorbing_boundary)
        self.dist_to_reflecting_boundary = float(dist_to_reflecting_boundary)
        self.estimate_boundary_normal_aligned = False
        self.channels = channels
        self.first_sphere_radius = 0.0

    def reset(self):
        self.first_sphere_radius = 0.0


class SampleStatistics:
    """SampleStatistics<T, DIM> (core/common.h:200-395): per-channel FP64 running statistics."""

    def __init__(self, *, dim=None, channels=1):
        self.channels = channels
        self.dim = dim
        self.n = 0
        self.mean = np.zeros(channels)
        self.m2 = np.zeros(channels)
        self.n_grad = 0
        self.grad_mean = np.zeros((channels, dim or 3))
        self.grad_m2 = np.zeros((channels, dim or 3))
        self.first_source = np.zeros(channels)
        self.derivative = np.zeros(channels)
        self.total_walk_length = 0
        self.total_splits = 0

    def _v(self, a):
        return float(a[0]) if self.channels == 1 else a.copy()

    def get_estimated_solution(self):
        return self._v(self.mean)

    def get_estimated_solution_variance(self):
        return self._v(self.m2 / (self.n - 1) if self.n > 1 else np.zeros(self.channels))

    def get_estimated_gradient(self):
        g = self.grad_mean[:, : self.dim]
        return g[0].copy() if self.channels == 1 else g.copy()

    def get_estimated_derivative(self):
        return self._v(self.derivative / self.n_grad if self.n_grad > 0 else np.zeros(self.channels))

    def get_solution_estimate_count(self):
        return self.n

    def get_gradient_estimate_count(self):
        return self.n_grad

    def get_mean_walk_length(self):
        return self.total_walk_length / self.n if self.n > 0 else 0.0


class SamplePointList(list):
    def __init__(self, *args, dim=None, channels=1):
        super().__init__(*args)


class SampleStatisticsList(list):
    def __init__(self, *args, dim=None, channels=1):
        super().__init__(*args)


def create_sample_statistics_list(size, *, dim, channels=1):
    _validate_dim(dim)
    _validate_channels(channels)
    return SampleStatisticsList([SampleStatistics(dim=dim, channels=channels) for _ in range(int(size))])


class _Solver:
    _wost = False

    def __init__(self, geometric_queries, walk_state_callback=None, terminal_contribution_callback=None, *, dim,
                 channels=1, seed=0x5EED):
        _validate_dim(dim)
        _validate_channels(channels)
        if walk_state_callback is not None or terminal_contribution_callback is not None:
            raise NotImplementedError("walk-state and terminal-contribution callbacks run on the host; "
                                      "the GPU solver has none")
        if geometric_queries.dim != dim:
            raise ValueError("geometric_queries has another dimension")
        self.queries = geometric_queries      # referenced, not copied (as the reference does)
        self.dim, self.channels = dim, channels
        self.seed = int(seed)
        self._stream_offset = 0

    def solve(self, pde, walk_settings, n_walks, sample_pts, statistics, run_single_threaded=False,
              report_progress=None):
        """Both of the reference's overloads: (pde, settings, int, SamplePoint, SampleStatistics)
        and (pde, settings, IntList, SamplePointList, SampleStatisticsList, ...)."""
        if isinstance(sample_pts, SamplePoint):
            return self.solve(pde, walk_settings, [int(n_walks)], [sample_pts], [statistics])
        torch = E._torch()
        n = len(sample_pts)
        if len(statistics) != n or len(n_walks) != n:
            raise ValueError("sample_pts, statistics and n_walks must have one entry per point")
        if n == 0:
            return
        if pde.dim != self.dim or pde.channels != self.channels:
            raise ValueError("the PDE's dim / channels differ from the solver's")
        d, C_ = self.dim, self.channels
        dev = "cuda"
        ws = walk_settings.to_c()
        scene = self.queries.scene(bool(walk_settings.solve_double_sided), float(walk_settings.silhouette_precision))
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
        u8 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.uint8)).to(dev)
        pts = f32(np.stack([p.pt for p in sample_pts]))
        normal = f32(np.stack([p.normal for p in sample_pts]))
        nw_host = np.asarray(n_walks, np.int32)
        q = np.array([int(p.estimation_quantity) for p in sample_pts], np.uint8)
        grad = bool((q == EstimationQuantity.SolutionAndGradient).any())
        fsr_host = np.array([p.first_sphere_radius for p in sample_pts], np.float32)
        common = dict(pt=pts, normal=normal, n_walks=torch.from_numpy(nw_host).to(dev),
                      n_walks_max=int(nw_host.max(initial=0)),
                      type=u8([int(p.type) for p in sample_pts]),
                      aligned=u8([1 if p.estimate_boundary_normal_aligned else 0 for p in sample_pts]),
                      dist_to_absorbing=f32([p.dist_to_absorbing_boundary for p in sample_pts]),
                      dist_to_reflecting=f32([p.dist_to_reflecting_boundary for p in sample_pts]),
                      estimation_quantity=u8(q) if (q != EstimationQuantity.Solution).any() else None,
                      direction_for_derivative=f32(np.stack([p.direction_for_derivative for p in sample_pts]))
                      if grad else None)
        opts = E.SolveOptions(seed=self.seed, stream_offset=self._stream_offset)
        for c in range(C_):
            st = E.DeviceStats.zeros(n, dim=d if grad else 0)
            # the caller's running statistics of this channel, merged into by the solve
            st.n.copy_(torch.tensor([s.n for s in statistics], dtype=torch.int64))
            st.mean.copy_(torch.tensor([s.mean[c] for s in statistics], dtype=torch.float64))
            st.m2.copy_(torch.tensor([s.m2[c] for s in statistics], dtype=torch.float64))
            st.walk_length.copy_(torch.tensor([s.total_walk_length for s in statistics], dtype=torch.int64))
            st.splits.copy_(torch.tensor([s.total_splits for s in statistics], dtype=torch.int64))
            if grad:
                st.n_grad.copy_(torch.tensor([s.n_grad for s in statistics], dtype=torch.int64))
                st.grad_mean.copy_(torch.tensor(np.stack([s.grad_mean[c, :d] for s in statistics])))
                st.grad_m2.copy_(torch.tensor(np.stack([s.grad_m2[c, :d] for s in statistics])))
                st.first_source.copy_(torch.tensor([s.first_source[c] for s in statistics], dtype=torch.float64))
                st.derivative.copy_(torch.tensor([s.derivative[c] for s in statistics], dtype=torch.float64))
            fsr = f32(fsr_host) if self._wost else None
            dp = E.DevicePoints(first_sphere_radius=fsr, **common)
            solve = E.wost_solve if self._wost else E.wos_solve
            solve(scene, pde.channel(c), ws, dp, st, opts)
            out = dict(n=st.n.cpu().numpy(), mean=st.mean.cpu().numpy(), m2=st.m2.cpu().numpy(),
                       len=st.walk_length.cpu().numpy(), splits=st.splits.cpu().numpy())
            if grad:
                out.update(n_grad=st.n_grad.cpu().numpy(), gm=st.grad_mean.cpu().numpy(),
                           g2=st.grad_m2.cpu().numpy(), fs=st.first_source.cpu().numpy(),
                           dd=st.derivative.cpu().numpy())
            for i, s in enumerate(statistics):
                s.mean[c], s.m2[c] = out["mean"][i], out["m2"][i]
                if grad:
                    s.grad_mean[c, :d], s.grad_m2[c, :d] = out["gm"][i], out["g2"][i]
                    s.first_source[c], s.derivative[c] = out["fs"][i], out["dd"][i]
                if c == C_ - 1:
                    s.n, s.total_walk_length, s.total_splits = int(out["n"][i]), int(out["len"][i]), int(out["splits"][i])
                    if grad:
                        s.n_grad = int(out["n_grad"][i])
            if fsr is not None and c == C_ - 1:
                for p, r in zip(sample_pts, fsr.cpu().numpy()):
                    p.first_sphere_radius = float(r)
            if fsr is None and c == C_ - 1:
                for p in sample_pts:
                    if p.estimation_quantity == EstimationQuantity.SolutionAndGradient:
                        p.first_sphere_radius = 0.99 * p.dist_to_absorbing_boundary
        self._stream_offset += int(nw_host.max(initial=0))


class WalkOnSpheres(_Solver):
    """WalkOnSpheres<T, DIM> (point_estimation/walk_on_spheres.h:22-168) on the GPU."""


class WalkOnStars(_Solver):
    """WalkOnStars<T, DIM> (point_estimation/walk_on_stars.h:22-187) on the GPU."""
    _wost = True


def get_empty_walk_state_callback(*, dim, channels=1):
    return None


def get_empty_terminal_contribution_callback(*, dim, channels=1):
    return None


Solvers = types.SimpleNamespace(
    SampleType=SampleType, EstimationQuantity=EstimationQuantity, WalkCompletionCode=WalkCompletionCode,
    WalkSettings=WalkSettings, SamplePoint=SamplePoint, SampleStatistics=SampleStatistics,
    SamplePointList=SamplePointList, SampleStatisticsList=SampleStatisticsList,
    create_sample_statistics_list=create_sample_statistics_list,
    WalkOnSpheres=WalkOnSpheres, WalkOnStars=WalkOnStars,
    get_empty_walk_state_callback=get_empty_walk_state_callback,
    get_empty_terminal_contribution_callback=get_empty_terminal_contribution_callback,
)


# ------------------------------------------------------------------------------------
# Utils
# ------------------------------------------------------------------------------------
class _Boundary:
    def __init__(self, dim):
        self.dim = dim
        self.positions = np.zeros((0, dim), np.float32)
        self.indices = np.zeros((0, dim), np.int32)
        self.min_coeff = None
        self.max_coeff = None


class FcpwDirichletBoundaryHandler(_Boundary):
    """FcpwDirichletBoundaryHandler<DIM> (utils/fcpw_geometric_queries.h:88-105): the absorbing
    boundary mesh, packed into the device scene's wide BVH on first use."""

    def __init__(self, *, dim):
        _validate_dim(dim)
        super().__init__(dim)

    def build_acceleration_structure(self, positions, indices):
        self.positions = np.asarray(positions, np.float32).reshape(-1, self.dim)
        self.indices = np.asarray(indices, np.int32).reshape(-1, self.dim)


class FcpwNeumannBoundaryHandler(FcpwDirichletBoundaryHandler):
    """FcpwNeumannBoundaryHandler<DIM> (:107-126): the reflecting boundary; silhouettes are built
    with the device scene (one-sided problems drop convex candidates, as
    getIgnoreCandidateSilhouetteCallback does)."""

    def build_acceleration_structure(self, positions, indices, ignore_candidate_silhouette=None):
        super().build_acceleration_structure(positions, indices)


class FcpwRobinBoundaryHandler(FcpwDirichletBoundaryHandler):
    """FcpwRobinBoundaryHandler<DIM> (:128-152): the reflecting boundary with per-primitive Robin
    coefficient bounds."""

    def build_acceleration_structure(self, positions, indices, ignore_candidate_silhouette=None,
                                     min_robin_coeff_values=(), max_robin_coeff_values=()):
        super().build_acceleration_structure(positions, indices)
        self.min_coeff = np.asarray(min_robin_coeff_values, np.float32)
        self.max_coeff = np.asarray(max_robin_coeff_values, np.float32)

    def update_coefficient_values(self, min_robin_coeff_values, max_robin_coeff_values):
        self.min_coeff = np.asarray(min_robin_coeff_values, np.float32)
        self.max_coeff = np.asarray(max_robin_coeff_values, np.float32)


def populate_geometric_queries_for_dirichlet_boundary(handler, geometric_queries, *, dim):
    geometric_queries._absorbing = handler
    geometric_queries._scene = None


def populate_geometric_queries_for_neumann_boundary(handler, branch_traversal_weight, geometric_queries, *, dim):
    geometric_queries._reflecting = handler
    geometric_queries._scene = None


populate_geometric_queries_for_robin_boundary = populate_geometric_queries_for_neumann_boundary


def get_ignore_candidate_silhouette_callback(solve_double_sided=False, silhouette_precision=1e-3):
    return ("ignore_candidate_silhouette", bool(solve_double_sided), float(silhouette_precision))


def get_branch_traversal_weight_callback(min_radial_dist=1e-4):
    return ("branch_traversal_weight", float(min_radial_dist))


def load_boundary_mesh(obj_file, positions, indices, *, dim):
    """loadBoundaryMesh<DIM> (fcpw_geometric_queries.h:159-211): `v` lines and, in 2D, `l`
    polylines split into segments, in 3D `f` triangles (0-based indices)."""
    _validate_dim(dim)
    P, I = [], []
    with open(obj_file) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                P.append([float(t) for t in tok[1:1 + dim]])
            elif dim == 2 and tok[0] == "l":
                idx = [int(t.split("/")[0]) - 1 for t in tok[1:]]
                I.extend([idx[k], idx[k + 1]] for k in range(len(idx) - 1))
            elif dim == 3 and tok[0] == "f":
                idx = [int(t.split("/")[0]) - 1 for t in tok[1:]]
                I.extend([idx[0], idx[k], idx[k + 1]] for k in range(1, len(idx) - 1))
    positions[:] = [np.asarray(p, np.float32) for p in P]
    indices[:] = [np.asarray(i, np.int32) for i in I]


def normalize(positions, *, dim):
    """normalize<DIM> (:255-273): centroid to the origin, largest vertex distance 1."""
    V = np.asarray(positions, np.float64).reshape(-1, dim)
    V = V - V.mean(0)
    V = V / np.linalg.norm(V, axis=1).max()
    positions[:] = list(V.astype(np.float32))


def apply_shift(shift, positions, *, dim):
    V = np.asarray(positions, np.float32).reshape(-1, dim) + np.asarray(shift, np.float32)
    positions[:] = list(V)


def flip_orientation(indices, *, dim):
    """flipOrientation<DIM> (:284-289): swaps the first two vertices of every primitive."""
    I = np.asarray(indices, np.int32).reshape(-1, dim).copy()
    I[:, [0, 1]] = I[:, [1, 0]]
    indices[:] = list(I)


def compute_bounding_box(positions, make_square, scale, *, dim):
    """computeBoundingBox<DIM> (:291-317)."""
    V = np.asarray(positions, np.float32).reshape(-1, dim)
    cm = V.mean(0)
    P = (V - cm) * np.float32(scale) + cm
    lo, hi = P.min(0), P.max(0)
    if make_square:
        c = 0.5 * (lo + hi)
        half = 0.5 * float((hi - lo).max())
        lo, hi = c - half, c + half
    return lo.astype(np.float32), hi.astype(np.float32)


def compute_signed_volume(positions, indices, *, dim):
    V = np.asarray(positions, np.float64).reshape(-1, dim)
    I = np.asarray(indices, np.int64).reshape(-1, dim)
    if dim == 2:
        a, b = V[I[:, 0]], V[I[:, 1]]
        return float(0.5 * np.sum(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]))
    a, b, c = V[I[:, 0]], V[I[:, 1]], V[I[:, 2]]
    return float(np.sum(np.einsum("ij,ij->i", np.cross(a, b), c)) / 6.0)


def partition_boundary_mesh(on_reflecting_boundary, positions, indices, absorbing_positions, absorbing_indices,
                            reflecting_positions, reflecting_indices, *, dim):
    """partitionBoundaryMesh<DIM> (:459-509): primitives whose midpoint satisfies the indicator
    go to the reflecting part; each part keeps its vertices in first-use order."""
    V = np.asarray(positions, np.float32).reshape(-1, dim)
    I = np.asarray(indices, np.int32).reshape(-1, dim)
    mid = V[I].mean(1).astype(np.float32)
    flags = np.array([bool(np.asarray(on_reflecting_boundary(m)).reshape(-1)[0]) for m in mid])
    for want, P_out, I_out in ((False, absorbing_positions, absorbing_indices),
                               (True, reflecting_positions, reflecting_indices)):
        remap, Pn, In = {}, [], []
        for prim, fl in zip(I, flags):
            if fl != want:
                continue
            row = []
            for v in prim:
                if v not in remap:
                    remap[v] = len(Pn)
                    Pn.append(V[v])
                row.append(remap[v])
            In.append(np.asarray(row, np.int32))
        P_out[:] = Pn
        I_out[:] = In


def compute_dist_to_boundary(geometric_queries, solve_locations, compute_signed_distance,
                             dist_to_absorbing_boundary, dist_to_reflecting_boundary, *, dim):
    """Distances of every location to both boundary parts (GPU queries)."""
    torch = E._torch()
    scene = geometric_queries.scene()
    P = torch.from_numpy(np.ascontiguousarray(np.asarray(solve_locations, np.float32).reshape(-1, dim))).cuda()
    for part, out in ((0, dist_to_absorbing_boundary), (1, dist_to_reflecting_boundary)):
        have = geometric_queries.has_non_empty_absorbing_boundary if part == 0 else \
            geometric_queries.has_non_empty_reflecting_boundary
        if have:
            d, _, _, _ = scene.closest_point(P, boundary=part, want_point=False)
            out[:] = d.cpu().numpy().tolist()
        else:
            out[:] = [float(np.finfo(np.float32).max)] * P.shape[0]


def inside_domain(geometric_queries, solve_locations, *, dim):
    """GeometricQueries::insideDomain at every location (parity of axis-ray hits on the GPU)."""
    torch = E._torch()
    scene = geometric_queries.scene()
    P = torch.from_numpy(np.ascontiguousarray(np.asarray(solve_locations, np.float32).reshape(-1, dim))).cuda()
    if not geometric_queries.domain_is_watertight:
        return np.ones(P.shape[0], bool)
    return scene.inside_domain(P).cpu().numpy().astype(bool)


def get_spatially_sorted_point_indices(points, *, dim):
    return IntList(E.spatially_sorted_point_indices(np.asarray(points, np.float32).reshape(-1, dim)).tolist())


def get_dense_grid_source_callback(*args, dim, channels=1):
    return _grid_callback(args, dim, channels, "source")


def get_dense_grid_dirichlet_callback(*args, dim, channels=1):
    return _grid_callback(args, dim, channels, "dirichlet")


def get_dense_grid_robin_callback(*args, dim, channels=1):
    return _grid_callback(args, dim, channels, "robin")


def get_dense_grid_robin_coefficient_callback(*args, dim):
    return _grid_callback(args, dim, 1, "robin coefficient")


def get_dense_grid_indicator_callback(*args, dim):
    """Host-evaluated indicator (used by partition_boundary_mesh): value > 0."""
    grid = args[0] if isinstance(args[0], DenseGrid) else DenseGrid(*args, dim=dim)
    return lambda x: bool(grid(np.asarray(x, np.float32))[0, 0] > 0)


class ProgressBar:
    def __init__(self, total=0):
        self.total = total

    def report(self, *a):
        pass

    def finish(self):
        pass


def get_report_progress_callback(progress_bar):
    return None


Utils = types.SimpleNamespace(
    FcpwDirichletBoundaryHandler=FcpwDirichletBoundaryHandler, FcpwNeumannBoundaryHandler=FcpwNeumannBoundaryHandler,
    FcpwRobinBoundaryHandler=FcpwRobinBoundaryHandler,
    populate_geometric_queries_for_dirichlet_boundary=populate_geometric_queries_for_dirichlet_boundary,
    populate_geometric_queries_for_neumann_boundary=populate_geometric_queries_for_neumann_boundary,
    populate_geometric_queries_for_robin_boundary=populate_geometric_queries_for_robin_boundary,
    get_ignore_candidate_silhouette_callback=get_ignore_candidate_silhouette_callback,
    get_branch_traversal_weight_callback=get_branch_traversal_weight_callback,
    load_boundary_mesh=load_boundary_mesh, normalize=normalize, apply_shift=apply_shift,
    flip_orientation=flip_orientation, compute_bounding_box=compute_bounding_box,
    compute_signed_volume=compute_signed_volume, partition_boundary_mesh=partition_boundary_mesh,
    compute_dist_to_boundary=compute_dist_to_boundary, inside_domain=inside_domain,
    get_spatially_sorted_point_indices=get_spatially_sorted_point_indices,
    FloatDenseGrid=DenseGrid, get_dense_grid_source_callback=get_dense_grid_source_callback,
    get_dense_grid_dirichlet_callback=get_dense_grid_dirichlet_callback,
    get_dense_grid_robin_callback=get_dense_grid_robin_callback,
    get_dense_grid_robin_coefficient_callback=get_dense_grid_robin_coefficient_callback,
    get_dense_grid_indicator_callback=get_dense_grid_indicator_callback,
    ProgressBar=ProgressBar, get_report_progress_callback=get_report_progress_callback,
)
