"""Array-level host API over the C-ABI (device arrays are torch tensors on the GPU).

PyTorch is used only for device memory and streams; every computation runs in
libzombie_amd.so.  The benchmark and the reference-shaped layers build on this one.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


def _torch():
    import torch  # deferred: importing torch is slow and the CPU tests do not need it
    return torch


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def check_array(name: str, t, dtype, shape: tuple, device=None):
    """The C-ABI takes raw device pointers and trusts their extents: refuse, on the host, an
    array of the wrong element type, shape or layout, or one on another device, before its
    pointer reaches a kernel (which would read or write past its end or misread its bytes)."""
    if t is None:
        return
    if not hasattr(t, "data_ptr"):
        raise TypeError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}; the C-ABI needs {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; the C-ABI needs {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} is not contiguous")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, the solve's arrays are on {device}")


def _stream_ptr(stream=None) -> int | None:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


# ------------------------------------------------------------------------------------
# descriptors
# ------------------------------------------------------------------------------------
def function(kind: int, eval_: int = A.ZG_EVAL_VALUE, params=()) -> A.zg_function:
    f = A.zg_function()
    f.kind = kind
    f.eval = eval_
    for i, v in enumerate(params):
        f.params[i] = float(v)
    return f


class GridFunction:
    """A ZG_FN_DENSE_GRID descriptor (DenseGrid<float,1,DIM>, utils/dense_grid.h) with its value
    array kept alive: values[i, j(, k)] is the value of the cell centred at
    lo + (index + 1/2) (hi - lo) / shape.  `fn` is the zg_function to put in a zg_pde; the object
    must outlive every solve that uses it.  device=None keeps the values in host memory (for the
    CPU oracle)."""

    def __init__(self, values, lo, hi, interpolate: bool = True, device="cuda"):
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.ndim not in (2, 3):
            raise ValueError("values must be a 2D or 3D array")
        if device is None:
            self.values = v
            ptr = v.ctypes.data
        else:
            torch = _torch()
            self.values = torch.from_numpy(v).to(device).contiguous()
            ptr = self.values.data_ptr()
        f = A.zg_function()
        f.kind = A.ZG_FN_DENSE_GRID
        f.eval = A.ZG_EVAL_VALUE
        f.grid = ptr
        for a in range(3):
            f.shape[a] = v.shape[a] if a < v.ndim else 1
            f.grid_min[a] = float(lo[a]) if a < v.ndim else 0.0
            f.grid_max[a] = float(hi[a]) if a < v.ndim else 1.0
        f.interpolate = 1 if interpolate else 0
        self.fn = f


def evaluate_function(dim: int, f: A.zg_function, pts, normals=None, aligned=None, absorption: float = 0.0,
                      stream=None):
    """f at the device points pts [n, dim] (zg_evaluate_function)."""
    torch = _torch()
    out = torch.empty(pts.shape[0], dtype=torch.float32, device=pts.device)
    A.check(A.lib().zg_evaluate_function(dim, C.byref(f), pts.shape[0], _ptr(pts), _ptr(normals), _ptr(aligned),
                                         float(absorption), _ptr(out), _stream_ptr(stream)))
    return out


def constant(value: float) -> A.zg_function:
    return function(A.ZG_FN_CONSTANT, A.ZG_EVAL_VALUE, (value,))


def zero() -> A.zg_function:
    return function(A.ZG_FN_ZERO)


def make_pde(source=None, dirichlet=None, robin=None, absorption: float = 0.0,
             is_source_constant: bool = False, robin_coeff=None) -> A.zg_pde:
    """PDE<float,DIM> (pde.h:19-45).  A robin_coeff descriptor makes the reflecting conditions
    Robin (are_robin_conditions_pure_neumann = 0; coefficients taken as non-negative)."""
    p = A.zg_pde()
    p.absorption_coeff = float(absorption)
    p.is_source_constant = 1 if is_source_constant else 0
    p.are_robin_conditions_pure_neumann = 1 if robin_coeff is None else 0
    p.are_robin_coeffs_nonnegative = 1
    p.source = source if source is not None else zero()
    p.dirichlet = dirichlet if dirichlet is not None else zero()
    p.robin = robin if robin is not None else zero()
    p.robin_coeff = robin_coeff if robin_coeff is not None else zero()
    return p


def walk_settings(eps_absorbing: float, eps_reflecting: float, max_walk_length: int,
                  solve_double_sided: bool = False, **overrides) -> A.zg_walk_settings:
    ws = A.zg_walk_settings()
    A.lib().zg_walk_settings_default(C.byref(ws), eps_absorbing, eps_reflecting, max_walk_length,
                                     1 if solve_double_sided else 0)
    for k, v in overrides.items():
        setattr(ws, k, v)
    return ws


def walk_settings_host(eps_absorbing: float, eps_reflecting: float, max_walk_length: int,
                       solve_double_sided: bool = False, **overrides) -> A.zg_walk_settings:
    """Same defaults as zg_walk_settings_default, without loading the library."""
    ws = A.zg_walk_settings()
    ws.epsilon_shell_for_absorbing_boundary = eps_absorbing
    ws.epsilon_shell_for_reflecting_boundary = eps_reflecting
    ws.silhouette_precision = 1e-3
    ws.russian_roulette_threshold = 0.0
    ws.splitting_threshold = A.FLT_MAX
    ws.max_walk_length = max_walk_length
    ws.steps_before_applying_tikhonov = 0
    ws.steps_before_using_maximal_spheres = max_walk_length
    ws.solve_double_sided = 1 if solve_double_sided else 0
    ws.use_gradient_control_variates = 1
    ws.use_gradient_antithetic_variates = 1
    for k, v in overrides.items():
        setattr(ws, k, v)
    return ws


@dataclass
class SolveOptions:
    seed: int = 0x5EED
    stream_offset: int = 0
    walks_per_lane: int = 8
    regroup_interval: int = 16
    step_interval: int = 4
    waves_per_cu: int = 0
    count_visits: int = 0   # 1: count the kernel's own BVH visits; 2: those of an exact traversal from the root
    distance_tolerance: float = 0.0
    wost_interleave_queries: bool = False  # walk on stars: interleave the lanes' queries item by item

    def to_c(self) -> A.zg_solve_options:
        o = A.zg_solve_options()
        o.seed = self.seed
        o.stream_offset = self.stream_offset
        o.walks_per_lane = self.walks_per_lane
        o.regroup_interval = self.regroup_interval
        o.step_interval = self.step_interval
        o.waves_per_cu = self.waves_per_cu
        o.count_visits = int(self.count_visits)
        o.distance_tolerance = self.distance_tolerance
        o.wost_interleave_queries = int(bool(self.wost_interleave_queries))
        return o


@dataclass
class SolveReport:
    completion_codes: tuple
    total_steps: int
    total_walks: int
    node_visits: int
    prim_visits: int
    queries: int
    kernel_ms: float
    # walk on stars, count_visits 1: reflecting-boundary and silhouette BVH reads
    reflecting_node_visits: int = 0
    reflecting_prim_visits: int = 0
    silhouette_node_visits: int = 0
    silhouette_prim_visits: int = 0

    @staticmethod
    def from_c(r: A.zg_solve_report) -> "SolveReport":
        return SolveReport(tuple(r.completion_codes), r.total_steps, r.total_walks, r.node_visits,
                           r.prim_visits, r.queries, r.kernel_ms, r.reflecting_node_visits,
                           r.reflecting_prim_visits, r.silhouette_node_visits, r.silhouette_prim_visits)


# ------------------------------------------------------------------------------------
# scene
# ------------------------------------------------------------------------------------
class Scene:
    """Device scene: packed wide BVHs of the absorbing and reflecting boundary parts."""

    def __init__(self, positions, indices, is_reflecting=None, domain_min=None, domain_max=None,
                 bvh_width: int = 8, leaf_size: int = 4, sah_bins: int = 16,
                 solve_double_sided: bool = False, silhouette_precision: float = 1e-3):
        V = np.ascontiguousarray(positions, dtype=np.float32)
        I = np.ascontiguousarray(indices, dtype=np.int32)
        self.dim = V.shape[1]
        if I.ndim != 2 or I.shape[1] != self.dim:
            raise ValueError("indices must be [n_prims, dim]")
        refl = None if is_reflecting is None else np.ascontiguousarray(is_reflecting, dtype=np.uint8)
        dmin = np.ascontiguousarray(domain_min if domain_min is not None else V.min(0) - 1.0, dtype=np.float32)
        dmax = np.ascontiguousarray(domain_max if domain_max is not None else V.max(0) + 1.0, dtype=np.float32)
        opts = A.zg_build_options(bvh_width, leaf_size, sah_bins, 0)
        h = C.c_void_p()
        A.check(A.lib().zg_scene_create(
            self.dim, V.ctypes.data, V.shape[0], I.ctypes.data, I.shape[0],
            None if refl is None else refl.ctypes.data, dmin.ctypes.data, dmax.ctypes.data,
            1 if solve_double_sided else 0, silhouette_precision, C.byref(opts), C.byref(h)))
        self._h = h
        self.domain_min, self.domain_max = dmin, dmax
        info = A.zg_scene_info()
        A.check(A.lib().zg_scene_get_info(self._h, C.byref(info)))
        self.info = info

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            A.lib().zg_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # batched queries (device tensors in, device tensors out)
    def closest_point(self, pts, boundary: int = 0, want_point=True, want_normal=False, want_prim=False):
        torch = _torch()
        n = pts.shape[0]
        dev = pts.device
        dist = torch.empty(n, dtype=torch.float32, device=dev)
        cp = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if want_point else None
        nr = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if want_normal else None
        pr = torch.empty(n, dtype=torch.int32, device=dev) if want_prim else None
        A.check(A.lib().zg_closest_point(self._h, boundary, n, _ptr(pts), _ptr(dist), _ptr(cp), _ptr(nr), _ptr(pr),
                                         _stream_ptr()))
        return dist, cp, nr, pr

    def inside_domain(self, pts):
        torch = _torch()
        out = torch.empty(pts.shape[0], dtype=torch.uint8, device=pts.device)
        A.check(A.lib().zg_inside_domain(self._h, pts.shape[0], _ptr(pts), _ptr(out), _stream_ptr()))
        return out

    def copy_bvh(self, boundary: int = 0):
        """The packed wide BVH of one boundary part on the host: (nodes, primitives) as float32
        arrays in the device layout (zg_scene_copy_bvh)."""
        info = self.info
        nn = info.n_absorbing_nodes if boundary == 0 else info.n_reflecting_nodes
        npr = info.n_absorbing_prims if boundary == 0 else info.n_reflecting_prims
        nodes = np.empty(max(1, nn * info.node_bytes // 4), np.float32)
        prims = np.empty(max(1, npr * info.prim_bytes // 4), np.float32)
        A.check(A.lib().zg_scene_copy_bvh(self._h, boundary, nodes.ctypes.data, nodes.size, prims.ctypes.data,
                                          prims.size))
        return nodes[: nn * info.node_bytes // 4], prims[: npr * info.prim_bytes // 4]

    def intersect_ray(self, origins, dirs, t_max, boundary: int = 0):
        torch = _torch()
        n = origins.shape[0]
        dev = origins.device
        hit = torch.empty(n, dtype=torch.uint8, device=dev)
        t = torch.empty(n, dtype=torch.float32, device=dev)
        p = torch.empty((n, self.dim), dtype=torch.float32, device=dev)
        nr = torch.empty((n, self.dim), dtype=torch.float32, device=dev)
        A.check(A.lib().zg_intersect_ray(self._h, boundary, n, _ptr(origins), _ptr(dirs), _ptr(t_max), _ptr(hit),
                                         _ptr(t), _ptr(p), _ptr(nr), _stream_ptr()))
        return hit, t, p, nr

    def set_robin_coefficients(self, min_coeff, max_coeff):
        """Per-primitive Robin coefficient bounds (host arrays over all primitives; None clears)."""
        if min_coeff is None:
            A.check(A.lib().zg_scene_set_robin_coefficients(self.handle, None, None))
            return
        lo = np.ascontiguousarray(min_coeff, np.float32)
        hi = np.ascontiguousarray(max_coeff, np.float32)
        n = self.info.n_absorbing_prims + self.info.n_reflecting_prims
        if lo.shape != (n,) or hi.shape != (n,):
            raise ValueError(f"coefficient bounds need one entry per primitive ({n}), got {lo.shape} and {hi.shape}")
        A.check(A.lib().zg_scene_set_robin_coefficients(self.handle, lo.ctypes.data, hi.ctypes.data))

    # walk-on-stars queries on the reflecting boundary
    def star_radius(self, pts, min_r, max_r, precision: float = 1e-3, flip=None):
        torch = _torch()
        out = torch.empty(pts.shape[0], dtype=torch.float32, device=pts.device)
        A.check(A.lib().zg_star_radius(self._h, pts.shape[0], _ptr(pts), _ptr(min_r), _ptr(max_r), float(precision),
                                       _ptr(flip), _ptr(out), _stream_ptr()))
        return out

    def line_of_sight(self, p1, p2):
        torch = _torch()
        out = torch.empty(p1.shape[0], dtype=torch.uint8, device=p1.device)
        A.check(A.lib().zg_line_of_sight(self._h, p1.shape[0], _ptr(p1), _ptr(p2), _ptr(out), _stream_ptr()))
        return out

    def sample_reflecting(self, x, radius, u):
        """u: [n, dim] uniform numbers (primitive choice, then the point on it)."""
        torch = _torch()
        n = x.shape[0]
        dev = x.device
        ok = torch.empty(n, dtype=torch.uint8, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=dev)
        nr = torch.empty((n, self.dim), dtype=torch.float32, device=dev)
        pdf = torch.empty(n, dtype=torch.float32, device=dev)
        A.check(A.lib().zg_sample_reflecting(self._h, n, _ptr(x), _ptr(radius), _ptr(u), _ptr(ok), _ptr(pt), _ptr(nr),
                                             _ptr(pdf), _stream_ptr()))
        return ok, pt, nr, pdf


# ------------------------------------------------------------------------------------
# points + statistics on the device
# ------------------------------------------------------------------------------------
@dataclass
class DevicePoints:
    pt: object
    dist_to_absorbing: object
    n_walks: object
    n_walks_max: int
    type: object = None
    normal: object = None
    aligned: object = None
    dist_to_reflecting: object = None
    first_sphere_radius: object = None
    cv_gradient: object = None
    cv_gradient_grid: object = None     # GridField
    estimation_quantity: object = None  # uint8 [n]: ZG_ESTIMATE_SOLUTION / _SOLUTION_AND_GRADIENT / _SKIP
    direction_for_derivative: object = None   # float32 [n, dim], default e_0

    @property
    def n(self) -> int:
        return int(self.pt.shape[0])

    def check_dim(self, dim: int):
        """The C-ABI takes plain pointers: points, normals and derivative directions must hold
        `dim` floats per point (the scene's dimension)."""
        for name in ("pt", "normal", "direction_for_derivative"):
            a = getattr(self, name)
            if a is not None and (a.dim() != 2 or a.shape[0] != self.n or a.shape[1] != dim):
                raise ValueError(f"{name} has shape {tuple(a.shape)}; the scene needs [{self.n}, {dim}]")

    def validate(self, dim: int, check_walk_bound: bool = True):
        """Full host-side check of every array a solve hands to the C-ABI (element type, shape,
        contiguity, one GPU device), required arrays present, and n_walks[] <= n_walks_max."""
        torch = _torch()
        self.check_dim(dim)
        n = self.n
        dev = self.pt.device
        if self.dist_to_absorbing is None or self.n_walks is None:
            raise ValueError("dist_to_absorbing and n_walks are required")
        f32, u8 = torch.float32, torch.uint8
        for name, dtype, shape in (("pt", f32, (n, dim)), ("normal", f32, (n, dim)), ("type", u8, (n,)),
                                   ("dist_to_absorbing", f32, (n,)), ("dist_to_reflecting", f32, (n,)),
                                   ("aligned", u8, (n,)), ("first_sphere_radius", f32, (n,)),
                                   ("n_walks", torch.int32, (n,)), ("cv_gradient", f32, (n, dim)),
                                   ("estimation_quantity", u8, (n,)), ("direction_for_derivative", f32, (n, dim))):
            check_array(name, getattr(self, name), dtype, shape, dev)
        if self.cv_gradient_grid is not None:
            g = self.cv_gradient_grid.data
            if g is None or g.dim() != 4 or g.shape[3] != 4:
                raise ValueError("cv_gradient_grid.data must be [res_z, res_y, res_x, 4]")
            check_array("cv_gradient_grid.data", g, f32, tuple(g.shape), dev)
        if int(self.n_walks_max) < 0:
            raise ValueError("n_walks_max must be >= 0")
        if check_walk_bound and n > 0:
            most = int(self.n_walks.max())
            if most > int(self.n_walks_max):
                raise ValueError(f"n_walks holds {most} walks for a point, above n_walks_max = {self.n_walks_max}")
        if dev.type != "cuda":
            raise ValueError(f"points must be on a GPU device, pt is on {dev}")

    def to_c(self) -> A.zg_sample_points:
        s = A.zg_sample_points()
        s.n = self.n
        s.pt = _ptr(self.pt)
        s.normal = _ptr(self.normal)
        s.type = _ptr(self.type)
        s.dist_to_absorbing_boundary = _ptr(self.dist_to_absorbing)
        s.dist_to_reflecting_boundary = _ptr(self.dist_to_reflecting)
        s.estimate_boundary_normal_aligned = _ptr(self.aligned)
        s.first_sphere_radius = _ptr(self.first_sphere_radius)
        s.n_walks = _ptr(self.n_walks)
        s.cv_gradient = _ptr(self.cv_gradient)
        if self.cv_gradient_grid is not None:
            self._grid_c = self.cv_gradient_grid.to_c()   # keep the struct alive for the call
            s.cv_gradient_grid = C.pointer(self._grid_c)
        s.n_walks_max = int(self.n_walks_max)
        s.estimation_quantity = _ptr(self.estimation_quantity)
        s.direction_for_derivative = _ptr(self.direction_for_derivative)
        return s


@dataclass
class GridField:
    """Vector field on a regular grid: data [res_z, res_y, res_x, 4] float32 on the device."""
    data: object
    lo: tuple
    hi: tuple

    def to_c(self) -> A.zg_grid_field:
        g = A.zg_grid_field()
        g.data = _ptr(self.data)
        rz, ry, rx = self.data.shape[0], self.data.shape[1], self.data.shape[2]
        g.res[0], g.res[1], g.res[2] = rx, ry, rz
        for i in range(3):
            g.lo[i] = float(self.lo[i]) if i < len(self.lo) else 0.0
            g.hi[i] = float(self.hi[i]) if i < len(self.hi) else 0.0
        return g


@dataclass
class DeviceStats:
    n: object
    mean: object
    m2: object
    walk_length: object
    splits: object
    # gradient members (SampleStatistics, common.h:226-247), present with zeros(..., dim=d)
    n_grad: object = None
    grad_mean: object = None     # [n, dim]
    grad_m2: object = None       # [n, dim]
    first_source: object = None
    derivative: object = None

    @staticmethod
    def zeros(count: int, device="cuda", dim: int = 0) -> "DeviceStats":
        """dim > 0 also allocates the gradient members (needed by gradient points)."""
        torch = _torch()
        z64 = lambda *s: torch.zeros(*s, dtype=torch.float64, device=device)
        i64 = lambda *s: torch.zeros(*s, dtype=torch.int64, device=device)
        st = DeviceStats(i64(count), z64(count), z64(count), i64(count), i64(count))
        if dim > 0:
            st.n_grad, st.grad_mean, st.grad_m2 = i64(count), z64(count, dim), z64(count, dim)
            st.first_source, st.derivative = z64(count), z64(count)
        return st

    def validate(self, count: int, dim: int, device, need_gradient: bool = False):
        """Element types, lengths, contiguity and device of every member the solve writes;
        need_gradient: the gradient members must be present (gradient points)."""
        torch = _torch()
        i64, f64 = torch.int64, torch.float64
        for name in ("n", "mean", "m2", "walk_length", "splits"):
            if getattr(self, name) is None:
                raise ValueError(f"statistics member {name} is required")
        grad = ("n_grad", "grad_mean", "grad_m2", "first_source", "derivative")
        if need_gradient and any(getattr(self, g) is None for g in grad):
            raise ValueError("gradient points need the gradient statistics (DeviceStats.zeros(..., dim=d))")
        for name, dtype, shape in (("n", i64, (count,)), ("mean", f64, (count,)), ("m2", f64, (count,)),
                                   ("walk_length", i64, (count,)), ("splits", i64, (count,)),
                                   ("n_grad", i64, (count,)), ("grad_mean", f64, (count, dim)),
                                   ("grad_m2", f64, (count, dim)), ("first_source", f64, (count,)),
                                   ("derivative", f64, (count,))):
            check_array(f"stats.{name}", getattr(self, name), dtype, shape, device)

    def to_c(self) -> A.zg_statistics:
        s = A.zg_statistics()
        s.n_solution_estimates = _ptr(self.n)
        s.solution_mean = _ptr(self.mean)
        s.solution_m2 = _ptr(self.m2)
        s.total_walk_length = _ptr(self.walk_length)
        s.total_splits = _ptr(self.splits)
        s.n_gradient_estimates = _ptr(self.n_grad)
        s.gradient_mean = _ptr(self.grad_mean)
        s.gradient_m2 = _ptr(self.grad_m2)
        s.total_first_source_contribution = _ptr(self.first_source)
        s.total_derivative_contribution = _ptr(self.derivative)
        return s

    def variance(self):
        torch = _torch()
        return self.m2 / torch.clamp(self.n - 1, min=1).to(torch.float64)

    def grad_variance(self):
        torch = _torch()
        return self.grad_m2 / torch.clamp(self.n_grad - 1, min=1).to(torch.float64)[:, None]

    def derivative_mean(self):
        """getEstimatedDerivative (common.h:353-357): total / max(1, solution estimates)"""
        torch = _torch()
        return self.derivative / torch.clamp(self.n, min=1).to(torch.float64)


class SdfScene(Scene):
    """SDF-grid Dirichlet geometry (utils/sdf_grid_geometric_queries.h): signed distances at the
    cell centres of a grid over [lo, hi] (values[shape], last axis fastest), trilinear.  Used like
    a Scene by closest_point (|sdf| + projection), inside_domain, intersect_ray (sphere tracing of
    the zero level set) and wos_solve."""

    def __init__(self, values, lo, hi, domain_min=None, domain_max=None):
        vals = np.ascontiguousarray(values, dtype=np.float32)
        self.dim = vals.ndim
        shape = np.ascontiguousarray(vals.shape, dtype=np.int32)
        lo = np.ascontiguousarray(lo, dtype=np.float32)
        hi = np.ascontiguousarray(hi, dtype=np.float32)
        dmin = None if domain_min is None else np.ascontiguousarray(domain_min, dtype=np.float32)
        dmax = None if domain_max is None else np.ascontiguousarray(domain_max, dtype=np.float32)
        h = C.c_void_p()
        A.check(A.lib().zg_sdf_scene_create(self.dim, vals.ctypes.data, shape.ctypes.data, lo.ctypes.data,
                                            hi.ctypes.data, None if dmin is None else dmin.ctypes.data,
                                            None if dmax is None else dmax.ctypes.data, C.byref(h)))
        self._h = h
        self.domain_min = lo if dmin is None else dmin
        self.domain_max = hi if dmax is None else dmax
        info = A.zg_scene_info()
        A.check(A.lib().zg_scene_get_info(self._h, C.byref(info)))
        self.info = info


def spatially_sorted_point_indices(points) -> np.ndarray:
    """getSpatiallySortedPointIndices (utils/nearest_neighbor_finder.h:112-128): the kd-tree
    leaf order of host points [n, dim] (int32 permutation); runs on the host."""
    pts = np.ascontiguousarray(points, np.float32)
    if pts.ndim != 2 or pts.shape[1] not in (2, 3):
        raise ValueError(f"points must be [n, 2] or [n, 3], got {pts.shape}")
    out = np.empty(len(pts), np.int32)
    A.check(A.lib().zg_spatially_sorted_indices(pts.shape[1], pts.ctypes.data, len(pts), out.ctypes.data))
    return out


def stratified_samples(dims: int, n_tables: int, n_samples: int, seed: int, stream_offset: int = 0,
                       device="cuda", stream=None):
    """The per-point Latin hypercube tables of gradient solves: [n_tables, n_samples, dims]."""
    torch = _torch()
    out = torch.empty((n_tables, n_samples, dims), dtype=torch.float32, device=device)
    A.check(A.lib().zg_stratified_samples(dims, n_tables, n_samples, seed, stream_offset, _ptr(out),
                                          _stream_ptr(stream)))
    return out


def _validate_solve(dim: int, points: DevicePoints, stats: DeviceStats, instance_of_point=None):
    points.validate(dim)
    need_grad = False
    if points.estimation_quantity is not None and points.n > 0:
        need_grad = bool((points.estimation_quantity == A.ZG_ESTIMATE_SOLUTION_AND_GRADIENT).any())
    stats.validate(points.n, dim, points.pt.device, need_grad)
    if instance_of_point is not None:
        check_array("instance_of_point", instance_of_point, _torch().int32, (points.n,), points.pt.device)


def wos_solve(scene: Scene, pde: A.zg_pde, ws: A.zg_walk_settings, points: DevicePoints, stats: DeviceStats,
              opts: SolveOptions | None = None, stream=None, report: bool = True) -> SolveReport | None:
    """zg_wos_solve.  report=False returns as soon as the work is enqueued on `stream` (no host
    wait; device-side errors surface at synchronize()) and returns None."""
    opts = opts or SolveOptions()
    _validate_solve(scene.dim, points, stats)
    sp = points.to_c()
    st = stats.to_c()
    oc = opts.to_c()
    rep = A.zg_solve_report() if report else None
    A.check(A.lib().zg_wos_solve(scene.handle, C.byref(pde), C.byref(ws), C.byref(sp), C.byref(st), C.byref(oc),
                                 C.byref(rep) if report else None, _stream_ptr(stream)))
    return SolveReport.from_c(rep) if report else None


def synchronize() -> None:
    """zg_synchronize: waits for the device and raises the errors of solves run without a report."""
    A.check(A.lib().zg_synchronize())


def wos_solve_noalloc(scene: Scene, pde: A.zg_pde, ws: A.zg_walk_settings, points: DevicePoints,
                      stats: DeviceStats, opts: SolveOptions | None = None, stream=None) -> SolveReport:
    """wos_solve through zg_wos_solve_noalloc: the workspace is the library's fixed per-GPU device
    array (no allocation by the call); solution estimates only, at most 1,048,572 points per call;
    calls that overlap on one GPU race.  The report's kernel_ms is 0."""
    opts = opts or SolveOptions()
    _validate_solve(scene.dim, points, stats)
    sp = points.to_c()
    st = stats.to_c()
    oc = opts.to_c()
    rep = A.zg_solve_report()
    A.check(A.lib().zg_wos_solve_noalloc(scene.handle, C.byref(pde), C.byref(ws), C.byref(sp), C.byref(st),
                                         C.byref(oc), C.byref(rep), _stream_ptr(stream)))
    return SolveReport.from_c(rep)


def probe_read_bandwidth(nbytes: int, reps: int, device="cuda", stream=None) -> float:
    """Read bandwidth in GB/s of a float4 streaming read of an nbytes buffer, repeated reps times
    (an L2-sized buffer measures the cache ceiling, a large one the HBM ceiling)."""
    torch = _torch()
    buf = torch.ones(nbytes // 4, dtype=torch.float32, device=device)
    ms = C.c_float()
    A.check(A.lib().zg_probe_read_bandwidth(_ptr(buf), buf.numel() * 4, reps, C.byref(ms), _stream_ptr(stream)))
    return buf.numel() * 4 * reps / (ms.value * 1e-3) / 1e9


class Batch:
    """Many independent 3D instances (scene + PDE each) solved by one persistent launch.

    Points of all instances are concatenated; `instance_of_point` (int32 device tensor) names
    each point's instance.  The scenes are referenced, not copied, and must outlive the batch.
    """

    def __init__(self, scenes, pdes):
        if len(scenes) != len(pdes) or not scenes:
            raise ValueError("need one PDE per scene and at least one instance")
        arr = (C.c_void_p * len(scenes))(*[s.handle.value for s in scenes])
        parr = (A.zg_pde * len(pdes))(*pdes)
        h = C.c_void_p()
        A.check(A.lib().zg_batch_create(len(scenes), arr, parr, C.byref(h)))
        self._h = h
        self.scenes = list(scenes)
        self.pdes = list(pdes)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            A.lib().zg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_wos_solve(batch: Batch, ws: A.zg_walk_settings, points: DevicePoints, instance_of_point,
                    stats: DeviceStats, opts: SolveOptions | None = None, stream=None) -> SolveReport:
    opts = opts or SolveOptions()
    if instance_of_point is None:
        raise ValueError("instance_of_point is required")
    _validate_solve(3, points, stats, instance_of_point)
    sp = points.to_c()
    st = stats.to_c()
    oc = opts.to_c()
    rep = A.zg_solve_report()
    A.check(A.lib().zg_batch_wos_solve(batch.handle, C.byref(ws), C.byref(sp), _ptr(instance_of_point),
                                       C.byref(st), C.byref(oc), C.byref(rep), _stream_ptr(stream)))
    return SolveReport.from_c(rep)


class SolveGraph:
    """A batched solve captured once in a hipGraph; every launch() replays it with fresh
    PCG32 streams and accumulates into the same statistics (kept alive by this object)."""

    def __init__(self, batch: Batch, ws: A.zg_walk_settings, points: DevicePoints, instance_of_point,
                 stats: DeviceStats, opts: SolveOptions | None = None):
        opts = opts or SolveOptions()
        self._keep = (batch, points, instance_of_point, stats)
        if instance_of_point is None:
            raise ValueError("instance_of_point is required")
        _validate_solve(3, points, stats, instance_of_point)
        sp = points.to_c()
        st = stats.to_c()
        oc = opts.to_c()
        h = C.c_void_p()
        A.check(A.lib().zg_batch_wos_graph_create(batch.handle, C.byref(ws), C.byref(sp), _ptr(instance_of_point),
                                                  C.byref(st), C.byref(oc), C.byref(h)))
        self._h = h

    def launch(self, stream=None):
        A.check(A.lib().zg_graph_launch(self._h, _stream_ptr(stream)))

    def report(self, stream=None) -> SolveReport:
        rep = A.zg_solve_report()
        A.check(A.lib().zg_graph_read_report(self._h, C.byref(rep), _stream_ptr(stream)))
        return SolveReport.from_c(rep)

    def close(self):
        if getattr(self, "_h", None):
            A.lib().zg_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wost_solve(scene: Scene, pde: A.zg_pde, ws: A.zg_walk_settings, points: DevicePoints, stats: DeviceStats,
               opts: SolveOptions | None = None, stream=None, report: bool = True) -> SolveReport | None:
    """zg_wost_solve; report=False as in wos_solve."""
    opts = opts or SolveOptions()
    _validate_solve(scene.dim, points, stats)
    sp = points.to_c()
    st = stats.to_c()
    oc = opts.to_c()
    rep = A.zg_solve_report() if report else None
    A.check(A.lib().zg_wost_solve(scene.handle, C.byref(pde), C.byref(ws), C.byref(sp), C.byref(st), C.byref(oc),
                                  C.byref(rep) if report else None, _stream_ptr(stream)))
    return SolveReport.from_c(rep) if report else None


# ------------------------------------------------------------------------------------
# boundary value caching splat
# ------------------------------------------------------------------------------------
@dataclass
class BvcSamples:
    """A sample cache on the device (SamplePoint fields the splat reads)."""
    pt: object
    normal: object
    type: object                     # uint8: ZG_SAMPLE_ON_ABSORBING_BOUNDARY / _ON_REFLECTING_ / _IN_DOMAIN
    pdf: object
    solution: object
    normal_derivative: object
    contribution: object = None      # Robin data h (reflecting) or source value (domain)
    robin_coeff: object = None
    normal_aligned: object = None

    def to_c(self) -> A.zg_bvc_samples:
        s = A.zg_bvc_samples()
        s.n = int(self.pt.shape[0])
        s.pt, s.normal, s.type, s.pdf = _ptr(self.pt), _ptr(self.normal), _ptr(self.type), _ptr(self.pdf)
        s.solution, s.normal_derivative = _ptr(self.solution), _ptr(self.normal_derivative)
        s.contribution, s.robin_coeff = _ptr(self.contribution), _ptr(self.robin_coeff)
        s.normal_aligned = _ptr(self.normal_aligned)
        return s


@dataclass
class BvcEvalPoints:
    pt: object
    dist_to_absorbing: object
    dist_to_reflecting: object = None
    type: object = None

    def to_c(self) -> A.zg_bvc_eval_points:
        e = A.zg_bvc_eval_points()
        e.n = int(self.pt.shape[0])
        e.pt, e.type = _ptr(self.pt), _ptr(self.type)
        e.dist_to_absorbing_boundary, e.dist_to_reflecting_boundary = _ptr(self.dist_to_absorbing), _ptr(self.dist_to_reflecting)
        return e


@dataclass
class BvcStats:
    """Five SampleStatistics per evaluation point (absorbing, absorbing aligned, reflecting,
    reflecting aligned, source) in FP64."""
    n: object
    mean: object
    m2: object
    grad_mean: object
    grad_m2: object

    @staticmethod
    def zeros(count: int, dim: int, device="cuda") -> "BvcStats":
        torch = _torch()
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=device)
        return BvcStats(torch.zeros(count, 5, dtype=torch.int64, device=device), z(count, 5), z(count, 5),
                        z(count, 5, dim), z(count, 5, dim))

    def to_c(self) -> A.zg_bvc_statistics:
        s = A.zg_bvc_statistics()
        s.n, s.solution_mean, s.solution_m2 = _ptr(self.n), _ptr(self.mean), _ptr(self.m2)
        s.gradient_mean, s.gradient_m2 = _ptr(self.grad_mean), _ptr(self.grad_m2)
        return s

    def solution(self):
        """EvaluationPoint::getEstimatedSolution: the sum of the five means"""
        return self.mean.sum(1)

    def gradient(self):
        return self.grad_mean.sum(1)


def bvc_splat(dim: int, samples: BvcSamples, evals: BvcEvalPoints, stats: BvcStats, radius_clamp: float = 1e-3,
              kernel_regularization: float = 0.0, robin_coeff_cutoff: float = float("inf"),
              cutoff_dist_to_absorbing: float = 0.0, cutoff_dist_to_reflecting: float = 0.0, absorption: float = 0.0,
              stream=None) -> None:
    """BoundaryValueCaching::splat (boundary_value_caching.h:370-402) on the GPU."""
    s, e, st = samples.to_c(), evals.to_c(), stats.to_c()
    A.check(A.lib().zg_bvc_splat(dim, absorption, C.byref(s), C.byref(e), radius_clamp, kernel_regularization,
                                 min(robin_coeff_cutoff, 3.0e38), cutoff_dist_to_absorbing, cutoff_dist_to_reflecting,
                                 C.byref(st), _stream_ptr(stream)))
