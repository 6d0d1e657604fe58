"""Boundary value caching on the GPU (variance_reduction/boundary_value_caching.h).

`BoundaryValueCachingSolver` mirrors the reference class of the same name
(boundary_value_caching.h:178-238): cache samples on the absorbing and reflecting boundaries
(both orientations for double-sided problems) and in the domain, estimate the boundary data of
the cached samples with the GPU walk solvers (solution, and the normal derivative through the
gradient estimator where it is unknown), then splat every cache into evaluation points with the
free-space kernels (`zg_bvc_splat`, one GPU thread per evaluation point).  Points near the
boundary can be solved directly (`estimate_solution_near_boundary`).

Samplers: `UniformBoundarySampler` restates UniformLineSegmentBoundarySampler /
UniformTriangleBoundarySampler (boundary_samplers.h:16-124, 216-400): primitives offset along the
averaged vertex normals (inwards for the default orientation, outwards for normal-aligned samples),
picked in proportion to offset area x weight, a uniform point on the picked primitive, density
weight / sampling mass; `UniformDomainSampler` restates domain_samplers.h:24-100.  The host draws
the sample positions (numpy PCG64, stratified in the primitive CDF and the box); all walks and
the splat run on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from . import engine as E


def _torch():
    import torch
    return torch


def _stratified(n: int, dims: int, rng: np.random.Generator) -> np.ndarray:
    """Latin hypercube samples in [0, 1)^dims (the role of generateStratifiedSamples)"""
    s = (np.arange(n)[:, None] + rng.random((n, dims))) / max(n, 1)
    for d in range(dims):
        s[:, d] = s[rng.permutation(n), d]
    return np.minimum(s, 1.0 - np.finfo(np.float32).eps)


class UniformBoundarySampler:
    """Area-uniform samples on a boundary mesh (segments in 2D, triangles in 3D)."""

    def __init__(self, positions, indices, weights=None, weights_aligned=None):
        self.V = np.asarray(positions, np.float64)
        self.I = np.asarray(indices, np.int64)
        self.dim = self.V.shape[1]
        m = len(self.I)
        self.w = np.ones(m) if weights is None else np.asarray(weights, np.float64)
        self.wa = np.ones(m) if weights_aligned is None else np.asarray(weights_aligned, np.float64)
        if np.any(~np.isfinite(self.w)) or np.any(self.w < 0) or np.any(~np.isfinite(self.wa)) or np.any(self.wa < 0):
            raise ValueError("weights must be non-negative and finite")
        # vertex normals: normalised sum of the adjacent unit primitive normals (computeNormals)
        n = self._prim_normals(self.V)
        vn = np.zeros_like(self.V)
        for k in range(self.dim):
            np.add.at(vn, self.I[:, k], n)
        self.vn = vn / np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
        self.tables = {}

    def _prim_normals(self, V, normalize=True):
        P = V[self.I]
        if self.dim == 2:
            s = P[:, 1] - P[:, 0]
            n = np.stack([s[:, 1], -s[:, 0]], 1)
        else:
            n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30) if normalize else n

    def _table(self, offset, weights):
        V = self.V + offset * self.vn
        a = np.linalg.norm(self._prim_normals(V, normalize=False), axis=1)
        if self.dim == 3:
            a = 0.5 * a
        mass_terms = a * weights
        mass = float(mass_terms.sum())
        cdf = np.cumsum(mass_terms) / mass if mass > 0 else np.zeros(len(a))
        return V, cdf, mass, weights

    def initialize(self, normal_offset: float, solve_double_sided: bool):
        """CDF tables over the inward-offset primitives (and the outward ones when double sided)"""
        self.tables = {False: self._table(-normal_offset, self.w)}
        if solve_double_sided:
            self.tables[True] = self._table(normal_offset, self.wa)

    def sample_count(self, n_total: int, aligned: bool = False) -> int:
        masses = {k: t[2] for k, t in self.tables.items()}
        total = sum(masses.values())
        if total <= 0 or aligned not in masses:
            return 0
        return int(math.ceil(n_total * masses[aligned] / total))

    def generate(self, n: int, aligned: bool, rng: np.random.Generator):
        """positions, unit normals, densities of n samples"""
        V, cdf, mass, weights = self.tables[aligned]
        if n <= 0 or mass <= 0:
            return np.zeros((0, self.dim)), np.zeros((0, self.dim)), np.zeros(0)
        u = _stratified(n, 1, rng)[:, 0]
        prim = np.minimum(np.searchsorted(cdf, u, side="right"), len(cdf) - 1)
        P = V[self.I[prim]]
        r = rng.random((n, self.dim - 1))
        if self.dim == 2:
            pt = P[:, 0] + r[:, :1] * (P[:, 1] - P[:, 0])
        else:
            u1 = np.sqrt(r[:, :1]); u2 = r[:, 1:2]
            a, b = 1 - u1, u2 * u1
            pt = P[:, 0] * a + P[:, 1] * b + P[:, 2] * (1 - a - b)
        nrm = self._prim_normals(V)[prim]
        pdf = weights[prim] / mass
        return pt, nrm, pdf


class UniformDomainSampler:
    """Uniform samples inside the solve region (stratified over its bounding box)."""

    def __init__(self, scene: E.Scene, region_min, region_max, region_volume: float):
        self.scene = scene
        self.lo = np.asarray(region_min, np.float64)
        self.hi = np.asarray(region_max, np.float64)
        self.volume = float(region_volume)

    def generate(self, n: int, rng: np.random.Generator):
        torch = _torch()
        ext = self.hi - self.lo
        m = int(min(n * np.prod(ext) / self.volume, 1e8)) if self.volume > 0 else n
        pts = (self.lo + ext * _stratified(m, len(ext), rng)).astype(np.float32)
        inside = self.scene.inside_domain(torch.from_numpy(pts).cuda()).bool().cpu().numpy()
        pts = pts[inside]
        return pts, np.full(len(pts), 1.0 / self.volume)


@dataclass
class SampleCache:
    """one cache: host arrays of its samples and their estimated boundary data"""
    kind: int                       # A.ZG_SAMPLE_*
    aligned: bool
    pt: np.ndarray
    normal: np.ndarray
    pdf: np.ndarray
    solution: np.ndarray = None
    normal_derivative: np.ndarray = None
    contribution: np.ndarray = None
    robin_coeff: np.ndarray = None

    def __len__(self):
        return len(self.pt)

    def device(self) -> E.BvcSamples:
        torch = _torch()
        n = len(self.pt)
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a if a is not None else np.zeros(n), np.float32)).cuda()
        return E.BvcSamples(pt=f(self.pt), normal=f(self.normal),
                            type=torch.full((n,), self.kind, dtype=torch.uint8, device="cuda"), pdf=f(self.pdf),
                            solution=f(self.solution), normal_derivative=f(self.normal_derivative),
                            contribution=f(self.contribution), robin_coeff=f(self.robin_coeff),
                            normal_aligned=torch.full((n,), 1 if self.aligned else 0, dtype=torch.uint8, device="cuda"))


class BoundaryValueCachingSolver:
    """BoundaryValueCachingSolver (boundary_value_caching.h:178-238, 811-946) over the GPU solvers."""

    def __init__(self, scene: E.Scene, absorbing_sampler: UniformBoundarySampler | None,
                 reflecting_sampler: UniformBoundarySampler | None, domain_sampler: UniformDomainSampler | None,
                 seed: int = 0x5EED):
        self.scene = scene
        self.dim = scene.dim
        self.abs_sampler, self.refl_sampler, self.dom_sampler = absorbing_sampler, reflecting_sampler, domain_sampler
        self.rng = np.random.default_rng(seed)
        self.caches: list[SampleCache] = []
        self.seed = seed
        self.calls = 0
        self.absorption = 0.0   # the PDE's, set by compute_sample_estimates: Yukawa kernels when > 0 (3D)

    # -------- generateSamples (:846-874) --------
    def generate_samples(self, absorbing_cache_size: int, reflecting_cache_size: int, domain_cache_size: int,
                         normal_offset_absorbing: float, normal_offset_reflecting: float, solve_double_sided: bool):
        self.caches = []
        for sampler, size, kind, off in ((self.abs_sampler, absorbing_cache_size, A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY,
                                          normal_offset_absorbing),
                                         (self.refl_sampler, reflecting_cache_size, A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY,
                                          normal_offset_reflecting)):
            if sampler is None:
                continue
            sampler.initialize(off, solve_double_sided)
            for aligned in ((False, True) if solve_double_sided else (False,)):
                pt, nrm, pdf = sampler.generate(sampler.sample_count(size, aligned), aligned, self.rng)
                self.caches.append(SampleCache(kind, aligned, pt, nrm, pdf))
        if self.dom_sampler is not None and domain_cache_size > 0:
            pt, pdf = self.dom_sampler.generate(domain_cache_size, self.rng)
            self.caches.append(SampleCache(A.ZG_SAMPLE_IN_DOMAIN, False, pt, np.zeros_like(pt), pdf))

    # -------- computeSampleEstimates (:876-901) --------
    def compute_sample_estimates(self, pde: A.zg_pde, ws: A.zg_walk_settings, n_walks_solution: int,
                                 n_walks_gradient: int, robin_coeff_cutoff: float = float("inf")):
        torch = _torch()
        dev = "cuda"
        # the splat uses the free-space kernels of the PDE (boundary_value_caching.h:386-387,422-423)
        self.absorption = float(pde.absorption_coeff)
        for c in self.caches:
            n = len(c)
            if n == 0:
                continue
            pt = torch.from_numpy(np.ascontiguousarray(c.pt, np.float32)).to(dev)
            nrm = torch.from_numpy(np.ascontiguousarray(c.normal, np.float32)).to(dev)
            if c.kind == A.ZG_SAMPLE_IN_DOMAIN:    # setSourceValues (:345-368)
                c.contribution = E.evaluate_function(self.dim, pde.source, pt).cpu().numpy()
                continue
            oriented = -c.normal if (ws.solve_double_sided and c.aligned) else c.normal
            aligned_t = torch.full((n,), 1 if c.aligned else 0, dtype=torch.uint8, device=dev)
            # setEstimationData (:563-626)
            kappa = np.zeros(n, np.float32)
            if c.kind == A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY and not pde.are_robin_conditions_pure_neumann:
                kappa = E.evaluate_function(self.dim, pde.robin_coeff, pt, nrm,
                                            aligned_t if ws.solve_double_sided else None).cpu().numpy()
            grad = (c.kind == A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY) | (np.abs(kappa) > robin_coeff_cutoff)
            walks = np.where(grad, n_walks_gradient, n_walks_solution).astype(np.int32)
            dA = self.scene.closest_point(pt, boundary=0, want_point=False)[0].contiguous()
            dR = None
            if self.scene.info.n_reflecting_prims > 0:
                dR = self.scene.closest_point(pt, boundary=1, want_point=False)[0].contiguous()
            dp = E.DevicePoints(
                pt=pt, dist_to_absorbing=dA, dist_to_reflecting=dR,
                n_walks=torch.from_numpy(walks).to(dev), n_walks_max=int(max(walks.max(), 1)),
                type=torch.full((n,), c.kind, dtype=torch.uint8, device=dev), normal=nrm, aligned=aligned_t,
                estimation_quantity=torch.from_numpy(np.where(grad, A.ZG_ESTIMATE_SOLUTION_AND_GRADIENT,
                                                              A.ZG_ESTIMATE_SOLUTION).astype(np.uint8)).to(dev),
                direction_for_derivative=torch.from_numpy(np.ascontiguousarray(oriented, np.float32)).to(dev))
            st = E.DeviceStats.zeros(n, dim=self.dim)
            opts = E.SolveOptions(seed=self.seed + 1000003 * (self.calls + 1))
            self.calls += 1
            if self.scene.info.n_reflecting_prims > 0:
                E.wost_solve(self.scene, pde, ws, dp, st, opts)
            else:
                E.wos_solve(self.scene, pde, ws, dp, st, opts)
            # setEstimatedBoundaryData (:628-681)
            c.solution = st.mean.cpu().numpy()
            c.robin_coeff = kappa
            deriv = st.derivative_mean().cpu().numpy()
            if c.kind == A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY:
                c.normal_derivative = np.zeros(n)
                if not ws.ignore_reflecting_boundary_contribution:
                    h = E.evaluate_function(self.dim, pde.robin, pt, nrm,
                                            aligned_t if ws.solve_double_sided else None).cpu().numpy()
                    if pde.are_robin_conditions_pure_neumann:
                        c.normal_derivative = h
                    else:
                        c.contribution = h
                        c.normal_derivative = np.where(np.abs(kappa) > robin_coeff_cutoff, deriv, 0.0)
            else:
                c.normal_derivative = deriv

    # -------- splat (:903-929) --------
    def splat(self, evals: E.BvcEvalPoints, stats: E.BvcStats, radius_clamp: float = 1e-3,
              kernel_regularization: float = 0.0, robin_coeff_cutoff: float = float("inf"),
              cutoff_dist_to_absorbing: float = 0.0, cutoff_dist_to_reflecting: float = 0.0):
        for c in self.caches:
            if len(c) == 0:
                continue
            E.bvc_splat(self.dim, c.device(), evals, stats, radius_clamp, kernel_regularization, robin_coeff_cutoff,
                        cutoff_dist_to_absorbing, cutoff_dist_to_reflecting, absorption=self.absorption)

    # -------- estimateSolutionNearBoundary (:931-945, :476-512) --------
    def estimate_solution_near_boundary(self, pde: A.zg_pde, ws: A.zg_walk_settings, evals: E.BvcEvalPoints,
                                        stats: E.BvcStats, cutoff_dist_to_absorbing: float,
                                        cutoff_dist_to_reflecting: float, n_walks: int):
        """Evaluation points within a cutoff get a direct walk estimate in place of the splat: the
        statistics of such a point are reset and its estimate stored in the category of its type
        (source for interior points)."""
        torch = _torch()
        near = evals.dist_to_absorbing < cutoff_dist_to_absorbing
        if evals.dist_to_reflecting is not None:
            near = near | (evals.dist_to_reflecting < cutoff_dist_to_reflecting)
        idx = torch.nonzero(near).flatten()
        if idx.numel() == 0:
            return
        n = idx.numel()
        ty = evals.type[idx] if evals.type is not None else torch.zeros(n, dtype=torch.uint8, device="cuda")
        dp = E.DevicePoints(pt=evals.pt[idx].contiguous(), dist_to_absorbing=evals.dist_to_absorbing[idx].contiguous(),
                            dist_to_reflecting=None if evals.dist_to_reflecting is None else
                            evals.dist_to_reflecting[idx].contiguous(),
                            n_walks=torch.full((n,), n_walks, dtype=torch.int32, device="cuda"), n_walks_max=n_walks,
                            type=ty.contiguous())
        st = E.DeviceStats.zeros(n)
        opts = E.SolveOptions(seed=self.seed + 7777777 + self.calls)
        self.calls += 1
        if self.scene.info.n_reflecting_prims > 0:
            E.wost_solve(self.scene, pde, ws, dp, st, opts)
        else:
            E.wos_solve(self.scene, pde, ws, dp, st, opts)
        cat = torch.where(ty == A.ZG_SAMPLE_ON_ABSORBING_BOUNDARY, 0,
                          torch.where(ty == A.ZG_SAMPLE_ON_REFLECTING_BOUNDARY, 2, 4)).long()
        for t in (stats.n, stats.mean, stats.m2, stats.grad_mean, stats.grad_m2):
            t[idx] = 0
        stats.n[idx, cat] = 1
        stats.mean[idx, cat] = st.mean


def signed_volume(positions, indices) -> float:
    """Signed area (2D polyline) or volume (3D triangle mesh) enclosed by a boundary mesh: the
    sum of the origin-based simplex volumes of its primitives (computeSignedVolume,
    fcpw_geometric_queries.h:430-456)."""
    V = np.asarray(positions, np.float64)
    P = V[np.asarray(indices, np.int64)]
    if V.shape[1] == 2:
        return float(0.5 * np.sum(P[:, 0, 0] * P[:, 1, 1] - P[:, 0, 1] * P[:, 1, 0]))
    return float(np.sum(np.einsum("ij,ij->i", np.cross(P[:, 0], P[:, 1]), P[:, 2])) / 6.0)


# demo/demo.cpp:211-248: the settings runBoundaryValueCaching reads, with its defaults
# (normalOffsetForAbsorbingBoundary defaults to 5 x the absorbing epsilon)
DEMO_DEFAULTS = {
    "nWalksForCachedSolutionEstimates": 128, "nWalksForCachedGradientEstimates": 640,
    "absorbingBoundaryCacheSize": 1024, "reflectingBoundaryCacheSize": 1024, "domainCacheSize": 1024,
    "robinCoeffCutoffForNormalDerivative": float("inf"), "normalOffsetForReflectingBoundary": 0.0,
    "radiusClampForKernels": 0.0, "regularizationForKernels": 0.0,
}


def solve_like_demo(scene: E.Scene, pde: A.zg_pde, ws: A.zg_walk_settings, absorbing_mesh, reflecting_mesh,
                    pt, dist_to_absorbing, dist_to_reflecting, solver_cfg: dict, domain_volume: float | None = None,
                    seed: int = 0x5EED):
    """runBoundaryValueCaching (demo/demo.cpp:199-320) on the GPU for interior points `pt`
    (device [n, dim]): boundary samplers on the absorbing and reflecting meshes ((positions,
    indices) pairs, None when empty), a domain sampler when the source counts, generateSamples,
    computeSampleEstimates, splat with the normal offsets as cutoffs, estimateSolutionNearBoundary
    for the points within them.  Returns the per-point solution (device, FP64)."""
    cfg = {**DEMO_DEFAULTS, **solver_cfg}
    eps_a = float(cfg.get("epsilonShellForAbsorbingBoundary", ws.epsilon_shell_for_absorbing_boundary))
    off_a = float(cfg.get("normalOffsetForAbsorbingBoundary", 5.0 * eps_a))
    off_r = float(cfg["normalOffsetForReflectingBoundary"])
    robin_cut = float(cfg["robinCoeffCutoffForNormalDerivative"])
    n_sol, n_grad = int(cfg["nWalksForCachedSolutionEstimates"]), int(cfg["nWalksForCachedGradientEstimates"])
    dom_size = 0 if ws.ignore_source_contribution else int(cfg["domainCacheSize"])
    double_sided = bool(ws.solve_double_sided)
    samplers = []
    for mesh in (absorbing_mesh, reflecting_mesh):
        ok = mesh is not None and len(mesh[1]) > 0
        samplers.append(UniformBoundarySampler(*mesh) if ok else None)
    dom = None
    if dom_size > 0:
        # the solve region: the bounding domain when double sided, else the domain itself
        # (demo/demo.cpp:250-262)
        lo, hi = np.asarray(scene.domain_min, np.float64), np.asarray(scene.domain_max, np.float64)
        if double_sided:
            vol = float(np.prod(hi - lo))
        elif domain_volume is not None:
            vol = float(domain_volume)
        else:
            vol = abs(sum(signed_volume(*m) for m in (absorbing_mesh, reflecting_mesh) if m is not None and len(m[1])))
        dom = UniformDomainSampler(scene, lo, hi, vol)
    solver = BoundaryValueCachingSolver(scene, samplers[0], samplers[1], dom, seed=seed)
    solver.generate_samples(int(cfg["absorbingBoundaryCacheSize"]), int(cfg["reflectingBoundaryCacheSize"]),
                            dom_size, off_a, off_r, double_sided)
    solver.compute_sample_estimates(pde, ws, n_sol, n_grad, robin_cut)
    ev = E.BvcEvalPoints(pt=pt, dist_to_absorbing=dist_to_absorbing, dist_to_reflecting=dist_to_reflecting)
    st = E.BvcStats.zeros(int(pt.shape[0]), scene.dim)
    solver.splat(ev, st, float(cfg["radiusClampForKernels"]), float(cfg["regularizationForKernels"]), robin_cut,
                 off_a, off_r)
    solver.estimate_solution_near_boundary(pde, ws, ev, st, off_a, off_r, n_sol)
    return st.solution()
