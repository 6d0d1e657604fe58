"""Model-problem driver: the reference's demo (demo/demo.cpp, demo/demo.py) for its walk-on-
spheres, walk-on-stars and boundary value caching solvers, run on the GPU through
zombie_amd.zombie (and zombie_amd.bvc).

    python -m zombie_amd.demo path/to/model_problem.json [--root DIR] [--walks N] [--seed S]

reads a model problem in the reference's JSON format (solverType "wos", "wost" or "bvc"; paths
relative to --root, default the current directory), sets it up as demo/model_problem.h does
(OBJ polyline normalised and flipped, square bounding box, PDE data from images as nearest-cell
dense grids, boundary partitioned by the reflecting-boundary image at segment midpoints), solves
on the config's gridRes x gridRes grid and writes the solution as a PFM at output.solutionFile
in Image::writePFM's layout (demo/grid.h:42-95, demo/image.h:265-290).

Images are read as the C++ demo reads them: PFM samples in file row order, three channels
reduced to 0.299 R + 0.587 G + 0.114 B (demo/image.h:150-225); a missing PFM is replaced by the
PNG of the same name converted by the rule that made the reference's PFMs (rows flipped, / 255,
demo/model_problems/image2pfm.py).  "bvc" follows runBoundaryValueCaching (demo/demo.cpp:
199-320; zombie_amd.bvc.solve_like_demo).  Exterior problems (Kelvin transform), SDF absorbing
boundaries and reverse walk on stars are outside this driver.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import zombie


def read_image(path: Path) -> np.ndarray:
    """[h, w] float32 grayscale in the C++ demo's storage order (file rows of the PFM)."""
    if path.suffix == ".pfm" and not path.exists() and path.with_suffix(".png").exists():
        from PIL import Image
        rgb = np.asarray(Image.open(path.with_suffix(".png")).convert("RGB"))[::-1].astype(np.float32) / 255.0
    elif path.suffix == ".pfm":
        with open(path, "rb") as f:
            kind = f.readline().strip()
            w, h = map(int, f.readline().split())
            scale = float(f.readline())
            c = 3 if kind == b"PF" else 1
            data = np.frombuffer(f.read(w * h * c * 4), dtype="<f4" if scale < 0 else ">f4").reshape(h, w, c)
        rgb = np.repeat(data, 3, axis=2) if c == 1 else data
    else:
        from PIL import Image
        rgb = np.asarray(Image.open(path).convert("RGB"))[::-1].astype(np.float32) / 255.0
    rgb = rgb.astype(np.float64)
    return (0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]).astype(np.float32)


def write_pfm(path: Path, values: np.ndarray) -> None:
    """values [res, res] indexed [x index, y index]: written as Image::writePFM writes the
    demo's grid (pixel (row j, column i) holds grid point i*res + j, rows stored bottom-up)."""
    path.parent.mkdir(parents=True, exist_ok=True)
    img = np.ascontiguousarray(values.T[::-1].astype("<f4"))
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1\n".encode())
        f.write(np.repeat(img[..., None], 3, axis=2).tobytes())


def image_grid(img: np.ndarray, lo, length, channels: int = 1):
    """The demo's image lookup (Image::get(uv), uv = (x - bMin) / maxLength, rows flipped) as a
    nearest-cell dense grid over [bMin, bMin + maxLength]^2: cell (i, j) = img[h - 1 - j, i]."""
    h, w = img.shape
    data = np.ascontiguousarray(img[::-1, :].T).reshape(-1)
    return data, [w, h], np.asarray(lo, np.float32), np.asarray(lo, np.float32) + np.float32(length)


def run(config_path: str, root: str = ".", n_walks: int | None = None, seed: int = 0x5EED, out: str | None = None):
    cfg = json.loads(Path(config_path).read_text())
    root = Path(root)
    solver_type = cfg["solverType"]
    if solver_type not in ("wos", "wost", "bvc"):
        raise SystemExit(f"solverType {solver_type!r}: this driver runs 'wos', 'wost' and 'bvc'")
    S, M, O = cfg["solver"], cfg["modelProblem"], cfg["output"]
    if M.get("solveExterior", False) or M.get("useSdfForAbsorbingBoundary", False):
        raise SystemExit("exterior problems and SDF absorbing boundaries are outside this driver")
    dim = 2
    U, Core, Solvers = zombie.Utils, zombie.Core, zombie.Solvers

    # geometry (model_problem.h:222-228)
    positions, indices = zombie.FloatNList(dim=dim), zombie.IntNList(dim=dim)
    U.load_boundary_mesh(str(root / M["geometry"]), positions, indices, dim=dim)
    if M.get("normalizeDomain", True):
        U.normalize(positions, dim=dim)
    if M.get("flipOrientation", True):
        U.flip_orientation(indices, dim=dim)
    lo, hi = U.compute_bounding_box(positions, True, 1.0, dim=dim)
    length = float((hi - lo).max())

    # PDE (model_problem.h:243-277)
    robin_coeff = float(M.get("robinCoeff", 0.0))
    solve_double_sided = bool(M.get("solveDoubleSided", False))
    pde = Core.PDE(dim=dim)
    pde.source = U.get_dense_grid_source_callback(*image_grid(read_image(root / M["sourceValue"]), lo, length), dim=dim)
    pde.dirichlet = U.get_dense_grid_dirichlet_callback(
        *image_grid(read_image(root / M["absorbingBoundaryValue"]), lo, length), dim=dim)
    pde.robin = U.get_dense_grid_robin_callback(
        *image_grid(read_image(root / M["reflectingBoundaryValue"]), lo, length), dim=dim)
    mask_grid = U.FloatDenseGrid(*image_grid(read_image(root / M["isReflectingBoundary"]), lo, length), dim=dim)
    pde.has_reflecting_boundary_conditions = U.get_dense_grid_indicator_callback(mask_grid, dim=dim)
    pde.robin_coeff = Core.get_constant_robin_coefficient_callback(robin_coeff, dim=dim)
    pde.absorption_coeff = float(M.get("absorptionCoeff", 0.0))
    pde.is_source_constant = False
    pde.are_robin_conditions_pure_neumann = robin_coeff == 0.0
    pde.are_robin_coeffs_nonnegative = robin_coeff >= 0.0

    # boundary parts and queries (model_problem.h:280-328)
    ap, ai, rp, ri = (zombie.FloatNList(dim=dim), zombie.IntNList(dim=dim), zombie.FloatNList(dim=dim),
                      zombie.IntNList(dim=dim))
    U.partition_boundary_mesh(pde.has_reflecting_boundary_conditions, positions, indices, ap, ai, rp, ri, dim=dim)
    queries = Core.GeometricQueries(bool(M.get("domainIsWatertight", True)), lo, hi, dim=dim)
    absorbing = U.FcpwDirichletBoundaryHandler(dim=dim)
    absorbing.build_acceleration_structure(ap, ai)
    U.populate_geometric_queries_for_dirichlet_boundary(absorbing, queries, dim=dim)
    ignore_sil = U.get_ignore_candidate_silhouette_callback(solve_double_sided)
    weight = U.get_branch_traversal_weight_callback()
    if robin_coeff == 0.0:
        reflecting = U.FcpwNeumannBoundaryHandler(dim=dim)
        reflecting.build_acceleration_structure(rp, ri, ignore_sil)
        U.populate_geometric_queries_for_neumann_boundary(reflecting, weight, queries, dim=dim)
    else:
        reflecting = U.FcpwRobinBoundaryHandler(dim=dim)
        coeff = [abs(robin_coeff)] * len(ri)
        reflecting.build_acceleration_structure(rp, ri, ignore_sil, coeff, coeff)
        U.populate_geometric_queries_for_robin_boundary(reflecting, weight, queries, dim=dim)

    # grid, distances, sample points (demo.cpp:8-45, grid.h:23-40)
    res = int(O["gridRes"])
    ext = hi - lo
    ii, jj = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    grid = np.stack([(ii / np.float32(res)).astype(np.float32) * ext[0] + lo[0],
                     (jj / np.float32(res)).astype(np.float32) * ext[1] + lo[1]], -1).reshape(-1, 2)
    grid = grid.astype(np.float32)
    inside = U.inside_domain(queries, grid, dim=dim)
    valid = inside | solve_double_sided
    dA, dR = zombie.FloatList(), zombie.FloatList()
    U.compute_dist_to_boundary(queries, grid, False, dA, dR, dim=dim)
    dA, dR = np.asarray(dA, np.float32), np.asarray(dR, np.float32)
    idx = np.nonzero(valid)[0]
    pts = [] if solver_type == "bvc" else Solvers.SamplePointList(
        [Solvers.SamplePoint(grid[k], np.zeros(dim, np.float32), Solvers.SampleType.InDomain,
                             Solvers.EstimationQuantity.Solution, 1.0, float(dA[k]), float(dR[k]), dim=dim)
         for k in idx], dim=dim)
    walks = int(n_walks if n_walks is not None else S.get("nWalks", 128))
    max_len = int(S.get("maxWalkLength", 1024))
    eps_a = float(S.get("epsilonShellForAbsorbingBoundary", 1e-3))
    settings = Solvers.WalkSettings(
        eps_a, float(S.get("epsilonShellForReflectingBoundary", 1e-3)) if solver_type != "wos" else 0.0,
        float(S.get("silhouettePrecision", 1e-3)) if solver_type != "wos" else 0.0,
        float(S.get("russianRouletteThreshold", 0.0)), float(S.get("splittingThreshold", np.inf)), max_len,
        int(S.get("stepsBeforeApplyingTikhonov", 0)),
        int(S.get("stepsBeforeUsingMaximalSpheres", max_len)) if solver_type != "wos" else 0,
        solve_double_sided, not S.get("disableGradientControlVariates", False),
        not S.get("disableGradientAntitheticVariates", False),
        bool(S.get("useCosineSamplingForDirectionalDerivatives", False)),
        bool(S.get("ignoreAbsorbingBoundaryContribution", False)),
        bool(S.get("ignoreReflectingBoundaryContribution", False)) if solver_type != "wos" else True,
        bool(S.get("ignoreSourceContribution", False)), False)
    if solver_type == "bvc":
        import torch
        from . import bvc as BV
        scene = queries.scene(solve_double_sided, settings.silhouette_precision)
        meshes = [(np.asarray(p_, np.float32).reshape(-1, dim), np.asarray(i_, np.int32).reshape(-1, dim))
                  for p_, i_ in ((ap, ai), (rp, ri))]
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        sol = BV.solve_like_demo(scene, pde.channel(0), settings.to_c(), meshes[0], meshes[1], dev(grid[idx]),
                                 dev(dA[idx]), dev(dR[idx]), S, seed=seed).cpu().numpy().astype(np.float32)
        var = np.zeros_like(sol)
    else:
        stats = Solvers.create_sample_statistics_list(len(pts), dim=dim)
        Solver = Solvers.WalkOnStars if solver_type == "wost" else Solvers.WalkOnSpheres
        Solver(queries, dim=dim, seed=seed).solve(pde, settings, zombie.IntList([walks] * len(pts)), pts, stats)
        sol = np.array([s.get_estimated_solution() for s in stats], np.float32)
        var = np.array([s.get_estimated_solution_variance() for s in stats], np.float32)

    # grid values (grid.h:55-80): masked pixels are 0
    mask = float(O.get("boundaryDistanceMask", 0.0))
    values = np.zeros(res * res, np.float32)
    variance = np.zeros(res * res, np.float32)
    keep = np.minimum(dA[idx], dR[idx]) >= mask
    values[idx[keep]] = sol[keep]
    variance[idx[keep]] = var[keep]
    out_path = Path(out) if out else root / O.get("solutionFile", "solution.pfm")
    write_pfm(out_path, values.reshape(res, res))
    return dict(values=values.reshape(res, res), variance=variance.reshape(res, res), walks=walks,
                solved=int(len(idx)), output=str(out_path))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config")
    ap.add_argument("--root", default=".")
    ap.add_argument("--walks", type=int, default=None)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    r = run(a.config, a.root, a.walks, a.seed, a.out)
    print(f"solved {r['solved']} grid points x {r['walks']} walks -> {r['output']}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
