"""The reference's own result as the yardstick: the engine demo problem
(demo/model_problems/engine/wost.json) solved by the GPU walk on stars and compared pixel by
pixel with solutions/wost.pfm, which the reference's WalkOnStars wrote (128 walks per pixel).

The problem is set up from tests/golden/engine_wost.npz exactly as the demo sets it up
(zombie_amd.problems.engine_wost).  Each pixel's reference value is a 128-walk mean of the same
estimator, so (GPU mean - reference) / sqrt(var/n_gpu + var/128), with var the per-walk
variance the GPU measured, is a z-score of two independent estimates of one quantity."""
from pathlib import Path

import numpy as np
import pytest

from zombie_amd import engine as E
from zombie_amd import problems as PB

from helpers import check_four_se

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "engine_wost.npz"


@pytest.fixture(scope="module")
def engine(gpu):
    torch = gpu
    p = PB.engine_wost(FIXTURE, device="cuda")
    sc = E.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, bvh_width=4, leaf_size=4)
    G = torch.from_numpy(p.meta["grid"]).cuda()
    inside = sc.inside_domain(G).bool()
    dA, _, _, _ = sc.closest_point(G, boundary=0, want_point=False)
    dR, _, _, _ = sc.closest_point(G, boundary=1, want_point=False)
    valid = inside & (torch.minimum(dA, dR) >= p.meta["mask_distance"])
    return p, sc, G, dA, dR, valid


def test_engine_setup_matches_reference_mask(engine):
    """The reference wrote 0 exactly where its point was outside the domain or within the mask
    distance of the boundary: our set-up must find the same pixels (up to the inside test of one
    grid row whose +x rays run along a horizontal boundary segment)."""
    p, sc, G, dA, dR, valid = engine
    ref = p.meta["solution"].reshape(-1)
    v = valid.cpu().numpy()
    assert (v & (ref == 0)).sum() == 0
    extra = np.nonzero(~v & (ref != 0))[0]
    assert len(extra) <= 256
    assert len(np.unique(extra % 256)) <= 1   # one grid row (y index) at most


def test_engine_wost_matches_reference_solution(gpu, engine):
    torch = gpu
    p, sc, G, dA, dR, valid = engine
    idx = torch.nonzero(valid).squeeze(1)
    n = idx.numel()
    walks = 1024
    print(f"engine: {n} pixels x {walks} walks", flush=True)
    dp = E.DevicePoints(pt=G[idx].contiguous(), dist_to_absorbing=dA[idx].contiguous(),
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device="cuda"), n_walks_max=walks,
                        dist_to_reflecting=dR[idx].contiguous())
    st = E.DeviceStats.zeros(n)
    rep = E.wost_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=2024))
    mean = st.mean.cpu().numpy()
    var = st.variance().cpu().numpy()
    cnt = st.n.cpu().numpy()
    ref = p.meta["solution"].reshape(-1)[idx.cpu().numpy()]
    assert rep.completion_codes[0] >= 0.999 * rep.total_walks   # walks end on the absorbing boundary
    se2 = var / np.maximum(cnt, 1) + var / p.meta["reference_walks"]
    ok = se2 > 0
    z = np.zeros_like(mean)
    z[ok] = (mean[ok] - ref[ok]) / np.sqrt(se2[ok])
    exceed = check_four_se(z, ok, "engine GPU vs reference wost.pfm")
    print(f"engine: z mean {z[ok].mean():+.4f} std {z[ok].std():.4f}, {exceed} beyond 4 SE", flush=True)
    # pixels without variance (every walk sees the same boundary value) agree exactly
    assert np.abs(mean[~ok] - ref[~ok]).max(initial=0.0) < 1e-5
    # 16 x 16 pixel regions: mean difference against its standard error
    xi = (idx.cpu().numpy() // 256) // 16
    yi = (idx.cpu().numpy() % 256) // 16
    region = xi * 16 + yi
    zr = []
    for r in np.unique(region):
        m = region == r
        if m.sum() < 8:
            continue
        diff = (mean[m] - ref[m]).mean()
        s = np.sqrt(se2[m].sum()) / m.sum()
        if s > 0:
            zr.append(diff / s)
    zr = np.array(zr)
    assert np.abs(zr).max() < 4.5, f"region z max {np.abs(zr).max():.2f}"
    assert abs(zr.mean()) < 0.3 and 0.7 < zr.std() < 1.3, f"region z mean {zr.mean():.3f} std {zr.std():.3f}"


def _write_pfm_gray(path, img):
    """img [h, w] in the demo reader's storage order (file rows), one channel."""
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"Pf\n{w} {h}\n-1\n".encode())
        f.write(np.ascontiguousarray(img, "<f4").tobytes())


def _write_engine_files(z, tmp_path):
    """The engine problem's data as the reference's demo reads them (OBJ, PFM images)."""
    data = tmp_path / "data"
    data.mkdir()
    with open(data / "geometry.obj", "w") as f:
        for p in z["positions"]:
            f.write(f"v {float(p[0])!r} {float(p[1])!r} 0\n")
        for a, b in z["segments"]:
            f.write(f"l {a + 1} {b + 1}\n")
    _write_pfm_gray(data / "absorbing_boundary_value.pfm", z["absorbing_value"])
    shape = tuple(int(v) for v in z["reflecting_mask_shape"])
    mask = np.unpackbits(z["reflecting_mask_bits"])[: shape[0] * shape[1]].reshape(shape).astype(np.float32)
    _write_pfm_gray(data / "is_reflecting_boundary.pfm", mask)
    _write_pfm_gray(data / "reflecting_boundary_value.pfm", np.zeros((8, 8), np.float32))
    _write_pfm_gray(data / "source_value.pfm", np.zeros((8, 8), np.float32))


def _model_problem(cfg):
    return {"geometry": "data/geometry.obj", "isReflectingBoundary": "data/is_reflecting_boundary.pfm",
            "absorbingBoundaryValue": "data/absorbing_boundary_value.pfm",
            "reflectingBoundaryValue": "data/reflecting_boundary_value.pfm",
            "sourceValue": "data/source_value.pfm", "robinCoeff": cfg["robinCoeff"],
            "absorptionCoeff": cfg["absorptionCoeff"], "solveDoubleSided": False,
            "solveExterior": False, "domainIsWatertight": True, "useSdfForAbsorbingBoundary": False}


def test_engine_json_problem_through_the_reference_python_surface(gpu, tmp_path):
    """The engine model problem as files in the reference's format (OBJ, PFM images, wost.json),
    run by zombie_amd.demo through the reference-shaped layer zombie_amd.zombie (the flow of
    demo/demo.py), compared with the reference's own solution."""
    import json
    from zombie_amd import demo

    z = np.load(FIXTURE, allow_pickle=False)
    cfg = dict(zip([str(n) for n in z["setting_names"]], z["settings"].tolist()))
    _write_engine_files(z, tmp_path)
    problem = {
        "solverType": "wost",
        "solver": {"nWalks": int(cfg["nWalks"]), "maxWalkLength": int(cfg["maxWalkLength"]),
                   "epsilonShellForAbsorbingBoundary": cfg["epsilonShellForAbsorbingBoundary"],
                   "epsilonShellForReflectingBoundary": cfg["epsilonShellForReflectingBoundary"],
                   "russianRouletteThreshold": cfg["russianRouletteThreshold"],
                   "splittingThreshold": cfg["splittingThreshold"],
                   "ignoreAbsorbingBoundaryContribution": bool(cfg["ignoreAbsorbingBoundaryContribution"]),
                   "ignoreReflectingBoundaryContribution": bool(cfg["ignoreReflectingBoundaryContribution"]),
                   "ignoreSourceContribution": bool(cfg["ignoreSourceContribution"])},
        "modelProblem": _model_problem(cfg),
        "output": {"solutionFile": "solutions/wost.pfm", "gridRes": int(cfg["gridRes"]),
                   "boundaryDistanceMask": cfg["boundaryDistanceMask"]},
    }
    (tmp_path / "wost.json").write_text(json.dumps(problem))
    print("engine demo driver: 1024 walks per pixel", flush=True)
    r = demo.run(str(tmp_path / "wost.json"), root=str(tmp_path), n_walks=1024, seed=77)
    out = demo.read_image(tmp_path / "solutions" / "wost.pfm")        # file rows, as the demo stores it
    ref_file = np.ascontiguousarray(z["solution"])
    got, var, ref = r["values"].reshape(-1), r["variance"].reshape(-1), PB.engine_wost(FIXTURE, device=None).meta["solution"].reshape(-1)
    # the written PFM is laid out as the reference's wost.pfm
    assert out.shape == ref_file.shape
    assert np.array_equal(out != 0, ref_file != 0) or ((out != 0) != (ref_file != 0)).sum() <= 256
    both = (got != 0) & (ref != 0)
    assert both.sum() > 0.95 * (ref != 0).sum()
    se2 = var[both] / r["walks"] + var[both] / 128.0
    ok = se2 > 0
    zz = (got[both][ok] - ref[both][ok]) / np.sqrt(se2[ok])
    check_four_se(zz, np.ones(len(zz), bool), "engine demo driver vs reference wost.pfm")


def test_engine_bvc_through_the_demo_driver_matches_reference_bvc(gpu, tmp_path):
    """The engine problem with the reference's boundary value caching settings (bvc.json: caches
    of 3,072 absorbing and 3,072 reflecting samples, 96 walks per solution and 960 per gradient
    estimate) run by the demo driver's "bvc" path, against solutions/bvc.pfm, which the
    reference's BoundaryValueCachingSolver wrote for the same problem.

    A BVC solution is one realisation of a spatially correlated estimator (every pixel sums the
    same random caches), so pixel z-scores do not apply.  Three runs with independent seeds
    measure the estimator's own run-to-run spread: if the reference's solution is another draw
    of the same estimator, its mean squared difference to our runs equals the mean squared
    difference between our runs (ratio 1; a bias or a different estimator raises it), and the
    difference of its pixel mean to ours stays within the spread of our runs' pixel means."""
    import json
    from zombie_amd import demo

    z = np.load(FIXTURE, allow_pickle=False)
    cfg = dict(zip([str(n) for n in z["setting_names"]], z["settings"].tolist()))
    bcfg = dict(zip([str(n) for n in z["bvc_setting_names"]], z["bvc_settings"].tolist()))
    _write_engine_files(z, tmp_path)
    solver = {k: (bool(v) if k.startswith("ignore") else int(v) if float(v).is_integer() and "Threshold" not in k
                  and "epsilon" not in k and "radius" not in k else v) for k, v in bcfg.items()}
    problem = {"solverType": "bvc", "solver": solver, "modelProblem": _model_problem(cfg),
               "output": {"solutionFile": "solutions/bvc.pfm", "gridRes": int(cfg["gridRes"]),
                          "boundaryDistanceMask": cfg["boundaryDistanceMask"]}}
    (tmp_path / "bvc.json").write_text(json.dumps(problem))
    runs = []
    for seed in (101, 202, 303):
        print(f"engine bvc: run with seed {seed}", flush=True)
        r = demo.run(str(tmp_path / "bvc.json"), root=str(tmp_path), seed=seed)
        runs.append(r["values"].reshape(-1).astype(np.float64))
    ref = np.ascontiguousarray(z["bvc_solution"][::-1, :].T).reshape(-1).astype(np.float64)
    both = (runs[0] != 0) & (ref != 0)
    assert both.sum() > 0.95 * (ref != 0).sum()
    ours = np.stack([r[both] for r in runs])
    R = ref[both]
    d_ours = np.mean([np.mean((ours[i] - ours[j]) ** 2) for i in range(3) for j in range(i + 1, 3)])
    d_ref = np.mean([np.mean((R - ours[i]) ** 2) for i in range(3)])
    means = ours.mean(1)
    spread = means.std(ddof=1)
    print(f"engine bvc: rms between our runs {np.sqrt(d_ours):.4f}, rms to the reference {np.sqrt(d_ref):.4f}, "
          f"ratio {d_ref / d_ours:.3f}; pixel means ours {means} reference {R.mean():.4f}", flush=True)
    assert d_ref / d_ours < 1.5, (d_ref, d_ours)
    assert abs(R.mean() - means.mean()) < 4.0 * spread * np.sqrt(1 + 1 / 3) + 2e-3, (R.mean(), means)
