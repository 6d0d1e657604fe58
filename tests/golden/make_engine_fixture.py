#!/usr/bin/env python3
"""Writes tests/golden/engine_wost.npz from the reference's engine demo problem.

The reference ships one model problem solved by its own Walk-on-Stars solver, with the result:
demo/model_problems/engine/wost.json (solver and output settings), data/geometry.obj (the
boundary polyline), data/absorbing_boundary_value.pfm (Dirichlet data), data/
is_reflecting_boundary.png (which segments are Neumann) and solutions/wost.pfm (the solution
on a 256 x 256 grid, 128 walks per pixel).  This script copies that data, as the reference's
demo reads it, into one compressed fixture:

- positions [V, 2] float32 and segments [L, 2] int32 (0-based) of the OBJ's `v` and `l`
  lines, before the demo's normalisation (model_problem.h:222-228);
- absorbing_value [h, w] float32: the Dirichlet image in the order Image<1>::readPFM stores
  it (file row order), reduced to one channel as setFromRGB does (0.299 R + 0.587 G + 0.114 B,
  demo/image.h:150-154);
- reflecting_mask [H, W] bool (packed): the reflecting-boundary image in the same storage
  order, > 0.  Only the PNG ships (the PFM the demo reads is missing from the reference
  checkout), so it is derived the way demo/model_problems/image2pfm.py makes the PFMs:
  rows flipped, values / 255.  The script checks that rule against the absorbing-value pair,
  for which both the PNG and the PFM ship;
- solution [256, 256] float32: wost.pfm's first channel (all three are equal), in file row
  order, as Image::writePFM wrote it (demo/image.h:265-290: file row h-1-j holds grid column
  j's y index);
- the wost.json settings the test needs;
- bvc_solution [256, 256] float32: solutions/bvc.pfm's first channel, the same problem solved
  by the reference's BoundaryValueCachingSolver (demo/demo.cpp:199-320 with bvc.json), and the
  bvc.json solver settings (bvc_settings / bvc_setting_names).

Run from the repository root:  python tests/golden/make_engine_fixture.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
from PIL import Image

ENGINE = Path("/root/reference/demo/model_problems/engine")
OUT = Path(__file__).resolve().parent / "engine_wost.npz"


def read_pfm(path: Path) -> np.ndarray:
    """[h, w, c] float32 in file row order (no flip), as Image::readPFM reads the samples."""
    with open(path, "rb") as f:
        kind = f.readline().strip()
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        c = 3 if kind == b"PF" else 1
        data = np.frombuffer(f.read(w * h * c * 4), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, c).astype(np.float32)


def gray(rgb: np.ndarray) -> np.ndarray:
    rgb = rgb.astype(np.float64)
    return (0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]).astype(np.float32)


def read_obj_polyline(path: Path):
    pos, seg = [], []
    for line in path.read_text().splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            pos.append((float(tok[1]), float(tok[2])))
        elif tok[0] == "l":
            idx = [int(t.split("/")[0]) - 1 for t in tok[1:]]
            seg.extend((idx[k], idx[k + 1]) for k in range(len(idx) - 1))
    return np.asarray(pos, np.float32), np.asarray(seg, np.int32)


def main() -> int:
    cfg = json.loads((ENGINE / "wost.json").read_text())
    positions, segments = read_obj_polyline(ENGINE / "data" / "geometry.obj")

    absorbing_pfm = read_pfm(ENGINE / "data" / "absorbing_boundary_value.pfm")
    absorbing_png = np.asarray(Image.open(ENGINE / "data" / "absorbing_boundary_value.png").convert("RGB"))
    # the PFM/PNG rule of image2pfm.py, checked on the pair that ships both
    derived = absorbing_png[::-1].astype(np.float32) / 255.0
    if np.abs(derived - absorbing_pfm).max() > 1e-6:
        raise SystemExit("absorbing_boundary_value.pfm is not the flipped PNG / 255: the mask rule is unproven")

    mask_png = np.asarray(Image.open(ENGINE / "data" / "is_reflecting_boundary.png").convert("RGB"))
    mask = gray(mask_png[::-1].astype(np.float32) / 255.0) > 0.0

    sol = read_pfm(ENGINE / "solutions" / "wost.pfm")
    if not (np.array_equal(sol[..., 0], sol[..., 1]) and np.array_equal(sol[..., 0], sol[..., 2])):
        raise SystemExit("wost.pfm channels differ")

    bvc_cfg = json.loads((ENGINE / "bvc.json").read_text())
    if bvc_cfg["modelProblem"] != cfg["modelProblem"] or bvc_cfg["output"]["gridRes"] != cfg["output"]["gridRes"]:
        raise SystemExit("bvc.json and wost.json describe different problems")
    bvc = read_pfm(ENGINE / "solutions" / "bvc.pfm")
    if not (np.array_equal(bvc[..., 0], bvc[..., 1]) and np.array_equal(bvc[..., 0], bvc[..., 2])):
        raise SystemExit("bvc.pfm channels differ")
    bvc_names = np.array(sorted(k for k, v in bvc_cfg["solver"].items() if isinstance(v, (int, float))))
    bvc_settings = np.array([float(bvc_cfg["solver"][k]) for k in bvc_names], np.float64)

    s, m, o = cfg["solver"], cfg["modelProblem"], cfg["output"]
    settings = np.array([s["nWalks"], s["maxWalkLength"], s["epsilonShellForAbsorbingBoundary"],
                         s["epsilonShellForReflectingBoundary"], s["russianRouletteThreshold"],
                         s["splittingThreshold"], float(s["ignoreAbsorbingBoundaryContribution"]),
                         float(s["ignoreReflectingBoundaryContribution"]), float(s["ignoreSourceContribution"]),
                         m["robinCoeff"], m["absorptionCoeff"], float(m["solveDoubleSided"]),
                         float(m["domainIsWatertight"]), o["gridRes"], o["boundaryDistanceMask"]], np.float64)
    names = np.array(["nWalks", "maxWalkLength", "epsilonShellForAbsorbingBoundary",
                      "epsilonShellForReflectingBoundary", "russianRouletteThreshold", "splittingThreshold",
                      "ignoreAbsorbingBoundaryContribution", "ignoreReflectingBoundaryContribution",
                      "ignoreSourceContribution", "robinCoeff", "absorptionCoeff", "solveDoubleSided",
                      "domainIsWatertight", "gridRes", "boundaryDistanceMask"])
    np.savez_compressed(OUT, positions=positions, segments=segments, absorbing_value=gray(absorbing_pfm),
                        reflecting_mask_bits=np.packbits(mask.reshape(-1)),
                        reflecting_mask_shape=np.array(mask.shape, np.int32),
                        solution=sol[..., 0].copy(), settings=settings, setting_names=names,
                        bvc_solution=bvc[..., 0].copy(), bvc_settings=bvc_settings, bvc_setting_names=bvc_names)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes): {len(positions)} vertices, {len(segments)} segments, "
          f"values {absorbing_pfm.shape[:2]}, mask {mask.shape}, solution {sol.shape[:2]}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
