"""The header-only C++ adapter (include/zombie_amd.hpp) called from C++ as a reference caller
calls WalkOnSpheres / WalkOnStars::solve, against the array API on the same walk streams.

tests/adapter_check.cpp (built by __graft_entry__.build() into zombie_amd/_build/adapter_check)
runs two progressive batch solve() calls and one single-point solve(); the same solves through
zombie_amd.engine (seed s, stream offsets 0 and n_walks) must give equal counts and walk lengths
and FP64 means equal up to the order of the per-walk atomics."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import _build
from zombie_amd import engine as E
from zombie_amd import problems as PB

pytestmark = pytest.mark.gpu


def _write_input(path, p, pts, dA, dR, walks, seed, wost):
    refl = p.is_reflecting if p.is_reflecting is not None else np.zeros(len(p.indices), np.uint8)
    ws = p.settings
    with open(path, "wb") as f:
        f.write(np.array([2, int(wost), len(p.positions), len(p.indices), len(pts), walks], "<i4").tobytes())
        f.write(np.array([seed], "<u8").tobytes())
        f.write(np.ascontiguousarray(p.positions, "<f4").tobytes())
        f.write(np.ascontiguousarray(p.indices, "<i4").tobytes())
        f.write(np.ascontiguousarray(refl, np.uint8).tobytes())
        f.write(np.ascontiguousarray(p.domain_min, "<f4").tobytes())
        f.write(np.ascontiguousarray(p.domain_max, "<f4").tobytes())
        f.write(np.ascontiguousarray(pts, "<f4").tobytes())
        f.write(np.ascontiguousarray(dA, "<f4").tobytes())
        f.write(np.ascontiguousarray(dR, "<f4").tobytes())
        f.write(C.string_at(C.addressof(p.pde), C.sizeof(p.pde)))
        f.write(np.array([ws.epsilon_shell_for_absorbing_boundary, ws.epsilon_shell_for_reflecting_boundary,
                          ws.silhouette_precision, ws.russian_roulette_threshold, ws.splitting_threshold],
                         "<f4").tobytes())
        f.write(np.array([ws.max_walk_length, ws.steps_before_applying_tikhonov,
                          ws.steps_before_using_maximal_spheres], "<i4").tobytes())
        f.write(np.array([ws.solve_double_sided], np.uint8).tobytes())


def _read_output(path, n):
    raw = open(path, "rb").read()
    rows = np.frombuffer(raw[: n * 32], dtype=np.dtype([("n", "<i8"), ("len", "<i8"), ("mean", "<f8"),
                                                       ("m2", "<f8")]))
    tail = np.frombuffer(raw[n * 32:], dtype=np.dtype([("n", "<i8"), ("mean", "<f8")]))
    return rows, tail[0]


@pytest.mark.parametrize("wost", [False, True])
def test_cpp_adapter_matches_array_api(gpu, tmp_path, wost):
    torch = gpu
    exe = _build.ADAPTER_CHECK
    assert exe.exists(), "zombie_amd/_build/adapter_check is built by __graft_entry__.build()"
    if wost:
        p = PB.mixed2d(n_segments=512, n_slits=4)
        V, I = p.meta["outer"]
        double_sided = True
    else:
        p = PB.star2d()
        V, I = p.positions, p.indices
        double_sided = False
    scene = E.Scene(p.positions, p.indices, p.is_reflecting, p.domain_min, p.domain_max, bvh_width=4, leaf_size=4,
                    solve_double_sided=double_sided)
    cand = PB.sample_box(p.domain_min, p.domain_max, 3000, 17)
    pts = cand[PB.points_in_polygon(cand, V, I)][:600].astype(np.float32)
    P = torch.from_numpy(pts).cuda()
    dA, _, _, _ = scene.closest_point(P, boundary=0, want_point=False)
    if wost:
        dR, _, _, _ = scene.closest_point(P, boundary=1, want_point=False)
    else:
        dR = torch.full_like(dA, float(np.finfo(np.float32).max))
    walks, seed = 64, 4242
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, p, pts, dA.cpu().numpy(), dR.cpu().numpy(), walks, seed, wost)
    print(f"adapter: {'stars' if wost else 'spheres'}, {len(pts)} points x {walks} walks x 2 calls", flush=True)
    r = subprocess.run([str(exe), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, single = _read_output(outp, len(pts))

    n = len(pts)
    st = E.DeviceStats.zeros(n)
    fsr = torch.zeros(n, dtype=torch.float32, device="cuda") if wost else None
    dp = E.DevicePoints(pt=P, dist_to_absorbing=dA.contiguous(), dist_to_reflecting=dR.contiguous(),
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device="cuda"), n_walks_max=walks,
                        type=torch.zeros(n, dtype=torch.uint8, device="cuda"),
                        normal=torch.zeros((n, 2), dtype=torch.float32, device="cuda"),
                        aligned=torch.zeros(n, dtype=torch.uint8, device="cuda"), first_sphere_radius=fsr)
    solve = E.wost_solve if wost else E.wos_solve
    for call in range(2):
        solve(scene, p.pde, p.settings, dp, st, E.SolveOptions(seed=seed, stream_offset=call * walks))
    assert np.array_equal(got["n"], st.n.cpu().numpy())
    assert np.array_equal(got["len"], st.walk_length.cpu().numpy())
    np.testing.assert_allclose(got["mean"], st.mean.cpu().numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(got["m2"], st.m2.cpu().numpy(), rtol=1e-8, atol=1e-10)
    assert (got["n"] > 0).all()
    # the single-point overload: point 0 alone, first call of a fresh solver
    one = E.DeviceStats.zeros(1)
    dp1 = E.DevicePoints(pt=P[:1].contiguous(), dist_to_absorbing=dA[:1].contiguous(),
                         dist_to_reflecting=dR[:1].contiguous(),
                         n_walks=torch.full((1,), walks, dtype=torch.int32, device="cuda"), n_walks_max=walks,
                         type=torch.zeros(1, dtype=torch.uint8, device="cuda"),
                         normal=torch.zeros((1, 2), dtype=torch.float32, device="cuda"),
                         aligned=torch.zeros(1, dtype=torch.uint8, device="cuda"),
                         first_sphere_radius=torch.zeros(1, dtype=torch.float32, device="cuda") if wost else None)
    solve(scene, p.pde, p.settings, dp1, one, E.SolveOptions(seed=seed, stream_offset=0))
    assert single["n"] == int(one.n.cpu()[0])
    assert abs(single["mean"] - float(one.mean.cpu()[0])) <= 1e-10 * max(1.0, abs(single["mean"]))
