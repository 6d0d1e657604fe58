"""Batched screened-Poisson solves: many instances (scene + PDE each) walked by one
persistent launch, eager and captured in a hipGraph.

The per-walk PCG32 stream depends only on (seed, global point index, walk index), so a
batched solve must reproduce, walk for walk, a single-scene solve of the same points: the
counts and walk lengths match exactly and the FP64 means up to the order of the atomics.
Against the oracle and the closed form exp(a . x) (|a|^2 = sigma) the check is the 4-SE rule.
"""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB

from helpers import check_four_se, z_scores

pytestmark = pytest.mark.gpu

N_INST = 6
PTS_PER_INST = 160


def _instance_points(torch, prob, scene, n, seed):
    cand = PB.sample_box(prob.domain_min, prob.domain_max, 6 * n, seed)
    ct = torch.from_numpy(cand).cuda()
    inside = scene.inside_domain(ct).cpu().numpy().astype(bool)
    pts = cand[inside][:n]
    assert len(pts) == n
    d, *_ = scene.closest_point(torch.from_numpy(pts).cuda(), want_point=False)
    return pts, d.cpu().numpy()


@pytest.fixture(scope="module")
def batch_setup(gpu):
    torch = gpu
    rng = np.random.default_rng(7)
    probs, scenes, pts, dist, inst = [], [], [], [], []
    for i in range(N_INST):
        sigma = float(rng.choice([0.5, 1.0, 4.0, 9.0]))
        direction = rng.normal(size=3)
        pr = PB.screened_instance(1280, sigma, direction, seed=100 + i)
        sc = E.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max, bvh_width=4, leaf_size=2)
        p, d = _instance_points(torch, pr, sc, PTS_PER_INST, 200 + i)
        probs.append(pr); scenes.append(sc); pts.append(p); dist.append(d)
        inst.append(np.full(len(p), i, np.int32))
    batch = E.Batch(scenes, [pr.pde for pr in probs])
    return dict(probs=probs, scenes=scenes, batch=batch, pts=np.concatenate(pts), dist=np.concatenate(dist),
                inst=np.concatenate(inst), settings=probs[0].settings)


def _points(torch, pts, dist, n_walks):
    nw = n_walks if isinstance(n_walks, np.ndarray) else np.full(len(pts), n_walks, np.int32)
    return E.DevicePoints(pt=torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda(),
                          dist_to_absorbing=torch.from_numpy(np.ascontiguousarray(dist, np.float32)).cuda(),
                          n_walks=torch.from_numpy(nw.astype(np.int32)).cuda(), n_walks_max=int(nw.max()))


def _host(st):
    return dict(n=st.n.cpu().numpy(), mean=st.mean.cpu().numpy(), m2=st.m2.cpu().numpy(),
                var=st.variance().cpu().numpy(), len=st.walk_length.cpu().numpy())


def _same(a, b, sel=slice(None), rtol=1e-9):
    """Identical walks: equal counts and lengths; the means agree to `rtol` (1e-9: only the order
    of the FP64 atomics differs; 2e-6: two kernel instantiations whose float32 arithmetic the
    compiler may contract into FMAs differently)."""
    assert (a["n"][sel] == b["n"][sel]).all()
    assert (a["len"][sel] == b["len"][sel]).all()
    np.testing.assert_allclose(a["mean"][sel], b["mean"][sel], rtol=rtol, atol=1e-12)
    m2 = b["m2"][sel]
    np.testing.assert_allclose(a["m2"][sel], m2, rtol=100 * rtol, atol=100 * rtol * float(np.abs(m2).max()))


@pytest.mark.parametrize("tol", [0.0, 0.1])
def test_batch_reproduces_single_scene_solves(gpu, batch_setup, tol):
    torch = gpu
    B = batch_setup
    opts = E.SolveOptions(seed=31, walks_per_lane=2, regroup_interval=8, distance_tolerance=tol)
    n = len(B["pts"])
    inst_t = torch.from_numpy(B["inst"]).cuda()
    st = E.DeviceStats.zeros(n)
    rep = E.batch_wos_solve(B["batch"], B["settings"], _points(torch, B["pts"], B["dist"], 64), inst_t, st, opts)
    g = _host(st)
    assert rep.total_walks == 64 * n
    walks = 0
    for i in range(N_INST):
        # the same global point indices: other instances' points get zero walks
        sel = B["inst"] == i
        nw = np.where(sel, 64, 0).astype(np.int32)
        dist = np.where(sel, B["dist"], 1.0).astype(np.float32)
        si = E.DeviceStats.zeros(n)
        r = E.wos_solve(B["scenes"][i], B["probs"][i].pde, B["settings"], _points(torch, B["pts"], dist, nw), si, opts)
        walks += r.total_walks
        _same(g, _host(si), sel, rtol=2e-6)
    assert walks == rep.total_walks


def test_batch_parity_with_oracle_and_closed_form(gpu, oracle_lib, batch_setup):
    torch = gpu
    O = oracle_lib
    B = batch_setup
    n = len(B["pts"])
    st = E.DeviceStats.zeros(n)
    E.batch_wos_solve(B["batch"], B["settings"], _points(torch, B["pts"], B["dist"], 256),
                      torch.from_numpy(B["inst"]).cuda(), st, E.SolveOptions(seed=5, distance_tolerance=0.1))
    g = _host(st)
    zs, oks = [], []
    for i in range(N_INST):
        sel = B["inst"] == i
        pr = B["probs"][i]
        osc = O.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max)
        pts = B["pts"][sel]
        ost = O.Stats(len(pts))
        O.solve(osc, pr.pde, pr.settings, pts, B["dist"][sel], 256, ost, seed=6 + i, rng_mode=O.RNG_PER_POINT)
        z, ok = z_scores(g["mean"][sel], g["var"][sel], g["n"][sel], ost.mean, ost.variance(), ost.n)
        zs.append(z); oks.append(ok)
        # closed form: pooled z of the instance (walk-count weighted errors)
        exact = PB.exp_linear(pts, pr.meta["a"])
        err = g["mean"][sel] - exact
        pooled = err.sum() / np.sqrt((g["var"][sel] / g["n"][sel]).sum())
        assert abs(pooled) < 4.0, f"instance {i} (sigma {pr.meta['sigma']}): pooled z {pooled:.2f}"
    check_four_se(np.concatenate(zs), np.concatenate(oks), "batched screened Poisson vs oracle")


def test_graph_replays_match_eager_solves(gpu, batch_setup):
    """Each replay draws fresh streams (stream offset advanced on the device) and must equal
    the eager solve with that offset; eager work runs between the replays."""
    torch = gpu
    B = batch_setup
    n = len(B["pts"])
    nw = 32
    dp = _points(torch, B["pts"], B["dist"], nw)
    inst_t = torch.from_numpy(B["inst"]).cuda()
    opts = E.SolveOptions(seed=77, stream_offset=1000, walks_per_lane=1, regroup_interval=16, distance_tolerance=0.1)
    sg = E.DeviceStats.zeros(n)
    graph = E.SolveGraph(B["batch"], B["settings"], dp, inst_t, sg, opts)
    se = E.DeviceStats.zeros(n)
    try:
        for k in range(3):
            graph.launch()
            rep = graph.report()
            assert rep.total_walks == nw * n
            # eager work between replays: a solve of the same batch and some torch kernels
            o = E.SolveOptions(seed=77, stream_offset=1000 + k * nw, walks_per_lane=1, regroup_interval=16,
                               distance_tolerance=0.1)
            E.batch_wos_solve(B["batch"], B["settings"], dp, inst_t, se, o)
            x = torch.randn(512, 512, device="cuda")
            (x @ x).sum().item()
            torch.cuda.synchronize()
            _same(_host(sg), _host(se))
            assert (_host(sg)["n"] == nw * (k + 1)).all()
    finally:
        graph.close()


def test_batch_argument_errors(gpu, batch_setup):
    torch = gpu
    B = batch_setup
    n = len(B["pts"])
    dp = _points(torch, B["pts"], B["dist"], 4)
    st = E.DeviceStats.zeros(n)
    with pytest.raises((A.ZombieError, ValueError)):        # instance_of_point missing (host check)
        E.batch_wos_solve(B["batch"], B["settings"], dp, None, st)
    bad = B["inst"].copy()
    bad[3] = N_INST                                          # index past the table: flagged, walks skipped
    with pytest.raises(A.ZombieError, match="instance_of_point"):
        E.batch_wos_solve(B["batch"], B["settings"], dp, torch.from_numpy(bad).cuda(), st)
    assert st.n[3].item() == 0
    # 2D scenes are not batched; widths must agree
    s2 = PB.star2d()
    sc2 = E.Scene(s2.positions, s2.indices, None, s2.domain_min, s2.domain_max)
    with pytest.raises(A.ZombieError):
        E.Batch([sc2], [s2.pde])
    pr = B["probs"][0]
    sc8 = E.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max, bvh_width=8)
    with pytest.raises(A.ZombieError):
        E.Batch([B["scenes"][0], sc8], [pr.pde, pr.pde])


def test_batch_autotuner_covers_its_grid(gpu, batch_setup):
    from zombie_amd import autotune as AT
    torch = gpu
    B = batch_setup
    dp = _points(torch, B["pts"], B["dist"], 8)
    res = AT.autotune_batch(B["probs"], dp, torch.from_numpy(B["inst"]).cuda(), walks=8)
    assert len(res.table) == 2 * 3 * 3
    assert {(r["bvh_width"], r["walks_per_lane"], r["regroup_interval"]) for r in res.table} == \
        {(w, l, g) for w in (4, 8) for l in (1, 2, 4) for g in (8, 16, 32)}
    assert res.walks_per_s == max(r["walks_per_s"] for r in res.table)
    assert res.bvh_width in (4, 8) and res.walks_per_lane in (1, 2, 4) and res.regroup_interval in (8, 16, 32)


def test_uneven_instances_reproduce_single_scene_solves_and_closed_form(gpu, oracle_lib):
    """Instances of very different sizes in one launch, as in the benchmarked batch (geodesic
    spheres of 80 to 20,480 triangles, 40 to 640 points each, sigma from 0 to 9): walk for walk
    equal to single-scene solves, each instance's pooled error within 4 SE of exp(a . x), and
    the largest instance within the 4-SE rule of the oracle."""
    torch = gpu
    O = oracle_lib
    sizes = [(80, 40, 0.0), (320, 80, 9.0), (1280, 160, 1.0), (5120, 320, 4.0), (20480, 640, 0.25)]
    rng = np.random.default_rng(17)
    probs, scenes, pts, dist, inst = [], [], [], [], []
    for i, (nt, npts, sigma) in enumerate(sizes):
        pr = PB.screened_instance(nt, sigma, rng.normal(size=3), seed=300 + i)
        sc = E.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max, bvh_width=4, leaf_size=2)
        p, d = _instance_points(torch, pr, sc, npts, 400 + i)
        probs.append(pr); scenes.append(sc); pts.append(p); dist.append(d)
        inst.append(np.full(len(p), i, np.int32))
    tris = [len(pr.indices) for pr in probs]
    assert min(tris) <= 100 and max(tris) >= 20000, tris
    batch = E.Batch(scenes, [pr.pde for pr in probs])
    P, D, I = np.concatenate(pts), np.concatenate(dist), np.concatenate(inst)
    n = len(P)
    opts = E.SolveOptions(seed=41, walks_per_lane=1, regroup_interval=16, distance_tolerance=0.1)
    st = E.DeviceStats.zeros(n)
    rep = E.batch_wos_solve(batch, probs[0].settings, _points(torch, P, D, 256), torch.from_numpy(I).cuda(), st, opts)
    g = _host(st)
    assert rep.total_walks == 256 * n
    for i, pr in enumerate(probs):
        sel = I == i
        nw = np.where(sel, 256, 0).astype(np.int32)
        di = np.where(sel, D, 1.0).astype(np.float32)
        si = E.DeviceStats.zeros(n)
        E.wos_solve(scenes[i], pr.pde, pr.settings, _points(torch, P, di, nw), si, opts)
        _same(g, _host(si), sel, rtol=2e-6)
        err = g["mean"][sel] - PB.exp_linear(P[sel], pr.meta["a"])
        if sizes[i][2] == 0.0:
            # harmonic with the constant boundary value 1: every walk returns exactly 1
            assert np.abs(err).max() < 1e-6 and np.abs(g["var"][sel]).max() < 1e-9
            continue
        pooled = err.sum() / np.sqrt((g["var"][sel] / g["n"][sel]).sum())
        assert abs(pooled) < 4.0, f"instance {i} ({tris[i]} triangles, sigma {sizes[i][2]}): pooled z {pooled:.2f}"
    big = len(sizes) - 1
    sel = I == big
    pr = probs[big]
    osc = O.Scene(pr.positions, pr.indices, None, pr.domain_min, pr.domain_max)
    ost = O.Stats(int(sel.sum()))
    O.solve(osc, pr.pde, pr.settings, P[sel], D[sel], 256, ost, seed=43, rng_mode=O.RNG_PER_POINT)
    z, ok = z_scores(g["mean"][sel], g["var"][sel], g["n"][sel], ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, "uneven batch: largest instance vs oracle")
