"""The numerator of the walk kernel's roofline: BVH nodes and primitives an exact nearest-first
closest-point query reads, counted on the GPU (zg_wos_solve with count_visits 2, the count the
bench takes at the walks' own positions) and by the oracle's restatement of the same traversal
over the packed BVH the GPU holds (oracle/packed_bvh.py, itself pinned by brute force)."""
import numpy as np
import pytest

from oracle import packed_bvh as PKB
from zombie_amd import engine as E
from zombie_amd import problems as PB

pytestmark = pytest.mark.gpu


def _brute_dist(V, F, x):
    best = np.full(len(x), np.inf)
    for t in range(len(F)):
        a, b, c = V[F[t, 0]], V[F[t, 1]], V[F[t, 2]]
        d2 = np.array([PKB._closest_on_triangle2(a, b, c, p) for p in x])
        best = np.minimum(best, d2)
    return np.sqrt(best)


@pytest.mark.parametrize("width,leaf", [(4, 2), (8, 4)])
def test_canonical_visit_count_matches_oracle_traversal(gpu, width, leaf):
    torch = gpu
    V, F = PB.icosphere(3)
    V = PB.sh_displace(V, 31).astype(np.float32)
    F = F.astype(np.int32)
    p = PB.sphere3d()
    lo, hi = V.min(0) - 1e-3, V.max(0) + 1e-3
    scene = E.Scene(V, F, None, lo, hi, bvh_width=width, leaf_size=leaf)
    nodes, prims = scene.copy_bvh(0)
    bvh = PKB.PackedBvh(nodes, prims, 3, width)
    rng = np.random.default_rng(5)
    x = rng.uniform(lo, hi, size=(240, 3)).astype(np.float32)
    res = [bvh.closest(q) for q in x]
    d_or = np.sqrt([r[0] for r in res])
    # the restated traversal is an exact closest-point query
    d_bf = _brute_dist(V.astype(np.float64), F, x.astype(np.float64)[:60])
    np.testing.assert_allclose(d_or[:60], d_bf, rtol=1e-6, atol=1e-7)
    # one exact projection per point: distances 0 put every point inside the epsilon shell
    n = len(x)
    dp = E.DevicePoints(pt=torch.from_numpy(x).cuda(), dist_to_absorbing=torch.zeros(n, device="cuda"),
                        n_walks=torch.ones(n, dtype=torch.int32, device="cuda"), n_walks_max=1)
    print(f"canonical count: width {width}, leaf {leaf}, {n} queries", flush=True)
    rep = E.wos_solve(scene, p.pde, p.settings, dp, E.DeviceStats.zeros(n), E.SolveOptions(seed=1, count_visits=2))
    nodes_or = sum(r[1] for r in res)
    prims_or = sum(r[2] for r in res)
    assert rep.total_walks == n
    assert abs(rep.node_visits - nodes_or) <= 0.002 * nodes_or, (rep.node_visits, nodes_or)
    assert abs(rep.prim_visits - prims_or) <= 0.002 * prims_or, (rep.prim_visits, prims_or)
    # the kernel's own count of the same projection queries (unseeded, exact) is that traversal too
    rep1 = E.wos_solve(scene, p.pde, p.settings, dp, E.DeviceStats.zeros(n), E.SolveOptions(seed=1, count_visits=1))
    assert rep1.node_visits == rep.node_visits and rep1.prim_visits == rep.prim_visits
