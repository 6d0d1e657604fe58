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
