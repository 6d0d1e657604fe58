"""PDE data on the GPU: the device evaluation of every function descriptor equals the oracle's,
and a walk-on-spheres solve with Dirichlet data given as a dense grid agrees with the oracle
(4-SE) and with the closed form the grid samples."""
import numpy as np
import pytest

from zombie_amd import _abi as A
from zombie_amd import engine as E
from zombie_amd import problems as PB

from helpers import check_four_se, z_scores

pytestmark = pytest.mark.gpu


def _t(torch, a, dtype=np.float32):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


@pytest.mark.parametrize("dim", [2, 3])
def test_function_descriptors_match_oracle(gpu, oracle_lib, dim):
    torch = gpu
    O = oracle_lib
    rng = np.random.default_rng(dim)
    n = 2000
    pts = rng.uniform(-1.2, 1.2, size=(n, dim)).astype(np.float32)
    nr = rng.normal(size=(n, dim)).astype(np.float32)
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    al = (rng.random(n) < 0.5).astype(np.uint8)
    kinds = [(A.ZG_FN_CONSTANT, (2.5,)), (A.ZG_FN_EXP_COS, (1.0, 1.0)), (A.ZG_FN_SIN_PRODUCT, (1.0, 2.0)),
             (A.ZG_FN_CUBIC_HARMONIC, (1.0,)), (A.ZG_FN_EXP_LINEAR, (1.0, 0.6, -0.8, 0.4))]
    for kind, prm in kinds:
        for ev in (A.ZG_EVAL_VALUE, A.ZG_EVAL_NORMAL_DERIVATIVE, A.ZG_EVAL_SCREENED_SOURCE):
            f = E.function(kind, ev, prm)
            g = E.evaluate_function(dim, f, _t(torch, pts), _t(torch, nr), _t(torch, al, np.uint8), 3.0).cpu().numpy()
            o = np.array([O.eval_function(dim, f, pts[i], nr[i], bool(al[i]), 3.0) for i in range(n)])
            # the device evaluates sin, cos and exp with the hardware instructions
            np.testing.assert_allclose(g, o, rtol=2e-5, atol=2e-5 * max(1.0, np.abs(o).max()),
                                       err_msg=f"kind {kind} eval {ev}")
    shape = (7, 5) if dim == 2 else (7, 5, 4)
    vals = rng.normal(size=shape).astype(np.float32)
    lo, hi = [-1.0] * dim, [1.0, 0.8, 1.1][:dim]
    for interp in (False, True):
        gd = E.GridFunction(vals, lo, hi, interpolate=interp)
        gh = E.GridFunction(vals, lo, hi, interpolate=interp, device=None)
        g = E.evaluate_function(dim, gd.fn, _t(torch, pts)).cpu().numpy()
        o = np.array([O.eval_function(dim, gh.fn, pts[i]) for i in range(n)])
        np.testing.assert_allclose(g, o, rtol=1e-6, atol=1e-6, err_msg=f"grid interpolate={interp}")


def test_wos_with_dense_grid_dirichlet(gpu, oracle_lib):
    torch = gpu
    O = oracle_lib
    p = PB.star2d()
    lo, hi = p.domain_min, p.domain_max
    res = 512
    c = [lo[a] + (np.arange(res) + 0.5) * (hi[a] - lo[a]) / res for a in range(2)]
    X, Y = np.meshgrid(c[0], c[1], indexing="ij")
    vals = np.exp(X) * np.cos(Y)
    gd = E.GridFunction(vals, lo, hi, interpolate=True)
    gh = E.GridFunction(vals, lo, hi, interpolate=True, device=None)
    pde_d = E.make_pde(dirichlet=gd.fn, is_source_constant=True)
    pde_h = E.make_pde(dirichlet=gh.fn, is_source_constant=True)
    pts = p.meta["points"][::20]
    osc = O.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    d, *_ = osc.closest_point(pts)
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    n = len(pts)
    st = E.DeviceStats.zeros(n)
    dp = E.DevicePoints(pt=_t(torch, pts), dist_to_absorbing=_t(torch, d),
                        n_walks=torch.full((n,), 256, dtype=torch.int32, device="cuda"), n_walks_max=256)
    E.wos_solve(sc, pde_d, p.settings, dp, st, E.SolveOptions(seed=3))
    g = dict(n=st.n.cpu().numpy(), mean=st.mean.cpu().numpy(), var=st.variance().cpu().numpy())
    ost = O.Stats(n)
    O.solve(osc, pde_h, p.settings, pts, d, 256, ost, seed=4, rng_mode=O.RNG_PER_POINT)
    z, ok = z_scores(g["mean"], g["var"], g["n"], ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, "dense-grid Dirichlet vs oracle")
    exact = p.exact(pts)
    pooled = (g["mean"] - exact).sum() / np.sqrt((g["var"] / np.maximum(g["n"], 1)).sum())
    assert abs(pooled) < 4.0, pooled
