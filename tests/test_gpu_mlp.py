"""GPU MLP kernels (MFMA bf16 forward/backward, Adam) against the FP64 numpy oracle, and the
learned control variate's unbiasedness and gain."""
import numpy as np
import pytest

from oracle import mlp as M
from zombie_amd import engine as E
from zombie_amd import learned_cv as LC
from zombie_amd import problems as PB

from helpers import check_four_se, z_scores

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net(gpu):
    cfg = LC.MLPConfig(in_dim=3, n_frequencies=32, hidden=128, n_hidden_layers=4, frequency_sigma=4.0, seed=7)
    return LC.MLP(cfg, max_batch=4096)


def test_mlp_loss_and_parameter_gradients_match_oracle(gpu, net):
    torch = gpu
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(1000, 3)).astype(np.float32)
    y = np.sin(2 * x).prod(1).astype(np.float32)
    loss, g = net.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
    theta, B = net.get_parameters(), net.frequencies()
    lb, gb = M.gradients(theta, B, x, y, 128, 4, bf16=True)
    l64, g64 = M.gradients(theta, B, x, y, 128, 4, bf16=False)
    assert abs(loss - lb) < 1e-3 * abs(lb) + 1e-6
    assert _rel(g, gb) < 0.02, _rel(g, gb)      # same bf16 rounding points, fp32 vs fp64 sums
    assert _rel(g, g64) < 0.05, _rel(g, g64)    # bf16 operands vs the exact model


def test_ragged_shapes_match_oracle(gpu):
    """Widths that are not multiples of the 64-column lane chunks or of the 32-deep MFMA k-tile
    (hidden 200, 24 Fourier features), a 2D input and a batch that is not a multiple of a tile."""
    torch = gpu
    net = LC.MLP(LC.MLPConfig(in_dim=2, n_frequencies=12, hidden=200, n_hidden_layers=2, frequency_sigma=3.0, seed=11),
                 max_batch=2048)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(999, 2)).astype(np.float32)
    y = np.sin(3 * x).prod(1).astype(np.float32)
    loss, g = net.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
    theta, B = net.get_parameters(), net.frequencies()
    lb, gb = M.gradients(theta, B, x, y, 200, 2, bf16=True)
    assert abs(loss - lb) < 1e-3 * abs(lb) + 1e-6
    assert _rel(g, gb) < 0.02, _rel(g, gb)
    v, gx = net.evaluate(torch.from_numpy(x).cuda())
    cb, gxb = M.input_gradient(theta, B, x, 200, 2, bf16=True)
    assert _rel(v.cpu().numpy(), cb) < 0.01
    assert _rel(gx.cpu().numpy(), gxb) < 0.02


def test_mlp_value_and_input_gradient_match_oracle(gpu, net):
    torch = gpu
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(3000, 3)).astype(np.float32)   # > max_batch: chunked
    v, gx = net.evaluate(torch.from_numpy(x).cuda())
    theta, B = net.get_parameters(), net.frequencies()
    cb, gb = M.input_gradient(theta, B, x, 128, 4, bf16=True)
    c64, g64 = M.input_gradient(theta, B, x, 128, 4, bf16=False)
    assert _rel(v.cpu().numpy(), cb) < 0.01 and _rel(v.cpu().numpy(), c64) < 0.03
    assert _rel(gx.cpu().numpy(), gb) < 0.02 and _rel(gx.cpu().numpy(), g64) < 0.05


def test_adam_training_fits_a_smooth_field(gpu):
    torch = gpu
    net = LC.MLP(LC.MLPConfig(seed=3), max_batch=8192)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.uniform(-1, 1, size=(65536, 3)).astype(np.float32)).cuda()
    y = torch.sin(2 * x).prod(1).contiguous()
    losses = LC.train(net, x, y, steps=400, batch=8192, lr=1e-3, log_every=50)
    assert losses[-1][1] < 0.05 * losses[0][1], losses


@pytest.fixture(scope="module")
def ball(oracle_lib):
    V, F = PB.icosphere(4)
    V = PB.sh_displace(V, 11)
    p = PB.sphere3d()
    p = PB.Problem("ball3", 3, V.astype(np.float32), F.astype(np.int32), None,
                   (V.min(0) - 1e-3).astype(np.float32), (V.max(0) + 1e-3).astype(np.float32),
                   p.pde, p.settings, 64, p.exact)
    sc = oracle_lib.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    cand = PB.sample_box(p.domain_min, p.domain_max, 8000, 5)
    pts = cand[sc.inside_domain(cand)][:2000]
    d, *_ = sc.closest_point(pts)
    return p, pts, d, sc


def _solve(torch, p, pts, d, walks, seed, grad=None):
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    n = len(pts)
    dp = E.DevicePoints(pt=torch.from_numpy(pts).cuda(), dist_to_absorbing=torch.from_numpy(d).cuda(),
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device="cuda"), n_walks_max=walks,
                        cv_gradient=None if grad is None else torch.as_tensor(grad, dtype=torch.float32).cuda())
    st = E.DeviceStats.zeros(n)
    E.wos_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=seed))
    return st.mean.cpu().numpy(), st.variance().cpu().numpy(), st.n.cpu().numpy()


def test_exact_gradient_control_variate_is_unbiased_and_reduces_variance(gpu, oracle_lib, ball):
    torch = gpu
    O = oracle_lib
    p, pts, d, osc = ball
    k = 2.0
    x = pts.astype(np.float64)
    s, c = np.sin(k * x), np.cos(k * x)
    grad = np.stack([k * c[:, 0] * s[:, 1] * s[:, 2], k * s[:, 0] * c[:, 1] * s[:, 2], k * s[:, 0] * s[:, 1] * c[:, 2]], 1)
    m0, v0, n0 = _solve(torch, p, pts, d, 256, 21)
    m1, v1, n1 = _solve(torch, p, pts, d, 256, 22, grad)
    ost = O.Stats(len(pts))
    O.solve(osc, p.pde, p.settings, pts, d, 256, ost, seed=23, rng_mode=O.RNG_PER_POINT)
    z, ok = z_scores(m1, v1, n1, ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, "cv(exact grad) vs oracle")
    assert v1.mean() < 0.9 * v0.mean(), (v0.mean(), v1.mean())


def test_learned_control_variate_is_unbiased_and_gains(gpu, oracle_lib, ball):
    torch = gpu
    O = oracle_lib
    p, pts, d, osc = ball
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    # training data: 20k interior points x 32 walks
    cand = PB.sample_box(p.domain_min, p.domain_max, 60000, 9)
    tx = cand[osc.inside_domain(cand)][:20000]
    td, *_ = osc.closest_point(tx)
    dp = E.DevicePoints(pt=torch.from_numpy(tx).cuda(), dist_to_absorbing=torch.from_numpy(td).cuda(),
                        n_walks=torch.full((len(tx),), 32, dtype=torch.int32, device="cuda"), n_walks_max=32)
    X, Y = LC.generate_training_data(sc, p, dp, walks=32)
    net = LC.MLP(LC.MLPConfig(seed=5), max_batch=8192)
    LC.train(net, X, Y, steps=600, batch=8192, lr=1e-3)
    _, g = net.evaluate(torch.from_numpy(pts).cuda(), value=False)
    m0, v0, n0 = _solve(torch, p, pts, d, 256, 31)
    m1, v1, n1 = _solve(torch, p, pts, d, 256, 32, g)
    ost = O.Stats(len(pts))
    O.solve(osc, p.pde, p.settings, pts, d, 256, ost, seed=33, rng_mode=O.RNG_PER_POINT)
    z, ok = z_scores(m1, v1, n1, ost.mean, ost.variance(), ost.n)
    check_four_se(z, ok, "learned cv vs oracle")
    assert v1.mean() < v0.mean(), (v0.mean(), v1.mean())


def test_all_steps_gradient_field_variate_is_unbiased(gpu, oracle_lib, ball):
    """Later-step terms -G(x_k).(x_{k+1}-x_k) with a fixed field G keep the estimator unbiased;
    with G = grad u on a grid the per-walk variance drops severalfold."""
    torch = gpu
    O = oracle_lib
    p, pts, d, osc = ball
    k = 2.0

    def exgrad(q):
        s, c = np.sin(k * q), np.cos(k * q)
        return np.stack([k * c[:, 0] * s[:, 1] * s[:, 2], k * s[:, 0] * c[:, 1] * s[:, 2],
                         k * s[:, 0] * s[:, 1] * c[:, 2]], 1)
    grid = LC.gradient_grid_from_function(exgrad, p.domain_min, p.domain_max, 48, 3)
    sc = E.Scene(p.positions, p.indices, None, p.domain_min, p.domain_max)
    n = len(pts)
    res = {}
    for label, g0, gg in (("plain", None, None), ("grid", exgrad(pts.astype(np.float64)), grid),
                          ("bad", np.zeros((n, 3)), LC.gradient_grid_from_function(
                              lambda q: np.ones_like(q) * 3.0, p.domain_min, p.domain_max, 8, 3))):
        dp = E.DevicePoints(pt=torch.from_numpy(pts).cuda(), dist_to_absorbing=torch.from_numpy(d).cuda(),
                            n_walks=torch.full((n,), 256, dtype=torch.int32, device="cuda"), n_walks_max=256,
                            cv_gradient=None if g0 is None else torch.as_tensor(g0, dtype=torch.float32).cuda(),
                            cv_gradient_grid=gg)
        st = E.DeviceStats.zeros(n)
        E.wos_solve(sc, p.pde, p.settings, dp, st, E.SolveOptions(seed=41 + len(res)))
        res[label] = (st.mean.cpu().numpy(), st.variance().cpu().numpy(), st.n.cpu().numpy())
    ost = O.Stats(n)
    O.solve(osc, p.pde, p.settings, pts, d, 256, ost, seed=47, rng_mode=O.RNG_PER_POINT)
    for label in ("grid", "bad"):
        m, v, c = res[label]
        z, ok = z_scores(m, v, c, ost.mean, ost.variance(), ost.n)
        check_four_se(z, ok, f"{label} field variate vs oracle")
    assert res["grid"][1].mean() < 0.5 * res["plain"][1].mean()
    # a poor field (constant) must not bias the estimate, only inflate the variance
    assert res["bad"][1].mean() > res["plain"][1].mean()
