"""The reference-shaped Python layer (zombie_amd.zombie): names, constructors and the host-side
set-up utilities, checked on the CPU against the reference's definitions and the engine demo
problem's fixture."""
from pathlib import Path

import numpy as np
import pytest

from zombie_amd import problems as PB
from zombie_amd import zombie

FIXTURE = Path(__file__).resolve().parent / "golden" / "engine_wost.npz"


def test_walk_settings_constructors():
    ws = zombie.Solvers.WalkSettings(1e-3, 2e-3, 512, True)
    assert ws.epsilon_shell_for_absorbing_boundary == 1e-3 and ws.epsilon_shell_for_reflecting_boundary == 2e-3
    assert ws.max_walk_length == 512 and ws.solve_double_sided
    # defaults of WalkSettings(float, float, int, bool) (core/common.h:21-40)
    assert ws.silhouette_precision == 1e-3 and ws.russian_roulette_threshold == 0.0
    assert ws.steps_before_using_maximal_spheres == 512 and ws.use_gradient_control_variates
    ws16 = zombie.Solvers.WalkSettings(1e-4, 0.0, 0.0, 0.5, 2.0, 100, 3, 7, False, False, True, True, True, False,
                                       True, False)
    c = ws16.to_c()
    assert c.max_walk_length == 100 and c.steps_before_applying_tikhonov == 3
    assert c.steps_before_using_maximal_spheres == 7 and c.use_gradient_control_variates == 0
    assert c.use_cosine_sampling_for_derivatives == 1 and c.ignore_source_contribution == 1
    assert abs(c.russian_roulette_threshold - 0.5) < 1e-7 and abs(c.splitting_threshold - 2.0) < 1e-7
    inf = zombie.Solvers.WalkSettings(1e-3, 1e-3, 1e-3, 0.0, np.inf, 10, 0, 10, False, True, True, False,
                                      False, False, False, False).to_c()
    assert inf.splitting_threshold > 3e38
    with pytest.raises(TypeError):
        zombie.Solvers.WalkSettings(1e-3, 1e-3, 10)
    kw = zombie.Solvers.WalkSettings(epsilon_shell_for_absorbing_boundary=1e-3,
                                     epsilon_shell_for_reflecting_boundary=1e-3, max_walk_length=64,
                                     solve_double_sided=False)
    assert kw.max_walk_length == 64


def test_factories_validate_dim_and_channels():
    with pytest.raises(ValueError):
        zombie.Core.PDE(dim=4)
    with pytest.raises(ValueError):
        zombie.Core.PDE(dim=2, channels=3)
    pde = zombie.Core.PDE(dim=3, channels=4)
    pde.dirichlet = zombie.Core.get_constant_dirichlet_callback(2.5, dim=3, channels=4)
    assert [pde.channel(c).dirichlet.params[0] for c in range(4)] == [2.5] * 4
    pde.source = lambda x: 1.0
    with pytest.raises(TypeError):
        pde.channel(0)
    st = zombie.Solvers.create_sample_statistics_list(5, dim=2)
    assert len(st) == 5 and st[0].get_solution_estimate_count() == 0
    sp = zombie.Solvers.SamplePoint([0.1, 0.2], [0.0, 0.0], zombie.Solvers.SampleType.InDomain,
                                    zombie.Solvers.EstimationQuantity.Solution, 1.0, 0.5, 0.7, dim=2)
    assert sp.dist_to_absorbing_boundary == 0.5 and sp.direction_for_derivative[0] == 1.0


def test_mesh_utilities_reproduce_the_engine_setup(tmp_path):
    """load_boundary_mesh / normalize / flip_orientation / compute_bounding_box /
    partition_boundary_mesh on the engine problem give what problems.engine_wost builds (whose
    pattern of masked pixels matches the reference's output, tests/test_oracle_engine.py)."""
    z = np.load(FIXTURE, allow_pickle=False)
    obj = tmp_path / "engine.obj"
    with open(obj, "w") as f:
        for p in z["positions"]:
            f.write(f"v {float(p[0])!r} {float(p[1])!r} 0\n")
        for a, b in z["segments"]:
            f.write(f"l {a + 1} {b + 1}\n")
    U = zombie.Utils
    pos, idx = zombie.FloatNList(dim=2), zombie.IntNList(dim=2)
    U.load_boundary_mesh(str(obj), pos, idx, dim=2)
    assert np.array_equal(np.asarray(idx), z["segments"])
    U.normalize(pos, dim=2)
    U.flip_orientation(idx, dim=2)
    lo, hi = U.compute_bounding_box(pos, True, 1.0, dim=2)
    ref = PB.engine_wost(FIXTURE, device=None)
    assert np.allclose(np.asarray(pos), ref.positions, atol=2e-7)
    assert np.array_equal(np.asarray(idx), ref.indices)
    assert np.allclose(lo, ref.domain_min, atol=2e-7) and np.allclose(hi, ref.domain_max, atol=2e-7)
    # the reflecting-boundary indicator as a dense grid over the square box (rows flipped as the
    # demo's image lookup does)
    shape = tuple(int(v) for v in z["reflecting_mask_shape"])
    mask = np.unpackbits(z["reflecting_mask_bits"])[: shape[0] * shape[1]].reshape(shape)
    L = float((hi - lo).max())
    grid = U.FloatDenseGrid(np.ascontiguousarray(mask[::-1].T).reshape(-1).astype(np.float32),
                            [shape[1], shape[0]], lo, lo + L, dim=2)
    indicator = U.get_dense_grid_indicator_callback(grid, dim=2)
    ap, ai, rp, ri = (zombie.FloatNList(dim=2), zombie.IntNList(dim=2), zombie.FloatNList(dim=2),
                      zombie.IntNList(dim=2))
    U.partition_boundary_mesh(indicator, pos, idx, ap, ai, rp, ri, dim=2)
    assert len(ri) == int(ref.is_reflecting.sum()) and len(ai) + len(ri) == len(idx)
    # every part's segments are the original segments, in order
    R = np.asarray(rp)[np.asarray(ri)]
    want = ref.positions[ref.indices[ref.is_reflecting.astype(bool)]]
    assert np.array_equal(R, want)
    assert abs(U.compute_signed_volume(pos, idx, dim=2)) > 0.1
