"""Dense-grid PDE data (DenseGrid<float,1,DIM>, utils/dense_grid.h:264-356) in the oracle, pinned
by what interpolation must reproduce: multilinear fields sampled at cell centres are
reproduced exactly between the centres, values clamp to the edge outside them, nearest lookup
returns the cell that contains the point, and the layout is last-axis-fastest."""
import numpy as np

from zombie_amd import _abi as A
from zombie_amd import engine as E


def _centres(lo, hi, shape):
    return [lo[a] + (np.arange(shape[a]) + 0.5) * (hi[a] - lo[a]) / shape[a] for a in range(len(shape))]


def test_bilinear_reproduces_linear_fields_and_clamps(oracle_lib):
    O = oracle_lib
    lo, hi, shape = (-1.0, 0.5), (2.0, 1.5), (5, 7)
    cx, cy = _centres(lo, hi, shape)
    X, Y = np.meshgrid(cx, cy, indexing="ij")
    g = E.GridFunction(2 * X - 3 * Y + 0.5, lo, hi, interpolate=True, device=None)
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = np.array([rng.uniform(cx[0], cx[-1]), rng.uniform(cy[0], cy[-1])], np.float32)
        assert abs(O.eval_function(2, g.fn, p) - (2 * p[0] - 3 * p[1] + 0.5)) < 1e-5
    # beyond the outermost centres the value is the edge value (no extrapolation)
    for p in ([hi[0] + 0.3, 1.0], [lo[0] - 5.0, 0.9], [0.2, hi[1] + 1.0]):
        q = np.array([np.clip(p[0], cx[0], cx[-1]), np.clip(p[1], cy[0], cy[-1])])
        assert abs(O.eval_function(2, g.fn, np.array(p, np.float32)) - (2 * q[0] - 3 * q[1] + 0.5)) < 1e-5


def test_trilinear_reproduces_multilinear_fields(oracle_lib):
    O = oracle_lib
    lo, hi, shape = (0.0, -2.0, 1.0), (1.0, 2.0, 4.0), (4, 6, 3)
    cx, cy, cz = _centres(lo, hi, shape)
    X, Y, Z = np.meshgrid(cx, cy, cz, indexing="ij")
    f = lambda x, y, z: (x + 1) * (2 * y - 1) * (z + 3)
    g = E.GridFunction(f(X, Y, Z), lo, hi, interpolate=True, device=None)
    rng = np.random.default_rng(2)
    for _ in range(200):
        p = np.array([rng.uniform(cx[0], cx[-1]), rng.uniform(cy[0], cy[-1]), rng.uniform(cz[0], cz[-1])], np.float32)
        v = O.eval_function(3, g.fn, p)
        assert abs(v - f(*p)) < 1e-4 * max(1.0, abs(f(*p)))


def test_nearest_cell_and_axis_order(oracle_lib):
    O = oracle_lib
    lo, hi, shape = (0.0, 0.0, 0.0), (2.0, 3.0, 4.0), (2, 3, 4)
    I, J, K = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    g = E.GridFunction(100 * I + 10 * J + K, lo, hi, interpolate=False, device=None)
    rng = np.random.default_rng(3)
    for _ in range(300):
        p = rng.uniform(-0.5, 4.5, 3).astype(np.float32)
        idx = [int(np.clip(np.floor(p[a]), 0, shape[a] - 1)) for a in range(3)]   # unit cells here
        assert O.eval_function(3, g.fn, p) == 100 * idx[0] + 10 * idx[1] + idx[2]
    assert g.fn.kind == A.ZG_FN_DENSE_GRID and list(g.fn.shape) == [2, 3, 4]
