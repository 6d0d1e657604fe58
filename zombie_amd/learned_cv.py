"""Learned control variate: a Fourier-feature MLP c(x) trained on the GPU from the solver's own
walk estimates; its input gradient gives the zero-mean first-step control variate
-grad c(x0).(x1 - x0) of every walk (x1 uniform on the first sphere, so the estimator stays
unbiased whatever c is).  Plain walk on spheres is the fallback and the yardstick its error is measured against.

Network (defaults): phi(x) = [sin(B x), cos(B x)] with B a 32 x d matrix of N(0, 4^2) entries
(64 features), 4 hidden layers of 128 SiLU units, 1 linear output; bf16 GEMM operands on MFMA
with fp32 accumulation; Adam (lr 1e-3) on fp32 master weights.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from . import engine as E


@dataclass
class MLPConfig:
    in_dim: int = 3
    n_frequencies: int = 32
    hidden: int = 128
    n_hidden_layers: int = 4
    frequency_sigma: float = 4.0
    seed: int = 0x5EED

    def to_c(self) -> A.zg_mlp_config:
        c = A.zg_mlp_config()
        c.in_dim, c.n_frequencies, c.hidden, c.n_hidden_layers = self.in_dim, self.n_frequencies, self.hidden, \
            self.n_hidden_layers
        c.frequency_sigma = self.frequency_sigma
        c.seed = self.seed
        return c


class MLP:
    """Device-resident MLP (libzombie_amd.so kernels)."""

    def __init__(self, config: MLPConfig = MLPConfig(), max_batch: int = 65536):
        self.config = config
        self.max_batch = max_batch
        h = C.c_void_p()
        A.check(A.lib().zg_mlp_create(C.byref(config.to_c()), max_batch, C.byref(h)))
        self._h = h
        n = C.c_int64()
        A.check(A.lib().zg_mlp_parameter_count(self._h, C.byref(n)))
        self.n_params = n.value

    def flops_per_sample(self) -> int:
        """Dense-layer FLOPs of one training sample: forward 2 * MACs (features -> hidden,
        hidden -> hidden, hidden -> 1), backward twice that (input and weight gradients)."""
        c = self.config
        feat = 2 * c.n_frequencies
        macs = feat * c.hidden + (c.n_hidden_layers - 1) * c.hidden * c.hidden + c.hidden
        return 3 * 2 * macs

    def close(self):
        if getattr(self, "_h", None):
            A.lib().zg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_parameters(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        A.check(A.lib().zg_mlp_get_parameters(self._h, out.ctypes.data, self.n_params))
        return out

    def set_parameters(self, theta: np.ndarray):
        theta = np.ascontiguousarray(theta, np.float32)
        A.check(A.lib().zg_mlp_set_parameters(self._h, theta.ctypes.data, self.n_params))

    def frequencies(self) -> np.ndarray:
        out = np.empty((self.config.n_frequencies, self.config.in_dim), np.float32)
        A.check(A.lib().zg_mlp_get_frequencies(self._h, out.ctypes.data, out.size))
        return out

    def gradients(self, x, y):
        """(loss, parameter gradient) of 0.5 mean (c(x) - y)^2 for device tensors x, y."""
        g = np.empty(self.n_params, np.float32)
        loss = np.empty(1, np.float32)
        A.check(A.lib().zg_mlp_gradients(self._h, x.data_ptr(), y.data_ptr(), x.shape[0], g.ctypes.data,
                                         loss.ctypes.data, E._stream_ptr()))
        return float(loss[0]), g

    def train_step(self, x, y, lr: float, want_loss: bool = False):
        loss = np.empty(1, np.float32)
        A.check(A.lib().zg_mlp_train_step(self._h, x.data_ptr(), y.data_ptr(), x.shape[0], lr,
                                          loss.ctypes.data if want_loss else None, E._stream_ptr()))
        return float(loss[0]) if want_loss else None

    def evaluate(self, x, value=True, gradient=True):
        import torch
        n = x.shape[0]
        v = torch.empty(n, dtype=torch.float32, device=x.device) if value else None
        g = torch.empty((n, self.config.in_dim), dtype=torch.float32, device=x.device) if gradient else None
        A.check(A.lib().zg_mlp_evaluate(self._h, x.data_ptr(), n, E._ptr(v), E._ptr(g), E._stream_ptr()))
        return v, g


@dataclass
class TrainReport:
    data_seconds: float
    train_seconds: float
    final_loss: float
    losses: list


def generate_training_data(scene, problem, points: E.DevicePoints, walks: int = 32, seed: int = 0xDA7A,
                           opts: E.SolveOptions | None = None):
    """(x, y) pairs: y = mean of `walks` plain walk-on-spheres estimates at x (GPU)."""
    import torch
    n = points.n
    dp = E.DevicePoints(pt=points.pt, dist_to_absorbing=points.dist_to_absorbing,
                        n_walks=torch.full((n,), walks, dtype=torch.int32, device=points.pt.device), n_walks_max=walks)
    st = E.DeviceStats.zeros(n, device=points.pt.device)
    o = opts or E.SolveOptions()
    o = E.SolveOptions(**{**o.__dict__, "seed": seed, "stream_offset": 0})
    E.wos_solve(scene, problem.pde, problem.settings, dp, st, o)
    return points.pt, st.mean.to(torch.float32)


def train(mlp: MLP, x, y, steps: int = 5000, batch: int = 65536, lr: float = 1e-3, seed: int = 0,
          log_every: int = 0) -> list:
    """Adam on random minibatches (sampled on the GPU) of the training pairs."""
    import torch
    gen = torch.Generator(device=x.device)
    gen.manual_seed(seed)
    n = x.shape[0]
    losses = []
    for s in range(steps):
        idx = torch.randint(0, n, (batch,), device=x.device, generator=gen)
        xb = x[idx].contiguous()
        yb = y[idx].contiguous()
        want = log_every and (s % log_every == 0 or s == steps - 1)
        loss = mlp.train_step(xb, yb, lr, want_loss=bool(want))
        if want:
            losses.append((s, loss))
    return losses


def bake_gradient_grid(mlp: MLP, lo, hi, res: int = 96) -> E.GridField:
    """Samples grad c on a res^d node grid over [lo, hi] (GPU), for the all-steps control
    variate of the walk kernel (trilinear lookup)."""
    import torch
    d = mlp.config.in_dim
    axes = [torch.linspace(float(lo[a]), float(hi[a]), res, device="cuda") for a in range(d)]
    if d == 3:
        zz, yy, xx = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        pts = torch.stack([xx, yy, zz], -1).reshape(-1, 3).contiguous()
        shape = (res, res, res)
    else:
        yy, xx = torch.meshgrid(axes[1], axes[0], indexing="ij")
        pts = torch.stack([xx, yy], -1).reshape(-1, 2).contiguous()
        shape = (1, res, res)
    _, g = mlp.evaluate(pts, value=False)
    data = torch.zeros(shape + (4,), dtype=torch.float32, device="cuda")
    data.view(-1, 4)[:, :d] = g
    return E.GridField(data=data.contiguous(), lo=tuple(float(v) for v in lo[:d]), hi=tuple(float(v) for v in hi[:d]))


def gradient_grid_from_function(fn, lo, hi, res: int, dim: int) -> E.GridField:
    """Same layout from a host callable fn(points [n, dim]) -> gradients [n, dim] (tests)."""
    import torch
    axes = [np.linspace(float(lo[a]), float(hi[a]), res) for a in range(dim)]
    if dim == 3:
        zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        pts = np.stack([xx, yy, zz], -1).reshape(-1, 3)
        shape = (res, res, res)
    else:
        yy, xx = np.meshgrid(axes[1], axes[0], indexing="ij")
        pts = np.stack([xx, yy], -1).reshape(-1, 2)
        shape = (1, res, res)
    g = np.zeros(shape + (4,), np.float32)
    g.reshape(-1, 4)[:, :dim] = fn(pts)
    return E.GridField(data=torch.from_numpy(g).cuda(), lo=tuple(lo[:dim]), hi=tuple(hi[:dim]))


def time_training(steps: int = 1000, batch: int = 65536, n: int = 2_000_000, seed: int = 1) -> dict:
    """Times `steps` Adam steps of the default network at `batch` on n synthetic pairs
    (x uniform in [-1, 1]^3, y = sin 2x sin 2y sin 2z) after a 20-step warm-up; returns the time
    per step and the achieved FLOP rate (6 x multiply-adds per sample and step)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = (torch.rand(n, 3, device="cuda", generator=g) * 2 - 1).contiguous()
    y = torch.sin(2 * x).prod(1).contiguous()
    net = MLP(MLPConfig(in_dim=3, seed=seed), max_batch=batch)
    train(net, x, y, steps=20, batch=batch, seed=seed)
    torch.cuda.synchronize()
    t = time.perf_counter()
    losses = train(net, x, y, steps=steps, batch=batch, seed=seed + 1, log_every=max(1, steps // 2))
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    flops = steps * batch * net.flops_per_sample()
    return {"steps": steps, "batch": batch, "ms_per_step": el / steps * 1e3, "tflops": flops / el / 1e12,
            "losses": losses}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="time the learned control variate's training on the GPU")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=65536)
    a = ap.parse_args()
    print(json.dumps(time_training(a.steps, a.batch)))
