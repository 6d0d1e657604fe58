"""CPU restatement of the Fourier-feature MLP control variate -- TEST INFRASTRUCTURE ONLY.

FP64 forward pass, loss 0.5 mean (c - y)^2, manual back-propagation to the parameters and to
the input.  With bf16=True the GEMM operands (Fourier features, hidden activations, hidden
weight matrices, back-propagated pre-activation gradients) are rounded to bfloat16 at the same
points as the GPU kernels, so the comparison can be tight; bf16=False is the exact FP64 model.
Parameter layout: per layer l = 0..L, W_l (out x in, row-major) then b_l.
"""
from __future__ import annotations

import numpy as np


def to_bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(theta, n_features, hidden, n_layers):
    theta = np.asarray(theta, np.float64)
    layers = []
    off = 0
    for l in range(n_layers + 1):
        fin = n_features if l == 0 else hidden
        fout = 1 if l == n_layers else hidden
        W = theta[off:off + fin * fout].reshape(fout, fin)
        off += fin * fout
        b = theta[off:off + fout]
        off += fout
        layers.append((W, b))
    assert off == len(theta)
    return layers


def forward(theta, B, x, hidden=128, n_layers=4, bf16=False):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    B = np.asarray(B, np.float64)
    x = np.asarray(x, np.float64)
    a = x @ B.T
    phi = r(np.concatenate([np.sin(a), np.cos(a)], 1))
    layers = unpack(theta, phi.shape[1], hidden, n_layers)
    acts, pres = [phi], []
    h = phi
    for l in range(n_layers):
        W, b = layers[l]
        z = h @ r(W).T + b
        pres.append(z)
        h = r(silu(z))
        acts.append(h)
    W, b = layers[n_layers]
    c = h @ W[0] + b[0]
    return c, dict(a=a, phi=phi, acts=acts, pres=pres, layers=layers, r=r)


def gradients(theta, B, x, y, hidden=128, n_layers=4, bf16=False):
    """(loss, d loss / d theta) for loss = 0.5 mean (c - y)^2."""
    c, s = forward(theta, B, x, hidden, n_layers, bf16)
    r, layers, acts, pres = s["r"], s["layers"], s["acts"], s["pres"]
    n = len(c)
    dc = (c - np.asarray(y, np.float64)) / n
    loss = 0.5 * np.sum((c - y) ** 2) / n
    grads = [None] * (n_layers + 1)
    W, b = layers[n_layers]
    grads[n_layers] = ((acts[n_layers].T @ dc)[None, :], np.array([dc.sum()]))
    dz = r(dc[:, None] * W[0][None, :] * silu_grad(pres[n_layers - 1]))
    for l in range(n_layers - 1, -1, -1):
        Wl, bl = layers[l]
        dW = dz.T @ acts[l]
        db = dz.sum(0)
        grads[l] = (dW, db)
        if l > 0:
            dz = r((dz @ r(Wl)) * silu_grad(pres[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return loss, flat


def input_gradient(theta, B, x, hidden=128, n_layers=4, bf16=False):
    """(c(x), grad_x c(x))."""
    c, s = forward(theta, B, x, hidden, n_layers, bf16)
    r, layers, acts, pres = s["r"], s["layers"], s["acts"], s["pres"]
    W, b = layers[n_layers]
    dz = r(np.ones((len(c), 1)) * W[0][None, :] * silu_grad(pres[n_layers - 1]))
    for l in range(n_layers - 1, -1, -1):
        Wl, _ = layers[l]
        dh = dz @ r(Wl)
        if l > 0:
            dz = r(dh * silu_grad(pres[l - 1]))
    B = np.asarray(B, np.float64)
    F = B.shape[0]
    a = s["a"]
    coef = dh[:, :F] * np.cos(a) - dh[:, F:] * np.sin(a)
    return c, coef @ B
