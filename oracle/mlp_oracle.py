"""CPU oracle of the spin-up surrogate network (test infrastructure only: the
product path never imports it).

Restates, in numpy, the forward and backward pass of the per-cell perceptron
n_in -> H -> H -> H -> n_out (ReLU) with the same BF16 quantisation points as the
HIP kernels (cama-flood_v4_amd/csrc/hip/cmf_mlp.hip):
  X_q = bf16(X), W_q = bf16(W); Z_l = H_{l-1} W_q,l + b_l; H_l = bf16(relu(Z_l))
  Y = H_3 W_q,4 + b_4;  L = mean(w (Y - T)^2);  G_Y = bf16(2 w (Y - T) / n),
  w = w_over where Y > T, else 1
  dW_l = H_{l-1}^T G_l;  db_l = sum_rows G_l;  G_{l-1} = bf16((G_l W_q,l^T) [H_{l-1} > 0])
With quant=False every bf16() is the identity and the arithmetic is float64: that
form is pinned by central finite differences of the loss (tests/test_mlp_oracle.py).
The network has no reference counterpart; the finite-difference check and the
closed-form single-layer cases are what pin it.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even); returns float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def bf16_bits(x):
    """BF16 bit patterns (uint16) of float32 values."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def layer_dims(n_in, hidden, n_out):
    return [(n_in, hidden), (hidden, hidden), (hidden, hidden), (hidden, n_out)]


def unpack(p, n_in, hidden, n_out):
    """Packed parameter vector -> [(W_l, b_l)], W_l of shape (d_in, d_out)."""
    out, o = [], 0
    for di, do in layer_dims(n_in, hidden, n_out):
        w = p[o:o + di * do].reshape(di, do)
        o += di * do
        b = p[o:o + do]
        o += do
        out.append((w, b))
    assert o == p.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers])


def init_params(n_in, hidden, n_out, seed):
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    layers = []
    for di, do in layer_dims(n_in, hidden, n_out):
        layers.append((rng.normal(0.0, np.sqrt(2.0 / di), (di, do)).astype(np.float32),
                       np.zeros(do, np.float32)))
    return pack(layers).astype(np.float32)


def forward(p, x, n_in, hidden, n_out, quant=True):
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    ft = np.float32 if quant else np.float64
    layers = unpack(np.asarray(p, ft), n_in, hidden, n_out)
    h = q(np.asarray(x, ft))
    hs = [h]
    for l, (w, b) in enumerate(layers):
        z = (h.astype(np.float64) @ q(w).astype(np.float64)).astype(ft) + b
        if l < 3:
            h = q(np.maximum(z, 0))
            hs.append(h)
        else:
            return z, hs


def loss_grad(p, x, t, n_in, hidden, n_out, quant=True, w_over=1.0):
    """(loss, packed gradient) of the weighted mean squared error: weight w_over
    where the prediction exceeds the target, 1 elsewhere."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    ft = np.float32 if quant else np.float64
    layers = unpack(np.asarray(p, ft), n_in, hidden, n_out)
    y, hs = forward(p, x, n_in, hidden, n_out, quant)
    d = y.astype(np.float64) - np.asarray(t, np.float64)
    n = d.size
    w = np.where(d > 0, w_over, 1.0)
    loss = float(np.mean(w * d * d))
    g = q((2.0 * w * d / n).astype(ft))
    grads = [None] * 4
    for l in range(3, -1, -1):
        hin = hs[l].astype(np.float64)
        dw = (hin.T @ g.astype(np.float64)).astype(ft)
        db = g.astype(np.float64).sum(axis=0).astype(ft)
        grads[l] = (dw, db)
        if l == 0:
            break
        dd = (g.astype(np.float64) @ q(layers[l][0]).astype(np.float64).T).astype(ft)
        g = q(np.where(hs[l] > 0, dd, 0).astype(ft))
    return loss, pack(grads)
