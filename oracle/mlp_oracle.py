"""CPU oracle of the learned-initialisation MLP (test infrastructure only).

The network is D -> 256 -> 256 -> O with ReLU after the two hidden layers and
a linear output; the loss is the mean over rows and output dimensions of the
squared error between the output and the standardised target,
    L = 1/(n O) sum_{r,o} (y_ro - (t_ro - ymean_o) / yscale)^2,
with inputs standardised per feature, x' = (x - xmean) * xinv.  The reference
has no learned initialisation, so nothing in it pins this network: the FP64
form here is pinned by central finite differences (tests/test_mlp_oracle.py),
and the device is checked against both forms below.

* forward_backward(params, ..., bf16=False): FP64 throughout.
* forward_backward(params, ..., bf16=True): the device's rounding points --
  x', the weights W1 and W2, the activations H1 and H2 and the gradients dH2
  and dH1 are rounded to bfloat16 (round to nearest even) where the device
  rounds them; sums are FP64, so only summation order differs from the device.
* adam(...): one Adam step (Kingma & Ba 2015, bias-corrected), f32.

Parameter vector layout (rows are output units): W1 [256 x D], b1 [256],
W2 [256 x 256], b2 [256], W3 [O x 256], b3 [O].
"""
from __future__ import annotations

import numpy as np

H = 256


def layout(D: int, O: int) -> dict:
    off, out = 0, {}
    for name, shape in (("W1", (H, D)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)), ("b3", (O,))):
        size = int(np.prod(shape))
        out[name] = (off, shape)
        off += size
    out["total"] = off
    return out


def unpack(p, D, O) -> dict:
    lay = layout(D, O)
    del lay["total"]
    return {k: np.asarray(p[o:o + int(np.prod(s))]).reshape(s) for k, (o, s) in lay.items()}


def pack(t: dict, D: int, O: int) -> np.ndarray:
    lay = layout(D, O)
    out = np.zeros(lay.pop("total"), np.float64)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = np.asarray(t[k], np.float64).ravel()
    return out


def bf16(a) -> np.ndarray:
    """Round f32 values to bfloat16 (nearest even), returned as f32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = r.astype(np.uint32)
    nan = np.isnan(a)
    out = r.view(np.float32).copy()
    out[nan] = np.nan
    return out


def forward(params, x, xmean, xinv, O, bf16_points=False):
    """Network output (standardised units) for rows x; returns (y, cache)."""
    x = np.asarray(x, np.float32)
    D = x.shape[1]
    P = unpack(np.asarray(params, np.float64), D, O)
    if bf16_points:
        xs = bf16((x - np.asarray(xmean, np.float32)) * np.asarray(xinv, np.float32)).astype(np.float64)
        W1 = bf16(P["W1"].astype(np.float32)).astype(np.float64)
        W2 = bf16(P["W2"].astype(np.float32)).astype(np.float64)
    else:
        xs = (x.astype(np.float64) - np.asarray(xmean, np.float64)) * np.asarray(xinv, np.float64)
        W1, W2 = P["W1"], P["W2"]
    W3 = P["W3"].astype(np.float32).astype(np.float64) if bf16_points else P["W3"]
    h1 = np.maximum(xs @ W1.T + P["b1"], 0.0)
    if bf16_points:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    h2 = np.maximum(h1 @ W2.T + P["b2"], 0.0)
    if bf16_points:
        h2 = bf16(h2.astype(np.float32)).astype(np.float64)
    y = h2 @ W3.T + P["b3"]
    return y, {"xs": xs, "h1": h1, "h2": h2, "W1": W1, "W2": W2, "W3": W3}


def forward_backward(params, x, t, xmean, xinv, ymean, yscale, bf16_points=False):
    """Loss and gradient (flat, FP64) of the batch (x, t)."""
    x = np.asarray(x, np.float32)
    t = np.asarray(t, np.float64)
    n, D = x.shape
    O = t.shape[1]
    y, c = forward(params, x, xmean, xinv, O, bf16_points)
    tn = (t - np.asarray(ymean, np.float64)) / float(yscale)
    diff = y - tn
    loss = float((diff ** 2).sum() / (n * O))
    dy = 2.0 * diff / (n * O)
    dh2 = (dy @ c["W3"]) * (c["h2"] > 0)
    if bf16_points:
        dh2 = bf16(dh2.astype(np.float32)).astype(np.float64)
    dh1 = (dh2 @ c["W2"]) * (c["h1"] > 0)
    if bf16_points:
        dh1 = bf16(dh1.astype(np.float32)).astype(np.float64)
    g = {
        "W3": dy.T @ c["h2"], "b3": dy.sum(0),
        "W2": dh2.T @ c["h1"], "b2": dh2.sum(0),
        "W1": dh1.T @ c["xs"], "b1": dh1.sum(0),
    }
    return loss, pack(g, D, O)


def adam(w, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step in f32 (t = step number, from 1)."""
    f = np.float32
    w, m, v, g = (np.asarray(a, np.float32) for a in (w, m, v, g))
    m = f(b1) * m + (f(1) - f(b1)) * g
    v = f(b2) * v + (f(1) - f(b2)) * (g * g)
    c1, c2 = f(1.0 - b1 ** t), f(1.0 - b2 ** t)
    w = w - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w, m, v
