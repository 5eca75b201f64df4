"""CPU oracle of the neural radiance cache MLP (TEST INFRASTRUCTURE ONLY).

The reference has no learned component; this restates the cache the HIP kernels in
crender_amd/csrc/nrc.hip implement, so their forward pass and gradients can be checked:
  * encoding of a query (13 floats: position in [0,1]^3, direction, normal, albedo,
    roughness) into 96 features: 12 frequencies of sin/cos (2^f * pi * p) per position
    axis (72), direction (3), normal (3), albedo (3), roughness (1), a constant 1 (bias
    of the first layer), zero padding to 96;
  * MLP 96 -> 64 x5 (ReLU) -> 32 (outputs 0..2 are RGB, the rest are padding), no biases;
  * relative L2 loss  sum_c (y_c - t_c)^2 / (stop_grad(y_c)^2 + 0.01) / (3 * batch);
  * analytic backward pass in FP64.
Matrices are row-major W[k][n] (y = x @ W).  bf16 rounding of weights, features and
activations is available to mirror the device arithmetic.
Its own correctness is pinned by central finite differences (tests/test_nrc_oracle.py).
"""
from __future__ import annotations

import numpy as np

N_IN = 96
N_HID = 64
N_OUT = 32
LAYERS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]
N_PARAMS = sum(k * n for k, n in LAYERS)
EPS_REL = 0.01


def offsets():
    off, o = [], 0
    for k, n in LAYERS:
        off.append(o)
        o += k * n
    return off


def split(params: np.ndarray):
    off = offsets()
    return [params[o:o + k * n].reshape(k, n) for o, (k, n) in zip(off, LAYERS)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(q: np.ndarray, round_bf16: bool = True) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64).reshape(-1, 13)
    n = q.shape[0]
    f = np.zeros((n, N_IN))
    freqs = (2.0 ** np.arange(12)) * np.pi
    for c in range(3):
        arg = np.float32(q[:, c:c + 1]) * np.float32(freqs[None, :])     # fp32 argument, as on the device
        arg = arg.astype(np.float64)
        f[:, c * 24:(c + 1) * 24:2] = np.sin(arg)
        f[:, c * 24 + 1:(c + 1) * 24:2] = np.cos(arg)
    f[:, 72:82] = q[:, 3:13]
    f[:, 82] = 1.0
    return bf16(f) if round_bf16 else f


def forward(params, x, round_bf16=True):
    """Returns (y (n, 32), activations [x, a1..a5])."""
    W = split(np.asarray(params, dtype=np.float64))
    if round_bf16:
        W = [bf16(w) for w in W]
    acts = [x]
    h = x
    for li, w in enumerate(W):
        z = h @ w
        if li < len(W) - 1:
            h = np.maximum(z, 0.0)
            if round_bf16:
                h = bf16(h)
            acts.append(h)
        else:
            h = z
    return h, acts


def loss_and_grad(params, q, targets, round_bf16=True):
    """Relative L2 loss and its gradient w.r.t. the parameters (FP64)."""
    x = encode(q, round_bf16)
    y, acts = forward(params, x, round_bf16)
    t = np.asarray(targets, dtype=np.float64).reshape(-1, 3)
    B = x.shape[0]
    e = y[:, :3] - t
    den = y[:, :3] ** 2 + EPS_REL
    loss = float(np.sum(e * e / den) / (3 * B))
    g = np.zeros_like(y)
    g[:, :3] = 2.0 * e / den / (3 * B)
    W = split(np.asarray(params, dtype=np.float64))
    if round_bf16:
        W = [bf16(w) for w in W]
    grads = [None] * len(W)
    for li in range(len(W) - 1, -1, -1):
        if round_bf16:
            g = bf16(g)
        grads[li] = acts[li].T @ g
        if li > 0:
            g = (g @ W[li].T) * (acts[li] > 0)
    return loss, np.concatenate([gr.reshape(-1) for gr in grads])


def loss_only(params, q, targets, round_bf16=False):
    x = encode(q, round_bf16)
    y, _ = forward(params, x, round_bf16)
    t = np.asarray(targets, dtype=np.float64).reshape(-1, 3)
    e = y[:, :3] - t
    den = y[:, :3] ** 2 + EPS_REL
    return float(np.sum(e * e / den) / (3 * x.shape[0]))
