"""FP64 numpy restatement of the learned coarse-step corrector (TEST
INFRASTRUCTURE: what the HIP forward/backward kernels are compared against).

No reference counterpart.  The network: four periodic 3x3 convolutions
3 -> 32 -> 32 -> 32 -> 2 with GELU (erf form) after the first three; the
input channels are (Re psi, Im psi, a V), the output channels a correction
to (Re psi, Im psi); the corrected field is psi_coarse + s net(x).  Loss:
mean over pixels and the two psi channels of
(psi_coarse + s net(x) - psi_fine)^2.

Layouts: activations [H][W][C]; weights W[co][ci][dy][dx] with tap
(dy, dx) reading x[(h + dy - 1) mod H, (w + dx - 1) mod W]; parameters
flattened layer by layer as (W, b).  Pinned by finite differences
(tests/test_corrector_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

CHANNELS = (3, 32, 32, 32, 2)


def param_shapes():
    out = []
    for ci, co in zip(CHANNELS[:-1], CHANNELS[1:]):
        out.append(((co, ci, 3, 3), (co,)))
    return out


def n_params():
    return sum(int(np.prod(w)) + int(np.prod(b)) for w, b in param_shapes())


def unflatten(theta):
    theta = np.asarray(theta, np.float64)
    layers, o = [], 0
    for ws, bs in param_shapes():
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        layers.append((theta[o:o + nw].reshape(ws), theta[o + nw:o + nw + nb]))
        o += nw + nb
    return layers


def flatten(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def init_params(seed=0, last_scale=1e-3):
    """He-style init; the last layer starts near zero so the corrector starts
    as the identity map on psi."""
    rng = np.random.default_rng(seed)
    layers = []
    for i, (ws, bs) in enumerate(param_shapes()):
        fan_in = ws[1] * 9
        w = rng.standard_normal(ws) * math.sqrt(2.0 / fan_in)
        if i == len(CHANNELS) - 2:
            w *= last_scale
        layers.append((w, np.zeros(bs)))
    return flatten(layers)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def shift(x, dy, dx):
    """s[h, w] = x[(h + dy - 1) mod H, (w + dx - 1) mod W]"""
    return np.roll(np.roll(x, -(dy - 1), axis=0), -(dx - 1), axis=1)


def conv(x, w, b):
    y = np.zeros(x.shape[:2] + (w.shape[0],))
    for dy in range(3):
        for dx in range(3):
            y += shift(x, dy, dx) @ w[:, :, dy, dx].T
    return y + b


def conv_dgrad(dz, w):
    """d/dx of sum(dz * conv(x)): taps read x at +offsets, so gradients flow back at -offsets"""
    dx_ = np.zeros(dz.shape[:2] + (w.shape[1],))
    for dy in range(3):
        for dx in range(3):
            g = dz @ w[:, :, dy, dx]
            dx_ += np.roll(np.roll(g, dy - 1, axis=0), dx - 1, axis=1)
    return dx_


def conv_wgrad(x, dz):
    co, ci = dz.shape[2], x.shape[2]
    gw = np.zeros((co, ci, 3, 3))
    for dy in range(3):
        for dx in range(3):
            s = shift(x, dy, dx).reshape(-1, ci)
            gw[:, :, dy, dx] = dz.reshape(-1, co).T @ s
    return gw, dz.reshape(-1, co).sum(0)


def _layers(theta, bf16):
    layers = unflatten(theta)
    if bf16:
        layers = [(round_bf16(w), b) for w, b in layers]
    return layers


def forward(theta, x, keep=False, bf16=False):
    """net(x).  bf16=True rounds the weights and the hidden activations to
    bfloat16 where the device kernels store them (accumulation stays exact)."""
    layers = _layers(theta, bf16)
    a = np.asarray(x, np.float64)
    cache = []
    for i, (w, b) in enumerate(layers):
        z = conv(a, w, b)
        cache.append((a, z))
        if i < len(layers) - 1:
            a = gelu(z)
            if bf16:
                a = round_bf16(a)
        else:
            a = z
    return (a, cache) if keep else a


def loss_and_grad(theta, x, target, bf16=False, scale=1.0):
    """loss = mean((x[..., :2] + scale net(x) - target)^2); gradient w.r.t. theta
    (with bf16=True the rounding is treated as the identity in the backward
    pass, as the device kernels do)"""
    layers = _layers(theta, bf16)
    out, cache = forward(theta, x, keep=True, bf16=bf16)
    r = np.asarray(x, np.float64)[..., :2] + scale * out - target
    n = r.size
    loss = float((r * r).sum() / n)
    d = 2.0 * scale * r / n
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        a_in, z = cache[i]
        dz = d if i == len(layers) - 1 else d * gelu_grad(z)
        gw, gb = conv_wgrad(a_in, dz)
        grads.append((gw, gb))
        d = conv_dgrad(dz, layers[i][0])
    return loss, flatten(grads[::-1])


def psi_to_channels(psi, V=None, a=1.0):
    """network input [H, W, 3] from psi and V (V = 0 when absent)"""
    v = np.zeros(psi.shape) if V is None else a * np.asarray(V, np.float64)
    return np.stack([psi.real, psi.imag, v], axis=-1)


def channels_to_psi(x):
    return x[..., 0] + 1j * x[..., 1]


def round_bf16(a):
    """round-to-nearest-even to bfloat16, returned as float64"""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
