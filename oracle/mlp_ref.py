"""CPU reference of the surrogate MLP (test infrastructure only).

Float64 forward/backward of the network in hybridmantis_amd/csrc/fd2_mlp.hip
(5 -> 256 -> 256 -> 256 -> 65, GELU tanh form, sigmoid + softmax heads,
BCE + cross-entropy loss), with optional rounding to bfloat16 at the points
where the GPU kernels round (inputs, weight copies, hidden activations and
backpropagated gradients), so the GPU's gradients can be compared with
FP32-accumulation tolerances.  The parameter vector layout is the kernels'
(W0 [32x256], b0, W1 [256x256], b1, W2, b2, W3 [256x80], b3 [80]).
"""
from __future__ import annotations

import numpy as np

FEAT, IN_PAD, HID, OUT, OUT_PAD = 5, 32, 256, 65, 80
W_OFF = [0, 8448, 74240, 140032]
B_OFF = [8192, 73984, 139776, 160512]
N_PARAMS = 160592
SHAPES = [(IN_PAD, HID), (HID, HID), (HID, HID), (HID, OUT_PAD)]
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(x) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    W = [p[o:o + r * c].reshape(r, c) for o, (r, c) in zip(W_OFF, SHAPES)]
    b = [p[o:o + c] for o, (_, c) in zip(B_OFF, SHAPES)]
    return W, b


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = C_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * x * x)


def forward(params, x, rnd=True):
    r = bf16 if rnd else (lambda v: np.asarray(v, np.float64))
    W, b = unpack(params)
    Wq = [r(w) for w in W]
    x0 = np.zeros((len(x), IN_PAD))
    x0[:, :FEAT] = x
    a = r(x0)
    acts, zs = [a], []
    for l in range(4):
        z = a @ Wq[l] + np.asarray(b[l], np.float32).astype(np.float64)
        zs.append(z)
        if l < 3:
            a = r(gelu(z))
            acts.append(a)
    return zs, acts, Wq


def probs(params, x, rnd=True):
    zs, _, _ = forward(params, x, rnd)
    z = zs[3]
    p0 = 1.0 / (1.0 + np.exp(-z[:, 0]))
    zr = z[:, 1:65]
    e = np.exp(zr - zr.max(1, keepdims=True))
    return np.concatenate([p0[:, None], e / e.sum(1, keepdims=True)], 1)


def loss_and_grads(params, x, y, rnd=True):
    """Mean loss and its gradient (N_PARAMS) for targets y [B x 65]."""
    r = bf16 if rnd else (lambda v: np.asarray(v, np.float64))
    zs, acts, Wq = forward(params, x, rnd)
    B = len(x)
    z = zs[3]
    y = np.asarray(y, np.float64)
    z0, y0 = z[:, 0], y[:, 0]
    p0 = 1.0 / (1.0 + np.exp(-z0))
    bce = np.maximum(z0, 0) - z0 * y0 + np.log1p(np.exp(-np.abs(z0)))
    zr = z[:, 1:65]
    m = zr.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(zr - m).sum(1))
    pr = np.exp(zr - lse[:, None])
    q = y[:, 1:65]
    wrad = (q.sum(1) > 0).astype(np.float64)
    ce = -(q * (zr - lse[:, None])).sum(1)
    loss = float(np.mean(bce + wrad * ce))
    dz = np.zeros((B, OUT_PAD))
    dz[:, 0] = (p0 - y0) / B
    dz[:, 1:65] = wrad[:, None] * (pr - q) / B
    dz = r(dz)
    g = np.zeros(N_PARAMS)
    for l in range(3, -1, -1):
        gW = acts[l].T @ dz
        g[W_OFF[l]:W_OFF[l] + gW.size] = gW.ravel()
        g[B_OFF[l]:B_OFF[l] + dz.shape[1]] = dz.sum(0)
        if l == 0:
            break
        da = dz @ Wq[l].T
        dz = r(da * gelu_grad(zs[l - 1]))
    # the output layer's bias only has the 65 real columns
    g[B_OFF[3] + OUT:B_OFF[3] + OUT_PAD] = 0.0
    return loss, g


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights (real rows/columns only), zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    ins = [FEAT, HID, HID, HID]
    outs = [HID, HID, HID, OUT]
    for l, (rr, cc) in enumerate(SHAPES):
        lim = np.sqrt(6.0 / (ins[l] + outs[l]))
        w = np.zeros((rr, cc), np.float32)
        w[:ins[l], :outs[l]] = rng.uniform(-lim, lim, (ins[l], outs[l]))
        p[W_OFF[l]:W_OFF[l] + rr * cc] = w.ravel()
    return p
