"""CPU oracle of the warm-start 3D U-Net (TEST INFRASTRUCTURE: only tests/
use it, as the checker of the HIP training and inference kernels).

Tensors are channels-last, [N, D, H, W, C].  Convolutions are 3x3x3 with
zero padding and unit stride; weights are [27, Cout, Cin] with the tap index
t = 9*kd + 3*kh + kw (kd, kh, kw in 0..2 covering offsets -1..1).  The
network (no counterpart in HemeLB; the learned warm start is new):

    enc1 = relu(conv(relu(conv(x))))               C1 channels at S^3
    enc2 = relu(conv(relu(conv(pool(enc1)))))      C2 at (S/2)^3
    enc3 = ... C3 at (S/4)^3;  bott = ... C4 at (S/8)^3
    dec3 = relu(conv(relu(conv([up(bott), enc3]))))   C3
    dec2 = ... [up(dec3), enc2] -> C2;  dec1 = ... [up(dec2), enc1] -> C1
    out  = head(dec1)                              1x1x1 conv to 4 channels

pool = 2x2x2 mean, up = 2x nearest, [a, b] = channel concatenation.  The
loss is the masked mean square error sum_c wc (out_c - t_c)^2 over fluid
voxels, divided by (voxels * channels).

`bf16=True` rounds conv inputs and weights to bfloat16 (round to nearest
even), as the MFMA kernels read them; accumulation stays float64 here.
Gradients are analytic; tests pin them with finite differences.
"""
from __future__ import annotations

import numpy as np

TAPS = [(kd, kh, kw) for kd in range(3) for kh in range(3) for kw in range(3)]


def to_bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def conv3d(x, w, b=None, bf16=False):
    """x [N,D,H,W,Ci], w [27,Co,Ci] -> [N,D,H,W,Co] (float64)."""
    if bf16:
        x, w = to_bf16(x), to_bf16(w)
    x = x.astype(np.float64)
    w = w.astype(np.float64)
    N, D, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((N, D, H, W, w.shape[1]))
    for t, (kd, kh, kw) in enumerate(TAPS):
        y += xp[:, kd:kd + D, kh:kh + H, kw:kw + W, :] @ w[t].T
    if b is not None:
        y += b
    return y


def conv3d_backward(x, w, dy, bf16=False):
    """Gradients of conv3d: (dx, dw, db)."""
    if bf16:
        x, w = to_bf16(x), to_bf16(w)
        dy = to_bf16(dy)
    x = x.astype(np.float64)
    w = w.astype(np.float64)
    dy = dy.astype(np.float64)
    N, D, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    dy2 = dy.reshape(-1, dy.shape[-1])
    for t, (kd, kh, kw) in enumerate(TAPS):
        win = xp[:, kd:kd + D, kh:kh + H, kw:kw + W, :]
        dw[t] = dy2.T @ np.ascontiguousarray(win).reshape(-1, Ci)
        dxp[:, kd:kd + D, kh:kh + H, kw:kw + W, :] += dy @ w[t]
    dx = dxp[:, 1:-1, 1:-1, 1:-1, :]
    db = dy.sum(axis=(0, 1, 2, 3))
    return dx, dw, db


def pool(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_backward(dy):
    N, D, H, W, C = dy.shape
    return dy.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


LEVELS = 4


def layer_shapes(cin: int, chans=(16, 32, 64, 128)):
    """(name, Cin, Cout) of the 3x3x3 convolutions, in forward order."""
    c1, c2, c3, c4 = chans
    enc = [("e1a", cin, c1), ("e1b", c1, c1), ("e2a", c1, c2), ("e2b", c2, c2),
           ("e3a", c2, c3), ("e3b", c3, c3), ("e4a", c3, c4), ("e4b", c4, c4)]
    dec = [("d3a", c4 + c3, c3), ("d3b", c3, c3), ("d2a", c3 + c2, c2), ("d2b", c2, c2),
           ("d1a", c2 + c1, c1), ("d1b", c1, c1)]
    return enc + dec


def init_params(cin: int, cout: int = 4, chans=(16, 32, 64, 128), seed: int = 0):
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in layer_shapes(cin, chans):
        p[name + ".w"] = (rng.standard_normal((27, co, ci)) *
                          np.sqrt(2.0 / (27 * ci))).astype(np.float32)
        p[name + ".b"] = np.zeros(co, np.float32)
    p["head.w"] = (rng.standard_normal((cout, chans[0])) *
                   np.sqrt(1.0 / chans[0])).astype(np.float32)
    p["head.b"] = np.zeros(cout, np.float32)
    return p


def forward(p, x, bf16=False, keep=False):
    """Returns the output [N,S,S,S,4] (and the saved activations)."""
    acts = {}

    def cr(name, a):
        acts[name + ".in"] = a
        y = np.maximum(conv3d(a, p[name + ".w"], p[name + ".b"], bf16), 0.0)
        acts[name + ".out"] = y
        return y

    e1 = cr("e1b", cr("e1a", x))
    e2 = cr("e2b", cr("e2a", pool(e1)))
    e3 = cr("e3b", cr("e3a", pool(e2)))
    e4 = cr("e4b", cr("e4a", pool(e3)))
    d3 = cr("d3b", cr("d3a", np.concatenate([up(e4), e3], -1)))
    d2 = cr("d2b", cr("d2a", np.concatenate([up(d3), e2], -1)))
    d1 = cr("d1b", cr("d1a", np.concatenate([up(d2), e1], -1)))
    acts["head.in"] = d1
    out = d1 @ p["head.w"].astype(np.float64).T + p["head.b"]
    return (out, acts) if keep else out


def loss_and_grad_out(out, target, mask, weights):
    """Masked weighted MSE and its gradient with respect to `out`."""
    m = mask[..., None].astype(np.float64)
    cnt = max(float(mask.sum()), 1.0) * out.shape[-1]
    d = (out - target) * m
    loss = float((weights * d * d).sum() / cnt)
    return loss, 2.0 * weights * d / cnt


def backward(p, acts, dout, bf16=False):
    """Parameter gradients of the loss whose gradient wrt the output is dout."""
    g = {}
    d1 = acts["head.in"]
    g["head.w"] = dout.reshape(-1, dout.shape[-1]).T @ d1.reshape(-1, d1.shape[-1])
    g["head.b"] = dout.sum(axis=(0, 1, 2, 3))
    dd1 = dout @ p["head.w"].astype(np.float64)

    def cr_back(name, dy):
        dy = dy * (acts[name + ".out"] > 0)
        dx, dw, db = conv3d_backward(acts[name + ".in"], p[name + ".w"], dy, bf16)
        g[name + ".w"] = dw
        g[name + ".b"] = db
        return dx

    def split(d, c_up):
        return d[..., :c_up], d[..., c_up:]

    dcat1 = cr_back("d1a", cr_back("d1b", dd1))
    c_up1 = acts["d2b.out"].shape[-1]
    dup2, de1 = split(dcat1, c_up1)
    dd2 = up_backward(dup2)
    dcat2 = cr_back("d2a", cr_back("d2b", dd2))
    c_up2 = acts["d3b.out"].shape[-1]
    dup3, de2 = split(dcat2, c_up2)
    dd3 = up_backward(dup3)
    dcat3 = cr_back("d3a", cr_back("d3b", dd3))
    c_up3 = acts["e4b.out"].shape[-1]
    dup4, de3 = split(dcat3, c_up3)
    de4 = up_backward(dup4)
    dp3 = cr_back("e4a", cr_back("e4b", de4))
    de3 = de3 + pool_backward(dp3)
    dp2 = cr_back("e3a", cr_back("e3b", de3))
    de2 = de2 + pool_backward(dp2)
    dp1 = cr_back("e2a", cr_back("e2b", de2))
    de1 = de1 + pool_backward(dp1)
    cr_back("e1a", cr_back("e1b", de1))
    return g
