"""The U-Net's HIP kernels against the CPU oracle (oracle/unet_oracle.py):
forward output and every parameter gradient, bfloat16 MFMA with binary32
accumulation against float64 with the same bfloat16 roundings of conv
inputs and weights.  Tolerances are relative L2 norms per tensor."""
import numpy as np
import pytest

from hemelb_amd import warmstart as WS
from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _case(side, batch, chans, seed):
    rng = np.random.default_rng(seed)
    v = WS.init_params(chans, 6, 4, seed)
    p = WS.unpack(v, chans)
    for k in p:      # non-zero biases exercise the bias paths
        if k.endswith(".b"):
            p[k] = (rng.standard_normal(p[k].shape) * 0.05).astype(np.float32)
    v = WS.pack_params(p, chans)
    x = rng.standard_normal((batch, side, side, side, 6)).astype(np.float32)
    t = rng.standard_normal((batch, side, side, side, 4)).astype(np.float32)
    mask = (rng.random((batch, side, side, side)) < 0.6).astype(np.uint8)
    return v, p, x, t, mask


@pytest.mark.parametrize("side,batch,chans", [
    (16, 1, (16, 16, 16, 16)),
    (16, 2, (16, 32, 32, 48)),
    (32, 1, (16, 32, 64, 128)),
    # 3 * 2^18 positions at the finest level: the weight gradient takes 384
    # positions per chunk in 2048 chunks
    (64, 3, (16, 16, 16, 16)),
    # the warm start's own shape: a 128^3 crop through the full 16-128
    # channel network (the oracle takes about a minute here)
    pytest.param(128, 1, (16, 32, 64, 128), marks=pytest.mark.timeout(900)),
])
def test_forward_and_gradients_match_oracle(side, batch, chans):
    print(f"[unet] side={side} batch={batch} chans={chans}", flush=True)
    v, p, x, t, mask = _case(side, batch, chans, side + batch)
    net = WS.UNet(side, batch, 6, 4, chans)
    net.set_params(v)
    np.testing.assert_array_equal(net.params(), v)
    y = net.forward(x)
    xp = np.concatenate([x, np.zeros(x.shape[:-1] + (10,), np.float32)], -1)
    pp = {k: a.astype(np.float64) for k, a in p.items()}
    out, acts = U.forward(pp, xp, bf16=True, keep=True)
    assert _rel(y, out) < 2e-2, _rel(y, out)
    wts = np.array([1.0, 1.0, 1.0, 0.5])
    loss = net.backward(t, mask, wts)
    # the oracle's gradient at the GPU's own output isolates the backward pass
    l_o, dout = U.loss_and_grad_out(y.astype(np.float64), t, mask, wts)
    assert abs(loss - l_o) <= 1e-4 * l_o
    g_o = U.backward(pp, acts, dout, bf16=True)
    g = WS.unpack(net.grads(), chans)
    worst = {}
    for name in g_o:
        worst[name] = _rel(g[name], g_o[name])
    print("[unet] relative gradient errors", {k: round(e, 4) for k, e in worst.items()},
          flush=True)
    assert max(worst.values()) < 5e-2, worst


def test_training_steps_reduce_the_loss():
    side, chans = 16, (16, 16, 32, 32)
    v, p, x, t, mask = _case(side, 1, chans, 7)
    net = WS.UNet(side, 1, 6, 4, chans)
    net.set_params(v)
    losses = []
    for _ in range(30):
        net.forward(x)
        losses.append(net.backward(t, mask))
        net.adam(lr=3e-3)
    print(f"[unet] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.7 * losses[0]
