"""Pins the CPU oracle of the warm-start U-Net: convolution against a direct
loop, pooling/upsampling adjointness, bfloat16 rounding, and every parameter
gradient against central finite differences."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv3d_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 4, 3, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    y = U.conv3d(x, w, b)
    want = np.zeros((1, 4, 3, 5, 3))
    for d in range(4):
        for h in range(3):
            for q in range(5):
                for t, (kd, kh, kw) in enumerate(U.TAPS):
                    dd, hh, qq = d + kd - 1, h + kh - 1, q + kw - 1
                    if 0 <= dd < 4 and 0 <= hh < 3 and 0 <= qq < 5:
                        want[0, d, h, q] += w[t] @ x[0, dd, hh, qq]
                want[0, d, h, q] += b
    np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12)


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((2, 4, 4, 4, 3))
    c = rng.standard_normal((2, 2, 2, 2, 3))
    # <pool(a), c> == <a, pool^T(c)> and <up(c), a> == <c, up^T(a)>
    assert np.isclose((U.pool(a) * c).sum(), (a * U.pool_backward(c)).sum())
    assert np.isclose((U.up(c) * a).sum(), (c * U.up_backward(a)).sum())


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    r = U.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0                      # tie to even
    assert r[2] == np.float32(1.0 + 2 ** -7)                  # rounds up
    assert r[3] == -2.5
    bits = r.view(np.uint32)
    assert (bits & 0xFFFF == 0).all()


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    chans = (2, 3, 2, 3)
    p = {k: v.astype(np.float64) for k, v in U.init_params(3, 4, chans, seed).items()}
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((1, 8, 8, 8, 3))
    t = rng.standard_normal((1, 8, 8, 8, 4))
    mask = rng.random((1, 8, 8, 8)) < 0.7
    wts = np.array([1.0, 1.0, 1.0, 0.5])

    def loss_of(pp):
        return U.loss_and_grad_out(U.forward(pp, x), t, mask, wts)[0]

    out, acts = U.forward(p, x, keep=True)
    _, dout = U.loss_and_grad_out(out, t, mask, wts)
    g = U.backward(p, acts, dout)
    eps = 1e-6
    for name in sorted(p):
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(3, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = loss_of(p)
            flat[idx] = old - eps
            lm = loss_of(p)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, idx, fd, an)
