"""Host-side pieces of the learned warm start: training geometries, network
inputs, the crop mapping, the parameter layout and the equilibrium that
seeds the LB run."""
import math

import numpy as np

from hemelb_amd import lb as L
from hemelb_amd import warmstart as WS


def test_equilibrium_has_the_requested_moments():
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal(50)
    u = 0.02 * rng.standard_normal((50, 3))
    f = WS.equilibrium(rho, u)
    c = np.asarray(L.D3Q19.VECTORS, float)
    np.testing.assert_allclose(f.sum(1), rho, rtol=1e-14)
    np.testing.assert_allclose(f @ c, rho[:, None] * u, rtol=1e-12, atol=1e-16)


def test_run_specs_are_seeded_and_fit_the_crop():
    side = 64
    for seed in (0, 1, 2, 3):
        a, b = WS.run_spec(seed, side), WS.run_spec(seed, side)
        assert (a.kind, a.tau, a.amplitude_mmHg, a.origin) == (b.kind, b.tau, b.amplitude_mmHg,
                                                               b.origin)
        assert a.kind == ("cylinder" if seed % 2 == 0 else "bifurcation")
        assert 0.6 <= a.tau <= 1.0 and 0.5 <= a.amplitude_mmHg <= 2.0
        c = a.geometry.links()[0] - np.asarray(a.origin)
        assert c.min() >= 0 and c.max() < side


def test_input_channels_and_crop_index():
    side = 64
    s = WS.run_spec(2, side)
    x = WS.input_field(s, side, 700, math.pi / 2)
    assert x.shape == (side, side, side, 6)
    assert np.allclose(x[..., 1], s.amplitude_mmHg) and np.allclose(x[..., 2], 0.7)
    assert np.allclose(x[..., 3], s.tau) and np.allclose(x[..., 4], 0, atol=1e-7)
    assert np.allclose(x[..., 5], 1.0)
    coords = s.geometry.links()[0]
    ix = WS._crop_index(s, coords, (0, 0, 0), side)
    sdf = x[..., 0].reshape(-1)
    assert (sdf[ix] > 0).all() and (sdf > 0).sum() == len(ix)


def test_parameter_layout_round_trip():
    chans = (16, 32, 32, 48)
    v = WS.init_params(chans, seed=3)
    p = WS.unpack(v, chans)
    assert p["e1a.w"].shape == (27, 16, 16) and np.all(p["e1a.w"][:, :, 6:] == 0)
    assert p["d3a.w"].shape == (27, 32, 48 + 32)
    np.testing.assert_array_equal(WS.pack_params(p, chans), v)
    n = sum(27 * co * ci + co for _, ci, co in WS.layer_shapes(chans)) + 4 * 16 + 4
    assert v.size == n
