"""Host-side pieces of the learned warm start: training geometries, network
inputs, the crop mapping, the parameter layout and the equilibrium that
seeds the LB run."""
import math

import numpy as np

from hemelb_amd import lb as L
from hemelb_amd import warmstart as WS


def test_run_specs_are_seeded_and_fit_the_crop():
    side = 64
    for regime in ("slow", "fast"):
        lo_t, hi_t = WS.REGIMES[regime]["tau"]
        lo_a, hi_a = WS.REGIMES[regime]["amp"]
        for seed in (0, 1, 2, 3):
            a, b = WS.run_spec(seed, side, regime), WS.run_spec(seed, side, regime)
            assert (a.kind, a.tau, a.amplitude, a.origin) == (b.kind, b.tau, b.amplitude,
                                                              b.origin)
            assert a.kind == ("cylinder" if seed % 2 == 0 else "bifurcation")
            assert lo_t <= a.tau <= hi_t and lo_a <= a.amplitude <= hi_a
            c = a.geometry.links()[0] - np.asarray(a.origin)
            assert c.min() >= 0 and c.max() < side


def test_lattice_amplitude_of_a_pressure_amplitude():
    """1 mmHg at tau 0.8 on 5 um voxels of blood: dt = (tau - 1/2) dx^2 rho /
    (3 mu), and the density amplitude is dp dt^2 / (rho dx^2 Cs2)."""
    from hemelb_amd import workloads as WL
    s = WS.RunSpec("cylinder", 0, 0.8, 1.0, "mmHg")
    dx, rho, mu = 5e-6, 1000.0, 0.004
    dt = 0.3 * dx * dx * rho / (3 * mu)
    want = 133.3223684 * dt * dt / (rho * dx * dx) * 3
    np.testing.assert_allclose(s.rho_amp, want, rtol=1e-6)
    np.testing.assert_allclose(WL.lattice_density_amplitude(0.8, 1.0), want, rtol=1e-6)
    assert WS.RunSpec("cylinder", 0, 0.8, 3e-4, "lattice").rho_amp == 3e-4


def test_input_channels_and_crop_index():
    side = 64
    s = WS.run_spec(2, side)
    x = WS.input_field(s, side, 700, math.pi / 2)
    assert x.shape == (side, side, side, 6)
    nu = (s.tau - 0.5) / 3
    assert np.allclose(x[..., 1], 10 * (s.tau - 0.5)) and np.allclose(x[..., 2], 0.7)
    assert np.allclose(x[..., 3], 0.5 * math.sqrt(nu * 700 / math.pi))
    assert np.allclose(x[..., 4], 0, atol=1e-7) and np.allclose(x[..., 5], 1.0)
    assert WS.output_scales(s) == (0.5 * s.amplitude, s.amplitude)
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
