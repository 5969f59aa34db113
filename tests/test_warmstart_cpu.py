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


def test_chapman_enskog_start_carries_the_viscous_stress():
    """A simple shear flow u = (g y, 0, 0): the non-equilibrium part of the
    start has no mass or momentum, and its second moment is the BGK viscous
    stress -tau rho Cs2 (du_x/dy + du_y/dx) = -tau rho g / 3 (fourth-order
    isotropy of the D3Q19 weights), away from the crop's faces."""
    from types import SimpleNamespace
    side, g, tau = 8, 1e-3, 0.8
    xyz = np.array([(x, y, z) for x in range(side) for y in range(side) for z in range(side)],
                   np.int32)
    spec = SimpleNamespace(origin=(0, 0, 0), tau=tau)
    inst = SimpleNamespace(sites=slice(0, len(xyz)), offset=(0, 0, 0))
    ens = SimpleNamespace(domain=SimpleNamespace(coords=xyz), instances=[inst])
    rho = np.full(len(xyz), 1.01)
    u = np.zeros((len(xyz), 3))
    u[:, 0] = g * xyz[:, 1]
    f = WS.chapman_enskog_f(ens, [spec], side, rho, u)
    from hemelb_amd.lb import D3Q19
    c = D3Q19.VECTORS.astype(float)
    w = D3Q19.EQMWEIGHTS
    cu = u @ c.T
    feq = w * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u * u).sum(1)[:, None])
    fneq = f - feq
    inner = ((xyz > 0) & (xyz < side - 1)).all(1)
    np.testing.assert_allclose(fneq.sum(1), 0.0, atol=1e-15)
    np.testing.assert_allclose(fneq @ c, 0.0, atol=1e-15)
    pxy = (fneq * c[:, 0] * c[:, 1]).sum(1)
    np.testing.assert_allclose(pxy[inner], -tau * 1.01 * g / 3.0, rtol=1e-12)
    np.testing.assert_allclose((fneq * c[:, 0] * c[:, 0]).sum(1)[inner], 0.0, atol=1e-15)
