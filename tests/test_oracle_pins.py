"""Pins of the CPU oracle (oracle/hlb_oracle.c).

Each building block is checked against something that does not restate it:
the reference's own code compiled from its sources (oracle/_ref, where it was
built), the reference tests' golden data (tests/golden), closed forms, or
independent textbook formulas written here in numpy.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from oracle import bindings as O
from tests import cases as K

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = O._p
needs_ref = pytest.mark.skipif(O.ref() is None, reason="reference glue (oracle/_ref) not built")

LAT = {15: K.L.D3Q15, 19: K.L.D3Q19}


def _lattice(Q):
    cx, cy, cz, inv = (np.zeros(Q, np.int32) for _ in range(4))
    w = np.zeros(Q)
    O.lib().orc_lattice(Q, P(cx), P(cy), P(cz), P(w), P(inv))
    return np.stack([cx, cy, cz], 1), w, inv


# ---------------------------------------------------------------------------
# Against the reference's own code (oracle/_ref)
# ---------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("Q", [15, 19])
def test_lattice_tables_equal_reference(Q):
    c, w, inv = _lattice(Q)
    rc, rw, rinv = (np.zeros(Q, np.int32) for _ in range(3)), np.zeros(Q), None
    cx, cy, cz, ri = (np.zeros(Q, np.int32) for _ in range(4))
    O.ref().ref_lattice(Q, P(cx), P(cy), P(cz), P(rw), P(ri))
    np.testing.assert_array_equal(c, np.stack([cx, cy, cz], 1))
    np.testing.assert_array_equal(w, rw)
    np.testing.assert_array_equal(inv, ri)


@needs_ref
@pytest.mark.parametrize("Q", [15, 19])
def test_moments_and_equilibrium_bitwise_equal_reference(Q):
    """Lattice::CalculateDensityMomentumFEq with HEMELB_USE_SSE3 (Lattice.h:119-321)."""
    rng = np.random.default_rng(Q)
    for k in range(3000):
        f = rng.uniform(1e-3, 1.0, Q) * rng.choice([1e-3, 1.0, 60.0])
        if k % 7 == 0:
            f = f * rng.choice([-1.0, 1.0], Q)
        rho, m = np.zeros(1), np.zeros(3)
        O.lib().orc_density_momentum(Q, P(f), P(rho), P(m))
        feq = np.zeros(Q)
        O.lib().orc_feq(Q, rho[0], P(m), P(feq))
        rr, rm, ru, rf = np.zeros(1), np.zeros(3), np.zeros(3), np.zeros(Q)
        O.ref().ref_density_momentum_feq(Q, P(f), P(rr), P(rm), P(ru), P(rf))
        assert rho[0] == rr[0]
        np.testing.assert_array_equal(m, rm)
        np.testing.assert_array_equal(feq, rf)


@needs_ref
@pytest.mark.parametrize("Q", [15, 19])
def test_mrt_basis_equals_reference(Q):
    nm = 11 if Q == 15 else 15
    M, d, s = np.zeros(nm * Q), np.zeros(nm), np.zeros(nm)
    rM, rd, rs = np.zeros(nm * Q), np.zeros(nm), np.zeros(nm)
    assert O.lib().orc_mrt_basis(Q, P(M), P(d), P(s), 0.7) == nm
    assert O.ref().ref_mrt_basis(Q, P(rM), P(rd), P(rs), 0.7) == nm
    np.testing.assert_array_equal(M, rM)
    np.testing.assert_array_equal(d, rd)
    np.testing.assert_array_equal(s, rs)


@needs_ref
def test_bessel_and_ipow_bitwise_equal_reference():
    ip, rip = np.zeros(2), np.zeros(2)
    O.lib().orc_ipow_three_halves(P(ip))
    O.ref().ref_ipow_three_halves(P(rip))
    np.testing.assert_array_equal(ip, rip)
    for x in np.linspace(0.0, 25.0, 501):
        for z in (ip * x, np.array([x, 0.0]), np.array([0.3 * x, -0.7 * x])):
            a, b = np.zeros(2), np.zeros(2)
            O.lib().orc_bessel_j0(z[0], z[1], 1e-12, P(a))
            O.ref().ref_bessel_j0(z[0], z[1], 1e-12, P(b))
            np.testing.assert_array_equal(a, b)


@needs_ref
def test_unit_conversions_bitwise_equal_reference():
    rng = np.random.default_rng(3)
    for _ in range(200):
        dt, dx, rho = rng.uniform(1e-6, 1e-3), rng.uniform(1e-5, 1e-2), rng.uniform(500, 2000)
        p, ref = rng.uniform(0, 120), rng.choice([0.0, 80.0])
        assert O.lib().orc_pressure_to_lattice(p, ref, dt, dx, rho) == \
            O.ref().ref_pressure_to_lattice(p, ref, dt, dx, rho)
        assert O.lib().orc_pressure_diff_to_lattice(p, dt, dx, rho) == \
            O.ref().ref_pressure_diff_to_lattice(p, dt, dx, rho)
        assert O.lib().orc_pressure_gradient_to_lattice(p, dt, dx, rho) == \
            O.ref().ref_pressure_gradient_to_lattice(p, dt, dx, rho)


# ---------------------------------------------------------------------------
# Golden data held by the reference's tests
# ---------------------------------------------------------------------------
def test_bessel_matches_mathematica_table():
    """Code/tests/util/BesselTests.cc: 201 Mathematica values of J0(i^{3/2} x)
    with Catch2's default relative tolerance."""
    g = json.load(open(os.path.join(GOLDEN, "bessel_j0_mathematica.json")))
    ip = np.zeros(2)
    O.lib().orc_ipow_three_halves(P(ip))
    eps = g["relative_tolerance"]
    for i, (re_, im_) in enumerate(zip(g["re"], g["im"])):
        z = ip * (i / 10.0)
        out = np.zeros(2)
        O.lib().orc_bessel_j0(z[0], z[1], 1e-12, P(out))
        for got, want in ((out[0], re_), (out[1], im_)):
            assert abs(got - want) <= eps * max(abs(got), abs(want)), (i, got, want)


@pytest.mark.parametrize("Q", [15, 19])
def test_mrt_basis_norms_match_kernel_tests(Q):
    """BASIS_TIMES_BASIS_TRANSPOSED expected by Code/tests/lb/KernelTests.cc:358-370."""
    g = json.load(open(os.path.join(GOLDEN, "mrt_basis_diag.json")))
    nm = 11 if Q == 15 else 15
    M, d, s = np.zeros(nm * Q), np.zeros(nm), np.zeros(nm)
    O.lib().orc_mrt_basis(Q, P(M), P(d), P(s), 0.8)
    np.testing.assert_array_equal(d, np.array(g[str(Q)]))


# ---------------------------------------------------------------------------
# Closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("Q", [15, 19])
def test_lattice_isotropy(Q):
    c, w, inv = _lattice(Q)
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(w @ c, 0.0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", w, c, c), np.eye(3) / 3, atol=1e-16)
    np.testing.assert_array_equal(c[inv], -c)
    np.testing.assert_array_equal(c, LAT[Q].VECTORS)
    np.testing.assert_array_equal(w, LAT[Q].EQMWEIGHTS)


def _textbook_feq(Q, rho, m):
    """Compressible second-order equilibrium written from the textbook form
    (as LbTestsHelper::CalculateLBGKEqmF does, LbTestsHelper.h:89-114)."""
    c, w, _ = _lattice(Q)
    cu = c @ m
    return w * (rho - 1.5 * (m @ m) / rho + 3 * cu + 4.5 * cu * cu / rho)


@pytest.mark.parametrize("Q", [15, 19])
def test_equilibrium_moments_and_textbook_form(Q):
    rng = np.random.default_rng(1)
    c, w, _ = _lattice(Q)
    for _ in range(200):
        rho = rng.uniform(0.5, 60)
        m = rng.normal(0, 0.05, 3) * rho
        feq = np.zeros(Q)
        O.lib().orc_feq(Q, rho, P(m), P(feq))
        np.testing.assert_allclose(feq.sum(), rho, rtol=1e-14)
        np.testing.assert_allclose(c.T @ feq, m, rtol=1e-12, atol=1e-13 * rho)
        np.testing.assert_allclose(feq, _textbook_feq(Q, rho, m), rtol=1e-13)


@pytest.mark.parametrize("Q", [15, 19])
@pytest.mark.parametrize("kernel", ["LBGK", "MRT"])
def test_collision_conserves_mass_and_momentum(Q, kernel):
    rng = np.random.default_rng(2)
    c, w, _ = _lattice(Q)
    for _ in range(100):
        f = w * rng.uniform(0.5, 2.0) * (1 + 0.05 * rng.normal(size=Q))
        fpc = np.zeros(Q)
        if kernel == "LBGK":
            O.lib().orc_collide_lbgk(Q, 0.7, P(f), P(fpc), None, None, None)
        else:
            O.lib().orc_collide_mrt(Q, 0.7, None, P(f), P(fpc))
        np.testing.assert_allclose(fpc.sum(), f.sum(), rtol=1e-14)
        np.testing.assert_allclose(c.T @ fpc, c.T @ f, atol=1e-14)


@pytest.mark.parametrize("Q", [15, 19])
def test_lbgk_collision_textbook(Q):
    """LBGK: f* = f + (f - feq) * (-1/tau) (LbTestsHelper::CalculateLBGKCollision)."""
    rng = np.random.default_rng(4)
    c, w, _ = _lattice(Q)
    for _ in range(100):
        f = w * (1 + 0.05 * rng.normal(size=Q))
        tau = rng.uniform(0.55, 2.0)
        fpc = np.zeros(Q)
        O.lib().orc_collide_lbgk(Q, tau, P(f), P(fpc), None, None, None)
        rho, m = f.sum(), c.T @ f
        np.testing.assert_allclose(fpc, f - (f - _textbook_feq(Q, rho, m)) / tau, rtol=1e-12,
                                   atol=1e-15)


@pytest.mark.parametrize("Q", [15, 19])
def test_mrt_with_single_rate_equals_lbgk(Q):
    """Code/tests/lb/KernelTests.cc:296-357: every moment relaxed with 1/tau
    reproduces LBGK (margin 1e-10) on anisotropic data."""
    nm = 11 if Q == 15 else 15
    tau = 0.8
    f = np.array([(d + 1) / 10.0 for d in range(Q)])   # InitialiseAnisotropicTestData(0)
    s = np.full(nm, 1.0 / tau)
    a, b = np.zeros(Q), np.zeros(Q)
    O.lib().orc_collide_mrt(Q, tau, P(s), P(f), P(a))
    O.lib().orc_collide_lbgk(Q, tau, P(f), P(b), None, None, None)
    np.testing.assert_allclose(a, b, atol=1e-10)


@pytest.mark.parametrize("Q", [15, 19])
def test_mrt_relaxes_each_moment_at_its_rate(Q):
    """m(f*) - m(feq) = (1 - S_k) (m(f) - m(feq)) for every kinetic moment."""
    nm = 11 if Q == 15 else 15
    rng = np.random.default_rng(5)
    c, w, _ = _lattice(Q)
    M, d, s = np.zeros(nm * Q), np.zeros(nm), np.zeros(nm)
    O.lib().orc_mrt_basis(Q, P(M), P(d), P(s), 0.7)
    M = M.reshape(nm, Q)
    f = w * (1 + 0.05 * rng.normal(size=Q))
    fpc = np.zeros(Q)
    O.lib().orc_collide_mrt(Q, 0.7, None, P(f), P(fpc))
    feq = _textbook_feq(Q, f.sum(), c.T @ f)
    np.testing.assert_allclose(M @ (fpc - feq), (1 - s) * (M @ (f - feq)), atol=1e-14)


def test_cosine_iolet_density_formula():
    """InOutLetCosine::GetDensity (InOutLetCosine.cc:26-43) incl. warm-up."""
    io = O.make_iolet(O.ORC_COSINE, (0, 0, 0), (0, 0, 1), density_mean=1.02,
                      density_amp=0.01, phase=0.3, period=200.0, warmup=50,
                      min_sim_density=0.97)
    for t in range(0, 400, 7):
        target = 1.02 + 0.01 * math.cos(2.0 * math.pi / 200.0 * t + 0.3)
        want = target if t >= 50 else (t / 50) * target + (1.0 - t / 50) * 0.97
        assert O.lib().orc_cosine_density(C.byref(io), t) == want


def test_womersley_limits_as_reference_tests():
    """Code/tests/lb/iolets/InOutLetTests.cc:203-235: zero velocity at the
    wall; at a small Womersley number the centreline velocity is the
    Poiseuille value G R^2/(4 eta) after a quarter period and its negative
    after three quarters."""
    R = 10.0
    G = 133.3223874 * 1e-6
    io = O.make_iolet(O.ORC_WOMERSLEY, (0, 0, 0), (0, 0, 1), radius=R,
                      pressure_gradient_amplitude=G, period=5.0, womersley_number=2.0)
    u = np.zeros(3)
    O.lib().orc_iolet_velocity(C.byref(io), P(np.array([R, 0.0, 0.0])), 0, P(u))
    np.testing.assert_allclose(u, 0.0, atol=1e-9)
    alpha = 1e-4
    io.womersley_number = alpha
    io.period = 2 * math.pi * R ** 2 / alpha ** 2
    pois = G * R ** 2 / 4.0
    for frac, sign in ((0.25, 1.0), (0.75, -1.0)):
        O.lib().orc_iolet_velocity(C.byref(io), P(np.zeros(3)), int(frac * io.period), P(u))
        assert u[0] == 0 and u[1] == 0
        assert abs(u[2] - sign * pois) <= 100 * 1.2e-7 * abs(pois)


def test_womersley_profile_against_scipy_bessel():
    """Velocity profile from the textbook Womersley solution using scipy's
    complex Bessel function (independent of the series in util/Bessel.cc).
    The reference's series stops when |term|^2 <= 1e-12 (Bessel.h:33), so J0
    carries an absolute error up to ~1e-6: the tolerance follows from that."""
    from scipy.special import jv
    R, alpha, T, G = 8.0, 4.0, 600.0, 2e-5
    io = O.make_iolet(O.ORC_WOMERSLEY, (1.0, 2.0, 3.0), (0, 0, 2.0), radius=R,
                      pressure_gradient_amplitude=G, period=T, womersley_number=alpha)
    io.normal[:] = [0.0, 0.0, 1.0]
    w = 2 * math.pi / T
    i32 = complex(math.cos(3 * math.pi / 4), math.sin(3 * math.pi / 4))
    u = np.zeros(3)
    for r in (0.0, 2.5, 5.0, 7.5):
        for t in (0, 37, 150, 444):
            O.lib().orc_iolet_velocity(C.byref(io), P(np.array([1.0 + r, 2.0, 3.5])), t, P(u))
            a = G / w * (1 - jv(0, i32 * alpha * r / R) / jv(0, i32 * alpha))
            want = -(a * complex(math.cos(w * t), math.sin(w * t))).real
            np.testing.assert_allclose(u[2], want, rtol=1e-6, atol=1e-6 * G / w)


def test_parabolic_iolet_with_warmup():
    io = O.make_iolet(O.ORC_PARABOLIC, (0, 0, 0), (0, 0, 1), radius=4.0, max_speed=0.05,
                      warmup=10)
    u = np.zeros(3)
    for t, r in ((3, 0.0), (10, 2.0), (20, 3.0)):
        O.lib().orc_iolet_velocity(C.byref(io), P(np.array([r, 0, 7.0])), t, P(u))
        want = 0.05 * min(1.0, t / 10) * (1 - r * r / 16.0)
        np.testing.assert_allclose(u, [0, 0, want], rtol=1e-15, atol=1e-18)


def test_units_for_the_small_cylinder_config():
    """80 mmHg at dx = 5e-5 m, tau = 0.8 (dt = 6.25e-5 s): rho0 close to 51."""
    dt = 0.3 * (1 / 3) * 5e-5 ** 2 * 1000.0 / 0.004
    assert abs(O.lib().orc_tau(dt, 5e-5, 1000.0, 0.004) - 0.8) < 1e-12
    rho0 = O.lib().orc_pressure_to_lattice(80.0, 0.0, dt, 5e-5, 1000.0) * 3.0
    lattice_pressure = 1000.0 * 5e-5 ** 3 / (5e-5 * dt * dt)
    np.testing.assert_allclose(rho0, 1.0 + 3 * 80 * 133.3223874 / lattice_pressure, rtol=1e-14)
    assert 50.5 < rho0 < 51.5


# ---------------------------------------------------------------------------
# Streamer known answers on the FourCube fixture (Code/tests/lb/StreamerTests.cc)
# ---------------------------------------------------------------------------
def _anisotropic(n, Q):
    """LbTestsHelper::InitialiseAnisotropicTestData (LbTestsHelper.h:164-171)."""
    return np.array([[(d + 1) / 10.0 + s / 100.0 for d in range(Q)] for s in range(n)])


def _collide_textbook(f, tau, Q):
    c, w, _ = _lattice(Q)
    return f - (f - _textbook_feq(Q, f.sum(), c.T @ f)) / tau


@pytest.mark.parametrize("wall", ["SBB", "BFL"])
def test_fourcube_streaming_known_answers(wall):
    """One StreamAndCollide + PostStep on anisotropic data; every slot is
    checked against independent formulas with the reference tests' 1e-10."""
    case = K.four_cube_case("LBGK", wall)
    dom, Q, tau = case.domain, 15, case.tau
    c, w, inv = _lattice(Q)
    sim = case.oracle()
    f0 = _anisotropic(dom.n, Q)
    sim.set_f(f0)
    sim.step(1)
    f1 = sim.get_f()
    index = {tuple(x): i for i, x in enumerate(dom.coords)}
    fpc = np.array([_collide_textbook(f0[i], tau, Q) for i in range(dom.n)])
    rho_in = case.inlets[0].params["density_mean"]
    rho_out = case.outlets[0].params["density_mean"]
    for i in range(dom.n):
        cut = int(dom.wall_mask[i] | dom.iolet_mask[i])
        for e in range(1, Q):
            dp = inv[e]
            if (dom.iolet_mask[i] >> (dp - 1)) & 1:
                # Nash: ghost equilibrium with the iolet density and the
                # normal projection of the site's momentum (normal = +-z)
                rho_g = rho_in if dom.site_type[i] == 2 else rho_out
                m = c.T @ f0[i]
                n = np.array([0.0, 0.0, 1.0 if dom.site_type[i] == 2 else -1.0])
                mg = n * (m @ n) / f0[i].sum() * rho_g
                want = _textbook_feq(Q, rho_g, mg)[e]
            elif (dom.wall_mask[i] >> (dp - 1)) & 1:
                q = float(dom.cut[i, dp - 1])
                if wall == "SBB" or (dom.wall_mask[i] >> (e - 1)) & 1:
                    want = fpc[i, dp]
                elif q >= 0.5:
                    want = (fpc[i, dp] + (2 * q - 1) * fpc[i, e]) / (2 * q)
                else:
                    # PostStep: 2q f*_dp(x) + (1-2q) (value streamed into slot dp)
                    src = index[tuple(dom.coords[i] - c[dp])] if not (cut >> (e - 1)) & 1 else None
                    streamed = fpc[src, dp] if src is not None else None
                    if streamed is None:
                        continue
                    want = 2 * q * fpc[i, dp] + (1 - 2 * q) * streamed
            else:
                src = index[tuple(dom.coords[i] - c[e])]
                want = fpc[src, e]
            assert abs(f1[i, e] - want) <= 1e-10, (i, e, f1[i, e], want)


def test_ladd_links_match_textbook_velocity():
    """LaddIoletLink (LaddIolet.h:29-66): f_new = f*_d - 2 w_d rho (u_w . c_d)/Cs2
    with the parabolic velocity half way along the link."""
    case = K.cylinder_case(5.0, 12, 19, "LBGK", "SBB", "parabolic", tau=0.9)
    dom, Q = case.domain, 19
    c, w, inv = _lattice(Q)
    io = case.inlets[0].params
    pos = np.array(case.inlets[0].position)
    sim = case.oracle()
    rng = np.random.default_rng(0)
    f0 = w * (1 + 0.01 * rng.normal(size=(dom.n, Q)))
    sim.set_f(f0)
    sim.set_time(25)
    sim.step(1)
    f1 = sim.get_f()
    t = 25
    umax = io["max_speed"] * min(1.0, t / io["warmup"])
    checked = 0
    for i in np.nonzero(dom.iolet_mask)[0]:
        if dom.site_type[i] != 2:
            continue
        fp = _collide_textbook(f0[i], case.tau, Q)
        rho = f0[i].sum()
        for e in range(1, Q):
            dp = inv[e]
            if not (dom.iolet_mask[i] >> (dp - 1)) & 1:
                continue
            half = dom.coords[i] + 0.5 * c[dp]
            d = half - pos
            r2 = d[0] ** 2 + d[1] ** 2
            u = np.array([0, 0, umax * (1 - r2 / io["radius"] ** 2)])
            want = fp[dp] - 2 * w[dp] * rho * (u @ c[dp]) * 3.0
            assert abs(f1[i, e] - want) <= 1e-12, (i, e)
            checked += 1
    assert checked > 50


# ---------------------------------------------------------------------------
# Whole-run physics
# ---------------------------------------------------------------------------
def test_pressure_driven_cylinder_approaches_poiseuille():
    """Steady pressure-driven flow in a cylinder (Code/tests/pythontests/
    poiseuilleflowtest.py:41-45 idea): the centreline velocity is within
    discretisation error of G R^2 / (4 nu rho)."""
    R, Lz, tau = 6.0, 30, 1.0
    g = K.Geometry.cylinder(R, Lz)
    dom = g.project(19)
    drho = 0.006
    ins = [K.IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                       dict(density_mean=1.0 + drho / 2, density_amp=0.0, period=1.0))
           for p in dom.inlets]
    outs = [K.IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                        dict(density_mean=1.0 - drho / 2, density_amp=0.0, period=1.0))
            for p in dom.outlets]
    case = K.Case(dom, "LBGK", "BFL", "NASH", "NASH", tau, ins, outs, rho0=1.0)
    sim = case.oracle()
    sim.step(1500)
    rho, u = sim.macro()
    nu = (tau - 0.5) / 3.0
    length = float(dom.outlets[0].position[2] - dom.inlets[0].position[2])
    G = drho / 3.0 / length
    z = dom.coords[:, 2]
    mid = np.abs(z - z.mean()) < 2
    cx, cy = dom.inlets[0].position[0], dom.inlets[0].position[1]
    r2 = (dom.coords[:, 0] - cx) ** 2 + (dom.coords[:, 1] - cy) ** 2
    centre = mid & (r2 < 1.0)
    u_c = u[centre, 2].mean()
    want = G * R ** 2 / (4 * nu * 1.0)
    assert u_c > 0
    assert abs(u_c - want) / want < 0.08, (u_c, want)
    # profile shape: parabolic in r
    ring = mid & (np.abs(np.sqrt(r2) - R / 2) < 0.5)
    assert abs(u[ring, 2].mean() / u_c - 0.75) < 0.06


# ---------------------------------------------------------------------------
# Wall shear stress (Lattice.h:652-727)
# ---------------------------------------------------------------------------
def _wss_inputs(Q, k, rng):
    rho = rng.uniform(0.8, 1.2)
    feq = np.zeros(Q)
    O.lib().orc_feq(Q, rho, P(rng.normal(0, 0.03, 3) * rho), P(feq))
    f = feq * (1 + rng.normal(0, 1e-2, Q))
    n = rng.normal(size=3)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    if k % 5 == 0:
        n = np.float32([0, 0, 1]) if k % 2 else np.float32([1, 0, 0])
    return f, n, rng.uniform(0.51, 2.0)


@needs_ref
@pytest.mark.parametrize("Q", [15, 19])
def test_wall_shear_stress_bitwise_equal_reference(Q):
    rng = np.random.default_rng(40 + Q)
    for k in range(2000):
        f, n, tau = _wss_inputs(Q, k, rng)
        a = O.lib().orc_wall_shear_stress(Q, tau, P(f), P(n))
        b = O.ref().ref_wall_shear_stress(Q, tau, P(f), P(n))
        assert a == b or (math.isnan(a) and math.isnan(b)), (k, a, b)


@pytest.mark.parametrize("Q", [15, 19])
def test_wall_shear_stress_is_the_tangential_traction(Q):
    """|t - (t.n) n| with t = sigma n and sigma = -(1 - 1/(2 tau)) Pi_neq
    (the deviatoric stress of Chen & Doolen 1998), computed in numpy.  The
    reference's sqrt(|t|^2 - (t.n)^2) assumes |n| = 1, which a float normal
    meets to ~1e-8: tight tolerance for axis normals, 1e-6 otherwise."""
    c, w, _ = _lattice(Q)
    rng = np.random.default_rng(7)
    for k in range(300):
        f, n, tau = _wss_inputs(Q, k, rng)
        rho, m = np.zeros(1), np.zeros(3)
        O.lib().orc_density_momentum(Q, P(f), P(rho), P(m))
        fneq = f - _textbook_feq(Q, rho[0], m)
        sigma = -(1 - 1 / (2 * tau)) * np.einsum("l,la,lb->ab", fneq, c, c)
        nd = n.astype(np.float64)
        t = sigma @ nd
        want = np.linalg.norm(t - (t @ nd) * nd)
        got = O.lib().orc_wall_shear_stress(Q, tau, P(f), P(n))
        tol = 1e-12 if k % 5 == 0 else 1e-6
        assert abs(got - want) <= tol * np.linalg.norm(t) + 1e-18, (k, got, want)


@pytest.mark.parametrize("Q", [15, 19])
def test_wall_shear_stress_of_a_simple_shear_flow(Q):
    """Chapman-Enskog non-equilibrium of u_x = gamma*y:
    f_neq_i = -(tau w_i rho / cs^2) c_ix c_iy gamma, whose wall shear stress
    on a wall with normal y is the viscous stress rho*nu*gamma,
    nu = cs^2 (tau - 1/2)."""
    c, w, _ = _lattice(Q)
    for tau, rho, gamma in [(0.8, 1.0, 1e-3), (0.55, 51.0, 3e-4), (1.7, 0.97, -2e-3)]:
        feq = np.zeros(Q)
        m = np.array([0.01, 0.0, 0.0]) * rho
        O.lib().orc_feq(Q, rho, P(m), P(feq))
        f = feq - tau * w * rho * 3.0 * c[:, 0] * c[:, 1] * gamma
        got = O.lib().orc_wall_shear_stress(Q, tau, P(f), P(np.float32([0, 1, 0])))
        want = rho * (tau - 0.5) / 3.0 * abs(gamma)
        assert abs(got - want) <= 1e-12 * want


def _stress(lib_fn, Q, tau, f, n):
    t6, tr, tg = np.zeros(6), np.zeros(3), np.zeros(3)
    vm, sr = np.zeros(1), np.zeros(1)
    lib_fn(Q, tau, P(f), P(n), P(t6), P(tr), P(tg), P(vm), P(sr))
    return t6, tr, tg, vm[0], sr[0]


@needs_ref
@pytest.mark.parametrize("Q", [15, 19])
def test_stress_outputs_bitwise_equal_reference(Q):
    """Stress tensor, traction, tangential traction, von Mises stress and
    shear rate (Lattice.h:508-907) against the reference's own functions."""
    rng = np.random.default_rng(60 + Q)
    for k in range(1500):
        f, n, tau = _wss_inputs(Q, k, rng)
        a = _stress(O.lib().orc_stress_fields, Q, tau, f, n)
        b = _stress(O.ref().ref_stress_fields, Q, tau, f, n)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("Q", [15, 19])
def test_stress_outputs_of_a_simple_shear_flow(Q):
    """Chapman-Enskog f_neq of u_x = gamma*y: shear rate |gamma|;
    CalculateStressTensor adds +(1 - 1/(2 tau)) Pi_neq (Lattice.h:630-645, the
    opposite sign of its comment), so sigma_xy = (1 - 1/(2 tau)) *
    (-tau rho cs^2 gamma) = -rho nu gamma; traction on a y wall
    (sigma_xy, p', 0) with p' = (rho - 1) cs^2; von Mises sqrt(6) times the
    stress parameter times |Pi_xy|."""
    c, w, _ = _lattice(Q)
    for tau, rho, gamma in [(0.8, 1.02, 1e-3), (0.6, 0.97, -4e-4)]:
        feq = np.zeros(Q)
        O.lib().orc_feq(Q, rho, P(np.zeros(3)), P(feq))
        f = feq - tau * w * rho * 3.0 * c[:, 0] * c[:, 1] * gamma
        t6, tr, tg, vm, sr = _stress(O.lib().orc_stress_fields, Q, tau, f,
                                     np.float32([0, 1, 0]))
        nu = (tau - 0.5) / 3.0
        assert abs(sr - abs(gamma)) <= 1e-12 * abs(gamma)
        assert abs(t6[1] + rho * nu * gamma) <= 1e-12 * abs(rho * nu * gamma)
        pr = (rho - 1) / 3.0
        np.testing.assert_allclose([t6[0], t6[3], t6[5]], [pr] * 3, atol=1e-15)
        np.testing.assert_allclose(tr, [t6[1], pr, 0.0], atol=1e-15)
        np.testing.assert_allclose(tg, [t6[1], 0.0, 0.0], atol=1e-15)
        pi_xy = -tau * rho * gamma / 3.0
        sp = (1 - 1 / (2 * tau)) / np.sqrt(2)
        assert abs(vm - sp * np.sqrt(6.0) * abs(pi_xy)) <= 1e-12 * abs(vm)


def test_threaded_step_equals_single_thread():
    """The oracle's step splits the sites over host threads; every fNew slot
    has one writer, so the result must not depend on the thread count
    (bit for bit, BFL PostStep and Ladd/Nash iolets included)."""
    case = K.tree_case(3, 6.0, seed=1)
    n0 = O.lib().orc_threads()
    try:
        O.lib().orc_set_threads(1)
        a = case.oracle()
        a.step(6)
        O.lib().orc_set_threads(max(2, n0))
        b = case.oracle()
        b.step(6)
    finally:
        O.lib().orc_set_threads(n0)
    np.testing.assert_array_equal(a.get_f(), b.get_f())


# ---------------------------------------------------------------------------
# The iolet time index inside a step
# ---------------------------------------------------------------------------
# SimulationState starts at t = 1 (SimulationState.cc:15-16).  During the
# stream of step t a Nash link takes the iolet density at the 0-indexed step
# t - 1 (BoundaryValues.cc:162-165), a Ladd link the iolet velocity at the
# 1-indexed t (LaddIolet.h:52, BoundaryValues.h:67-70).  Each test below
# checks every iolet slot of one step against a textbook formula evaluated
# at that index, and that the neighbouring indices would miss by far more
# than the tolerance, so an off-by-one in either direction fails.
def _perturbed(dom, Q, seed):
    c, w, _ = _lattice(Q)
    rng = np.random.default_rng(seed)
    return w * (1.0 + 0.02 * rng.normal(size=(dom.n, Q))) + w * 0.01 * (c @ [0.0, 0.0, 1.0])


def _nash_expected(case, f0, rho_of_site):
    """Slot values of every Nash link after one step from f0: the ghost
    equilibrium (textbook form) at the iolet density, with the momentum's
    projection on the float-rounded normal."""
    dom, Q = case.domain, case.domain.Q
    c, w, inv = _lattice(Q)
    out = {}
    for i in np.nonzero(dom.iolet_mask)[0]:
        spec = case.inlets[0] if dom.site_type[i] == 2 else case.outlets[0]
        n = np.asarray(spec.normal, float)
        nf = (n / np.linalg.norm(n)).astype(np.float32).astype(float)
        rho_g = rho_of_site(i)
        m = c.T @ f0[i]
        mg = nf * ((m @ nf) / f0[i].sum()) * rho_g
        feq = _textbook_feq(Q, rho_g, mg)
        for dp in range(1, Q):
            if (dom.iolet_mask[i] >> (dp - 1)) & 1:
                out[(i, inv[dp])] = feq[inv[dp]]
    return out


@pytest.mark.parametrize("t0", [None, 2, 5, 8, 13])
def test_nash_link_uses_density_of_step_t_minus_1(t0):
    """A short-period, large-amplitude cosine inlet (T = 8 steps): the Nash
    slots after the stream of step t equal the ghost equilibrium at
    rho(t - 1).  t0 = None checks a fresh run, which starts at t = 1 and so
    streams with rho(0) = mean + amplitude."""
    Q = 15
    dom = K.Geometry.cylinder(5.0, 12).project(Q)
    T, mean, amp, phase = 8.0, 1.02, 0.01, 0.4
    ins = [K.IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                       dict(density_mean=mean, density_amp=amp, period=T, phase=phase))
           for p in dom.inlets]
    outs = [K.IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                        dict(density_mean=1.0, density_amp=0.0, period=T))
            for p in dom.outlets]
    case = K.Case(dom, "LBGK", "SBB", "NASH", "NASH", 0.8, ins, outs, rho0=1.0)
    sim = case.oracle()
    f0 = _perturbed(dom, Q, 11)
    sim.set_f(f0)
    if t0 is None:
        assert sim.time() == 1
        t = 1
    else:
        sim.set_time(t0)
        t = t0
    sim.step(1)
    assert sim.time() == t + 1
    f1 = sim.get_f()

    def rho_at(tt):
        def rho(i):
            if dom.site_type[i] != 2:
                return 1.0
            return mean + amp * math.cos(2.0 * math.pi / T * tt + phase)
        return rho

    want = _nash_expected(case, f0, rho_at(t - 1))
    assert sum(1 for (i, _) in want if dom.site_type[i] == 2) > 20
    for (i, e), v in want.items():
        assert abs(f1[i, e] - v) <= 1e-12, (i, e, f1[i, e], v)
    # the neighbouring time indices give values far outside the tolerance
    for other in (t, t - 2):
        alt = _nash_expected(case, f0, rho_at(other))
        gaps = [abs(alt[k] - want[k]) for k in want if dom.site_type[k[0]] == 2]
        assert min(gaps) > 1e-7, (other, min(gaps))


def _ladd_expected(case, f0, u_of):
    """Slot values of every Ladd link after one step from f0:
    f*_d - 2 w_d rho (u_w . c_d) / Cs2 with u_w at the link's midpoint."""
    dom, Q = case.domain, case.domain.Q
    c, w, inv = _lattice(Q)
    out = {}
    for i in np.nonzero(dom.iolet_mask)[0]:
        if dom.site_type[i] != 2:
            continue
        fp = _collide_textbook(f0[i], case.tau, Q)
        rho = f0[i].sum()
        for dp in range(1, Q):
            if not (dom.iolet_mask[i] >> (dp - 1)) & 1:
                continue
            u = u_of(dom.coords[i] + 0.5 * c[dp])
            out[(i, inv[dp])] = fp[dp] - 2 * w[dp] * rho * (u @ c[dp]) * 3.0
    return out


def _ladd_case(inlet):
    Q = 19
    dom = K.Geometry.cylinder(5.0, 12).project(Q)
    _, outs = K.pressure_iolets(dom, 1.0, 0.0, 40.0, 1.0)
    return K.Case(dom, "LBGK", "SBB", "LADD", "NASH", 0.9, [inlet(dom.inlets[0])], outs,
                  rho0=1.0)


@pytest.mark.parametrize("t0", [None, 25, 99])
def test_ladd_link_uses_one_indexed_step_during_warmup(t0):
    """Parabolic inlet inside its warm-up (100 steps): the profile scales
    with t / warmup at the 1-indexed step t being streamed."""
    umax, warm = 0.02, 100

    def inlet(p):
        return K.IoletSpec("parabolic", tuple(p.position), tuple(p.normal),
                           dict(radius=p.radius, max_speed=umax, warmup=warm))
    case = _ladd_case(inlet)
    dom = case.domain
    spec = case.inlets[0]
    pos, n = np.asarray(spec.position, float), np.asarray(spec.normal, float)
    sim = case.oracle()
    f0 = _perturbed(dom, 19, 12)
    sim.set_f(f0)
    t = 1 if t0 is None else t0
    if t0 is not None:
        sim.set_time(t0)
    sim.step(1)
    f1 = sim.get_f()

    def u_at(tt):
        def u(x):
            d = x - pos
            r2 = d @ d - (d @ n) ** 2
            return n * umax * min(1.0, tt / warm) * (1 - r2 / spec.params["radius"] ** 2)
        return u

    want = _ladd_expected(case, f0, u_at(t))
    assert len(want) > 50
    for (i, e), v in want.items():
        assert abs(f1[i, e] - v) <= 1e-13, (i, e, f1[i, e], v)
    for other in (t - 1, t + 1):
        alt = _ladd_expected(case, f0, u_at(other))
        assert max(abs(alt[k] - want[k]) for k in want) > 1e-7


@pytest.mark.parametrize("t0", [None, 7, 30])
def test_ladd_link_uses_one_indexed_step_womersley(t0):
    """Womersley inlet with a 40-step period: the velocity at the link's
    midpoint is the textbook solution (scipy's complex Bessel function) at
    the 1-indexed step t.  The reference's J0 series carries ~1e-6 relative
    error (Bessel.h:33), which sets the tolerance; one step of phase moves
    the slots by ~1e-4."""
    from scipy.special import jv
    T, G, alpha = 40.0, 2e-4, 3.0

    def inlet(p):
        return K.IoletSpec("womersley", tuple(p.position), tuple(p.normal),
                           dict(radius=p.radius, pressure_gradient_amplitude=G, period=T,
                                womersley_number=alpha))
    case = _ladd_case(inlet)
    dom = case.domain
    spec = case.inlets[0]
    R = spec.params["radius"]
    pos, n = np.asarray(spec.position, float), np.asarray(spec.normal, float)
    sim = case.oracle()
    f0 = _perturbed(dom, 19, 13)
    sim.set_f(f0)
    t = 1 if t0 is None else t0
    if t0 is not None:
        sim.set_time(t0)
    sim.step(1)
    f1 = sim.get_f()
    w_ = 2 * math.pi / T
    i32 = complex(math.cos(3 * math.pi / 4), math.sin(3 * math.pi / 4))

    def u_at(tt):
        def u(x):
            d = x - pos
            r = math.sqrt(max(0.0, d @ d - (d @ n) ** 2))
            a = G / w_ * (1 - jv(0, i32 * alpha * r / R) / jv(0, i32 * alpha))
            return n * -(a * complex(math.cos(w_ * tt), math.sin(w_ * tt))).real
        return u

    want = _ladd_expected(case, f0, u_at(t))
    assert len(want) > 50
    scale = 2 * (1 / 18) * 3 * G / w_
    for (i, e), v in want.items():
        assert abs(f1[i, e] - v) <= 1e-5 * scale, (i, e, f1[i, e], v)
    for other in (t - 1, t + 1):
        alt = _ladd_expected(case, f0, u_at(other))
        assert max(abs(alt[k] - want[k]) for k in want) > 1e-2 * scale
