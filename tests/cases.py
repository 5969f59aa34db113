"""Domains and configurations shared by the CPU and GPU tests.

Every case can be run both by the CPU oracle (oracle/bindings.py) and by the
HIP implementation (hemelb_amd.LBM) from the same arrays.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from hemelb_amd import _native as N
from hemelb_amd import lb as L
from hemelb_amd.geometry import Domain, Geometry
from oracle import bindings as O

KERNELS = {"LBGK": (O.ORC_LBGK, L.LBGK), "MRT": (O.ORC_MRT, L.MRT)}
WALLS = {"SBB": (O.ORC_WALL_SBB, L.SimpleBounceBack),
         "BFL": (O.ORC_WALL_BFL, L.BouzidiFirdaousLallemand)}
BCS = {"NASH": (O.ORC_IOLET_NASH, L.NashZerothOrderPressure),
       "LADD": (O.ORC_IOLET_LADD, L.LaddIolet)}


@dataclass
class IoletSpec:
    kind: str                    # "cosine" | "parabolic" | "womersley"
    position: tuple
    normal: tuple
    params: dict = field(default_factory=dict)

    def to_oracle(self) -> O.orc_iolet:
        kind = {"cosine": O.ORC_COSINE, "parabolic": O.ORC_PARABOLIC,
                "womersley": O.ORC_WOMERSLEY}[self.kind]
        return O.make_iolet(kind, self.position, self.normal, **self.params)

    def to_product(self):
        cls = {"cosine": L.InOutLetCosine, "parabolic": L.InOutLetParabolicVelocity,
               "womersley": L.InOutLetWomersleyVelocity}[self.kind]
        return cls(position=tuple(self.position), normal=tuple(self.normal), **self.params)


@dataclass
class Case:
    domain: Domain
    kernel: str = "LBGK"
    wall: str = "SBB"
    inlet_bc: str = "NASH"
    outlet_bc: str = "NASH"
    tau: float = 0.8
    inlets: list = field(default_factory=list)
    outlets: list = field(default_factory=list)
    rho0: float = 1.0

    def oracle(self) -> O.OracleSim:
        d = self.domain
        sim = O.OracleSim(d.Q, KERNELS[self.kernel][0], WALLS[self.wall][0],
                          BCS[self.inlet_bc][0], BCS[self.outlet_bc][0], self.tau,
                          d.coords, d.site_type, d.wall_mask, d.iolet_mask, d.iolet_id,
                          d.cut, [i.to_oracle() for i in self.inlets],
                          [o.to_oracle() for o in self.outlets])
        sim.set_equilibrium(self.rho0)
        return sim

    def traits(self) -> L.Traits:
        return L.Traits(L.LATTICES[self.domain.Q], KERNELS[self.kernel][1],
                        WALLS[self.wall][1], BCS[self.inlet_bc][1], BCS[self.outlet_bc][1])

    def product(self, storage="fp64", device=0, layout="blocks") -> L.LBM:
        lbm = L.LBM(self.traits(), self.domain, L.LbmParameters(tau=self.tau),
                    [i.to_product() for i in self.inlets],
                    [o.to_product() for o in self.outlets], storage=storage,
                    rho0=self.rho0, device=device, layout=layout)
        lbm.SetInitialConditions(self.rho0)
        return lbm


# ---------------------------------------------------------------------------
# FourCube (Code/tests/helpers/FourCubeLatticeData.cc:53-170)
# ---------------------------------------------------------------------------
def glibc_rand(seed: int = 1):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    while True:
        yield int(libc.rand())


def four_cube(Q: int = 15, sites_per_block: int = 6, seed: int = 1) -> Domain:
    """The 4^3 fluid cube of the reference's unit tests: z = min plane has
    inlet links (iolet 0), z = max outlet links (iolet 0), x/y faces walls,
    cut distances float(rand() % 10000) / 10000.0 drawn per link in
    (i, j, k, direction) order; sites in the reference's local order
    (collision type, then block traversal with z fastest)."""
    lat = L.LATTICES[Q]
    rnd = glibc_rand(seed)
    bs = sites_per_block
    lo, hi = 1, bs - 2
    rows = []
    for i in range(bs):
        for j in range(bs):
            for k in range(bs):
                if not (lo <= i <= hi and lo <= j <= hi and lo <= k <= hi):
                    continue
                wm = im = 0
                iid = -1
                had_in = had_out = False
                cut = np.full(Q - 1, -1.0, np.float32)
                for d in range(1, Q):
                    ni, nj, nk = i + lat.VECTORS[d][0], j + lat.VECTORS[d][1], k + lat.VECTORS[d][2]
                    r = next(rnd)
                    dist = np.float32(float(np.float32(r % 10000)) / 10000.0)
                    if nk < lo:
                        im |= 1 << (d - 1); iid = 0; had_in = True; cut[d - 1] = dist
                    elif nk > hi:
                        im |= 1 << (d - 1); iid = 0; had_out = True; cut[d - 1] = dist
                    elif ni < lo or nj < lo or ni > hi or nj > hi:
                        wm |= 1 << (d - 1); cut[d - 1] = dist
                st = N.HLB_SITE_INLET if had_in else (N.HLB_SITE_OUTLET if had_out else N.HLB_SITE_FLUID)
                nrm = np.zeros(3, np.float32)
                if i == lo: nrm = np.array([-1, 0, 0], np.float32)
                if i == hi: nrm = np.array([1, 0, 0], np.float32)
                if j == lo: nrm = np.array([0, -1, 0], np.float32)
                if j == hi: nrm = np.array([0, 1, 0], np.float32)
                ct = (1 if wm else 0) if st == N.HLB_SITE_FLUID else (
                    (4 if wm else 2) if st == N.HLB_SITE_INLET else (5 if wm else 3))
                rows.append((ct, (i * bs + j) * bs + k, (i, j, k), st, wm, im, iid, cut, nrm))
    rows.sort(key=lambda r: (r[0], r[1]))
    dom = Domain(Q,
                 np.array([r[2] for r in rows], np.int32),
                 np.array([r[3] for r in rows], np.uint32),
                 np.array([r[4] for r in rows], np.uint32),
                 np.array([r[5] for r in rows], np.uint32),
                 np.array([r[6] for r in rows], np.int32),
                 np.array([r[7] for r in rows], np.float32),
                 np.array([r[8] for r in rows], np.float32), [], [])
    return dom


def four_cube_case(kernel="LBGK", wall="SBB", Q=15, tau=0.8) -> Case:
    """four_cube.xml iolets: cosine pressure, inlet mean 80.1 mmHg, outlet
    80.0 mmHg, period 0.6 s, with the reference's unit conversion."""
    dt, dx = 0.0857, 0.01
    uc = L.UnitConverter(dt, dx, 1000.0, 0.0)
    mean_in = uc.ConvertPressureToLatticeUnits(80.1) / L.Cs2
    mean_out = uc.ConvertPressureToLatticeUnits(80.0) / L.Cs2
    period = uc.ConvertTimeToLatticeUnits(0.6)
    ins = [IoletSpec("cosine", (2.5, 2.5, 0.5), (0, 0, 1),
                     dict(density_mean=mean_in, density_amp=0.0, period=period))]
    outs = [IoletSpec("cosine", (2.5, 2.5, 4.5), (0, 0, -1),
                      dict(density_mean=mean_out, density_amp=0.0, period=period))]
    return Case(four_cube(Q), kernel, wall, "NASH", "NASH", tau, ins, outs,
                rho0=uc.ConvertPressureToLatticeUnits(80.0) / L.Cs2)


# ---------------------------------------------------------------------------
# Vessels
# ---------------------------------------------------------------------------
def pressure_iolets(geom_domain: Domain, rho_in_mean, rho_in_amp, period, rho_out):
    ins = [IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                     dict(density_mean=rho_in_mean, density_amp=rho_in_amp, period=period))
           for p in geom_domain.inlets]
    outs = [IoletSpec("cosine", tuple(p.position), tuple(p.normal),
                      dict(density_mean=rho_out, density_amp=0.0, period=period))
            for p in geom_domain.outlets]
    return ins, outs


def womersley_inlets(geom_domain: Domain, peak_speed, period, tau):
    """Womersley inlets whose centreline velocity amplitude is `peak_speed`
    (lattice units) for the lattice viscosity of `tau`."""
    ins = []
    nu = (tau - 0.5) / 3.0
    for p in geom_domain.inlets:
        omega = 2 * math.pi / period
        alpha = p.radius * math.sqrt(omega / nu)
        # |A| at r = 0 for unit pressure gradient: |(1 - 1/J0(i^{3/2} a))| / omega
        z = complex(-math.sqrt(0.5), math.sqrt(0.5)) * alpha
        j0 = _bessel_j0(z)
        amp = abs(1 - 1 / j0) / omega
        ins.append(IoletSpec("womersley", tuple(p.position), tuple(p.normal),
                             dict(radius=p.radius, pressure_gradient_amplitude=peak_speed / amp,
                                  period=period, womersley_number=alpha)))
    return ins


def parabolic_inlets(geom_domain: Domain, max_speed, warmup=0):
    return [IoletSpec("parabolic", tuple(p.position), tuple(p.normal),
                      dict(radius=p.radius, max_speed=max_speed, warmup=warmup))
            for p in geom_domain.inlets]


def _bessel_j0(z: complex) -> complex:
    s, term, k = 1 + 0j, 1 + 0j, 0
    while abs(term) > 1e-16 * abs(s) or k < 5:
        k += 1
        term *= -(z * z) / (4 * k * k)
        s += term
    return s


def cylinder_case(radius=6.0, length=24, Q=15, kernel="LBGK", wall="SBB", inlet="cosine",
                  tau=0.8, steps_period=200.0) -> Case:
    g = Geometry.cylinder(radius, length)
    dom = g.project(Q)
    if inlet == "cosine":
        ins, outs = pressure_iolets(dom, 1.01, 0.005, steps_period, 1.0)
        return Case(dom, kernel, wall, "NASH", "NASH", tau, ins, outs, rho0=1.0)
    _, outs = pressure_iolets(dom, 1.0, 0.0, steps_period, 1.0)
    if inlet == "womersley":
        ins = womersley_inlets(dom, 0.02, steps_period, tau)
    else:
        ins = parabolic_inlets(dom, 0.02, warmup=40)
    return Case(dom, kernel, wall, "LADD", "NASH", tau, ins, outs, rho0=1.0)


def tree_case(generations=3, root_radius=8.0, seed=0, Q=19, kernel="LBGK", wall="BFL",
              tau=0.7, period=400.0) -> Case:
    g = Geometry.vessel_tree(seed, generations, root_radius)
    dom = g.project(Q)
    _, outs = pressure_iolets(dom, 1.0, 0.0, period, 1.0)
    ins = womersley_inlets(dom, 0.05, period, tau)
    return Case(dom, kernel, wall, "LADD", "NASH", tau, ins, outs, rho0=1.0)


# ---------------------------------------------------------------------------
# Oracle runs of the benchmark workloads (hemelb_amd.workloads)
# ---------------------------------------------------------------------------
def _oracle_iolet(io) -> O.orc_iolet:
    if isinstance(io, L.InOutLetCosine):
        return O.make_iolet(O.ORC_COSINE, io.position, io.normal,
                            density_mean=io.density_mean, density_amp=io.density_amp,
                            phase=io.phase, period=io.period, warmup=int(io.warmup))
    if isinstance(io, L.InOutLetWomersleyVelocity):
        return O.make_iolet(O.ORC_WOMERSLEY, io.position, io.normal, radius=io.radius,
                            pressure_gradient_amplitude=io.pressure_gradient_amplitude,
                            period=io.period, womersley_number=io.womersley_number)
    if isinstance(io, L.InOutLetParabolicVelocity):
        return O.make_iolet(O.ORC_PARABOLIC, io.position, io.normal, radius=io.radius,
                            max_speed=io.max_speed, warmup=int(io.warmup))
    raise TypeError(type(io))


_ORC_BC = {L.NashZerothOrderPressure: O.ORC_IOLET_NASH, L.LaddIolet: O.ORC_IOLET_LADD}
_ORC_WALL = {L.SimpleBounceBack: O.ORC_WALL_SBB, L.BouzidiFirdaousLallemand: O.ORC_WALL_BFL}
_ORC_KERNEL = {L.LBGK: O.ORC_LBGK, L.MRT: O.ORC_MRT}


def workload_oracle(w) -> O.OracleSim:
    """The oracle set up like ``w.solver()``: the same domain, traits,
    relaxation time and iolets, rest equilibrium at w.rho0."""
    d, tr = w.domain, w.traits
    sim = O.OracleSim(d.Q, _ORC_KERNEL[tr.Kernel], _ORC_WALL[tr.WallStreamer],
                      _ORC_BC[tr.InletStreamer], _ORC_BC[tr.OutletStreamer], w.tau,
                      d.coords, d.site_type, d.wall_mask, d.iolet_mask, d.iolet_id, d.cut,
                      [_oracle_iolet(i) for i in w.inlets],
                      [_oracle_iolet(o) for o in w.outlets])
    sim.set_equilibrium(w.rho0)
    return sim
