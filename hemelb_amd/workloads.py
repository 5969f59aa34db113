"""The synthetic workloads the solver is measured and checked on.

* ``small_cylinder``: straight cylinder, radius 12 voxels, length 96 along z
  (voxel 5e-5 m), D3Q15 LBGK tau = 0.8, simple bounce-back; Nash inlet at
  80 + 1 cos(2 pi t / 2000) mmHg, Nash outlet at 80 mmHg; rest equilibrium at
  80 mmHg; 6000 steps.
* ``vessel_tree``: the seeded 5-generation bifurcating tree (root radius 40
  voxels, Murray's-law splits), D3Q19 at tau = 0.7 (LBGK, or MRT with the
  d'Humieres rates), BFL walls, a Ladd/Womersley root inlet whose centreline
  velocity amplitude is 0.05, 16 zero-gauge Nash outlets; 10000 steps.

Physical parameters follow the reference's defaults: blood density
1000 kg/m^3 and viscosity 0.004 Pa s (Code/constants.h, SimConfig.cc).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from . import lb as L
from .geometry import Domain, Geometry

BLOOD_DENSITY = 1000.0       # kg/m^3
BLOOD_VISCOSITY = 0.004      # Pa s


@dataclass
class Workload:
    name: str
    domain: Domain
    traits: L.Traits
    tau: float
    inlets: list
    outlets: list
    rho0: float
    steps: int
    period: float                      # steps per cycle of the inlet
    meta: dict = field(default_factory=dict)

    def solver(self, storage: str = "fp64", layout: str = "blocks", device: int = 0) -> L.LBM:
        lbm = L.LBM(self.traits, self.domain, L.LbmParameters(tau=self.tau), self.inlets,
                    self.outlets, storage=storage, rho0=self.rho0, device=device,
                    layout=layout)
        lbm.SetInitialConditions(self.rho0)
        return lbm


def _j0(z: complex) -> complex:
    s, term, k = 1 + 0j, 1 + 0j, 0
    while abs(term) > 1e-17 * abs(s) or k < 5:
        k += 1
        term *= -(z * z) / (4 * k * k)
        s += term
    return s


def womersley_inlet(plane, peak_speed: float, period: float, tau: float):
    """Womersley inlet (InOutLetWomersleyVelocity) on `plane` whose centreline
    velocity amplitude is `peak_speed`; alpha = R sqrt(omega / nu) from the
    lattice viscosity nu = Cs2 (tau - 1/2)."""
    nu = (tau - 0.5) * L.Cs2
    omega = 2 * math.pi / period
    alpha = plane.radius * math.sqrt(omega / nu)
    amp = abs(1 - 1 / _j0(complex(-math.sqrt(0.5), math.sqrt(0.5)) * alpha)) / omega
    return L.InOutLetWomersleyVelocity(tuple(plane.position), tuple(plane.normal),
                                       radius=plane.radius,
                                       pressure_gradient_amplitude=peak_speed / amp,
                                       period=period, womersley_number=alpha)


def small_cylinder(radius: float = 12.0, length: int = 96) -> Workload:
    dx, tau = 5e-5, 0.8
    dt = (tau - 0.5) * L.Cs2 * dx * dx * BLOOD_DENSITY / BLOOD_VISCOSITY
    units = L.UnitConverter(dt, dx, BLOOD_DENSITY, 0.0)
    dom = Geometry.cylinder(radius, length).project(15)
    period = 2000.0
    to_rho = lambda p: units.ConvertPressureToLatticeUnits(p) / L.Cs2  # noqa: E731
    rho_mean = to_rho(80.0)
    rho_amp = units.ConvertPressureDifferenceToLatticeUnits(1.0) / L.Cs2
    ins = [L.InOutLetCosine(tuple(p.position), tuple(p.normal), density_mean=rho_mean,
                            density_amp=rho_amp, period=period) for p in dom.inlets]
    outs = [L.InOutLetCosine(tuple(p.position), tuple(p.normal), density_mean=rho_mean,
                             density_amp=0.0, period=period) for p in dom.outlets]
    traits = L.Traits(L.D3Q15, L.LBGK, L.SimpleBounceBack, L.NashZerothOrderPressure,
                      L.NashZerothOrderPressure)
    return Workload("small_cylinder_R12_L96_D3Q15_LBGK_SBB_Nash", dom, traits, tau, ins,
                    outs, rho_mean, 6000, period,
                    dict(time_step_s=dt, voxel_size_m=dx, units=units))


def vessel_tree(kernel: str = "LBGK", seed: int = 0, generations: int = 5,
                root_radius: float = 40.0, geometry: Geometry | None = None) -> Workload:
    tau, period, peak = 0.7, 5000.0, 0.05
    g = geometry if geometry is not None else Geometry.vessel_tree(seed, generations,
                                                                   root_radius)
    dom = g.project(19)
    ins = [womersley_inlet(p, peak, period, tau) for p in dom.inlets]
    outs = [L.InOutLetCosine(tuple(p.position), tuple(p.normal), density_mean=1.0,
                             density_amp=0.0, period=period) for p in dom.outlets]
    traits = L.Traits(L.D3Q19, {"LBGK": L.LBGK, "MRT": L.MRT}[kernel],
                      L.BouzidiFirdaousLallemand, L.LaddIolet, L.NashZerothOrderPressure)
    name = f"vessel_tree_{generations}gen_R{root_radius:g}_D3Q19_{kernel}_BFL_LaddWomersley_Nash"
    return Workload(name, dom, traits, tau, ins, outs, 1.0, 10000, period,
                    dict(womersley=ins[0].womersley_number,
                         pressure_gradient=ins[0].pressure_gradient_amplitude,
                         peak_speed=peak))
