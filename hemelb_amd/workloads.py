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


# ---------------------------------------------------------------------------
# Batched ensemble of stenosed cylinders
# ---------------------------------------------------------------------------
@dataclass
class Instance:
    radius: float
    length: int
    severity: float            # area reduction of the Gaussian stenosis
    width: float               # its width (voxels)
    z0: float                  # its axial position (voxels)
    tau: float
    amplitude_mmHg: float      # inlet amplitude (mmHg; lattice density if built so)
    offset: tuple              # placement of the instance in the ensemble
    sites: slice               # its rows of the ensemble domain
    inlets: slice = slice(0, 0)   # its entries of the ensemble's inlet list
    outlets: slice = slice(0, 0)  # and outlet list


@dataclass
class Ensemble:
    name: str
    domain: Domain
    traits: L.Traits
    instances: list
    instance_of_site: "object"   # int32 [n]
    inlets: list
    outlets: list
    steps: int
    period: float
    voxel_size_m: float
    meta: dict = field(default_factory=dict)

    @property
    def taus(self):
        return [i.tau for i in self.instances]

    def solver(self, storage: str = "fp32", device: int = 0,
               layout: str = "blocks") -> L.LBM:
        lbm = L.LBM(self.traits, self.domain, L.LbmParameters(tau=self.taus[0]), self.inlets,
                    self.outlets, storage=storage, rho0=1.0, device=device, layout=layout,
                    instances=self.instance_of_site, instance_tau=self.taus)
        lbm.SetInitialConditions(1.0)
        return lbm

    def instance_domain(self, k: int) -> Domain:
        """Instance k alone (its own iolet ids), for checks against a plain run."""
        inst = self.instances[k]
        d, sl = self.domain, inst.sites
        ids = d.iolet_id[sl].copy()
        ids[d.site_type[sl] == 2] -= inst.inlets.start
        ids[d.site_type[sl] == 3] -= inst.outlets.start
        return Domain(d.Q, d.coords[sl].copy(), d.site_type[sl].copy(), d.wall_mask[sl].copy(),
                      d.iolet_mask[sl].copy(), ids, d.cut[sl].copy(), d.normal[sl].copy(),
                      [], [])


def lattice_density_amplitude(tau: float, amplitude_mmHg: float,
                              voxel_size_m: float = 5e-6) -> float:
    """Lattice density amplitude of an inlet pressure amplitude for blood at
    relaxation time tau on voxels of voxel_size_m (time step from the lattice
    viscosity, UnitConverter.cc)."""
    dt = (tau - 0.5) * L.Cs2 * voxel_size_m ** 2 * BLOOD_DENSITY / BLOOD_VISCOSITY
    units = L.UnitConverter(dt, voxel_size_m, BLOOD_DENSITY, 80.0)
    return units.ConvertPressureDifferenceToLatticeUnits(amplitude_mmHg) / L.Cs2


def make_ensemble(parts: list, period: float, voxel_size_m: float = 5e-6,
                  steps: int = 4000, name: str = "ensemble",
                  amplitude_units: str = "mmHg") -> Ensemble:
    """An ensemble of independent vessels: parts = [(geometry, Instance
    fields radius, length, severity, width, z0, tau, amplitude)] with one
    inlet and any number of outlets each.  Inlet pressure 80 + A cos(2 pi t /
    period) mmHg, outlets 80 mmHg (the reference pressure, density 1), D3Q19
    LBGK, BFL walls, Nash iolets; amplitude_units "lattice" gives A as a
    lattice density amplitude instead.  Each vessel sits in its own aligned
    cube of 8^3 blocks, so that the Morton order keeps it contiguous."""
    import numpy as np

    doms = [g.project(19) for g, *_ in parts]
    extent = max(int((d.coords.max(0) - d.coords.min(0)).max()) + 1 for d in doms)
    side = 8
    while side < extent + 8:
        side *= 2
    per_row = max(1, int(np.ceil(len(parts) ** (1 / 3))))
    pieces, instances, ins, outs = [], [], [], []
    start = 0
    n_in = n_out = 0
    for k, ((g, R, length, sev, width, z0, tau, amp), d) in enumerate(zip(parts, doms)):
        cube = (k % per_row, (k // per_row) % per_row, k // (per_row * per_row))
        off = np.array([c * side - int(d.coords[:, a].min()) + 8 for a, c in enumerate(cube)],
                       np.int32)
        ids = d.iolet_id.copy()
        ids[d.site_type == 2] += n_in
        ids[d.site_type == 3] += n_out
        pieces.append(Domain(19, d.coords + off, d.site_type, d.wall_mask, d.iolet_mask, ids,
                             d.cut, d.normal, [], []))
        if amplitude_units == "lattice":
            rho_amp = float(amp)
        elif amplitude_units == "mmHg":
            rho_amp = lattice_density_amplitude(tau, amp, voxel_size_m)
        else:
            raise ValueError(f"unknown amplitude units {amplitude_units!r}")
        for p in d.inlets:
            ins.append(L.InOutLetCosine(tuple(np.asarray(p.position) + off), tuple(p.normal),
                                        density_mean=1.0, density_amp=rho_amp, period=period))
        for p in d.outlets:
            outs.append(L.InOutLetCosine(tuple(np.asarray(p.position) + off), tuple(p.normal),
                                         density_mean=1.0, density_amp=0.0, period=period))
        instances.append(Instance(R, length, sev, width, z0, tau, amp,
                                  tuple(int(v) for v in off), slice(start, start + d.n),
                                  slice(n_in, n_in + len(d.inlets)),
                                  slice(n_out, n_out + len(d.outlets))))
        n_in += len(d.inlets)
        n_out += len(d.outlets)
        start += d.n
    cat = lambda f: np.concatenate([getattr(p, f) for p in pieces])  # noqa: E731
    dom = Domain(19, cat("coords"), cat("site_type"), cat("wall_mask"), cat("iolet_mask"),
                 cat("iolet_id"), cat("cut"), cat("normal"), [], [])
    inst = np.concatenate([np.full(p.n, k, np.int32) for k, p in enumerate(pieces)])
    traits = L.Traits(L.D3Q19, L.LBGK, L.BouzidiFirdaousLallemand, L.NashZerothOrderPressure,
                      L.NashZerothOrderPressure)
    return Ensemble(name, dom, traits, instances, inst, ins, outs, steps, period,
                    voxel_size_m, dict(cube_side=side, sites=int(dom.n)))


def stenosed_ensemble(count: int = 96, seed: int = 0, radius_range=(8.0, 40.0),
                      voxel_size_m: float = 5e-6, period: float = 2000.0,
                      steps: int = 4000) -> Ensemble:
    """`count` seeded stenosed cylinders: radius ~U(radius_range), length
    10 x radius, Gaussian stenosis of area reduction ~U(0.2, 0.7) and width
    ~U(1, 3) x radius at a random axial position, tau ~U(0.6, 1.0), inlet
    pressure amplitude ~U(0.5, 2) mmHg."""
    import numpy as np

    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(count):
        R = float(rng.uniform(*radius_range))
        length = int(round(10 * R))
        sev = float(rng.uniform(0.2, 0.7))
        width = float(rng.uniform(1.0, 3.0) * R)
        z0 = float(rng.uniform(0.3, 0.7) * length)
        tau = float(rng.uniform(0.6, 1.0))
        amp = float(rng.uniform(0.5, 2.0))
        g = Geometry.stenosed_cylinder(R, length, sev, z0, width)
        parts.append((g, R, length, sev, width, z0, tau, amp))
    return make_ensemble(parts, period, voxel_size_m, steps,
                         f"stenosed_ensemble_{count}_seed{seed}_D3Q19_LBGK_BFL_Nash")
