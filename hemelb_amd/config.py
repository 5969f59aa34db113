"""HemeLB XML configuration (``hemelbsettings`` version 5) for the LB path:
reading and writing the elements the timestep uses, and building a solver
from them as the reference's SimBuilder does.

Elements read (Code/configuration/SimConfig.cc):
* ``simulation``: steps, step_length, extra_warmup_steps, voxel_size,
  origin, fluid_density, fluid_viscosity, reference_pressure, stresstype
  (``DoIOForSimulation``, :104-170);
* ``geometry/datafile``;
* ``initialconditions``: ``pressure/uniform`` or ``checkpoint``;
* ``inlets`` / ``outlets``: position and normal (``DoIOForBaseInOutlet``,
  :510-525), with pressure/cosine (:669-683), velocity/parabolic (:715-727)
  or velocity/womersley (:729-743) conditions.

Every dimensional value is checked against the units the reference requires
(``GetDimensionalValue``, SimConfig.h:44-58).  Other elements (properties,
visualisation, monitoring) are left to the caller.  ``build`` converts to
lattice units with the reference's rules (Code/configuration/SimBuilder.cc:
49-160): tau from the step length, voxel size and fluid properties, iolet
positions (x - origin) / dx, pressures through UnitConverter, periods in
steps.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

from . import lb as L

DEFAULT_FLUID_DENSITY = 1000.0    # kg/m^3 (Code/constants.h:21)
DEFAULT_FLUID_VISCOSITY = 0.004   # Pa s (Code/constants.h:22)


class ConfigError(ValueError):
    pass


@dataclass
class IoletConfig:
    """One inlet or outlet as written in the file (physical units)."""
    kind: str                       # "cosine" | "parabolic" | "womersley"
    position_m: tuple
    normal: tuple
    # cosine pressure
    mean_mmHg: float = 0.0
    amplitude_mmHg: float = 0.0
    phase_rad: float = 0.0
    period_s: float = 0.0
    # velocity iolets
    radius_m: float = 0.0
    maximum_ms: float = 0.0
    pressure_gradient_amplitude_mmHgm: float = 0.0
    womersley_number: float = 0.0


@dataclass
class SimConfig:
    steps: int
    step_length_s: float
    voxel_size_m: float
    origin_m: tuple = (0.0, 0.0, 0.0)
    fluid_density: float = DEFAULT_FLUID_DENSITY
    fluid_viscosity: float = DEFAULT_FLUID_VISCOSITY
    reference_pressure_mmHg: float = 0.0
    warmup_steps: int = 0
    stress_type: int = 1
    datafile: str = ""
    initial_pressure_mmHg: float | None = 0.0
    checkpoint: str | None = None
    initial_time: int | None = None   # <initialconditions><time>
    inlets: list = field(default_factory=list)
    outlets: list = field(default_factory=list)
    path: str = ""                  # the file read, for relative paths

    @property
    def total_steps(self) -> int:
        """Steps including the extra warm-up (SimConfig.cc:130-138)."""
        return self.steps + self.warmup_steps

    def datafile_path(self) -> str:
        base = os.path.dirname(os.path.abspath(self.path)) if self.path else os.getcwd()
        return os.path.normpath(os.path.join(base, self.datafile))


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
def _child(el, name, required=True):
    c = el.find(name)
    if c is None and required:
        raise ConfigError(f"missing element <{name}> in <{el.tag}>")
    return c


def _value(el, name, units, cast=float, required=True, default=None):
    c = _child(el, name, required)
    if c is None:
        return default
    got = c.get("units")
    if got != units:
        raise ConfigError(f"<{name}> has units {got!r}, expected {units!r}")
    v = c.get("value")
    if v is None:
        raise ConfigError(f"<{name}> has no value")
    return cast(v)


def _vector(text: str) -> tuple:
    t = text.strip()
    if not (t.startswith("(") and t.endswith(")")):
        raise ConfigError(f"bad vector {text!r}")
    parts = [float(p) for p in t[1:-1].split(",")]
    if len(parts) != 3:
        raise ConfigError(f"bad vector {text!r}")
    return tuple(parts)


def _iolet(el) -> IoletConfig:
    cond = _child(el, "condition")
    ctype, sub = cond.get("type"), cond.get("subtype")
    pos = _value(el, "position", "m", _vector)
    nrm = _value(el, "normal", "dimensionless", _vector)
    if ctype == "pressure" and sub == "cosine":
        return IoletConfig("cosine", pos, nrm,
                           mean_mmHg=_value(cond, "mean", "mmHg"),
                           amplitude_mmHg=_value(cond, "amplitude", "mmHg"),
                           phase_rad=_value(cond, "phase", "rad"),
                           period_s=_value(cond, "period", "s"))
    if ctype == "velocity" and sub == "parabolic":
        return IoletConfig("parabolic", pos, nrm, radius_m=_value(cond, "radius", "m"),
                           maximum_ms=_value(cond, "maximum", "m/s"))
    if ctype == "velocity" and sub == "womersley":
        return IoletConfig("womersley", pos, nrm, radius_m=_value(cond, "radius", "m"),
                           pressure_gradient_amplitude_mmHgm=_value(
                               cond, "pressure_gradient_amplitude", "mmHg/m"),
                           period_s=_value(cond, "period", "s"),
                           womersley_number=_value(cond, "womersley_number", "dimensionless"))
    raise ConfigError(f"iolet condition {ctype}/{sub} is not supported by the GPU timestep")


def read_xml(path: str) -> SimConfig:
    root = ET.parse(path).getroot()
    if root.tag != "hemelbsettings":
        raise ConfigError("root element must be <hemelbsettings>")
    if root.get("version") != "5":
        raise ConfigError(f"unsupported settings version {root.get('version')!r}")
    sim = _child(root, "simulation")
    stress = _child(sim, "stresstype")
    cfg = SimConfig(
        steps=_value(sim, "steps", "lattice", int),
        step_length_s=_value(sim, "step_length", "s"),
        voxel_size_m=_value(sim, "voxel_size", "m"),
        origin_m=_value(sim, "origin", "m", _vector),
        fluid_density=_value(sim, "fluid_density", "kg/m3", required=False,
                             default=DEFAULT_FLUID_DENSITY),
        fluid_viscosity=_value(sim, "fluid_viscosity", "Pa.s", required=False,
                               default=DEFAULT_FLUID_VISCOSITY),
        reference_pressure_mmHg=_value(sim, "reference_pressure", "mmHg", required=False,
                                       default=0.0),
        warmup_steps=_value(sim, "extra_warmup_steps", "lattice", int, required=False,
                            default=0),
        stress_type=int(stress.get("value")),
        path=path)
    geo = _child(root, "geometry")
    cfg.datafile = _child(geo, "datafile").get("path")
    ic = root.find("initialconditions")
    if ic is not None:
        # exactly one of <pressure> and <checkpoint> (SimConfig.cc:630-668)
        cfg.initial_time = _value(ic, "time", "lattice", int, required=False)
        pr = ic.find("pressure")
        cp = ic.find("checkpoint")
        if pr is not None and cp is not None:
            raise ConfigError("<initialconditions> holds both <pressure> and <checkpoint>")
        if pr is not None:
            cfg.initial_pressure_mmHg = _value(pr, "uniform", "mmHg")
        elif cp is not None:
            cfg.initial_pressure_mmHg = None
            cfg.checkpoint = cp.get("file")
            if cfg.checkpoint is None:
                raise ConfigError("<checkpoint> has no file attribute")
        else:
            raise ConfigError("<initialconditions> holds no known initial condition")
    for tag, lst in (("inlets", cfg.inlets), ("outlets", cfg.outlets)):
        grp = root.find(tag)
        if grp is not None:
            for el in grp.findall(tag[:-1]):
                lst.append(_iolet(el))
    return cfg


# ---------------------------------------------------------------------------
# writing
# ---------------------------------------------------------------------------
def _fmt(v) -> str:
    return repr(float(v))


def _vec(v) -> str:
    return "(" + ",".join(_fmt(x) for x in v) + ")"


def _put(parent, name, value, units):
    return ET.SubElement(parent, name, {"value": value, "units": units})


def write_xml(cfg: SimConfig, path: str) -> None:
    """The settings file of `cfg` with the reference's element names."""
    root = ET.Element("hemelbsettings", {"version": "5"})
    sim = ET.SubElement(root, "simulation")
    ET.SubElement(sim, "stresstype", {"value": str(int(cfg.stress_type))})
    _put(sim, "steps", str(int(cfg.steps)), "lattice")
    _put(sim, "step_length", _fmt(cfg.step_length_s), "s")
    if cfg.warmup_steps:
        _put(sim, "extra_warmup_steps", str(int(cfg.warmup_steps)), "lattice")
    _put(sim, "voxel_size", _fmt(cfg.voxel_size_m), "m")
    _put(sim, "origin", _vec(cfg.origin_m), "m")
    _put(sim, "fluid_density", _fmt(cfg.fluid_density), "kg/m3")
    _put(sim, "fluid_viscosity", _fmt(cfg.fluid_viscosity), "Pa.s")
    if cfg.reference_pressure_mmHg:
        _put(sim, "reference_pressure", _fmt(cfg.reference_pressure_mmHg), "mmHg")
    geo = ET.SubElement(root, "geometry")
    ET.SubElement(geo, "datafile", {"path": cfg.datafile})
    ic = ET.SubElement(root, "initialconditions")
    if cfg.initial_time is not None:
        _put(ic, "time", str(int(cfg.initial_time)), "lattice")
    if cfg.checkpoint:
        ET.SubElement(ic, "checkpoint", {"file": cfg.checkpoint})
    else:
        _put(ET.SubElement(ic, "pressure"), "uniform", _fmt(cfg.initial_pressure_mmHg or 0.0),
             "mmHg")
    for tag, lst in (("inlets", cfg.inlets), ("outlets", cfg.outlets)):
        grp = ET.SubElement(root, tag)
        for io in lst:
            el = ET.SubElement(grp, tag[:-1])
            if io.kind == "cosine":
                c = ET.SubElement(el, "condition", {"type": "pressure", "subtype": "cosine"})
                _put(c, "amplitude", _fmt(io.amplitude_mmHg), "mmHg")
                _put(c, "mean", _fmt(io.mean_mmHg), "mmHg")
                _put(c, "phase", _fmt(io.phase_rad), "rad")
                _put(c, "period", _fmt(io.period_s), "s")
            elif io.kind == "parabolic":
                c = ET.SubElement(el, "condition", {"type": "velocity", "subtype": "parabolic"})
                _put(c, "radius", _fmt(io.radius_m), "m")
                _put(c, "maximum", _fmt(io.maximum_ms), "m/s")
            elif io.kind == "womersley":
                c = ET.SubElement(el, "condition", {"type": "velocity", "subtype": "womersley"})
                _put(c, "pressure_gradient_amplitude",
                     _fmt(io.pressure_gradient_amplitude_mmHgm), "mmHg/m")
                _put(c, "period", _fmt(io.period_s), "s")
                _put(c, "womersley_number", _fmt(io.womersley_number), "dimensionless")
                _put(c, "radius", _fmt(io.radius_m), "m")
            else:
                raise ConfigError(f"unknown iolet kind {io.kind!r}")
            _put(el, "normal", _vec(io.normal), "dimensionless")
            _put(el, "position", _vec(io.position_m), "m")
    ET.indent(root)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


# ---------------------------------------------------------------------------
# building (SimBuilder)
# ---------------------------------------------------------------------------
class SimBuilder:
    """Lattice-unit objects of a configuration (Code/configuration/SimBuilder.cc)."""

    def __init__(self, cfg: SimConfig):
        self.cfg = cfg
        self.units = L.UnitConverter(cfg.step_length_s, cfg.voxel_size_m, cfg.fluid_density,
                                     cfg.reference_pressure_mmHg)
        self.params = L.LbmParameters(cfg.step_length_s, cfg.voxel_size_m, cfg.fluid_density,
                                      cfg.fluid_viscosity)

    def position(self, x_m) -> tuple:
        """UnitConverter::ConvertPositionToLatticeUnits (UnitConverter.cc:69-72)."""
        o, dx = self.cfg.origin_m, self.cfg.voxel_size_m
        return tuple((x_m[a] - o[a]) / dx for a in range(3))

    def iolet(self, io: IoletConfig):
        u = self.units
        pos = self.position(io.position_m)
        warm = int(self.cfg.warmup_steps)
        if io.kind == "cosine":
            return L.InOutLetCosine(
                pos, tuple(io.normal),
                density_mean=u.ConvertPressureToLatticeUnits(io.mean_mmHg) / L.Cs2,
                density_amp=u.ConvertPressureDifferenceToLatticeUnits(io.amplitude_mmHg) / L.Cs2,
                phase=io.phase_rad, period=u.ConvertTimeToLatticeUnits(io.period_s),
                warmup=warm)
        if io.kind == "parabolic":
            return L.InOutLetParabolicVelocity(
                pos, tuple(io.normal), radius=u.ConvertDistanceToLatticeUnits(io.radius_m),
                max_speed=u.ConvertSpeedToLatticeUnits(io.maximum_ms), warmup=warm)
        if io.kind == "womersley":
            return L.InOutLetWomersleyVelocity(
                pos, tuple(io.normal), radius=u.ConvertDistanceToLatticeUnits(io.radius_m),
                pressure_gradient_amplitude=u.ConvertPressureGradientToLatticeUnits(
                    io.pressure_gradient_amplitude_mmHgm),
                period=u.ConvertTimeToLatticeUnits(io.period_s),
                womersley_number=io.womersley_number)
        raise ConfigError(f"unknown iolet kind {io.kind!r}")

    def inlets(self) -> list:
        return [self.iolet(io) for io in self.cfg.inlets]

    def outlets(self) -> list:
        return [self.iolet(io) for io in self.cfg.outlets]

    def initial_density(self) -> float:
        """EquilibriumInitialCondition density: the uniform pressure in
        lattice units over Cs2 (SimBuilder.cc:64-67)."""
        if self.cfg.initial_pressure_mmHg is None:
            raise ConfigError("the configuration starts from a checkpoint")
        return self.units.ConvertPressureToLatticeUnits(self.cfg.initial_pressure_mmHg) / L.Cs2

    def build(self, traits: L.Traits, storage: str = "fp64", layout: str = "blocks",
              device: int = 0, geometry=None) -> L.LBM:
        """The solver of the configuration: its .gmy projected on the traits'
        lattice, the iolets in lattice units and the initial condition
        (uniform-pressure equilibrium or the checkpoint file)."""
        from .geometry import Geometry
        g = geometry if geometry is not None else Geometry.read_gmy(self.cfg.datafile_path())
        dom = g.project(traits.Lattice.NUMVECTORS)
        rho0 = (self.initial_density() if self.cfg.initial_pressure_mmHg is not None
                else 1.0)
        lbm = L.LBM(traits, dom, self.params, self.inlets(), self.outlets(), storage=storage,
                    rho0=rho0, device=device, layout=layout)
        if self.cfg.checkpoint:
            from . import extraction as X
            base = os.path.dirname(os.path.abspath(self.cfg.path)) if self.cfg.path else ""
            X.load_checkpoint(lbm, os.path.join(base, self.cfg.checkpoint))
        else:
            lbm.SetInitialConditions(rho0)
        if self.cfg.initial_time is not None:
            lbm.SetTimeStep(int(self.cfg.initial_time))
        return lbm


def config_for_cylinder_workload(steps: int, step_length_s: float, voxel_size_m: float,
                                 datafile: str, inlet_planes, outlet_planes,
                                 mean_mmHg: float, amplitude_mmHg: float, period_s: float,
                                 outlet_mmHg: float, initial_mmHg: float,
                                 viscosity: float = DEFAULT_FLUID_VISCOSITY) -> SimConfig:
    """Settings for a generated geometry whose iolet planes are given in
    lattice units (origin 0): cosine pressure inlets and constant-pressure
    outlets, as the small-cylinder workload uses."""
    dx = voxel_size_m
    to_m = lambda p: tuple(float(v) * dx for v in p)  # noqa: E731
    ins = [IoletConfig("cosine", to_m(p.position), tuple(map(float, p.normal)),
                       mean_mmHg=mean_mmHg, amplitude_mmHg=amplitude_mmHg, period_s=period_s)
           for p in inlet_planes]
    outs = [IoletConfig("cosine", to_m(p.position), tuple(map(float, p.normal)),
                        mean_mmHg=outlet_mmHg, amplitude_mmHg=0.0, period_s=period_s)
            for p in outlet_planes]
    return SimConfig(steps=steps, step_length_s=step_length_s, voxel_size_m=voxel_size_m,
                     fluid_viscosity=viscosity, datafile=datafile,
                     initial_pressure_mmHg=initial_mmHg, inlets=ins, outlets=outs)


__all__ = ["SimConfig", "IoletConfig", "ConfigError", "read_xml", "write_xml", "SimBuilder",
           "config_for_cylinder_workload"]
