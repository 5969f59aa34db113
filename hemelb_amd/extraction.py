"""Property extraction and the periodic-state monitor.

The fields are those the reference's property output writes
(Code/extraction/LbDataSourceIterator.cc:33-47): pressure = density * Cs2,
velocity = momentum / density, wall shear stress magnitude at wall sites
(Code/lb/streamers/Common.h:37-50), all from fOld in binary64 and converted to
physical units with util::UnitConverter when asked.

The periodic-state monitor generalises StabilityTester's steady-state check
(Code/lb/StabilityTester.h:160-183) to pulsatile flow: the velocity field is
sampled at the same phase of consecutive cycles and the run is periodic once
max_sites |u(t) - u(t - T)| / u_ref drops below the tolerance.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .lb import LBM, Cs2, UnitConverter

NO_VALUE = np.finfo(np.float64).max   # constants.h:47


@dataclass
class Fields:
    density: np.ndarray            # [n]
    velocity: np.ndarray           # [n, 3]
    pressure: np.ndarray           # [n], lattice units (density * Cs2)
    wall_shear_stress: np.ndarray  # [n], NO_VALUE at sites without wall links
    wall: np.ndarray               # [n] bool, sites with wall links

    def physical(self, units: UnitConverter) -> "Fields":
        wss = np.where(self.wall, units.ConvertStressToPhysicalUnits(
            np.where(self.wall, self.wall_shear_stress, 0.0)), NO_VALUE)
        return Fields(self.density, units.ConvertVelocityToPhysicalUnits(self.velocity),
                      units.ConvertPressureToPhysicalUnits(self.pressure), wss, self.wall)


def extract(lbm: LBM, wall_shear_stress: bool = True) -> Fields:
    rho, u = lbm.macroscopic()
    wall = lbm.domain.wall_mask != 0
    wss = lbm.wall_shear_stress() if wall_shear_stress else np.full(rho.shape, NO_VALUE)
    return Fields(rho, u, rho * Cs2, wss, wall)


def _norms(diff: np.ndarray, ref: np.ndarray) -> dict:
    """L2 and Linf of `diff`, absolute and relative to the reference field.
    Vector fields ([n, 3]) use the per-site Euclidean norm."""
    if diff.ndim == 2:
        d, r = np.sqrt((diff ** 2).sum(1)), np.sqrt((ref ** 2).sum(1))
    else:
        d, r = np.abs(diff), np.abs(ref)
    l2, l2r = float(np.sqrt((d ** 2).sum())), float(np.sqrt((r ** 2).sum()))
    li, lir = float(d.max(initial=0.0)), float(r.max(initial=0.0))
    return {"l2": l2 / max(np.sqrt(d.size), 1.0), "linf": li,
            "l2_rel": l2 / l2r if l2r > 0 else float("nan"),
            "linf_rel": li / lir if lir > 0 else float("nan")}


def field_errors(test: Fields, ref: Fields) -> dict:
    """Errors of `test` against `ref`.

    velocity: per-site vector error, relative to the reference's L2 / max |u|;
    pressure: relative to the reference's pressure range (max - min), since
    the absolute pressure (~Cs2) would hide every error;
    wall_shear_stress: over the wall sites, relative like velocity.
    `l2` entries are root-mean-square values."""
    out = {"velocity": _norms(test.velocity - ref.velocity, ref.velocity)}
    dp = test.pressure - ref.pressure
    p = _norms(dp, ref.pressure)
    span = float(ref.pressure.max() - ref.pressure.min())
    p["l2_rel"] = p["l2"] / span if span > 0 else float("nan")
    p["linf_rel"] = p["linf"] / span if span > 0 else float("nan")
    out["pressure"] = p
    w = ref.wall
    if w.any() and (ref.wall_shear_stress[w] != NO_VALUE).all():
        out["wall_shear_stress"] = _norms(test.wall_shear_stress[w] - ref.wall_shear_stress[w],
                                          ref.wall_shear_stress[w])
    return out


@dataclass
class PeriodicStateMonitor:
    """Runs an LBM cycle by cycle until the velocity at the same phase of two
    consecutive cycles differs by less than `tolerance` (relative to
    `reference_speed`), or `max_cycles` have run."""
    lbm: LBM
    period: float                  # steps per cycle
    reference_speed: float
    tolerance: float = 1e-4
    history: list = field(default_factory=list)

    def run(self, max_cycles: int) -> dict:
        t0 = time.perf_counter()
        start = self.lbm.GetTimeStep()
        self.lbm.snapshot_velocity()
        converged = False
        for k in range(1, max_cycles + 1):
            target = start + int(round(k * self.period))
            self.lbm.step(target - self.lbm.GetTimeStep())
            change = self.lbm.velocity_change(self.reference_speed)
            self.history.append(change)
            self.lbm.snapshot_velocity()
            if change < self.tolerance:
                converged = True
                break
        self.lbm.synchronize()
        return {"converged": converged, "cycles": len(self.history),
                "steps": self.lbm.GetTimeStep() - start,
                "seconds": time.perf_counter() - t0,
                "last_change": self.history[-1] if self.history else float("nan")}


def cycles_to_periodic(lbm: LBM, period: int, tolerance: float = 1e-4,
                       max_cycles: int = 12, stop_on_plateau: bool = False):
    """Runs `lbm` one inlet cycle at a time until the cycle-to-cycle change
    max |u(t) - u(t - T)| / max |u| drops below `tolerance`, or, with
    stop_on_plateau, until a cycle no longer halves it (the rounding floor
    of a reduced-precision form).  Returns the cycles run when it stopped
    for either reason (None after max_cycles) and every cycle's change."""
    lbm.snapshot_velocity()
    history, cycles, done = [], 0, False
    while cycles < max_cycles and not done:
        lbm.step(int(period))
        cycles += 1
        umax = lbm.monitor()["velocity_max"]
        history.append(lbm.velocity_change(max(umax, 1e-30)))
        lbm.snapshot_velocity()
        done = history[-1] < tolerance or (stop_on_plateau and len(history) > 1 and
                                           history[-1] > 0.5 * history[-2])
    return (cycles if done else None), history


def periodic_state_staged(stages: list, period: int, tolerance: float = 1e-4,
                          stage_tolerance: float = 1e-2, max_cycles: int = 12) -> dict:
    """The periodic state of the last solver of `stages` with the transient
    run in the earlier, faster (reduced-precision) ones, all of the same
    domain and iolets.

    Each earlier stage runs from its current state until its cycle-to-cycle
    change is below `stage_tolerance` or stops halving.  Its distributions,
    exact in binary64, and its time step then start the next stage.  The
    last stage runs until its own change is below `tolerance`; it ends in
    its own periodic state (to `tolerance`) at a whole number of cycles from
    the first stage's start.  A middle stage absorbs the error that the
    handover from a coarser form brings into the first cycle of the next,
    so that the last stage can meet the tolerance in one cycle.  Returns the
    cycles, changes and wall times of every stage."""
    for lbm in stages:
        lbm.synchronize()
    t0 = time.perf_counter()
    out, handover = [], 0.0
    for k, lbm in enumerate(stages):
        last = k == len(stages) - 1
        if k:
            t = time.perf_counter()
            lbm.set_distributions(stages[k - 1].get_distributions())
            lbm.SetTimeStep(stages[k - 1].GetTimeStep())
            lbm.synchronize()
            handover += time.perf_counter() - t
        t = time.perf_counter()
        cycles, history = cycles_to_periodic(lbm, period, tolerance if last else stage_tolerance,
                                             max_cycles, stop_on_plateau=not last)
        lbm.synchronize()
        out.append({"cycles": cycles, "changes": history, "seconds": time.perf_counter() - t})
    return {"stages": out, "handover_seconds": handover, "seconds": time.perf_counter() - t0,
            "converged": out[-1]["cycles"] is not None}


# ---------------------------------------------------------------------------
# Extraction (.xtr) files and checkpoints
# ---------------------------------------------------------------------------
TYPES = {"float": 0, "double": 1, "int32": 2, "uint32": 3, "int64": 4, "uint64": 5}


@dataclass
class OutputField:
    """One field of an extraction file (extraction::OutputField)."""
    name: str
    type: str = "float"
    n_elements: int = 1
    offsets: tuple = ()


class ExtractionWriter:
    """An .xtr v5 file and its .off offset file for one rank
    (extraction::LocalPropertyOutput, Code/extraction/LocalPropertyOutput.cc).
    `append` writes one record: the time step then, per site, its coordinates
    and each field's values."""

    def __init__(self, path: str, voxel_size_m: float, origin_m, coords: np.ndarray,
                 fields: list):
        import ctypes as C

        from . import _native as N
        self._N = N
        self.coords = np.ascontiguousarray(coords, np.int32)
        self.fields = list(fields)
        self._offs = [np.ascontiguousarray(f.offsets, np.float64) for f in self.fields]
        specs = (N.hlbx_field * max(1, len(self.fields)))()
        for k, f in enumerate(self.fields):
            specs[k].name = f.name.encode()
            specs[k].type = TYPES[f.type]
            specs[k].n_elements = int(f.n_elements)
            specs[k].n_offsets = len(f.offsets)
            specs[k].offsets = self._offs[k].ctypes.data_as(C.POINTER(C.c_double))
        origin = np.ascontiguousarray(origin_m, np.float64)
        h = C.c_void_p()
        N.check(N.lib().hlbx_create(path.encode(), float(voxel_size_m),
                                    origin.ctypes.data_as(C.c_void_p), self.coords.shape[0],
                                    self.coords.ctypes.data_as(C.c_void_p), len(self.fields),
                                    specs, C.byref(h)), io=True)
        self._h = h

    def append(self, time_step: int, arrays: list) -> None:
        import ctypes as C
        n = self.coords.shape[0]
        data = [np.ascontiguousarray(a, np.float64).reshape(n, f.n_elements)
                for a, f in zip(arrays, self.fields)]
        ptrs = (C.c_void_p * max(1, len(data)))(*[d.ctypes.data for d in data])
        self._N.check(self._N.lib().hlbx_append(self._h, int(time_step), ptrs), io=True)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._N.check(self._N.lib().hlbx_close(self._h), io=True)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def checkpoint_writer(lbm: LBM, path: str, voxel_size_m: float = 1.0,
                      origin_m=(0.0, 0.0, 0.0)) -> ExtractionWriter:
    """A distribution checkpoint file: one binary64 field of Q values per
    site named "distributions" (SimConfig.cc:304-326)."""
    return ExtractionWriter(path, voxel_size_m, origin_m, lbm.domain.coords,
                            [OutputField("distributions", "double", lbm.Q)])


def write_checkpoint(writer: ExtractionWriter, lbm: LBM) -> int:
    """Appends the record of the current time step: fOld of the step about to
    run, as PropertyActor::EndIteration writes it (PropertyActor.cc:77-82)."""
    t = lbm.GetTimeStep()
    writer.append(t, [lbm.get_distributions()])
    return t


def load_checkpoint(lbm: LBM, path: str, time_step: int | None = None) -> int:
    """CheckpointInitialCondition (InitialCondition.cc:15-22,
    LocalDistributionInput.cc:38-170): the record of `time_step` (default:
    the last one) becomes fOld and the simulation time."""
    import ctypes as C

    from . import _native as N
    f = np.empty((lbm.domain.n, lbm.Q), np.float64)
    t = C.c_uint64()
    N.check(N.lib().hlbx_read_distributions(
        path.encode(), lbm.domain.n, lbm.domain.coords.ctypes.data_as(C.c_void_p), lbm.Q,
        -1 if time_step is None else int(time_step), f.ctypes.data_as(C.c_void_p),
        C.byref(t)), io=True)
    lbm.set_distributions(f)
    lbm.SetTimeStep(int(t.value))
    return int(t.value)


def property_writer(lbm: LBM, path: str, units: UnitConverter, origin_m=(0.0, 0.0, 0.0),
                    reference_pressure_mmHg: float | None = None) -> ExtractionWriter:
    """Pressure (mmHg, stored as float minus the reference pressure), velocity
    (m/s) and wall shear stress (Pa) as float fields, the reference's default
    property output types (SimConfig.cc:457).  The pressure field's offset is
    the reference pressure, by default the one `units` converts with (both
    come from the fluid's reference pressure in the reference: SimConfig.cc:465
    and UnitConverter.cc:33), so that stored value + offset is the pressure
    write_properties converted."""
    ref = units.ref if reference_pressure_mmHg is None else float(reference_pressure_mmHg)
    return ExtractionWriter(path, units.dx, origin_m, lbm.domain.coords, [
        OutputField("pressure", "float", 1, (ref,)),
        OutputField("velocity", "float", 3),
        OutputField("shearstress", "float", 1)])


def write_properties(writer: ExtractionWriter, lbm: LBM, units: UnitConverter) -> int:
    f = extract(lbm).physical(units)
    t = lbm.GetTimeStep()
    writer.append(t, [f.pressure, f.velocity, f.wall_shear_stress])
    return t
