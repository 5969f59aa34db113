"""Host-side mirror of the reference's ``hemelb::lb`` API for the timestep.

Names and argument meanings follow Code/lb: ``Traits`` (Code/Traits.h:17-39)
selects lattice, kernel, wall and iolet streamers; ``LbmParameters``
(Code/lb/LbmParameters.h:32-167) derives tau and omega; the iolet classes
mirror Code/lb/iolets/InOutLet*.h; ``LBM`` is the driver that
``SimulationMaster`` steps (Code/lb/lb.h:30-142).  All computation happens in
``libhemelb_amd.so`` on the GPU; this module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .geometry import Domain


# ---------------------------------------------------------------------------
# Lattices (Code/lb/lattices/D3Q15.h:15-56, D3Q19.h:14-49)
# ---------------------------------------------------------------------------
class _Lattice:
    NUMVECTORS: int
    VECTORS: np.ndarray
    EQMWEIGHTS: np.ndarray

    @classmethod
    def inverse(cls) -> np.ndarray:
        """INVERSEDIRECTIONS (Lattice.h:49-69)."""
        inv = np.full(cls.NUMVECTORS, -1, np.int32)
        for i in range(cls.NUMVECTORS):
            for j in range(i, cls.NUMVECTORS):
                if (cls.VECTORS[i] == -cls.VECTORS[j]).all():
                    inv[i], inv[j] = j, i
        return inv


class D3Q15(_Lattice):
    NUMVECTORS = 15
    VECTORS = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                        (0, 0, 1), (0, 0, -1), (1, 1, 1), (-1, -1, -1), (1, 1, -1),
                        (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1)],
                       np.int32)
    EQMWEIGHTS = np.array([2.0 / 9.0] + [1.0 / 9.0] * 6 + [1.0 / 72.0] * 8)


class D3Q19(_Lattice):
    NUMVECTORS = 19
    VECTORS = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                        (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
                        (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
                        (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], np.int32)
    EQMWEIGHTS = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)


LATTICES = {15: D3Q15, 19: D3Q19}


# Collision kernels and streamers: tags that select the compiled variants.
class LBGK:  # Code/lb/kernels/LBGK.h
    CODE = N.HLB_LBGK


class MRT:   # Code/lb/kernels/MRT.h with the d'Humieres basis of the lattice
    CODE = N.HLB_MRT


class SimpleBounceBack:           # Code/lb/streamers/SimpleBounceBack.h
    CODE = N.HLB_WALL_SBB


class BouzidiFirdaousLallemand:   # Code/lb/streamers/BouzidiFirdaousLallemand.h
    CODE = N.HLB_WALL_BFL


class NashZerothOrderPressure:    # Code/lb/streamers/NashZerothOrderPressure.h
    CODE = N.HLB_IOLET_NASH


class LaddIolet:                  # Code/lb/streamers/LaddIolet.h
    CODE = N.HLB_IOLET_LADD


@dataclass
class Traits:
    """Compile-time choices of the reference build (Code/Traits.h:17-39,
    CMake/HemeLbOptions.cmake HEMELB_LATTICE/KERNEL/WALL_BOUNDARY/
    INLET_BOUNDARY/OUTLET_BOUNDARY); defaults are the reference defaults."""
    Lattice: type = D3Q15
    Kernel: type = LBGK
    WallStreamer: type = SimpleBounceBack
    InletStreamer: type = NashZerothOrderPressure
    OutletStreamer: type = NashZerothOrderPressure


# ---------------------------------------------------------------------------
# Parameters and units
# ---------------------------------------------------------------------------
Cs2 = 1.0 / 3.0                 # constants.h:41
mmHg_TO_PASCAL = 133.3223874    # constants.h:20


class LbmParameters:
    """tau = 0.5 + (dt*eta/rho)/(Cs2*dx^2), omega = -1/tau (LbmParameters.h:158-163)."""

    def __init__(self, time_step_s=1.0, voxel_size_m=1.0, fluid_density_kgm3=1000.0,
                 viscosity_pas=0.004, tau=None):
        self.time_step = time_step_s
        self.voxel_size = voxel_size_m
        self.fluid_density = fluid_density_kgm3
        self.eta = viscosity_pas
        self.tau = (N.lib().hlb_tau_from_physical(time_step_s, voxel_size_m,
                                                  fluid_density_kgm3, viscosity_pas)
                    if tau is None else float(tau))

    def GetTau(self) -> float:
        return self.tau

    def GetOmega(self) -> float:
        return -1.0 / self.tau


class UnitConverter:
    """util::UnitConverter pressure conversions (Code/util/UnitConverter.cc:14-48)."""

    def __init__(self, time_step_s, voxel_size_m, fluid_density_kgm3=1000.0,
                 reference_pressure_mmHg=0.0):
        self.dt, self.dx, self.rho = time_step_s, voxel_size_m, fluid_density_kgm3
        self.ref = reference_pressure_mmHg

    def ConvertPressureToLatticeUnits(self, p_mmhg):
        return N.lib().hlb_pressure_to_lattice(p_mmhg, self.ref, self.dt, self.dx, self.rho)

    def ConvertPressureDifferenceToLatticeUnits(self, dp_mmhg):
        return N.lib().hlb_pressure_difference_to_lattice(dp_mmhg, self.dt, self.dx, self.rho)

    def ConvertPressureGradientToLatticeUnits(self, pg):
        return N.lib().hlb_pressure_gradient_to_lattice(pg, self.dt, self.dx, self.rho)

    def ConvertTimeToLatticeUnits(self, t_s):
        return t_s / self.dt

    def ConvertDistanceToLatticeUnits(self, x_m):
        return x_m / self.dx

    def ConvertSpeedToLatticeUnits(self, v):
        return v / (self.dx / self.dt)

    # Output conversions (UnitConverter.cc:31-46, UnitConverter.h:58-120),
    # vectorised over numpy arrays.
    @property
    def lattice_pressure(self):
        lattice_mass = self.rho * self.dx * self.dx * self.dx
        return lattice_mass / (self.dx * self.dt * self.dt)

    def ConvertPressureToPhysicalUnits(self, p_lattice):
        return self.ref + (p_lattice - Cs2) * self.lattice_pressure / mmHg_TO_PASCAL

    def ConvertVelocityToPhysicalUnits(self, u):
        return u * (self.dx / self.dt)

    def ConvertStressToPhysicalUnits(self, s):
        return s * self.lattice_pressure


# ---------------------------------------------------------------------------
# Iolets (Code/lb/iolets)
# ---------------------------------------------------------------------------
@dataclass
class InOutLet:
    position: tuple = (0.0, 0.0, 0.0)   # lattice units
    normal: tuple = (0.0, 0.0, 1.0)     # normalised by the library

    def _fill(self, s: N.hlb_iolet) -> None:
        s.position[:] = list(map(float, self.position))
        s.normal[:] = list(map(float, self.normal))


@dataclass
class InOutLetCosine(InOutLet):
    """Density mean + amp*cos(2 pi t/period + phase) (InOutLetCosine.cc:26-43)."""
    density_mean: float = 1.0
    density_amp: float = 0.0
    phase: float = 0.0
    period: float = 1.0
    warmup: int = 0

    def to_c(self) -> N.hlb_iolet:
        s = N.hlb_iolet()
        s.kind = N.HLB_IOLET_COSINE
        self._fill(s)
        s.density_mean, s.density_amp = self.density_mean, self.density_amp
        s.phase, s.period, s.warmup = self.phase, self.period, int(self.warmup)
        return s

    def GetDensity(self, time_step: int, minimum_simulation_density: float = 0.0) -> float:
        w = 2.0 * math.pi / self.period
        target = self.density_mean + self.density_amp * math.cos(w * time_step + self.phase)
        if time_step >= self.warmup:
            return target
        f = time_step / self.warmup
        return f * target + (1.0 - f) * minimum_simulation_density


@dataclass
class InOutLetParabolicVelocity(InOutLet):
    radius: float = 1.0
    max_speed: float = 0.0
    warmup: int = 0

    def to_c(self) -> N.hlb_iolet:
        s = N.hlb_iolet()
        s.kind = N.HLB_IOLET_PARABOLIC
        self._fill(s)
        s.radius, s.max_speed, s.warmup = self.radius, self.max_speed, int(self.warmup)
        return s


@dataclass
class InOutLetWomersleyVelocity(InOutLet):
    radius: float = 1.0
    pressure_gradient_amplitude: float = 0.0
    period: float = 1.0
    womersley_number: float = 1.0

    def to_c(self) -> N.hlb_iolet:
        s = N.hlb_iolet()
        s.kind = N.HLB_IOLET_WOMERSLEY
        self._fill(s)
        s.radius = self.radius
        s.pressure_gradient_amplitude = self.pressure_gradient_amplitude
        s.period = self.period
        s.womersley_number = self.womersley_number
        return s


def _iolet_array(lst):
    arr = (N.hlb_iolet * max(1, len(lst)))()
    for i, io in enumerate(lst):
        arr[i] = io.to_c()
    return arr


STORAGE = {"fp64": N.HLB_STORE_FP64, "fp32": N.HLB_STORE_FP32, "fp16": N.HLB_STORE_FP16}
LAYOUT = {"blocks": N.HLB_LAYOUT_BLOCKS, "pull_table": N.HLB_LAYOUT_PULL_TABLE,
          "compact_blocks": N.HLB_LAYOUT_COMPACT_BLOCKS}


class LBM:
    """The LB actor of one rank on one GPU (Code/lb/lb.h:30-142)."""

    def __init__(self, traits: Traits, domain: Domain, params: LbmParameters,
                 inlets=(), outlets=(), storage: str = "fp64", rho0: float = 1.0,
                 device: int = 0, layout: str = "blocks", instances=None,
                 instance_tau=None, ghosts: Domain | None = None):
        """`instances` ([n] int, optional) makes an ensemble: independent
        vessels in one solver, instance k relaxed with instance_tau[k].
        `ghosts` (a Domain of the neighbouring subdomains' sites this one
        pulls from) makes a subdomain whose ghost slots a HaloPlan fills."""
        if domain.Q != traits.Lattice.NUMVECTORS:
            raise ValueError("domain was projected for a different lattice")
        self.traits, self.domain, self.params = traits, domain, params
        self.Q = domain.Q
        p = N.hlb_params()
        p.lattice = self.Q
        p.kernel = traits.Kernel.CODE
        p.wall = traits.WallStreamer.CODE
        p.inlet_bc = traits.InletStreamer.CODE
        p.outlet_bc = traits.OutletStreamer.CODE
        p.storage = STORAGE[storage]
        p.tau = params.GetTau()
        p.rho0 = rho0
        p.device = device
        p.layout = LAYOUT[layout]
        self._inlets, self._outlets = list(inlets), list(outlets)
        self._cdom = domain.as_c()
        h = C.c_void_p()
        if ghosts is not None:
            if instances is not None:
                raise ValueError("ensembles take no ghost sites")
            self._ghosts = ghosts
            self._cghost = ghosts.as_c()
            N.check(N.lib().hlb_create_subdomain(
                C.byref(p), C.byref(self._cdom), C.byref(self._cghost),
                _iolet_array(self._inlets), len(self._inlets),
                _iolet_array(self._outlets), len(self._outlets), C.byref(h)))
        elif instances is None:
            N.check(N.lib().hlb_create(C.byref(p), C.byref(self._cdom),
                                       _iolet_array(self._inlets), len(self._inlets),
                                       _iolet_array(self._outlets), len(self._outlets),
                                       C.byref(h)))
        else:
            self._inst = np.ascontiguousarray(instances, np.int32)
            self._inst_tau = np.ascontiguousarray(instance_tau, np.float64)
            N.check(N.lib().hlb_create_ensemble(
                C.byref(p), C.byref(self._cdom), self._inst.ctypes.data_as(C.c_void_p),
                len(self._inst_tau), self._inst_tau.ctypes.data_as(C.c_void_p),
                _iolet_array(self._inlets), len(self._inlets),
                _iolet_array(self._outlets), len(self._outlets), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().hlb_destroy(self._h)
            self._h = C.c_void_p()

    def close(self):
        self.__del__()

    # InitialCondition / SimulationState
    def SetInitialConditions(self, density: float, momentum=(0.0, 0.0, 0.0)) -> None:
        m = np.asarray(momentum, np.float64)
        N.check(N.lib().hlb_set_equilibrium(self._h, float(density),
                                            m.ctypes.data_as(C.c_void_p)))

    def SetTimeStep(self, t: int) -> None:
        N.check(N.lib().hlb_set_time(self._h, int(t)))

    def GetTimeStep(self) -> int:
        return int(N.lib().hlb_get_time(self._h))

    def step(self, n: int = 1) -> None:
        N.check(N.lib().hlb_step(self._h, int(n)))

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        N.check(N.lib().hlb_time_steps(self._h, int(n), C.byref(ms)))
        return ms.value

    def copy_ceiling(self, reps: int = 20) -> dict:
        """Fastest device copy of the state arrays (hlb_copy_ceiling): read +
        write GB/s, ms and bytes per copy, and the copy kernel's launch."""
        r = N.hlb_copy_result()
        N.check(N.lib().hlb_copy_ceiling(self._h, int(reps), C.byref(r)))
        return {"GBs": r.bytes_per_s / 1e9, "ms": r.ms, "bytes": r.bytes,
                "threads": r.threads, "xcd_remap": r.xcd_remap, "width": r.width,
                "order": "blocks" if r.order else "arrays"}

    # FieldData export/import
    def get_distributions(self) -> np.ndarray:
        f = np.empty((self.domain.n, self.Q), np.float64)
        N.check(N.lib().hlb_get_distributions(self._h, f.ctypes.data_as(C.c_void_p)))
        return f

    def set_distributions(self, f: np.ndarray) -> None:
        f = np.ascontiguousarray(f, np.float64)
        assert f.shape == (self.domain.n, self.Q)
        N.check(N.lib().hlb_set_distributions(self._h, f.ctypes.data_as(C.c_void_p)))

    def set_equilibrium_fields(self, density, momentum) -> None:
        """fOld = feq(density, momentum) per site (caller's order).  numpy
        arrays, or float64 device tensors (anything with ``data_ptr`` and
        ``is_cuda``) used in place."""
        n = self.domain.n
        if getattr(density, "is_cuda", False) or getattr(momentum, "is_cuda", False):
            import torch
            if not (density.is_cuda and momentum.is_cuda):
                raise ValueError("density and momentum must both be device tensors")
            if density.dtype != torch.float64 or momentum.dtype != torch.float64:
                raise ValueError("device fields must be float64")
            if tuple(density.shape) != (n,) or tuple(momentum.shape) != (n, 3):
                raise ValueError("fields must have shapes (n,) and (n, 3)")
            if not (density.is_contiguous() and momentum.is_contiguous()):
                raise ValueError("device fields must be contiguous")
            torch.cuda.current_stream().synchronize()
            N.check(N.lib().hlb_set_equilibrium_fields(self._h, density.data_ptr(),
                                                       momentum.data_ptr(), 1))
            return
        rho = np.ascontiguousarray(density, np.float64)
        m = np.ascontiguousarray(momentum, np.float64)
        if rho.shape != (n,) or m.shape != (n, 3):
            raise ValueError("fields must have shapes (n,) and (n, 3)")
        N.check(N.lib().hlb_set_equilibrium_fields(self._h, rho.ctypes.data_as(C.c_void_p),
                                                   m.ctypes.data_as(C.c_void_p), 0))

    def macroscopic(self):
        rho = np.empty(self.domain.n, np.float64)
        u = np.empty((self.domain.n, 3), np.float64)
        N.check(N.lib().hlb_get_macroscopic(self._h, rho.ctypes.data_as(C.c_void_p),
                                            u.ctypes.data_as(C.c_void_p)))
        return rho, u

    def wall_shear_stress(self) -> np.ndarray:
        """Per-site wall shear stress magnitude of fOld (lattice units);
        float max (the reference's NO_VALUE) at sites without wall links."""
        out = np.empty(self.domain.n, np.float64)
        N.check(N.lib().hlb_get_wall_shear_stress(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def stress(self) -> dict:
        """Stress tensor [n,6] (xx xy xz yy yz zz), von Mises stress, shear
        rate, traction and tangential traction [n,3] of fOld (lattice units)."""
        n = self.domain.n
        out = {"tensor": np.empty((n, 6)), "von_mises": np.empty(n),
               "shear_rate": np.empty(n), "traction": np.empty((n, 3)),
               "tangential_traction": np.empty((n, 3))}
        p = lambda k: out[k].ctypes.data_as(C.c_void_p)  # noqa: E731
        N.check(N.lib().hlb_get_stress(self._h, p("tensor"), p("von_mises"), p("shear_rate"),
                                       p("traction"), p("tangential_traction")))
        return out

    def monitor(self) -> dict:
        m = N.hlb_monitor()
        N.check(N.lib().hlb_reduce(self._h, C.byref(m)))
        return {"mass": m.mass, "density_min": m.density_min,
                "density_max": m.density_max, "velocity_max": m.velocity_max,
                "stable": bool(m.stable), "iolet_flux": list(m.iolet_flux[:m.n_iolets])}

    @property
    def instance_count(self) -> int:
        return int(N.lib().hlb_instance_count(self._h))

    def monitor_instances(self) -> list:
        k = self.instance_count
        arr = (N.hlb_instance_monitor * k)()
        N.check(N.lib().hlb_reduce_instances(self._h, arr))
        return [{"mass": m.mass, "density_min": m.density_min, "density_max": m.density_max,
                 "velocity_max": m.velocity_max, "stable": bool(m.stable), "sites": m.sites}
                for m in arr]

    def velocity_change_instances(self, reference_speed) -> np.ndarray:
        ref = np.ascontiguousarray(np.broadcast_to(reference_speed, (self.instance_count,)),
                                   np.float64)
        out = np.empty(self.instance_count)
        N.check(N.lib().hlb_velocity_change_instances(
            self._h, ref.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
        return out

    def set_launch_shape(self, threads=256, planes=1, groups=0, xcd_remap=0, split=0,
                         wide=0) -> None:
        """xcd_remap: 0 hardware workgroup order, 1 one contiguous range per
        XCD, K in 2..4096 runs of K workgroups per XCD; split: planes with
        and without boundary sites in two launches (one-pass shapes); wide:
        64-bit element addressing even for states below 4 GiB."""
        sh = N.hlb_launch_shape(threads, planes, groups, int(xcd_remap), int(split), int(wide))
        N.check(N.lib().hlb_set_launch_shape(self._h, C.byref(sh)))

    @property
    def launch_shape(self) -> dict:
        sh = N.hlb_launch_shape()
        N.check(N.lib().hlb_get_launch_shape(self._h, C.byref(sh)))
        return {"threads": sh.threads, "planes": sh.planes, "groups": sh.groups,
                "xcd_remap": sh.xcd_remap, "split": sh.split, "wide": sh.wide}

    def autotune(self, steps: int = 20) -> dict:
        """Times the candidate launch shapes on the current state (which is
        restored afterwards) and keeps the fastest."""
        sh, ms = N.hlb_launch_shape(), C.c_double()
        N.check(N.lib().hlb_autotune(self._h, int(steps), C.byref(sh), C.byref(ms)))
        return {"threads": sh.threads, "planes": sh.planes, "groups": sh.groups,
                "xcd_remap": sh.xcd_remap, "split": sh.split, "wide": sh.wide,
                "ms_per_step": ms.value}

    def snapshot_velocity(self) -> None:
        N.check(N.lib().hlb_snapshot_velocity(self._h))

    def velocity_change(self, reference_speed: float) -> float:
        v = C.c_double()
        N.check(N.lib().hlb_velocity_change(self._h, reference_speed, C.byref(v)))
        return v.value

    def halo_pack(self, pairs: np.ndarray) -> np.ndarray:
        pairs = np.ascontiguousarray(pairs, np.int64)
        out = np.empty(pairs.shape[0], np.float64)
        N.check(N.lib().hlb_halo_pack(self._h, pairs.ctypes.data_as(C.c_void_p),
                                      pairs.shape[0], out.ctypes.data_as(C.c_void_p)))
        return out

    def halo_unpack(self, pairs: np.ndarray, values: np.ndarray) -> None:
        pairs = np.ascontiguousarray(pairs, np.int64)
        values = np.ascontiguousarray(values, np.float64)
        assert values.shape[0] == pairs.shape[0]
        N.check(N.lib().hlb_halo_unpack(self._h, pairs.ctypes.data_as(C.c_void_p),
                                        pairs.shape[0], values.ctypes.data_as(C.c_void_p)))

    def device_bytes(self) -> int:
        return int(N.lib().hlb_device_bytes(self._h))

    def synchronize(self) -> None:
        N.check(N.lib().hlb_synchronize(self._h))


class HaloPlan:
    """The values `dst`'s ghost sites need from `src` every step (the
    reference's FieldData::SendAndReceive + CopyReceived between two ranks),
    moved on the device: exchange() after each step of both subdomains."""

    def __init__(self, src: LBM, dst: LBM):
        h = C.c_void_p()
        N.check(N.lib().hlb_halo_plan_create(src._h, dst._h, C.byref(h)))
        self._h, self._keep = h, (src, dst)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().hlb_halo_plan_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def count(self) -> int:
        return int(N.lib().hlb_halo_plan_count(self._h))

    def exchange(self) -> None:
        N.check(N.lib().hlb_halo_exchange(self._h))

    def last_times(self) -> tuple:
        a, b = C.c_double(), C.c_double()
        N.check(N.lib().hlb_halo_last_times(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value
