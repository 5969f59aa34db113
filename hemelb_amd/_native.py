"""ctypes bindings of ``libhemelb_amd.so`` (declarations in ``include/hemelb_amd.h``
and ``include/hemelb_geometry.h``).

The library is built in-tree by ``hemelb_amd.build``; importing this module
fails with an explicit error when it is missing, there is no other code path.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhemelb_amd.so")

# hlb_status
HLB_OK, HLB_E_INVALID, HLB_E_DEVICE, HLB_E_NOMEM, HLB_E_STATE = 0, -1, -2, -3, -4
HLB_D3Q15, HLB_D3Q19 = 15, 19
HLB_LBGK, HLB_MRT = 0, 1
HLB_WALL_SBB, HLB_WALL_BFL = 0, 1
HLB_IOLET_NASH, HLB_IOLET_LADD = 0, 1
HLB_IOLET_COSINE, HLB_IOLET_PARABOLIC, HLB_IOLET_WOMERSLEY = 0, 1, 2
HLB_SITE_FLUID, HLB_SITE_INLET, HLB_SITE_OUTLET = 1, 2, 3
HLB_STORE_FP64, HLB_STORE_FP32, HLB_STORE_FP16 = 0, 1, 2
HLB_LAYOUT_BLOCKS, HLB_LAYOUT_PULL_TABLE, HLB_LAYOUT_COMPACT_BLOCKS = 0, 1, 2
HLBG_CUT_NONE, HLBG_CUT_WALL, HLBG_CUT_INLET, HLBG_CUT_OUTLET = 0, 1, 2, 3


class hlb_params(C.Structure):
    _fields_ = [("lattice", C.c_int32), ("kernel", C.c_int32), ("wall", C.c_int32),
                ("inlet_bc", C.c_int32), ("outlet_bc", C.c_int32),
                ("storage", C.c_int32), ("tau", C.c_double), ("rho0", C.c_double),
                ("device", C.c_int32), ("layout", C.c_int32)]


class hlb_domain(C.Structure):
    _fields_ = [("n_sites", C.c_int64), ("coords", C.c_void_p),
                ("site_type", C.c_void_p), ("wall_mask", C.c_void_p),
                ("iolet_mask", C.c_void_p), ("iolet_id", C.c_void_p),
                ("cut_distance", C.c_void_p), ("wall_normal", C.c_void_p)]


class hlb_iolet(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32),
                ("position", C.c_double * 3), ("normal", C.c_double * 3),
                ("density_mean", C.c_double), ("density_amp", C.c_double),
                ("phase", C.c_double), ("period", C.c_double),
                ("warmup", C.c_int64), ("radius", C.c_double),
                ("max_speed", C.c_double),
                ("pressure_gradient_amplitude", C.c_double),
                ("womersley_number", C.c_double)]


class hlb_monitor(C.Structure):
    _fields_ = [("mass", C.c_double), ("density_min", C.c_double),
                ("density_max", C.c_double), ("velocity_max", C.c_double),
                ("stable", C.c_int32), ("n_iolets", C.c_int32),
                ("iolet_flux", C.c_double * 64)]


class hlb_instance_monitor(C.Structure):
    _fields_ = [("mass", C.c_double), ("density_min", C.c_double),
                ("density_max", C.c_double), ("velocity_max", C.c_double),
                ("stable", C.c_int32), ("reserved", C.c_int32), ("sites", C.c_int64)]


class hlb_launch_shape(C.Structure):
    _fields_ = [("threads", C.c_int32), ("planes", C.c_int32), ("groups", C.c_int32),
                ("xcd_remap", C.c_int32), ("split", C.c_int32), ("wide", C.c_int32)]


class hlb_copy_result(C.Structure):
    _fields_ = [("bytes_per_s", C.c_double), ("ms", C.c_double), ("bytes", C.c_int64),
                ("threads", C.c_int32), ("xcd_remap", C.c_int32), ("width", C.c_int32),
                ("order", C.c_int32)]


HLBX_FLOAT, HLBX_DOUBLE, HLBX_INT32, HLBX_UINT32, HLBX_INT64, HLBX_UINT64 = range(6)


class hlbx_field(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("n_elements", C.c_int32),
                ("n_offsets", C.c_int32), ("offsets", C.POINTER(C.c_double))]


class hlbg_iolet(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("normal", C.c_double * 3),
                ("radius", C.c_double)]


_P = C.c_void_p
_SIGS = {
    # solver
    "hlb_abi_version": (C.c_int, []),
    "hlb_last_error": (C.c_char_p, []),
    "hlb_create": (C.c_int, [C.POINTER(hlb_params), C.POINTER(hlb_domain),
                             C.POINTER(hlb_iolet), C.c_int32, C.POINTER(hlb_iolet),
                             C.c_int32, C.POINTER(_P)]),
    "hlb_destroy": (None, [_P]),
    "hlb_set_equilibrium": (C.c_int, [_P, C.c_double, _P]),
    "hlb_set_time": (C.c_int, [_P, C.c_uint64]),
    "hlb_get_time": (C.c_uint64, [_P]),
    "hlb_step": (C.c_int, [_P, C.c_int64]),
    "hlb_get_distributions": (C.c_int, [_P, _P]),
    "hlb_set_distributions": (C.c_int, [_P, _P]),
    "hlb_set_equilibrium_fields": (C.c_int, [_P, _P, _P, C.c_int32]),
    "hlb_get_macroscopic": (C.c_int, [_P, _P, _P]),
    "hlb_get_wall_shear_stress": (C.c_int, [_P, _P]),
    "hlb_get_stress": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "hlb_reduce": (C.c_int, [_P, C.POINTER(hlb_monitor)]),
    "hlb_reduce_noalloc": (C.c_int, [_P, C.POINTER(hlb_monitor)]),
    "hlb_copy_ceiling": (C.c_int, [_P, C.c_int32, C.POINTER(hlb_copy_result)]),
    "hlb_create_ensemble": (C.c_int, [C.POINTER(hlb_params), C.POINTER(hlb_domain), _P,
                                      C.c_int32, _P, C.POINTER(hlb_iolet), C.c_int32,
                                      C.POINTER(hlb_iolet), C.c_int32, C.POINTER(_P)]),
    "hlb_instance_count": (C.c_int32, [_P]),
    "hlb_reduce_instances": (C.c_int, [_P, C.POINTER(hlb_instance_monitor)]),
    "hlb_velocity_change_instances": (C.c_int, [_P, _P, _P]),
    "hlb_halo_unpack": (C.c_int, [_P, _P, C.c_int64, _P]),
    "hlb_set_launch_shape": (C.c_int, [_P, C.POINTER(hlb_launch_shape)]),
    "hlb_get_launch_shape": (C.c_int, [_P, C.POINTER(hlb_launch_shape)]),
    "hlb_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(hlb_launch_shape),
                               C.POINTER(C.c_double)]),
    "hlb_snapshot_velocity": (C.c_int, [_P]),
    "hlb_velocity_change": (C.c_int, [_P, C.c_double, C.POINTER(C.c_double)]),
    "hlb_halo_pack": (C.c_int, [_P, _P, C.c_int64, _P]),
    "hlb_create_subdomain": (C.c_int, [C.POINTER(hlb_params), C.POINTER(hlb_domain),
                                       C.POINTER(hlb_domain), C.POINTER(hlb_iolet), C.c_int32,
                                       C.POINTER(hlb_iolet), C.c_int32, C.POINTER(_P)]),
    "hlb_halo_plan_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "hlb_halo_plan_destroy": (None, [_P]),
    "hlb_halo_plan_count": (C.c_int64, [_P]),
    "hlb_halo_exchange": (C.c_int, [_P]),
    "hlb_halo_pack_device": (C.c_int, [_P, _P]),
    "hlb_halo_unpack_device": (C.c_int, [_P, _P]),
    "hlb_halo_last_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hlb_time_steps": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_double)]),
    "hlb_site_count": (C.c_int64, [_P]),
    "hlb_lattice_q": (C.c_int32, [_P]),
    "hlb_device_bytes": (C.c_int64, [_P]),
    "hlb_stream": (_P, [_P]),
    "hlb_synchronize": (C.c_int, [_P]),
    "hlb_tau_from_physical": (C.c_double, [C.c_double] * 4),
    "hlb_pressure_to_lattice": (C.c_double, [C.c_double] * 5),
    "hlb_pressure_difference_to_lattice": (C.c_double, [C.c_double] * 4),
    "hlb_pressure_gradient_to_lattice": (C.c_double, [C.c_double] * 4),
    # geometry
    "hlbg_cylinder": (C.c_int, [C.c_double, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "hlbg_stenosed_cylinder": (C.c_int, [C.c_double, C.c_int32, C.c_double, C.c_double,
                                         C.c_double, C.c_int32, C.POINTER(_P)]),
    "hlbg_vessel_tree": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_int32,
                                   C.POINTER(_P)]),
    "hlbg_read_gmy": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "hlbg_write_gmy": (C.c_int, [_P, C.c_char_p]),
    "hlbg_set_iolets": (C.c_int, [_P, C.POINTER(hlbg_iolet), C.c_int32,
                                  C.POINTER(hlbg_iolet), C.c_int32]),
    "hlbg_free": (None, [_P]),
    "hlbg_fluid_site_count": (C.c_int64, [_P]),
    "hlbg_block_info": (C.c_int, [_P, _P, _P]),
    "hlbg_iolet_count": (C.c_int32, [_P, C.c_int32]),
    "hlbg_get_iolet": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(hlbg_iolet)]),
    "hlbg_get_links": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "hlbg_project": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "hlbg_signed_distance": (C.c_int, [_P, _P, _P, _P]),
    "hlbg_last_error": (C.c_char_p, []),
    # extraction files
    "hlbx_create": (C.c_int, [C.c_char_p, C.c_double, _P, C.c_int64, _P, C.c_int32,
                              C.POINTER(hlbx_field), C.POINTER(_P)]),
    "hlbx_append": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "hlbx_close": (C.c_int, [_P]),
    "hlbx_read_distributions": (C.c_int, [C.c_char_p, C.c_int64, _P, C.c_int32, C.c_int64,
                                          _P, C.POINTER(C.c_uint64)]),
    "hlbx_last_error": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m hemelb_amd.build` "
                "(the HIP implementation is the only code path)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


class HemeLBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code: int, geometry: bool = False, io: bool = False) -> None:
    if code != HLB_OK:
        fn = (lib().hlbx_last_error if io else
              lib().hlbg_last_error if geometry else lib().hlb_last_error)
        raise HemeLBError(code, (fn() or b"").decode())


def declared_functions(header_paths) -> list[str]:
    """Names of the functions declared in the given C headers."""
    import re
    names = []
    for p in header_paths:
        txt = open(p).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(hlb[gxn]?_\w+)\s*\(", txt, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))
