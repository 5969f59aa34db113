"""ctypes bindings of the CPU oracle (``liboracle.so``) and of the reference
glue (``_ref/libhlbref.so``).

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libhlbref.so")

ORC_LBGK, ORC_MRT = 0, 1
ORC_WALL_SBB, ORC_WALL_BFL = 0, 1
ORC_IOLET_NASH, ORC_IOLET_LADD = 0, 1
ORC_COSINE, ORC_PARABOLIC, ORC_WOMERSLEY = 0, 1, 2


class orc_iolet(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32),
                ("position", C.c_double * 3), ("normal", C.c_double * 3),
                ("density_mean", C.c_double), ("density_amp", C.c_double),
                ("phase", C.c_double), ("period", C.c_double), ("warmup", C.c_int64),
                ("radius", C.c_double), ("max_speed", C.c_double),
                ("pressure_gradient_amplitude", C.c_double),
                ("womersley_number", C.c_double), ("min_sim_density", C.c_double)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_D = C.c_double
_V = C.c_void_p
_SIGS = {
    "orc_lattice": (None, [C.c_int, _V, _V, _V, _V, _V]),
    "orc_density_momentum": (None, [C.c_int, _V, _V, _V]),
    "orc_feq": (None, [C.c_int, _D, _V, _V]),
    "orc_collide_lbgk": (None, [C.c_int, _D, _V, _V, _V, _V, _V]),
    "orc_collide_mrt": (None, [C.c_int, _D, _V, _V, _V]),
    "orc_mrt_basis": (C.c_int, [C.c_int, _V, _V, _V, _D]),
    "orc_bessel_j0": (None, [_D, _D, _D, _V]),
    "orc_ipow_three_halves": (None, [_V]),
    "orc_cosine_density": (_D, [C.POINTER(orc_iolet), C.c_uint64]),
    "orc_iolet_velocity": (None, [C.POINTER(orc_iolet), _V, C.c_uint64, _V]),
    "orc_tau": (_D, [_D, _D, _D, _D]),
    "orc_pressure_to_lattice": (_D, [_D] * 5),
    "orc_pressure_diff_to_lattice": (_D, [_D] * 4),
    "orc_pressure_gradient_to_lattice": (_D, [_D] * 4),
    "orc_create": (_V, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _D, C.c_int64,
                        _V, _V, _V, _V, _V, _V, C.POINTER(orc_iolet), C.c_int,
                        C.POINTER(orc_iolet), C.c_int]),
    "orc_destroy": (None, [_V]),
    "orc_set_equilibrium": (None, [_V, _D, _V]),
    "orc_set_time": (None, [_V, C.c_uint64]),
    "orc_get_time": (C.c_uint64, [_V]),
    "orc_step": (C.c_int, [_V, C.c_int64]),
    "orc_threads": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
    "orc_get_f": (None, [_V, _V]),
    "orc_set_f": (None, [_V, _V]),
    "orc_macro": (None, [_V, _V, _V]),
    "orc_wall_shear_stress": (_D, [C.c_int, _D, _V, _V]),
    "orc_wall_shear_stress_all": (None, [_V, _V, _V]),
    "orc_stress_fields": (None, [C.c_int, _D, _V, _V, _V, _V, _V, _V, _V]),
}
_REF_SIGS = {
    "ref_lattice": (C.c_int, [C.c_int, _V, _V, _V, _V, _V]),
    "ref_density_momentum_feq": (C.c_int, [C.c_int, _V, _V, _V, _V, _V]),
    "ref_feq": (C.c_int, [C.c_int, _D, _V, _V]),
    "ref_mrt_basis": (C.c_int, [C.c_int, _V, _V, _V, _D]),
    "ref_bessel_j0": (None, [_D, _D, _D, _V]),
    "ref_ipow_three_halves": (None, [_V]),
    "ref_pressure_to_lattice": (_D, [_D] * 5),
    "ref_wall_shear_stress": (_D, [C.c_int, _D, _V, _V]),
    "ref_stress_fields": (C.c_int, [C.c_int, _D, _V, _V, _V, _V, _V, _V, _V]),
    "ref_pressure_diff_to_lattice": (_D, [_D] * 4),
    "ref_pressure_gradient_to_lattice": (_D, [_D] * 4),
}

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `python oracle/build_oracle.py`")
        h = C.CDLL(ORACLE_LIB)
        for n, (r, a) in _SIGS.items():
            getattr(h, n).restype = r
            getattr(h, n).argtypes = a
        _lib = h
    return _lib


def ref():
    """The reference glue library, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        h = C.CDLL(REF_LIB)
        for n, (r, a) in _REF_SIGS.items():
            getattr(h, n).restype = r
            getattr(h, n).argtypes = a
        _ref = h
    return _ref


def make_iolet(kind, position, normal, **kw) -> orc_iolet:
    s = orc_iolet()
    s.kind = kind
    s.position[:] = list(map(float, position))
    s.normal[:] = list(map(float, normal))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _iolet_array(lst):
    arr = (orc_iolet * max(1, len(lst)))()
    for i, io in enumerate(lst):
        arr[i] = io
    return arr


class OracleSim:
    """The reference's push-scheme timestep on the CPU, for one domain."""

    def __init__(self, Q, kernel, wall, inlet_bc, outlet_bc, tau, coords, site_type,
                 wall_mask, iolet_mask, iolet_id, cut, inlets=(), outlets=()):
        self.Q = Q
        self.tau = float(tau)
        self.n = int(coords.shape[0])
        self._wall = np.ascontiguousarray(wall_mask, np.uint32)
        self._keep = [np.ascontiguousarray(coords, np.int32),
                      np.ascontiguousarray(site_type, np.uint32),
                      np.ascontiguousarray(wall_mask, np.uint32),
                      np.ascontiguousarray(iolet_mask, np.uint32),
                      np.ascontiguousarray(iolet_id, np.int32),
                      None if cut is None else np.ascontiguousarray(cut, np.float32)]
        ins, outs = _iolet_array(list(inlets)), _iolet_array(list(outlets))
        self._h = lib().orc_create(Q, kernel, wall, inlet_bc, outlet_bc, tau, self.n,
                                   *[_p(a) for a in self._keep], ins, len(inlets),
                                   outs, len(outlets))
        if not self._h:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_equilibrium(self, rho, m=(0.0, 0.0, 0.0)):
        mm = np.asarray(m, np.float64)
        lib().orc_set_equilibrium(self._h, float(rho), _p(mm))

    def set_time(self, t):
        lib().orc_set_time(self._h, int(t))

    def time(self):
        return int(lib().orc_get_time(self._h))

    def step(self, n=1) -> int:
        return int(lib().orc_step(self._h, int(n)))

    def get_f(self):
        f = np.empty((self.n, self.Q), np.float64)
        lib().orc_get_f(self._h, _p(f))
        return f

    def set_f(self, f):
        f = np.ascontiguousarray(f, np.float64)
        lib().orc_set_f(self._h, _p(f))

    def macro(self):
        rho = np.empty(self.n)
        u = np.empty((self.n, 3))
        lib().orc_macro(self._h, _p(rho), _p(u))
        return rho, u

    def stress(self, normal):
        """Per-site stress outputs of fOld (the oracle's orc_stress_fields)."""
        n, Q = self.n, self.Q
        f = self.get_f()
        nrm = np.ascontiguousarray(normal, np.float32)
        out = {"tensor": np.empty((n, 6)), "von_mises": np.empty(n),
               "shear_rate": np.empty(n), "traction": np.zeros((n, 3)),
               "tangential_traction": np.zeros((n, 3))}
        vm, sr = np.zeros(1), np.zeros(1)
        for i in range(n):
            t6, tr, tg = np.zeros(6), np.zeros(3), np.zeros(3)
            wall = self._wall[i] != 0
            lib().orc_stress_fields(Q, self.tau, _p(np.ascontiguousarray(f[i])),
                                    _p(nrm[i]) if wall else None, _p(t6), _p(tr), _p(tg),
                                    _p(vm), _p(sr))
            out["tensor"][i], out["von_mises"][i], out["shear_rate"][i] = t6, vm[0], sr[0]
            if wall:
                out["traction"][i], out["tangential_traction"][i] = tr, tg
        return out

    def wall_shear_stress(self, normal):
        """Per-site WSS magnitude of fOld; float max at non-wall sites."""
        nrm = np.ascontiguousarray(normal, np.float32)
        out = np.empty(self.n)
        lib().orc_wall_shear_stress_all(self._h, _p(nrm), _p(out))
        return out
