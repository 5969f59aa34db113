"""The C ABI: every declared symbol is exported, argument validation happens
before any device work (so it runs without a GPU), and the parameter helpers
agree with the reference's formulas."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from hemelb_amd import _native as N
from hemelb_amd import lb as L
from oracle import bindings as O
from tests import cases as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_declared_function_is_exported():
    headers = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
    names = N.declared_functions(headers)
    assert len(names) > 30
    lib = C.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert N.lib().hlb_abi_version() == 1


def _create(case, **over):
    p = N.hlb_params()
    p.lattice, p.kernel, p.wall = case.domain.Q, N.HLB_LBGK, N.HLB_WALL_SBB
    p.inlet_bc = p.outlet_bc = N.HLB_IOLET_NASH
    p.storage, p.tau, p.rho0, p.device, p.layout = 0, 0.8, 1.0, 0, 0
    for k, v in over.items():
        if hasattr(p, k):
            setattr(p, k, v)
    dom = over.get("domain", case.domain)
    cd = dom.as_c()
    ins = L._iolet_array([i.to_product() for i in case.inlets])
    outs = L._iolet_array([o.to_product() for o in case.outlets])
    h = C.c_void_p()
    rc = N.lib().hlb_create(C.byref(p), C.byref(cd), ins, len(case.inlets), outs,
                            len(case.outlets), C.byref(h))
    return rc, N.lib().hlb_last_error().decode()


@pytest.fixture(scope="module")
def case():
    return K.cylinder_case(4.0, 10, 15, "LBGK", "SBB", "cosine")


def _copy(dom):
    import copy
    d = copy.copy(dom)
    for f in ("coords", "site_type", "wall_mask", "iolet_mask", "iolet_id", "cut", "normal"):
        setattr(d, f, getattr(dom, f).copy())
    return d


@pytest.mark.parametrize("field,value,msg", [
    ("lattice", 27, "lattice"), ("kernel", 7, "kernel"), ("wall", 5, "wall"),
    ("inlet_bc", 9, "iolet streamer"), ("storage", 4, "storage"), ("layout", 3, "layout"),
    ("tau", -1.0, "tau")])
def test_bad_parameters_are_rejected(case, field, value, msg):
    rc, err = _create(case, **{field: value})
    assert rc == N.HLB_E_INVALID and msg in err


def test_inconsistent_geometry_is_rejected(case):
    d = _copy(case.domain)
    d.coords[1] = d.coords[0]
    rc, err = _create(case, domain=d)
    assert rc == N.HLB_E_INVALID and "same coordinates" in err
    d = _copy(case.domain)
    i = int(np.nonzero(d.wall_mask)[0][0])
    d.wall_mask[i] = 0          # a solid neighbour without a cut link
    rc, err = _create(case, domain=d)
    assert rc == N.HLB_E_INVALID and "no writer" in err
    d = _copy(case.domain)
    j = int(np.nonzero(d.iolet_mask)[0][0])
    d.site_type[j] = N.HLB_SITE_FLUID
    rc, err = _create(case, domain=d)
    assert rc == N.HLB_E_INVALID and "INLET/OUTLET" in err
    d = _copy(case.domain)
    d.iolet_id[j] = 5
    rc, err = _create(case, domain=d)
    assert rc == N.HLB_E_INVALID and "iolet id" in err


def test_empty_and_missing_site_arrays_are_rejected(case):
    p = N.hlb_params()
    p.lattice, p.kernel, p.wall = case.domain.Q, N.HLB_LBGK, N.HLB_WALL_SBB
    p.inlet_bc = p.outlet_bc = N.HLB_IOLET_NASH
    p.storage, p.tau, p.rho0, p.device, p.layout = 0, 0.8, 1.0, 0, 0
    ins = L._iolet_array([i.to_product() for i in case.inlets])
    outs = L._iolet_array([o.to_product() for o in case.outlets])
    for n, drop in ((0, None), (-3, None), (case.domain.n, "coords"),
                    (case.domain.n, "wall_mask")):
        cd = case.domain.as_c()
        cd.n_sites = n
        if drop:
            setattr(cd, drop, None)
        h = C.c_void_p()
        rc = N.lib().hlb_create(C.byref(p), C.byref(cd), ins, len(case.inlets), outs,
                                len(case.outlets), C.byref(h))
        assert rc == N.HLB_E_INVALID, (n, drop)
        assert not h.value


def test_ladd_needs_a_velocity_iolet(case):
    rc, err = _create(case, inlet_bc=N.HLB_IOLET_LADD)
    assert rc == N.HLB_E_INVALID and "velocity" in err


def test_parameter_helpers_match_oracle():
    lib = N.lib()
    rng = np.random.default_rng(0)
    for _ in range(100):
        dt, dx, rho, eta = rng.uniform(1e-6, 1e-3), rng.uniform(1e-5, 1e-2), 1000.0, 0.004
        p = rng.uniform(0, 100)
        assert lib.hlb_tau_from_physical(dt, dx, rho, eta) == O.lib().orc_tau(dt, dx, rho, eta)
        assert lib.hlb_pressure_to_lattice(p, 0.0, dt, dx, rho) == \
            O.lib().orc_pressure_to_lattice(p, 0.0, dt, dx, rho)
        assert lib.hlb_pressure_difference_to_lattice(p, dt, dx, rho) == \
            O.lib().orc_pressure_diff_to_lattice(p, dt, dx, rho)
        assert lib.hlb_pressure_gradient_to_lattice(p, dt, dx, rho) == \
            O.lib().orc_pressure_gradient_to_lattice(p, dt, dx, rho)


def _create_ensemble(e, instance):
    p = N.hlb_params()
    p.lattice, p.kernel, p.wall = 19, N.HLB_LBGK, N.HLB_WALL_BFL
    p.inlet_bc = p.outlet_bc = N.HLB_IOLET_NASH
    p.storage, p.tau, p.rho0, p.device, p.layout = 1, 0.8, 1.0, 0, 0
    cd = e.domain.as_c()
    inst = np.ascontiguousarray(instance, np.int32)
    taus = np.ascontiguousarray(e.taus, np.float64)
    ins, outs = L._iolet_array(e.inlets), L._iolet_array(e.outlets)
    h = C.c_void_p()
    rc = N.lib().hlb_create_ensemble(C.byref(p), C.byref(cd), inst.ctypes.data_as(C.c_void_p),
                                     len(taus), taus.ctypes.data_as(C.c_void_p), ins,
                                     len(e.inlets), outs, len(e.outlets), C.byref(h))
    err = N.lib().hlb_last_error().decode()
    if rc == N.HLB_OK:
        N.lib().hlb_destroy(h)
    return rc, err


def test_ensemble_validation():
    """Instances must be separate vessels: linked sites or a shared 8^3
    block are rejected before any device work."""
    from hemelb_amd import workloads as WL
    e = WL.stenosed_ensemble(3, seed=1, radius_range=(4.0, 6.0))
    assert len(e.inlets) == len(e.outlets) == 3
    rc, err = _create_ensemble(e, e.instance_of_site)
    assert rc in (N.HLB_OK, N.HLB_E_DEVICE), err     # valid (no GPU here)
    bad = e.instance_of_site.copy()
    bad[e.instances[0].sites.start:e.instances[0].sites.start + 40] = 1
    rc, err = _create_ensemble(e, bad)
    assert rc == N.HLB_E_INVALID and "instances" in err
    bad = e.instance_of_site.copy()
    bad[0] = 7
    rc, err = _create_ensemble(e, bad)
    assert rc == N.HLB_E_INVALID and "instance out of range" in err
