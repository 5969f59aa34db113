"""Workload definitions and the host-side property/error helpers."""
import math

import numpy as np

from hemelb_amd import extraction as X
from hemelb_amd import lb as L
from hemelb_amd import workloads as WL


def test_small_cylinder_matches_its_description():
    w = WL.small_cylinder()
    assert abs(w.domain.n - math.pi * 144 * 96) / (math.pi * 144 * 96) < 0.03
    assert w.traits.Lattice is L.D3Q15 and w.tau == 0.8 and w.steps == 6000
    # dt from tau = 0.5 + dt eta / rho / (Cs2 dx^2)  (LbmParameters.h:34-39)
    assert abs(w.meta["time_step_s"] - 6.25e-5) < 1e-12
    units = w.meta["units"]
    # rest density at 80 mmHg (UnitConverter.cc:26-40): Cs2 + 80 * 133.3223874 / (1000 * 0.64)
    assert abs(w.rho0 - (1 / 3 + 80 * 133.3223874 / 640.0) * 3) < 1e-9
    assert abs(units.ConvertPressureToPhysicalUnits(w.rho0 * L.Cs2) - 80.0) < 1e-9
    assert abs(w.inlets[0].density_amp - 133.3223874 / 640.0 * 3) < 1e-12
    assert w.inlets[0].period == 2000.0


def test_vessel_tree_workload_small_variant():
    w = WL.vessel_tree("MRT", seed=3, generations=3, root_radius=6.0)
    assert w.traits.Kernel is L.MRT and len(w.inlets) == 1 and len(w.outlets) == 4
    wom = w.inlets[0]
    nu = (0.7 - 0.5) / 3
    assert abs(wom.womersley_number - wom.radius * math.sqrt(2 * math.pi / 5000 / nu)) < 1e-12


def test_field_errors_norms():
    rng = np.random.default_rng(0)
    n = 1000
    u = rng.normal(size=(n, 3))
    rho = 1 + 0.01 * rng.normal(size=n)
    wall = rng.random(n) < 0.3
    wss = np.where(wall, rng.random(n), X.NO_VALUE)
    ref = X.Fields(rho, u, rho * L.Cs2, wss, wall)
    du = np.zeros_like(u)
    du[7] = [3e-6, 0, 4e-6]
    test = X.Fields(rho + 1e-9, u + du, (rho + 1e-9) * L.Cs2,
                    np.where(wall, np.where(wall, wss, 0.0) * (1 + 1e-3), X.NO_VALUE), wall)
    e = X.field_errors(test, ref)
    umag = np.sqrt((u ** 2).sum(1))
    assert abs(e["velocity"]["linf"] - 5e-6) < 1e-15
    assert abs(e["velocity"]["linf_rel"] - 5e-6 / umag.max()) < 1e-15
    assert abs(e["velocity"]["l2_rel"] - 5e-6 / np.sqrt((umag ** 2).sum())) < 1e-15
    span = (rho.max() - rho.min()) * L.Cs2
    assert abs(e["pressure"]["linf_rel"] - 1e-9 * L.Cs2 / span) < 1e-12
    assert abs(e["wall_shear_stress"]["linf_rel"] - 1e-3) < 1e-9


def test_physical_fields_use_reference_units():
    units = L.UnitConverter(6.25e-5, 5e-5, 1000.0, 0.0)
    f = X.Fields(np.array([3.0]), np.array([[0.01, 0.0, 0.0]]), np.array([1.0]),
                 np.array([0.002]), np.array([True]))
    p = f.physical(units)
    lattice_pressure = 1000.0 * 5e-5 ** 2 / 6.25e-5 ** 2
    assert abs(p.pressure[0] - (1.0 - 1 / 3) * lattice_pressure / 133.3223874) < 1e-9
    assert abs(p.velocity[0, 0] - 0.01 * 0.8) < 1e-15
    assert abs(p.wall_shear_stress[0] - 0.002 * lattice_pressure) < 1e-12


def test_traffic_records_name_a_full_launch_shape():
    """bench.py attaches a profiled traffic figure only when the record's
    launch shape equals the solver's, field for field."""
    import json
    import os
    from hemelb_amd import _native as N
    path = os.path.join(os.path.dirname(__file__), "..", "profiles", "traffic.json")
    recs = json.load(open(path))
    fields = {name for name, _ in N.hlb_launch_shape._fields_}
    for key in ("tree5_fp64", "tree5_fp32", "tree5_fp16"):
        rec = recs[key]
        assert set(rec["launch"]) == fields, key
        assert rec["bytes_per_launch"] == rec["fetch_bytes"] + rec["write_bytes"]
        assert rec["sites"] == 7929603
