"""Reduced-precision storage: f_i - w_i rho0 held in binary32 or binary16,
collision arithmetic in binary32, iolet densities and every reduction in
binary64.  Checked against the binary64 method (itself bit-exact with the
CPU oracle) on the same inputs; the bounds are written in each test."""
import numpy as np
import pytest

from hemelb_amd import extraction as X
from tests import cases as K

pytestmark = pytest.mark.gpu


def _fields(case, storage, steps, layout="blocks"):
    lbm = case.product(storage=storage, layout=layout)
    lbm.step(steps)
    return lbm, X.extract(lbm)


def _report(label, storage, err):
    v, p = err["velocity"], err["pressure"]
    w = err.get("wall_shear_stress", {"l2_rel": float("nan"), "linf_rel": float("nan")})
    print(f"[precision] {label} {storage}: u l2_rel {v['l2_rel']:.3e} linf_rel "
          f"{v['linf_rel']:.3e}; p l2_rel {p['l2_rel']:.3e} linf_rel {p['linf_rel']:.3e}; "
          f"wss l2_rel {w['l2_rel']:.3e} linf_rel {w['linf_rel']:.3e}", flush=True)


CASES = [
    ("cyl19-lbgk-bfl-ladd-wom",
     lambda: K.cylinder_case(8.0, 40, 19, "LBGK", "BFL", "womersley", tau=0.7), 600),
    ("cyl15-lbgk-sbb-nash", lambda: K.cylinder_case(8.0, 40, 15, "LBGK", "SBB", "cosine"), 600),
    ("tree19-mrt-bfl-ladd", lambda: K.tree_case(3, 8.0, seed=2, kernel="MRT"), 400),
    ("cyl15-mrt-bfl-nash", lambda: K.cylinder_case(8.0, 40, 15, "MRT", "BFL", "cosine"), 600),
]

# relative L2 velocity error bounds against binary64
BOUND = {"fp32": 1e-5, "fp16": 2e-2}


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("label,make,steps", CASES, ids=[c[0] for c in CASES])
def test_reduced_storage_error_against_binary64(label, make, steps, storage):
    case = make()
    print(f"[precision] {label} n={case.domain.n} steps={steps} {storage}", flush=True)
    _, ref = _fields(case, "fp64", steps)
    lbm, got = _fields(case, storage, steps)
    err = X.field_errors(got, ref)
    _report(label, storage, err)
    assert lbm.monitor()["stable"]
    assert err["velocity"]["l2_rel"] <= BOUND[storage]
    assert err["velocity"]["linf_rel"] <= 10 * BOUND[storage]
    assert err["pressure"]["linf_rel"] <= 10 * BOUND[storage]
    assert err["wall_shear_stress"]["l2_rel"] <= 10 * BOUND[storage]


LAYOUT_CASES = [
    ("tree19-lbgk-bfl-ladd", lambda: K.tree_case(3, 6.0, seed=1)),
    ("cyl15-mrt-bfl-nash", lambda: K.cylinder_case(6.0, 24, 15, "MRT", "BFL", "cosine")),
]


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("label,make", LAYOUT_CASES, ids=[c[0] for c in LAYOUT_CASES])
def test_reduced_storage_layouts_agree_bitwise(label, make, storage):
    """The block layouts (binary16 with 32-bit offsets issues all of a
    site's loads in one batch) against the pull table, bit for bit."""
    case = make()
    print(f"[precision] layouts {label} {storage}", flush=True)
    a, _ = _fields(case, storage, 30, "blocks")
    for layout in ("pull_table", "compact_blocks"):
        b, _ = _fields(case, storage, 30, layout)
        np.testing.assert_array_equal(a.get_distributions(), b.get_distributions(), layout)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_reduced_storage_reductions_are_binary64(storage):
    """hlb_reduce sums the binary64 densities of the exported state; the
    exported state is exact (stored value + w_i rho0 in binary64)."""
    case = K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)
    lbm, f = _fields(case, storage, 50)
    m = lbm.monitor()
    np.testing.assert_allclose(m["mass"], f.density.sum(), rtol=1e-13)
    assert m["density_min"] == f.density.min() and m["density_max"] == f.density.max()


def test_fp32_storage_export_import_resume_is_exact():
    """The exported fOld (stored value + w_i rho0 in binary64) converts back
    to the same binary32 value, so a resumed run continues bit for bit.
    (Binary16 storage cannot resume bit for bit: the post-stream values of
    boundary links carry binary32 arithmetic that binary16 cannot hold.)"""
    storage = "fp32"
    case = K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)
    a = case.product(storage=storage)
    a.step(20)
    f = a.get_distributions()
    b = case.product(storage=storage)
    b.set_distributions(f)
    b.SetTimeStep(a.GetTimeStep())
    np.testing.assert_array_equal(b.get_distributions(), f)
    a.step(15)
    b.step(15)
    np.testing.assert_array_equal(a.get_distributions(), b.get_distributions())
