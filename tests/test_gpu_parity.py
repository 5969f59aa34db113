"""Parity of the HIP timestep with the CPU oracle (plain binary64 method).

The bar is bit-exact equality of every distribution (np.array_equal, so
+0.0 == -0.0) after each checked step, on the same seeded inputs.
"""
import numpy as np
import pytest

from tests import cases as K

pytestmark = pytest.mark.gpu


def _run_and_compare(case: K.Case, steps, chunks=None, label="", layout="blocks"):
    print(f"[parity] {label} layout={layout} n={case.domain.n} Q={case.domain.Q} "
          f"{case.kernel} {case.wall} {case.inlet_bc}/{case.outlet_bc} steps={steps}",
          flush=True)
    orc = case.oracle()
    gpu = case.product(layout=layout)
    np.testing.assert_array_equal(gpu.get_distributions(), orc.get_f())
    done = 0
    for n in (chunks or [steps]):
        orc.step(n)
        gpu.step(n)
        done += n
        assert gpu.GetTimeStep() == orc.time() == 1 + done
        fo, fg = orc.get_f(), gpu.get_distributions()
        if not np.array_equal(fo, fg):
            bad = np.argwhere(fo != fg)
            s, d = bad[0]
            raise AssertionError(
                f"{label}: {len(bad)} mismatches after {done} steps; first site {s} "
                f"{case.domain.coords[s]} dir {d}: oracle {fo[s, d]!r} gpu {fg[s, d]!r}")
    return orc, gpu


PARITY_CASES = [
    ("fourcube-lbgk-sbb", lambda: K.four_cube_case("LBGK", "SBB"), 12),
    ("fourcube-lbgk-bfl", lambda: K.four_cube_case("LBGK", "BFL"), 12),
    ("fourcube-mrt-bfl", lambda: K.four_cube_case("MRT", "BFL"), 12),
    ("fourcube-mrt-sbb", lambda: K.four_cube_case("MRT", "SBB"), 12),
    ("cyl15-lbgk-sbb-nash", lambda: K.cylinder_case(6.0, 24, 15, "LBGK", "SBB", "cosine"), 60),
    ("cyl15-mrt-bfl-nash", lambda: K.cylinder_case(6.0, 24, 15, "MRT", "BFL", "cosine"), 40),
    ("cyl19-lbgk-bfl-ladd-wom", lambda: K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7), 60),
    ("cyl19-mrt-bfl-ladd-par", lambda: K.cylinder_case(6.0, 24, 19, "MRT", "BFL", "parabolic", tau=0.7), 60),
    ("cyl19-lbgk-sbb-ladd-par", lambda: K.cylinder_case(5.0, 20, 19, "LBGK", "SBB", "parabolic", tau=0.9), 50),
    ("cyl15-lbgk-bfl-ladd-wom", lambda: K.cylinder_case(6.0, 20, 15, "LBGK", "BFL", "womersley", tau=0.75), 40),
    ("tree19-lbgk-bfl-ladd", lambda: K.tree_case(3, 6.0, seed=1), 30),
    ("tree19-mrt-bfl-ladd", lambda: K.tree_case(3, 5.0, seed=2, kernel="MRT"), 20),
]


LAYOUTS = ["blocks", "pull_table", "compact_blocks"]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("label,make,steps", PARITY_CASES, ids=[c[0] for c in PARITY_CASES])
def test_parity_bitwise(label, make, steps, layout):
    _run_and_compare(make(), steps, chunks=[1, 2, steps - 3], label=label, layout=layout)


def test_macroscopic_and_monitor_match_oracle():
    case = K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)
    orc, gpu = _run_and_compare(case, 25, label="macro")
    rho_o, u_o = orc.macro()
    rho_g, u_g = gpu.macroscopic()
    np.testing.assert_array_equal(rho_g, rho_o)
    np.testing.assert_array_equal(u_g, u_o)
    m = gpu.monitor()
    assert m["stable"]
    np.testing.assert_allclose(m["mass"], rho_o.sum(), rtol=1e-13)
    assert m["density_min"] == rho_o.min() and m["density_max"] == rho_o.max()
    np.testing.assert_allclose(m["velocity_max"], np.sqrt((u_o ** 2).sum(1)).max(), rtol=1e-15)
    # flux through each iolet: sum over its sites of rho (u . n)
    dom = case.domain
    for k, (io, sign) in enumerate([(case.inlets[0], 0), (case.outlets[0], 1)]):
        sel = (dom.site_type == (2 if k == 0 else 3))
        n = np.asarray(io.normal, float) / np.linalg.norm(io.normal)
        expect = (rho_o[sel] * (u_o[sel] @ n)).sum()
        np.testing.assert_allclose(m["iolet_flux"][k], expect, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("label,make", [
    ("fourcube-mrt-bfl", lambda: K.four_cube_case("MRT", "BFL")),
    ("cyl19-lbgk-bfl-ladd-wom",
     lambda: K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)),
    ("cyl15-lbgk-sbb-nash", lambda: K.cylinder_case(6.0, 24, 15, "LBGK", "SBB", "cosine"))])
def test_wall_shear_stress_matches_oracle(label, make):
    case = make()
    orc, gpu = _run_and_compare(case, 20, label="wss " + label)
    w_o = orc.wall_shear_stress(case.domain.normal)
    w_g = gpu.wall_shear_stress()
    np.testing.assert_array_equal(w_g, w_o)
    wall = case.domain.wall_mask != 0
    assert (w_g[~wall] == np.finfo(np.float64).max).all()
    assert np.isfinite(w_g[wall]).all() and (w_g[wall] > 0).any()


def test_export_import_resume_is_exact():
    case = K.cylinder_case(6.0, 24, 19, "MRT", "BFL", "womersley", tau=0.7)
    orc, gpu = _run_and_compare(case, 10, label="resume-a")
    f = gpu.get_distributions()
    t = gpu.GetTimeStep()
    gpu2 = case.product()
    gpu2.set_distributions(f)
    gpu2.SetTimeStep(t)
    orc.step(15)
    gpu2.step(15)
    np.testing.assert_array_equal(gpu2.get_distributions(), orc.get_f())


def test_checkpoint_file_resume_is_exact(tmp_path):
    """Write fOld to an .xtr checkpoint, start a new solver from it
    (CheckpointInitialCondition) and continue in step with the oracle."""
    from hemelb_amd import extraction as X
    case = K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)
    orc, gpu = _run_and_compare(case, 12, label="checkpoint")
    path = str(tmp_path / "cp.xtr")
    w = X.checkpoint_writer(gpu, path, 5e-5)
    X.write_checkpoint(w, gpu)
    w.close()
    gpu2 = case.product()
    assert X.load_checkpoint(gpu2, path) == gpu.GetTimeStep() == orc.time()
    orc.step(10)
    gpu2.step(10)
    np.testing.assert_array_equal(gpu2.get_distributions(), orc.get_f())


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_instability_is_reported(storage):
    # The monitor's reduction tests every f > 0 on the state (the shifted
    # forms add w_i rho0 back first); the step itself does not test.
    case = K.cylinder_case(5.0, 16, 15, "LBGK", "SBB", "cosine")
    gpu = case.product(storage=storage)
    gpu.step(3)
    assert gpu.monitor()["stable"]
    f = gpu.get_distributions()
    f[7, 3] = -1e-3
    gpu.set_distributions(f)
    assert not gpu.monitor()["stable"]


def test_convergence_monitor():
    case = K.cylinder_case(5.0, 16, 19, "LBGK", "SBB", "cosine")
    gpu = case.product()
    gpu.step(5)
    gpu.snapshot_velocity()
    assert gpu.velocity_change(0.01) == 0.0
    _, u0 = gpu.macroscopic()
    gpu.step(1)
    _, u1 = gpu.macroscopic()
    expect = np.sqrt(((u1 - u0) ** 2).sum(1)).max() / 0.01
    np.testing.assert_allclose(gpu.velocity_change(0.01), expect, rtol=1e-14)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_halo_pack_returns_pending_post_collision_values(layout):
    """After n steps the state holds the post-collision values of step n whose
    streaming is still pending: they equal the collision of the oracle's state
    before its n-th step, i.e. what the reference sends in step n."""
    from oracle import bindings as O
    case = K.cylinder_case(5.0, 16, 19, "LBGK", "SBB", "cosine")
    orc, gpu = case.oracle(), case.product(layout=layout)
    orc.step(3)
    gpu.step(4)
    f3 = orc.get_f()          # fOld at the start of step 4
    rng = np.random.default_rng(0)
    sites = rng.choice(case.domain.n, 50, replace=False)
    pairs = np.array([(s, d) for s in sites for d in range(19)], np.int64)
    packed = gpu.halo_pack(pairs).reshape(len(sites), 19)
    fpc = np.empty(19)
    for k, s in enumerate(sites):
        O.lib().orc_collide_lbgk(19, case.tau, O._p(np.ascontiguousarray(f3[s])),
                                 O._p(fpc), None, None, None)
        np.testing.assert_array_equal(packed[k], fpc)


@pytest.mark.parametrize("layout,storage", [("blocks", "fp64"), ("compact_blocks", "fp64"),
                                            ("blocks", "fp32")])
def test_halo_unpack_restores_received_values(layout, storage):
    """Unpacking another solver's packed values makes the pending state equal
    again: two identical runs, one perturbed at the pairs through unpack and
    then repaired with the other's packed values, stay identical after more
    steps; unpack followed by pack returns the unpacked values."""
    case = K.cylinder_case(5.0, 16, 19, "LBGK", "BFL", "womersley", tau=0.7)
    a, b = case.product(storage, layout=layout), case.product(storage, layout=layout)
    a.step(4)
    b.step(4)
    rng = np.random.default_rng(1)
    sites = rng.choice(case.domain.n, 40, replace=False)
    pairs = np.array([(s, d) for s in sites for d in range(19)], np.int64)
    sent = a.halo_pack(pairs)
    noise = sent * (1 + 0.01 * rng.standard_normal(sent.shape))
    b.halo_unpack(pairs, noise)
    if storage == "fp64":
        np.testing.assert_array_equal(b.halo_pack(pairs), noise)
    b.halo_unpack(pairs, sent)
    np.testing.assert_array_equal(b.halo_pack(pairs), sent)
    a.step(5)
    b.step(5)
    np.testing.assert_array_equal(a.get_distributions(), b.get_distributions())
    with pytest.raises(Exception, match="twice"):
        b.halo_unpack(np.concatenate([pairs[:2], pairs[:1]]), sent[:3])


@pytest.mark.parametrize("label,make", [
    ("fourcube-mrt-bfl", lambda: K.four_cube_case("MRT", "BFL")),
    ("cyl19-lbgk-bfl-ladd-wom",
     lambda: K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7))])
def test_stress_outputs_match_oracle(label, make):
    case = make()
    orc, gpu = _run_and_compare(case, 15, label="stress " + label)
    want = orc.stress(case.domain.normal)
    got = gpu.stress()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def _oracle_feq(Q, rho, m):
    from oracle import bindings as O
    P = lambda a: a.ctypes.data_as(O.C.c_void_p)  # noqa: E731
    out = np.empty((len(rho), Q))
    row = np.empty(Q)
    for i in range(len(rho)):
        O.lib().orc_feq(Q, float(rho[i]), P(np.ascontiguousarray(m[i])), P(row))
        out[i] = row
    return out


@pytest.mark.parametrize("layout", LAYOUTS)
def test_equilibrium_fields_match_oracle(layout):
    """hlb_set_equilibrium_fields: per-site feq bit-exact with the oracle's
    Lattice::CalculateFeq, from host arrays and from device tensors, and the
    run continues in step with the oracle started from the same fOld."""
    from hemelb_amd import _native as N
    from tests.devmem import DeviceArray
    case = K.cylinder_case(6.0, 24, 19, "MRT", "BFL", "womersley", tau=0.7)
    n = case.domain.n
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.97, 1.03, n)
    m = rng.normal(0.0, 0.01, (n, 3)) * rho[:, None]
    want = _oracle_feq(19, rho, m)
    print(f"[parity] equilibrium fields layout={layout} n={n}", flush=True)
    gpu = case.product(layout=layout)
    gpu.SetTimeStep(4)
    gpu.set_equilibrium_fields(rho, m)
    assert gpu.GetTimeStep() == 4
    np.testing.assert_array_equal(gpu.get_distributions(), want)
    dev = case.product(layout=layout)
    d_rho, d_m = DeviceArray(rho), DeviceArray(m)
    N.check(N.lib().hlb_set_equilibrium_fields(dev._h, d_rho.ptr, d_m.ptr, 1))
    np.testing.assert_array_equal(dev.get_distributions(), want)
    with pytest.raises(ValueError):
        dev.set_equilibrium_fields(rho[:-1], m)
    orc = case.oracle()
    orc.set_f(want)
    orc.step(9)
    dev.step(9)
    np.testing.assert_array_equal(dev.get_distributions(), orc.get_f())


def test_equilibrium_fields_reduced_precision():
    case = K.cylinder_case(6.0, 24, 19, "LBGK", "BFL", "womersley", tau=0.7)
    n = case.domain.n
    rng = np.random.default_rng(8)
    rho = rng.uniform(0.97, 1.03, n)
    m = rng.normal(0.0, 0.01, (n, 3)) * rho[:, None]
    want = _oracle_feq(19, rho, m)
    for storage, tol in (("fp32", 2e-7), ("fp16", 2e-3)):
        g = case.product(storage=storage)
        g.set_equilibrium_fields(rho, m)
        err = np.abs(g.get_distributions() - want).max() / np.abs(want).max()
        print(f"[parity] equilibrium fields {storage}: max rel error {err:.2e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("layout", ["blocks", "compact_blocks"])
def test_split_launch_is_bitexact(layout):
    """Planes with and without boundary sites in two launches of separately
    compiled kernels give the oracle's results bit for bit."""
    case = K.tree_case(3, 6.0, seed=1)
    print(f"[parity] split launch layout={layout} n={case.domain.n}", flush=True)
    orc = case.oracle()
    gpu = case.product(layout=layout)
    gpu.set_launch_shape(threads=128, xcd_remap=8, split=1)
    for n in (1, 2, 17):
        orc.step(n)
        gpu.step(n)
        np.testing.assert_array_equal(gpu.get_distributions(), orc.get_f())


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("layout", ["blocks", "compact_blocks"])
def test_split_launch_reduced_storage_is_bitexact(storage, layout):
    """The split launch (a tuner candidate for every storage form) against
    the single launch in the reduced-precision forms, whose boundary path
    differs (binary16 loads the BFL own values on every lane)."""
    case = K.tree_case(3, 6.0, seed=1)
    print(f"[parity] split launch {storage} layout={layout} n={case.domain.n}", flush=True)
    one = case.product(storage=storage, layout=layout)
    one.set_launch_shape(threads=128, xcd_remap=8, split=0)
    two = case.product(storage=storage, layout=layout)
    two.set_launch_shape(threads=128, xcd_remap=8, split=1)
    for n in (1, 2, 17):
        one.step(n)
        two.step(n)
        np.testing.assert_array_equal(two.get_distributions(), one.get_distributions())


def test_wide_addressing_is_bitexact():
    """The block layout's step with 64-bit element indices (the path of
    states of 4 GiB and more) against the oracle, in every storage form
    against the default 32-bit byte offsets: the test domains are small, so
    the launch shape's `wide` flag selects that path here."""
    case = K.tree_case(3, 6.0, seed=1)
    print(f"[parity] wide addressing n={case.domain.n}", flush=True)
    orc = case.oracle()
    wide = case.product(storage="fp64")
    wide.set_launch_shape(threads=128, xcd_remap=2, wide=1)
    assert wide.launch_shape["wide"] == 1
    for n in (1, 2, 17):
        orc.step(n)
        wide.step(n)
        np.testing.assert_array_equal(wide.get_distributions(), orc.get_f())
    for storage in ("fp32", "fp16"):
        narrow = case.product(storage=storage)
        wide = case.product(storage=storage)
        wide.set_launch_shape(threads=128, xcd_remap=8, wide=1)
        for n in (1, 2, 17):
            narrow.step(n)
            wide.step(n)
            np.testing.assert_array_equal(wide.get_distributions(),
                                          narrow.get_distributions(), storage)
