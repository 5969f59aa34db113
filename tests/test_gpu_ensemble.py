"""Batched ensemble: independent vessels with their own relaxation times in
one solver.  Each instance must evolve exactly as it does alone in a plain
solver, and every launch shape must give the same bits."""
import numpy as np
import pytest

from hemelb_amd import lb as L
from hemelb_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ensemble():
    return WL.stenosed_ensemble(5, seed=3, radius_range=(4.0, 7.0), period=200.0)


def _alone(e, k, storage):
    inst = e.instances[k]
    lbm = L.LBM(e.traits, e.instance_domain(k), L.LbmParameters(tau=inst.tau),
                e.inlets[inst.inlets], e.outlets[inst.outlets], storage=storage, rho0=1.0)
    lbm.SetInitialConditions(1.0)
    return lbm


@pytest.mark.parametrize("storage,layout", [("fp64", "blocks"), ("fp32", "blocks"),
                                            ("fp32", "compact_blocks")])
def test_ensemble_matches_independent_runs(ensemble, storage, layout):
    e = ensemble
    print(f"[ensemble] n={e.domain.n} instances={len(e.instances)} {storage} {layout}",
          flush=True)
    ens = e.solver(storage=storage, layout=layout)
    ens.step(40)
    f = ens.get_distributions()
    for k, inst in enumerate(e.instances):
        one = _alone(e, k, storage)
        one.step(40)
        np.testing.assert_array_equal(f[inst.sites], one.get_distributions(),
                                      err_msg=f"instance {k} tau {inst.tau}")


def test_instance_monitors(ensemble):
    e = ensemble
    ens = e.solver(storage="fp64")
    ens.step(30)
    ens.snapshot_velocity()
    ens.step(7)
    rho, u = ens.macroscopic()
    mons = ens.monitor_instances()
    change = ens.velocity_change_instances(0.01)
    assert len(mons) == len(e.instances) == ens.instance_count
    u0 = None
    for k, (inst, m) in enumerate(zip(e.instances, mons)):
        sl = inst.sites
        assert m["sites"] == sl.stop - sl.start and m["stable"]
        np.testing.assert_allclose(m["mass"], rho[sl].sum(), rtol=1e-13)
        assert m["density_min"] == rho[sl].min() and m["density_max"] == rho[sl].max()
        np.testing.assert_allclose(m["velocity_max"], np.sqrt((u[sl] ** 2).sum(1)).max(),
                                   rtol=1e-15)
    assert (change > 0).all()


def test_launch_shapes_agree_bitwise(ensemble):
    e = ensemble
    ref = e.solver(storage="fp32")
    ref.step(12)
    want = ref.get_distributions()
    shapes = [dict(threads=64, xcd_remap=1), dict(threads=128), dict(planes=2, xcd_remap=1),
              dict(threads=128, planes=2), dict(groups=7, xcd_remap=1),
              dict(planes=3, groups=5), dict(xcd_remap=2), dict(xcd_remap=16),
              dict(threads=128, xcd_remap=64), dict(groups=9, xcd_remap=4),
              dict(planes=2, xcd_remap=8), dict(threads=64, xcd_remap=3),
              dict(groups=11, xcd_remap=5), dict(threads=128, xcd_remap=8, split=0),
              dict(threads=256, split=0), dict(threads=64, xcd_remap=16, split=1),
              dict(threads=128, xcd_remap=2, wide=1), dict(planes=2, wide=1),
              dict(threads=64, split=1, wide=1)]
    for bad in (dict(xcd_remap=4097), dict(xcd_remap=-1), dict(split=2), dict(wide=2)):
        with pytest.raises(Exception):
            ref.set_launch_shape(**bad)
    for sh in shapes:
        lbm = e.solver(storage="fp32")
        lbm.set_launch_shape(**sh)
        lbm.step(12)
        np.testing.assert_array_equal(lbm.get_distributions(), want, err_msg=str(sh))


def test_autotune_restores_the_state(ensemble):
    e = ensemble
    lbm = e.solver(storage="fp32")
    lbm.step(5)
    before, t = lbm.get_distributions(), lbm.GetTimeStep()
    best = lbm.autotune(steps=4)
    print(f"[ensemble] autotuned shape {best}", flush=True)
    assert lbm.GetTimeStep() == t
    np.testing.assert_array_equal(lbm.get_distributions(), before)
    assert lbm.launch_shape == {k: best[k] for k in ("threads", "planes", "groups", "xcd_remap",
                                                     "split", "wide")}
    other = e.solver(storage="fp32")
    other.step(5)
    lbm.step(10)
    other.step(10)
    np.testing.assert_array_equal(lbm.get_distributions(), other.get_distributions())
