"""Parity at the benchmark's full size: the 5-generation vessel tree
(7.93 M fluid sites, D3Q19 LBGK, BFL walls, Ladd/Womersley inlet, 16 Nash
outlets) is bit-exact with the oracle over a few steps, and the three
layouts agree bit for bit over more steps."""
import time

import numpy as np
import pytest

from hemelb_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tree():
    t0 = time.time()
    w = WL.vessel_tree("LBGK")
    print(f"[fullsize] tree: {w.domain.n} sites ({time.time() - t0:.1f}s)", flush=True)
    return w


def _oracle(w):
    from oracle import bindings as O
    d = w.domain
    ins = [O.make_iolet(O.ORC_WOMERSLEY, i.position, i.normal, radius=i.radius,
                        pressure_gradient_amplitude=i.pressure_gradient_amplitude,
                        period=i.period, womersley_number=i.womersley_number)
           for i in w.inlets]
    outs = [O.make_iolet(O.ORC_COSINE, o.position, o.normal, density_mean=o.density_mean,
                         density_amp=o.density_amp, period=o.period) for o in w.outlets]
    sim = O.OracleSim(19, O.ORC_LBGK, O.ORC_WALL_BFL, O.ORC_IOLET_LADD, O.ORC_IOLET_NASH,
                      w.tau, d.coords, d.site_type, d.wall_mask, d.iolet_mask, d.iolet_id,
                      d.cut, ins, outs)
    sim.set_equilibrium(w.rho0)
    return sim


@pytest.mark.timeout(600)
def test_vessel_tree_bitexact_with_oracle(tree):
    t0 = time.time()
    orc = _oracle(tree)
    gpu = tree.solver(storage="fp64")
    print(f"[fullsize] set-up {time.time() - t0:.1f}s", flush=True)
    for n in (1, 2):
        orc.step(n)
        gpu.step(n)
        fo, fg = orc.get_f(), gpu.get_distributions()
        bad = int(np.count_nonzero(fo != fg))
        print(f"[fullsize] step {orc.time() - 1}: {bad} mismatches", flush=True)
        assert bad == 0


@pytest.mark.timeout(600)
def test_layouts_agree_at_full_size(tree):
    ref = None
    for layout in ("blocks", "compact_blocks", "pull_table"):
        lbm = tree.solver(storage="fp64", layout=layout)
        lbm.step(25)
        f = lbm.get_distributions()
        del lbm
        if ref is None:
            ref = f
        else:
            assert np.array_equal(f, ref), layout
        print(f"[fullsize] {layout}: 25 steps", flush=True)
