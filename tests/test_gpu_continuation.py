"""Binary64 periodic state with the transient run in faster forms
(hemelb_amd.extraction.periodic_state_staged): each earlier solver runs
until its cycle-to-cycle change is below 1e-2 or stops halving, then hands
its state (exact in binary64) to the next at the same time step; the last,
binary64, runs to the 1e-4 periodic-state tolerance.  The chains binary16 >
binary64 and binary16 > binary32 > binary64 must end at the same inlet
phase as the plain binary64 run from rest and in the same periodic state:
within 1e-3 of max |u| (both are within about 1e-4 of it)."""
import numpy as np
import pytest

from hemelb_amd import extraction as X
from hemelb_amd import workloads as WL

pytestmark = pytest.mark.gpu

CASES = [
    ("small-cylinder-d3q15-sbb-nash", WL.small_cylinder),
    ("tree3-d3q19-bfl-ladd-nash",
     lambda: WL.vessel_tree("LBGK", seed=1, generations=3, root_radius=8.0)),
]


CHAINS = [["fp16", "fp64"], ["fp16", "fp32", "fp64"]]


@pytest.mark.parametrize("chain", CHAINS, ids=[">".join(c) for c in CHAINS])
@pytest.mark.parametrize("label,make", CASES, ids=[c[0] for c in CASES])
def test_faster_transient_reaches_the_binary64_periodic_state(label, make, chain):
    w = make()
    period = int(round(w.period))
    print(f"[continuation] {label} {'>'.join(chain)}: {w.domain.n} sites, period {period}",
          flush=True)
    plain = w.solver(storage="fp64")
    cycles, history = X.cycles_to_periodic(plain, period, 1e-4, 20)
    stages = [w.solver(storage=st) for st in chain]
    r = X.periodic_state_staged(stages, period, 1e-4, 1e-2, 20)
    final = stages[-1]
    print(f"[continuation] {label}: plain {cycles} cycles {history}; stages "
          f"{[(x['cycles'], x['changes']) for x in r['stages']]}", flush=True)
    assert cycles is not None and r["converged"]
    assert all(x["cycles"] is not None for x in r["stages"])
    assert (final.GetTimeStep() - plain.GetTimeStep()) % period == 0
    u0, u1 = plain.macroscopic()[1], final.macroscopic()[1]
    diff = float(np.abs(u1 - u0).max()) / float(np.abs(u0).max())
    print(f"[continuation] {label}: periodic fields differ by {diff:.2e} of max |u|",
          flush=True)
    assert diff < 1e-3
    assert final.monitor()["stable"]
