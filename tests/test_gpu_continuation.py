"""Binary64 periodic state with the transient run in binary16
(hemelb_amd.extraction.periodic_state_continued): the binary16 solver runs
from rest until its cycle-to-cycle change is below 1e-2 or stops halving,
hands its state (exact in binary64) to a binary64 solver at the same time
step, which runs to the 1e-4 periodic-state tolerance.  It must end at the
same inlet phase as the plain binary64 run from rest and in the same
periodic state: within 1e-3 of max |u| (both are within about 1e-4 of it)."""
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


@pytest.mark.parametrize("label,make", CASES, ids=[c[0] for c in CASES])
def test_binary16_transient_reaches_the_binary64_periodic_state(label, make):
    w = make()
    period = int(round(w.period))
    print(f"[continuation] {label}: {w.domain.n} sites, period {period}", flush=True)
    plain = w.solver(storage="fp64")
    cycles, history = X.cycles_to_periodic(plain, period, 1e-4, 20)
    final = w.solver(storage="fp64")
    r = X.periodic_state_continued(w.solver(storage="fp16"), final, period, 1e-4, 1e-2, 20)
    print(f"[continuation] {label}: plain {cycles} cycles {history}; fp16 "
          f"{r['cycles_first']} cycles {r['changes_first']}, fp64 {r['cycles']} cycles "
          f"{r['changes']}", flush=True)
    assert cycles is not None and r["converged"] and r["cycles_first"] is not None
    assert (final.GetTimeStep() - plain.GetTimeStep()) % period == 0
    u0, u1 = plain.macroscopic()[1], final.macroscopic()[1]
    diff = float(np.abs(u1 - u0).max()) / float(np.abs(u0).max())
    print(f"[continuation] {label}: periodic fields differ by {diff:.2e} of max |u|",
          flush=True)
    assert diff < 1e-3
    assert final.monitor()["stable"]
