"""The periodic-state loop's stopping rules (hemelb_amd.extraction
cycles_to_periodic), on a scripted stand-in for a solver: the change per
cycle is given, so the rules are checked without a GPU."""
from hemelb_amd import extraction as X


class ScriptedSolver:
    """Reports a given cycle-to-cycle change after each cycle of steps."""

    def __init__(self, changes):
        self.changes = list(changes)
        self.t = 1
        self.cycles = 0

    def snapshot_velocity(self):
        pass

    def step(self, n):
        self.t += n
        self.cycles += 1

    def monitor(self):
        return {"velocity_max": 1.0}

    def velocity_change(self, reference_speed):
        return self.changes[self.cycles - 1] / reference_speed


def test_stops_below_the_tolerance():
    s = ScriptedSolver([1.0, 0.1, 0.02, 2e-3, 2e-4, 2e-5, 1e-6])
    cycles, history = X.cycles_to_periodic(s, 100, 1e-4, 12)
    assert cycles == 6 and history == [1.0, 0.1, 0.02, 2e-3, 2e-4, 2e-5]
    assert s.t == 1 + 6 * 100


def test_reports_no_convergence_after_max_cycles():
    s = ScriptedSolver([1.0, 0.5, 0.3, 0.2])
    cycles, history = X.cycles_to_periodic(s, 10, 1e-4, 4)
    assert cycles is None and len(history) == 4


def test_plateau_stops_a_reduced_precision_transient():
    # halving or better continues; the first cycle that does not halve stops
    s = ScriptedSolver([1.0, 0.11, 0.024, 0.0069, 0.0061, 0.006])
    cycles, history = X.cycles_to_periodic(s, 10, 1e-3, 12, stop_on_plateau=True)
    assert cycles == 5 and history[-1] == 0.0061
    # without the rule the same run does not converge
    s = ScriptedSolver([1.0, 0.11, 0.024, 0.0069, 0.0061, 0.006])
    cycles, _ = X.cycles_to_periodic(s, 10, 1e-3, 6)
    assert cycles is None
