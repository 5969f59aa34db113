"""CPU oracle of the HemeLB timestep -- test infrastructure only (see bindings.py)."""
