"""Meta solvers combining the thermal and electrical solvers."""
