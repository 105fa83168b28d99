"""Electrical solvers (PLaSK category ``electrical``)."""
