"""Thermal solvers (PLaSK category ``thermal``)."""
