"""MI355X-native FEM thermal / electrical steady-state solvers (PLaSK thermal.static
and electrical.shockley path): C++ host over a C ABI over gfx950 HIP kernels."""
__version__ = "0.1.0"
