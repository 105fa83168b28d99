"""ctypes binding of the C ABI in ``include/plask_fem.h`` (``libplaskfem.so``).

The library is the only compute path: if it is missing or no GPU is visible,
``FemContext`` raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplaskfem.so")

PF_OK, PF_EINVAL, PF_ESTATE, PF_EDEVICE, PF_ESHAPE, PF_ENOCONV, PF_EBREAK, PF_ENONFINITE = 0, -1, -2, -3, -4, -5, -6, -8
PF_GEOM_CARTESIAN, PF_GEOM_CYLINDRICAL = 0, 1
PF_THERMAL, PF_ELECTRICAL = 0, 1
PF_FP64, PF_FP32_REFINE, PF_FP32_RELIABLE, PF_FP64_PREC32 = 0, 1, 2, 3
PF_X0_GIVEN, PF_X0_ZERO, PF_X0_DEVICE, PF_X0_UNIFORM, PF_X0_DEVICE_CHECKED = 0, 1, 2, 3, 4
PF_NORM_INIT, PF_NORM_B, PF_NORM_B_FULL, PF_NORM_START = 0, 1, 2, 3
PRECOND_JACOBI, PRECOND_LINE, PRECOND_MG = 1, 2, 3
PF_HELD_FIXED, PF_HELD_REFERENCE = 0, 1
PF_POST_FLUX, PF_POST_CURRENT, PF_POST_JOULE = 0, 1, 2
KCAT_APPLY, KCAT_PRECOND, KCAT_MG_RESIDUAL, KCAT_MG_SMOOTH, KCAT_MG_TRANSFER, KCAT_MG_COARSE = range(6)

EXPORTED = [
    "pf_ctx_create", "pf_ctx_destroy", "pf_last_error", "pf_grid_set", "pf_coeffs_set", "pf_source_set",
    "pf_dirichlet_set", "pf_load_set", "pf_interp_nodes", "pf_solve", "pf_get_solution", "pf_apply", "pf_apply_host", "pf_rhs_host",
    "pf_solution_device", "pf_num_nodes", "pf_num_elements", "pf_kernel_times", "pf_time_apply", "pf_set_profile",
    "pf_set_launch", "pf_set_line_segment", "pf_set_prefetch", "pf_stream_copy", "pf_set_apply", "pf_set_graphs", "pf_time_kernels", "pf_set_block_order", "pf_solution_prolong",
    "pf_set_multigrid", "pf_multigrid_levels", "pf_precond_apply_host", "pf_graphs_enabled", "pf_kernel_stats",
    "pf_coeffs_powerlaw", "pf_coeffs_update", "pf_solution_prolong_dev", "pf_net_output_device",
    "pf_capacity_set", "pf_dynamic_step", "pf_boundary_matrix_set", "pf_postprocess", "pf_interp_elements",
    "pf_junction_state", "pf_junction_cond_set", "pf_set_multigrid_fusion", "pf_coeffs_update_from",
    "pf_source_from", "pf_set_multigrid_xcd",
]

# kernel ids of pf_kernel_stats (PF_K_*) and the kernel functions they stand for
KERNEL_NAMES = {0: "apply (k_apply3d/k_apply2d MODE_OP, fused PCG step)",
                1: "residual (k_apply3d/k_apply2d MODE_RESG, MODE_RESP)",
                2: "line sweep (k_mg_smooth)", 3: "restriction (k_mg_restrict)",
                4: "prolongation (k_mg_prolong_add)", 5: "line preconditioner (k_cg_part)",
                6: "jacobi update (k_cg_update)", 7: "k_cg_r_to_f32", 8: "k_cg_z_from_f32"}


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("precond", C.c_int), ("x0", C.c_int), ("maxit", C.c_int),
                ("rtol", C.c_double), ("inner_rtol", C.c_double), ("max_outer", C.c_int), ("check_every", C.c_int),
                ("norm", C.c_int), ("init_value", C.c_double)]


class KernelStat(C.Structure):
    _fields_ = [("kernel", C.c_int), ("level", C.c_int), ("value_bytes", C.c_int), ("launches", C.c_int),
                ("seconds", C.c_double), ("bytes", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer", C.c_int), ("rel_resid", C.c_double), ("bnorm", C.c_double),
                ("seconds", C.c_double), ("status", C.c_int), ("start", C.c_int), ("fallback", C.c_int)]


class FemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class NoConvergence(FemError):
    pass


_lib = None


def load_library(path=LIB_PATH):
    """Load libplaskfem.so and declare the C signatures.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libplaskfem.so not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    P, i64, I = C.c_void_p, C.c_int64, C.c_int
    L.pf_ctx_create.argtypes = [C.POINTER(P), I, I]
    L.pf_ctx_destroy.argtypes = [P]
    L.pf_ctx_destroy.restype = None
    L.pf_last_error.argtypes = [P]
    L.pf_last_error.restype = C.c_char_p
    L.pf_grid_set.argtypes = [P, I, I, P, i64, P, i64, P, i64]
    L.pf_coeffs_set.argtypes = [P, I, P, P, i64]
    L.pf_source_set.argtypes = [P, P, i64]
    L.pf_dirichlet_set.argtypes = [P, P, P, i64]
    L.pf_load_set.argtypes = [P, P, P, i64]
    L.pf_interp_nodes.argtypes = [P, P, i64, C.c_double, P]
    L.pf_solution_set.argtypes = [P, P]
    L.pf_solve.argtypes = [P, C.POINTER(SolveOpts), P, C.POINTER(SolveStats)]
    L.pf_get_solution.argtypes = [P, P]
    L.pf_apply.argtypes = [P, I, P, P]
    L.pf_apply_host.argtypes = [P, P, P]
    L.pf_rhs_host.argtypes = [P, P]
    L.pf_solution_device.argtypes = [P]
    L.pf_solution_device.restype = P
    L.pf_num_nodes.argtypes = [P]
    L.pf_num_nodes.restype = i64
    L.pf_num_elements.argtypes = [P]
    L.pf_num_elements.restype = i64
    L.pf_kernel_times.argtypes = [P, P, P, I]
    L.pf_kernel_stats.argtypes = [P, P, I]
    L.pf_coeffs_powerlaw.argtypes = [P, I, P, P, P, P, i64]
    L.pf_coeffs_update.argtypes = [P]
    L.pf_time_apply.argtypes = [P, I, I, C.POINTER(C.c_double)]
    L.pf_set_profile.argtypes = [P, I]
    L.pf_set_launch.argtypes = [P, I, I]
    L.pf_set_line_segment.argtypes = [P, I]
    L.pf_set_prefetch.argtypes = [P, I]
    L.pf_set_apply.argtypes = [P, I, I, I]
    L.pf_set_graphs.argtypes = [P, I]
    L.pf_graphs_enabled.argtypes = [P]
    L.pf_set_block_order.argtypes = [P, I]
    L.pf_time_kernels.argtypes = [P, I, I, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pf_stream_copy.argtypes = [P, i64, I, C.POINTER(C.c_double)]
    L.pf_solution_prolong.argtypes = [P, P, P, I, P, I, P, I, C.c_double, C.c_double]
    L.pf_solution_prolong_dev.argtypes = [P, P, P, I, P, I, P, I, C.c_double, C.c_double]
    L.pf_set_multigrid.argtypes = [P, I, I, C.c_double]
    L.pf_multigrid_levels.argtypes = [P]
    L.pf_precond_apply_host.argtypes = [P, I, I, P, P]
    L.pf_capacity_set.argtypes = [P, P, i64, C.c_double]
    L.pf_boundary_matrix_set.argtypes = [P, P, P, P, i64]
    L.pf_postprocess.argtypes = [P, I, P, P, P]
    L.pf_set_multigrid_fusion.argtypes = [P, I]
    L.pf_set_multigrid_xcd.argtypes = [P, I]
    L.pf_coeffs_update_from.argtypes = [P, P]
    L.pf_currents.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pf_get_currents.argtypes = [P, P, P]
    L.pf_get_coeffs.argtypes = [P, P, P]
    L.pf_source_from.argtypes = [P, P, I]
    L.pf_interp_elements.argtypes = [P, I, P, i64, C.c_double, P]
    L.pf_dynamic_step.argtypes = [P, C.POINTER(SolveOpts), I, C.POINTER(SolveStats)]
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class FemContext:
    """One device context: a grid, element coefficients, sources and Dirichlet nodes.

    Node order is natural (axis 0 fastest), element order likewise.
    """

    def __init__(self, device=0, profile=False):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.pf_ctx_create(C.byref(h), int(device), 1 if profile else 0)
        if rc != PF_OK:
            raise FemError(rc, "pf_ctx_create failed: no usable GPU device %d" % device)
        self.h = h
        self.dim = 0
        self.shape = ()

    def close(self):
        if getattr(self, "h", None):
            self.L.pf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != PF_OK:
            msg = self.L.pf_last_error(self.h).decode(errors="replace")
            cls = NoConvergence if rc == PF_ENOCONV else FemError
            raise cls(rc, "%s: %s" % (what, msg))

    # -- setup ---------------------------------------------------------------
    def set_grid(self, axes, cylindrical=False):
        axes = [_f64(a) for a in axes]
        self._axes = axes
        self.dim = len(axes)
        self.shape = tuple(len(a) for a in axes)
        a2 = axes[2] if self.dim == 3 else None
        n2 = len(a2) if a2 is not None else 1
        rc = self.L.pf_grid_set(self.h, self.dim, PF_GEOM_CYLINDRICAL if cylindrical else PF_GEOM_CARTESIAN,
                                _p(axes[0]), len(axes[0]), _p(axes[1]), len(axes[1]), _p(a2), n2)
        self._ck(rc, "pf_grid_set")

    @property
    def nnodes(self):
        return int(self.L.pf_num_nodes(self.h))

    @property
    def nelems(self):
        return int(self.L.pf_num_elements(self.h))

    def set_coeffs(self, k_inplane, k_vert, equation=PF_THERMAL):
        self._kh = _f64(np.ravel(k_inplane))
        self._kv = _f64(np.ravel(k_vert))
        self._ck(self.L.pf_coeffs_set(self.h, equation, _p(self._kh), _p(self._kv), len(self._kh)), "pf_coeffs_set")

    def set_source(self, q):
        if q is None:
            self._ck(self.L.pf_source_set(self.h, None, 0), "pf_source_set")
        else:
            self._q = _f64(np.ravel(q))
            self._ck(self.L.pf_source_set(self.h, _p(self._q), len(self._q)), "pf_source_set")

    def set_dirichlet(self, nodes, values):
        self._dn = np.ascontiguousarray(nodes, dtype=np.int64)
        self._dv = _f64(np.broadcast_to(values, self._dn.shape))
        self._ck(self.L.pf_dirichlet_set(self.h, _p(self._dn), _p(self._dv), len(self._dn)), "pf_dirichlet_set")

    def set_solution(self, x):
        """Set the device solution (node values, natural order)."""
        self._x0 = _f64(x)
        self._ck(self.L.pf_solution_set(self.h, _p(self._x0)), "pf_solution_set")

    def interp(self, points, fill=np.nan):
        """The solution linearly interpolated at points (n x dim, um); outside: fill."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros(len(pts))
        self._ck(self.L.pf_interp_nodes(self.h, _p(pts), len(pts), float(fill), _p(out)), "pf_interp_nodes")
        return out

    def set_load(self, nodes, values):
        """Nodal load added to the element load vector (boundary terms); repeated nodes
        are summed; empty clears it."""
        self._ln = np.ascontiguousarray(nodes, dtype=np.int64)
        self._lv = _f64(np.broadcast_to(values, self._ln.shape))
        self._ck(self.L.pf_load_set(self.h, _p(self._ln), _p(self._lv), len(self._ln)), "pf_load_set")

    # -- compute -------------------------------------------------------------
    def solve(self, x0=None, precision=PF_FP64, rtol=1e-8, maxit=200000, inner_rtol=None, max_outer=50,
              check_every=32, keep_on_device=False, raise_noconv=True, norm=PF_NORM_INIT, init_value=None,
              precond=PRECOND_LINE):
        """Solve; the relative residual is ||B - A X|| / ||B - A X_init|| (norm=PF_NORM_INIT,
        X_init = init_value, default the first entry of x0 or 0) unless norm says otherwise."""
        if init_value is None:
            init_value = float(np.ravel(x0)[0]) if x0 is not None else 0.0
        o = SolveOpts(precision, precond, PF_X0_GIVEN if x0 is not None else PF_X0_ZERO, int(maxit), float(rtol),
                      float(inner_rtol or 0.0), int(max_outer), int(check_every), int(norm), float(init_value))
        st = SolveStats()
        if x0 is not None:
            x = _f64(np.array(x0, dtype=np.float64, copy=True))
        else:
            x = None if keep_on_device else np.zeros(self.nnodes)
        rc = self.L.pf_solve(self.h, C.byref(o), _p(x), C.byref(st))
        if rc != PF_OK and not (rc == PF_ENOCONV and not raise_noconv):
            self._ck(rc, "pf_solve")
        return x, st

    def solve_warm(self, precision=PF_FP64, rtol=1e-8, maxit=200000, inner_rtol=None, check_every=32,
                   norm=PF_NORM_INIT, init_value=0.0, precond=PRECOND_LINE, checked=False):
        """Start from the solution already on the device (no host transfers); checked:
        fall back to the uniform init_value start if that has the smaller residual."""
        o = SolveOpts(precision, precond, PF_X0_DEVICE_CHECKED if checked else PF_X0_DEVICE, int(maxit), float(rtol),
                      float(inner_rtol or 0.0), 50, int(check_every), int(norm), float(init_value))
        st = SolveStats()
        rc = self.L.pf_solve(self.h, C.byref(o), None, C.byref(st))
        self._ck(rc, "pf_solve")
        return st

    def solve_uniform(self, init_value, precision=PF_FP64, rtol=1e-8, maxit=200000, inner_rtol=None, check_every=32,
                      norm=PF_NORM_INIT, max_outer=50, precond=PRECOND_LINE):
        """Start from init_value everywhere (device fill); the solution stays on the device."""
        o = SolveOpts(precision, precond, PF_X0_UNIFORM, int(maxit), float(rtol), float(inner_rtol or 0.0), int(max_outer),
                      int(check_every), int(norm), float(init_value))
        st = SolveStats()
        rc = self.L.pf_solve(self.h, C.byref(o), None, C.byref(st))
        self._ck(rc, "pf_solve")
        return st

    def heat_flux(self, k_h=None, k_v=None):
        """Element heat flux -k grad T (W/m^2), E x dim (pf_postprocess(PF_POST_FLUX))."""
        out = np.zeros((self.nelems, self.dim))
        if k_h is None:
            self._ck(self.L.pf_postprocess(self.h, PF_POST_FLUX, None, None, _p(out)), "pf_postprocess")
        else:
            a, b = _f64(k_h), _f64(k_v)
            self._ck(self.L.pf_postprocess(self.h, PF_POST_FLUX, _p(a), _p(b), _p(out)), "pf_postprocess")
        return out

    def interp_elements(self, what, points, fill=0.0):
        """An element field of pf_postprocess at points (n x dim)."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        nc = 1 if what == PF_POST_JOULE else self.dim
        out = np.zeros((len(pts), nc))
        self._ck(self.L.pf_interp_elements(self.h, int(what), _p(pts), len(pts), float(fill), _p(out)),
                 "pf_interp_elements")
        return out if nc > 1 else out[:, 0]

    def set_boundary_matrix(self, rows, cols, vals):
        """Boundary terms of the operator, A(row, col) += val for col <= row (both triangles
        for off-diagonal entries); empty lists clear them."""
        self._br = np.ascontiguousarray(rows, dtype=np.int64)
        self._bc = np.ascontiguousarray(cols, dtype=np.int64)
        self._bv = _f64(vals)
        n = len(self._bv)
        self._ck(self.L.pf_boundary_matrix_set(self.h, _p(self._br) if n else None, _p(self._bc) if n else None,
                                               _p(self._bv) if n else None, n), "pf_boundary_matrix_set")

    def set_capacity(self, c_elem, theta=0.5):
        """Transient operator: per-element lumped capacities c (stiffness units, each of the
        element's nodes takes c) and the theta of the theta-method; None: steady operator."""
        if c_elem is None:
            self._ck(self.L.pf_capacity_set(self.h, None, 0, float(theta)), "pf_capacity_set")
            return
        self._ce = _f64(np.ravel(c_elem))
        self._ck(self.L.pf_capacity_set(self.h, _p(self._ce), len(self._ce), float(theta)), "pf_capacity_set")

    def dynamic_step(self, precision=PF_FP64, rtol=1e-8, maxit=200000, inner_rtol=None, check_every=32,
                     norm=PF_NORM_START, init_value=0.0, precond=PRECOND_MG, raise_noconv=True, held=PF_HELD_FIXED):
        """One theta-method step from the device temperature: (theta K + C) T' = (C - (1 - theta) K) T + F;
        held = PF_HELD_REFERENCE gives Dirichlet nodes the reference's value + (B T)_r."""
        o = SolveOpts(precision, precond, PF_X0_DEVICE, int(maxit), float(rtol), float(inner_rtol or 0.0), 50,
                      int(check_every), int(norm), float(init_value))
        st = SolveStats()
        rc = self.L.pf_dynamic_step(self.h, C.byref(o), int(held), C.byref(st))
        if rc != PF_OK and not (rc == PF_ENOCONV and not raise_noconv):
            self._ck(rc, "pf_dynamic_step")
        return st

    def solution(self):
        x = np.zeros(self.nnodes)
        self._ck(self.L.pf_get_solution(self.h, _p(x)), "pf_get_solution")
        return x

    def apply(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        self._ck(self.L.pf_apply_host(self.h, _p(x), _p(y)), "pf_apply_host")
        return y

    def rhs(self):
        b = np.zeros(self.nnodes)
        self._ck(self.L.pf_rhs_host(self.h, _p(b)), "pf_rhs_host")
        return b

    def currents(self):
        """Element current densities (E x dim, kA/cm^2) and Joule heat (E, W/m^3) of the
        device solution taken as the potential (pf_currents, pf_get_currents)."""
        dj2, j2 = C.c_double(), C.c_double()
        self._ck(self.L.pf_currents(self.h, C.byref(dj2), C.byref(j2)), "pf_currents")
        cur = np.zeros(self.nelems * self.dim)
        heat = np.zeros(self.nelems)
        self._ck(self.L.pf_get_currents(self.h, _p(cur), _p(heat)), "pf_get_currents")
        return cur.reshape(-1, self.dim), heat

    def coeffs(self):
        """The element conductivities on the device (in-plane, vertical)."""
        kh, kv = np.zeros(self.nelems), np.zeros(self.nelems)
        self._ck(self.L.pf_get_coeffs(self.h, _p(kh), _p(kv)), "pf_get_coeffs")
        return kh, kv

    def kernel_times(self, n=2):
        """(average seconds, timed launches) per kernel class KCAT_* of the last profiled solve."""
        s = np.zeros(n)
        c = np.zeros(n, dtype=np.int64)
        self._ck(self.L.pf_kernel_times(self.h, _p(s), _p(c), int(n)), "pf_kernel_times")
        return s, c

    def kernel_stats(self):
        """Per (kernel, level, precision) records of the last profiled solve: dicts with
        kernel (PF_K_* id), name, level, value_bytes, launches, seconds (summed), bytes
        (summed algorithmic HBM bytes)."""
        n = self.L.pf_kernel_stats(self.h, None, 0)
        if n < 0:
            self._ck(n, "pf_kernel_stats")
        arr = (KernelStat * max(n, 1))()
        self.L.pf_kernel_stats(self.h, C.cast(arr, C.c_void_p), n)
        return [dict(kernel=a.kernel, name=KERNEL_NAMES.get(a.kernel, str(a.kernel)), level=a.level,
                     value_bytes=a.value_bytes, launches=a.launches, seconds=a.seconds, bytes=a.bytes)
                for a in arr[:n]]

    def set_powerlaw(self, k0_h, k0_v, a_h=None, a_v=None, equation=PF_THERMAL):
        """Per-element conductivity law k = k0 (300 / T_avg)^a, evaluated on the device by
        coeffs_update() from the current solution."""
        E = len(np.ravel(k0_h))
        self._law = [_f64(np.ravel(v)) if v is not None else np.zeros(E) for v in (k0_h, k0_v, a_h, a_v)]
        self._ck(self.L.pf_coeffs_powerlaw(self.h, equation, *[_p(v) for v in self._law], E), "pf_coeffs_powerlaw")

    def coeffs_update(self):
        """Element conductivities from the law of set_powerlaw at the current element temperatures."""
        self._ck(self.L.pf_coeffs_update(self.h), "pf_coeffs_update")

    def coeffs_update_from(self, other):
        """The law of set_powerlaw at the element temperatures of another context's solution
        (same grid and device), device to device."""
        self._ck(self.L.pf_coeffs_update_from(self.h, other.h), "pf_coeffs_update_from")

    def source_from(self, other, what=PF_POST_JOULE):
        """Element heat source from another context's Joule heat (same grid and device)."""
        self._ck(self.L.pf_source_from(self.h, other.h, int(what)), "pf_source_from")

    def set_launch(self, rows, planes):
        self._ck(self.L.pf_set_launch(self.h, int(rows), int(planes)), "pf_set_launch")

    def set_line_segment(self, seg):
        self._ck(self.L.pf_set_line_segment(self.h, int(seg)), "pf_set_line_segment")

    def set_apply(self, kind, planes=16, precision=-1):
        """3D operator kernel: kind 0 tile blocks, 1 row blocks of `planes` planes;
        precision PF_FP64, PF_FP32_REFINE or -1 (both)."""
        self._ck(self.L.pf_set_apply(self.h, int(precision), int(kind), int(planes)), "pf_set_apply")

    def set_multigrid(self, max_levels=0, coarse_sweeps=1, omega=0.65):
        """Multigrid preconditioner (precond=PRECOND_MG, 3D) options; levels are rebuilt
        by the next multigrid solve."""
        self._ck(self.L.pf_set_multigrid(self.h, int(max_levels), int(coarse_sweeps), float(omega)),
                 "pf_set_multigrid")

    def set_multigrid_xcd(self, mode):
        """3D multigrid: residual tiles and sweep groups in XCD-aligned order on the coarse
        levels (1, the default), on every level (2), or nowhere (0)."""
        self._ck(self.L.pf_set_multigrid_xcd(self.h, int(mode)), "pf_set_multigrid_xcd")

    def set_multigrid_fusion(self, on):
        """3D: grid transfers inside the smoother / residual kernels (True) or separate."""
        self._ck(self.L.pf_set_multigrid_fusion(self.h, 1 if on else 0), "pf_set_multigrid_fusion")

    def precond_apply(self, r, precision=PF_FP64, precond=PRECOND_LINE):
        """z = M^-1 r for one preconditioner (test hook; host buffers)."""
        r = _f64(r)
        z = np.zeros_like(r)
        self._ck(self.L.pf_precond_apply_host(self.h, int(precision), int(precond), _p(r), _p(z)),
                 "pf_precond_apply_host")
        return z

    def multigrid_levels(self):
        return int(self.L.pf_multigrid_levels(self.h))

    def set_block_order(self, xcd):
        self._ck(self.L.pf_set_block_order(self.h, 1 if xcd else 0), "pf_set_block_order")

    def set_graphs(self, on):
        self._ck(self.L.pf_set_graphs(self.h, 1 if on else 0), "pf_set_graphs")

    def graphs_enabled(self):
        """Whether CG chunks are replayed from hipGraphs (off under a kernel-tracing profiler)."""
        return bool(self.L.pf_graphs_enabled(self.h))

    def set_prefetch(self, depth):
        self._ck(self.L.pf_set_prefetch(self.h, int(depth)), "pf_set_prefetch")

    def prolong(self, coarse, axes, offset=0.0, scale=1.0):
        """Device start vector: offset + scale * coarse (node values on the 3D grid
        `axes`), trilinear onto this grid, Dirichlet values imposed."""
        c = np.ascontiguousarray(coarse, dtype=np.float32)
        ax = [_f64(a) for a in axes]
        assert c.size == np.prod([len(a) for a in ax])
        self._ck(self.L.pf_solution_prolong(self.h, _p(c), _p(ax[0]), len(ax[0]), _p(ax[1]), len(ax[1]), _p(ax[2]),
                                            len(ax[2]), float(offset), float(scale)), "pf_solution_prolong")

    def prolong_device(self, coarse_dev, axes, offset=0.0, scale=1.0):
        """As prolong, for a coarse field already in device memory (address coarse_dev,
        float32 node values on the 3D grid `axes`); allocates no device memory."""
        ax = [_f64(a) for a in axes]
        self._ck(self.L.pf_solution_prolong_dev(self.h, C.c_void_p(int(coarse_dev)), _p(ax[0]), len(ax[0]), _p(ax[1]),
                                                len(ax[1]), _p(ax[2]), len(ax[2]), float(offset), float(scale)),
                 "pf_solution_prolong_dev")

    def stream_copy(self, nbytes, reps=20):
        """Seconds per device-to-device copy of nbytes (read nbytes, write nbytes)."""
        t = C.c_double(0.)
        self._ck(self.L.pf_stream_copy(self.h, int(nbytes), int(reps), C.byref(t)), "pf_stream_copy")
        return t.value

    def set_profile(self, on):
        self._ck(self.L.pf_set_profile(self.h, 1 if on else 0), "pf_set_profile")

    def time_kernels(self, precision=PF_FP64, precond=PRECOND_LINE, reps=50):
        """(apply, update) seconds per launch, back-to-back launches after a solve."""
        a, u = C.c_double(0.), C.c_double(0.)
        self._ck(self.L.pf_time_kernels(self.h, int(precision), int(precond), int(reps), C.byref(a), C.byref(u)),
                 "pf_time_kernels")
        return a.value, u.value

    def time_apply(self, precision=PF_FP64, reps=50):
        t = C.c_double(0.)
        self._ck(self.L.pf_time_apply(self.h, precision, int(reps), C.byref(t)), "pf_time_apply")
        return t.value
