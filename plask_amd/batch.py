"""ctypes binding of the batched 2D Shockley solver (pf_batch_* in include/plask_fem.h)."""
import ctypes as C

import numpy as np

from .fem import FemError, load_library


PROGRESS = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p)


class BatchOpts(C.Structure):
    _fields_ = [("maxloops", C.c_int), ("maxit", C.c_int), ("maxerr", C.c_double), ("rtol", C.c_double),
                ("stable", C.c_int), ("block", C.c_int), ("nwg", C.c_int), ("precond", C.c_int),
                ("loop_budget", C.c_int), ("progress", PROGRESS), ("progress_user", C.c_void_p)]


def print_progress(launches, finished, total, seconds, user):
    """Progress callback: one line per kernel launch of a batch solve."""
    print("batch: launch %d, %d/%d instances finished, %.2f s" % (launches, finished, total, seconds), flush=True)


_PRINT_PROGRESS = PROGRESS(print_progress)
_NO_PROGRESS = PROGRESS()


class BatchStat(C.Structure):
    _fields_ = [("loops", C.c_int), ("iters", C.c_int), ("status", C.c_int), ("restarts", C.c_int), ("err", C.c_double),
                ("maxj", C.c_double), ("rel", C.c_double), ("current", C.c_double), ("bwd", C.c_double)]


def _decl(L):
    if getattr(L, "_batch_decl", False):
        return
    P, I = C.c_void_p, C.c_int
    L.pf_batch_create.argtypes = [C.POINTER(P), I]
    L.pf_batch_destroy.argtypes = [P]
    L.pf_batch_destroy.restype = None
    L.pf_batch_last_error.argtypes = [P]
    L.pf_batch_last_error.restype = C.c_char_p
    L.pf_batch_set.argtypes = [P, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P]
    L.pf_batch_autotune.argtypes = [P, C.POINTER(BatchOpts), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_double)]
    L.pf_batch_solve.argtypes = [P, C.POINTER(BatchOpts), P, P, C.POINTER(C.c_double)]
    L.pf_batch_conductivities.argtypes = [P, P, P]
    L._batch_decl = True


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class ShockleyBatch:
    """Many 2D Shockley problems solved by one persistent kernel launch."""

    def __init__(self, instances, cylindrical=True, device=0):
        self.L = load_library()
        _decl(self.L)
        h = C.c_void_p()
        rc = self.L.pf_batch_create(C.byref(h), device)
        if rc:
            raise FemError(rc, "pf_batch_create failed")
        self.h = h
        self.instances = instances
        n0 = np.array([len(p["axes"][0]) for p in instances], np.int32)
        n1 = np.array([len(p["axes"][1]) for p in instances], np.int32)
        self.sizes = n0.astype(np.int64) * n1
        self.esizes = (n0.astype(np.int64) - 1) * (n1 - 1)
        keep = [np.concatenate([p["axes"][0] for p in instances]), np.concatenate([p["axes"][1] for p in instances]),
                np.concatenate([p["sh"] for p in instances]), np.concatenate([p["sv"] for p in instances]),
                np.array([len(p["dir_nodes"]) for p in instances], np.int64),
                np.concatenate([p["dir_nodes"] for p in instances]).astype(np.int64),
                np.concatenate([p["dir_vals"] for p in instances]).astype(np.float64),
                np.array([p["jlo"] for p in instances], np.int32), np.array([p["jhi"] for p in instances], np.int32),
                np.array([p["beta"] for p in instances], np.float64), np.array([p["js"] for p in instances], np.float64)]
        keep = [np.ascontiguousarray(k) for k in keep]
        rc = self.L.pf_batch_set(h, len(instances), 1 if cylindrical else 0, _p(n0), _p(n1), *[_p(k) for k in keep])
        self._ck(rc, "pf_batch_set")

    def _ck(self, rc, what):
        if rc:
            raise FemError(rc, "%s: %s" % (what, self.L.pf_batch_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            self.L.pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def opts(self, maxloops=30, maxit=100000, maxerr=0.05, rtol=1e-8, stable=False, block=0, nwg=0, precond=3,
             loop_budget=5, progress=False):
        """precond 3: multigrid V-cycle per instance (the default), 2: line preconditioner.
        loop_budget: Picard loops per kernel launch and instance (0: one launch for all);
        progress: print a line after every launch."""
        return BatchOpts(maxloops, maxit, maxerr, rtol, 1 if stable else 0, block, nwg, int(precond),
                         int(loop_budget), _PRINT_PROGRESS if progress else _NO_PROGRESS, None)

    def autotune(self, **kw):
        o = self.opts(**kw)
        b, n, t = C.c_int(), C.c_int(), C.c_double()
        self._ck(self.L.pf_batch_autotune(self.h, C.byref(o), C.byref(b), C.byref(n), C.byref(t)), "pf_batch_autotune")
        return b.value, n.value, t.value

    def solve(self, want_v=True, **kw):
        o = self.opts(**kw)
        V = np.zeros(int(self.sizes.sum())) if want_v else None
        st = (BatchStat * len(self.instances))()
        t = C.c_double()
        self._ck(self.L.pf_batch_solve(self.h, C.byref(o), _p(V) if want_v else None, C.cast(st, C.c_void_p),
                                       C.byref(t)), "pf_batch_solve")
        stats = [dict(loops=s.loops, iters=s.iters, status=s.status, restarts=s.restarts, err=s.err, maxj=s.maxj, rel=s.rel, bwd=s.bwd,
                      current=s.current) for s in st]
        parts = None
        if want_v:
            parts = np.split(V, np.cumsum(self.sizes)[:-1])
        return parts, stats, t.value

    def conductivities(self):
        """Per instance (in-plane, vertical) element conductivities of the last loop of the
        last solve (junction elements: the last junction update)."""
        E = int(self.esizes.sum())
        sh, sv = np.zeros(E), np.zeros(E)
        self._ck(self.L.pf_batch_conductivities(self.h, _p(sh), _p(sv)), "pf_batch_conductivities")
        cut = np.cumsum(self.esizes)[:-1]
        return list(zip(np.split(sh, cut), np.split(sv, cut)))
