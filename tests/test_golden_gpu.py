"""The GPU solve against golden vectors of the reference's own iterative solver (NSPCG's
IC(0)-CG compiled from the reference sources, iterative_matrix.hpp:141-268, on the same
assembled systems; tests/golden/): FP64 PCG with the multigrid and the line
preconditioner, and the FP32 forms, to NSPCG's solutions."""
import numpy as np
import pytest

from golden_cases import cases
from plask_amd.fem import (PF_FP32_REFINE, PF_FP64, PF_FP64_PREC32, PRECOND_LINE, PRECOND_MG, FemContext)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,precond,tol", [(PF_FP64, PRECOND_MG, 1e-9), (PF_FP64, PRECOND_LINE, 1e-9),
                                                   (PF_FP64_PREC32, PRECOND_MG, 1e-9), (PF_FP32_REFINE, PRECOND_MG, 1e-8)])
def test_gpu_solution_equals_the_references_nspcg(precision, precond, tol):
    for c in cases():
        p = c["p"]
        print("golden", c["name"], c["seed"], c["shape"], precision, precond, flush=True)
        ctx = FemContext()
        ctx.set_grid(p["axes"], cylindrical=c["cyl"])
        ctx.set_coeffs(p["k_h"], p["k_v"])
        ctx.set_source(p["q"])
        ctx.set_dirichlet(p["dir_nodes"], p["dir_vals"])
        st = ctx.solve_uniform(300.0, precision=precision, rtol=1e-10,
                               precond=precond)
        T = ctx.solution()
        ctx.close()
        rise = np.abs(c["T"] - 300.0).max()
        d = np.abs(T - c["T"]).max() / rise
        print("  iterations %d, max diff / rise %.2e" % (st.iters, d), flush=True)
        assert d <= tol
