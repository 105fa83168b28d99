"""The oracle against golden vectors of the reference's own iterative solver (NSPCG's
IC(0)-CG compiled from the reference sources, iterative_matrix.hpp:141-268; fixture and
generator under tests/golden/): on the same assembled systems the oracle's band Cholesky
(cholesky_matrix.hpp:90-111) and NSPCG agree to NSPCG's stopping accuracy (zeta 1e-11)."""
import numpy as np

import oracle as O
from golden_cases import cases


def test_band_cholesky_equals_the_references_nspcg():
    for c in cases():
        p = c["p"]
        g = O.Grid(p["axes"], cyl=c["cyl"])
        A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_BAND, p["dir_nodes"], p["dir_vals"])
        T = O.band_cholesky_solve(A)
        rise = np.abs(T - 300.0).max()
        d = np.abs(c["T"] - T).max()
        print(c["name"], c["seed"], c["shape"], "NSPCG iterations", c["iters"], "rise %.3g K" % rise,
              "max diff / rise %.2e" % (d / rise))
        assert c["iters"] > 1 and rise > 1.0
        assert d <= 1e-10 * rise
