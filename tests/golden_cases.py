"""Loader of the golden vectors of the reference's NSPCG (tests/golden/nspcg_golden.npz,
made by tests/golden/make_nspcg_golden.py): the workload inputs regenerated from their
seeds, with NSPCG's solution, iteration count and stopping value."""
import os

import numpy as np

from plask_amd import workloads

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nspcg_golden.npz")


def cases():
    with np.load(PATH, allow_pickle=False) as z:
        out = []
        for c in range(int(z["ncases"])):
            name, seed = str(z["case%d_name" % c]), int(z["case%d_seed" % c])
            shape = tuple(int(v) for v in z["case%d_shape" % c])
            if name == "thermal3d":
                p, cyl = workloads.thermal_3d(seed=seed, n=shape), False
            else:
                p, cyl = workloads.thermal_2d_cyl(seed=seed, nr=shape[0], nz=shape[1]), True
            out.append(dict(name=name, seed=seed, shape=shape, p=p, cyl=cyl, T=z["case%d_T" % c].copy(),
                            iters=int(z["case%d_iters" % c])))
        return out
