from plask_amd.electrical.shockley import (ActiveCond2D, ActiveCond3D, ActiveCondCyl, Shockley2D,  # noqa: F401
                                           Shockley3D, ShockleyCyl)
