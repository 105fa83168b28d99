from plask_amd.electrical.shockley import Shockley2D, ShockleyCyl, Shockley3D  # noqa: F401
