from plask_amd.meta.shockley import ThermoElectric2D, ThermoElectric3D, ThermoElectricCyl  # noqa: F401
