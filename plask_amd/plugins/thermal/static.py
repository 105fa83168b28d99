from plask_amd.thermal.static import Static2D, StaticCyl, Static3D, Convection, Radiation  # noqa: F401
