from plask_amd.thermal.dynamic import Dynamic2D, DynamicCyl, Dynamic3D  # noqa: F401
