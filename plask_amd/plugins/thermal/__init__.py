"""thermal category (see plask_amd.thermal)."""
