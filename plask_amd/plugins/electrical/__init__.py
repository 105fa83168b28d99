"""electrical category (see plask_amd.electrical)."""
