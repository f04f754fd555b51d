"""Native modules of the gfx950 engine (the counterpart of DASP's ``cmod`` package)."""
