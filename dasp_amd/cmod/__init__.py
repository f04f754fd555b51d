"""Native modules of the gfx950 engine (the counterpart of DASP's ``cmod`` package).

``cent`` is the CPython extension built by ``dasp_amd/setup.py`` (PyInit_cent). One HIP runtime per
process: PyTorch ships its own libamdhip64 with the same soname, so torch, when installed, is
imported before the extension loads; the extension then binds to that runtime and torch tensors
and the engine share the device context (loaded first, the system runtime would leave torch
unable to initialise HIP).
"""
try:
    import torch  # noqa: F401
except ImportError:
    pass
