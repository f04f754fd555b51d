"""MI355X-native Shack-Hartmann image formation and centroiding for DASP.

A drop-in for DASP's ``util.centcmod.centcmod`` construct/run/update contract. The compute path is
hand-written HIP for gfx950 in ``dasp_amd/csrc`` behind a C ABI (``include/dasp_cent.h``); the Python
modules here mirror the reference's interface on top of it.
"""
