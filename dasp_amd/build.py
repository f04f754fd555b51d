"""Build the gfx950 engine: hipcc objects, the C-ABI library and the ``cent`` extension module.

1. every HIP/C++ source of the engine is compiled for gfx950 into ``dasp_amd/csrc/_build``;
2. the objects are linked into ``dasp_amd/libdasp_cent.so``, the plain C ABI of
   include/dasp_cent.h for callers in other languages;
3. ``dasp_amd/setup.py build_ext --inplace`` builds the CPython extension ``cent``
   (``dasp_amd/cmod/cent*.so``, PyInit_cent) over the same objects as ``extra_objects``, the way
   DASP builds its cmod extensions (cmod/src/setup.py:55-63, :76-82).
``python -m dasp_amd.build`` rebuilds everything.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libdasp_cent.so")
ARCH = "gfx950"

SOURCES = ["cent_form.hip", "cent_backend.hip", "cent_fused.hip", "cent_dft.hip", "cent_dft32.hip", "cent_batch.hip", "cent_batch32.hip", "cent_corr.hip", "cent_interp.hip", "cent_noise.hip", "cent_noalloc.hip", "recon_mvm.hip", "atmos.hip",
           "cent_api.cpp", "host_math.cpp", "recon_api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP translation unit.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


# Per-file extras. The DFT-GEMM kernels keep their matrix-core accumulators in ordinary VGPRs
# (no AGPR copies before the accumulators are reused as operands; same occupancy).
EXTRA = {"cent_dft.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
         "cent_dft32.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(verbose=False):
    """Compile every HIP/C++ source for gfx950 and link the shared library. Returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    root = os.path.dirname(HERE)
    r = subprocess.run([sys.executable, os.path.join(HERE, "setup.py"), "build_ext", "--inplace", "--force",
                        "--build-temp", os.path.join(OBJDIR, "ext"),
                        "--build-lib", os.path.join(OBJDIR, "ext", "lib")], cwd=root, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("extension build failed:\n" + (r.stdout + r.stderr)[-6000:])
    if verbose:
        print("built the cent extension module in", os.path.join(HERE, "cmod"))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
