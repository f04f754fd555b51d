"""setup.py for the ``cent`` extension module (``dasp_amd/cmod/cent*.so``, import name ``cent``).

Built the way DASP builds its cmod extensions: a distutils/setuptools ``Extension`` over the
module's C source, with numpy's include directory, and the engine compiled separately and handed
over as ``extra_objects`` (cmod/src/setup.py:55-63 for centmodule, :76-82 for extra_objects).
Here the separately compiled objects are the hipcc gfx950 objects under ``dasp_amd/csrc/_build``
that ``dasp_amd/build.py`` writes first. Run from the repository root:

    python dasp_amd/setup.py build_ext --inplace

(``python -m dasp_amd.build`` does both steps.)
"""
import glob
import importlib.util
import os

import numpy
from setuptools import Extension, setup

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OBJDIR = os.path.join(HERE, "csrc", "_build")
ROCM_LIB = os.path.realpath("/opt/rocm/lib")
CLANG_RT = glob.glob(os.path.join(ROCM_LIB, "llvm", "lib", "clang", "*", "lib", "linux", "libclang_rt.builtins-x86_64.a"))

# the engine's objects: exactly the hipcc-compiled sources of dasp_amd/build.py (never a glob, so
# that an object left from an older source list cannot be linked in)
_spec = importlib.util.spec_from_file_location("_dasp_build", os.path.join(HERE, "build.py"))
_build = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_build)
OBJECTS = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in _build.SOURCES]

cent = Extension(
    "dasp_amd.cmod.cent",
    sources=[os.path.relpath(os.path.join(HERE, "csrc", "centmodule.cpp"), ROOT)],
    include_dirs=[numpy.get_include(), os.path.join(ROOT, "include")],
    library_dirs=[ROCM_LIB],
    libraries=["amdhip64"],
    extra_compile_args=["-std=c++17", "-O2"],
    extra_link_args=["-Wl,-rpath," + ROCM_LIB] + CLANG_RT[:1],
    extra_objects=OBJECTS,
    language="c++",
)

if __name__ == "__main__":
    os.chdir(ROOT)
    setup(name="dasp_amd_cent", version="2", ext_modules=[cent], script_args=None)
