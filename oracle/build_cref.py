"""Build oracle/_build/libcentref.so, the FP32 C restatement used as a checker and CPU baseline.

Test infrastructure (see oracle/__init__.py). The reference's own cmod/src/centmodule.c cannot be
built here: it needs FFTW3f, GSL and the Python 2 C-API, none of which exist in this image.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libcentref.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    src = os.path.join(HERE, "cref", "centref.c")
    tmp = OUT + ".tmp"
    subprocess.run(["gcc", "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-pthread", src, "-o", tmp, "-lm"],
                   check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
