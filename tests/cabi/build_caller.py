"""Builds tests/cabi/foam_sequence (the compiled C++ caller of the C ABI) with
g++ against include/gis.h and the wrappers' bridge header only -- no ROCm or
OpenFOAM include path -- linked to cufoam_amd/libgis_amd.so."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "foam_sequence.cpp")
EXE = os.path.join(HERE, "_build", "foam_sequence")


def build(force: bool = False) -> str:
    """Compile (when missing or older than its sources) and return the executable path."""
    deps = [SRC, os.path.join(ROOT, "include", "gis.h"),
            os.path.join(ROOT, "cufoam_amd", "openfoam", "gisFoam", "gisFoamBridge.H")]
    if (not force and os.path.exists(EXE)
            and os.path.getmtime(EXE) >= max(os.path.getmtime(d) for d in deps)):
        return EXE
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    cmd = ["g++", "-std=c++11", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "cufoam_amd", "openfoam", "gisFoam"),
           SRC, "-o", EXE + ".tmp", "-L", os.path.join(ROOT, "cufoam_amd"), "-lgis_amd",
           "-Wl,-rpath,$ORIGIN/../../../cufoam_amd"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(EXE + ".tmp", EXE)
    return EXE


if __name__ == "__main__":
    print(build(force=True))
