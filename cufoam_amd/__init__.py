"""gis for MI355X: cuFoam's sparse iterative solver path, rebuilt for gfx950.

Package layout:
  csrc/       HIP kernels and the extern "C" ABI (include/gis.h) -> libgis_amd.so
  build.py    hipcc build of libgis_amd.so (in-tree)
  gis.py      ctypes binding (numpy host arrays, torch device tensors)
  foam.py     host mirror of the cuFoam lduMatrix::solver wrappers
  mesh.py     synthetic LDU systems (structured hex meshes)
"""
from . import mesh  # noqa: F401

__all__ = ["mesh", "gis", "foam", "build"]
