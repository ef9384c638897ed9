"""MI355X (gfx950) implementation of the MELODY 2D explicit dynamics step.

Package layout
  csrc/           HIP kernels, host orchestration and the C-ABI of include/melody2d.h
  native.py       ctypes binding of that C-ABI (no CPU fallback)
  solver.py       Model / Body / Solver façade with the reference's call structure (include/melody2d.hpp
                  is the same for C++)
  preprocessor/   model-file generator (stands in for the reference's MATLAB preprocessor)
  build.py        in-tree build of the native library
"""
__all__ = ["native", "solver", "preprocessor", "build"]
