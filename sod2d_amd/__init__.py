"""sod2d_amd -- MI355X (gfx950) implementation of SOD2D's compressible
spectral-element right-hand side and Runge-Kutta stage.

Product layout:
  csrc/       HIP kernels, runtime and C++ host set-up behind the C-ABI
              (include/sod2d_hip.h), built into libsod2d_hip.so
  fortran/    ISO_C_BINDING shim keeping the reference's subroutine interfaces
  case.py     host-side case set-up (tables, box meshes, geometry)
  device.py   device context and the reference-named operators on torch tensors
  solver.py   the time-step sequencer of rk_4_ls_main on the device
"""
from ._lib import SodhipError, lib  # noqa: F401

__all__ = ["SodhipError", "lib"]
