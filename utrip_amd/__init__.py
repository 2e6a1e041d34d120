"""utrip_amd -- MI355X-native CANS MHD time step.

The product is ``libcans_gpu.so`` (HIP kernels for gfx950 behind the C-ABI in
``include/cans_gpu.h``) plus the Fortran drop-in libraries
``libcans{1,2,3}d_gpu.so`` that export the CANS subroutine names.  The Python
module :mod:`utrip_amd.cans` mirrors the CANS routine surface over torch
device tensors for harnesses, tests and benchmarks.
"""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_GPU = PKG_DIR / "libcans_gpu.so"

__all__ = ["PKG_DIR", "LIB_GPU"]
