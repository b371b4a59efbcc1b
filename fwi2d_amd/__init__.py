"""fwi2d_amd -- MI355X-native 2-D acoustic modeling, RTM and FWI.

The engine is the C-ABI library ``lib/libfwi2d.so`` (hand-written gfx950 HIP
kernels + C++ host code, ``csrc/``); the drop-in executables
``bin/modeling``, ``bin/migration`` and ``bin/inversion`` (``host/``) read the
reference's ``parameters.txt``.  This Python package only binds the ABI for
tests and benchmarks.
"""
from .abi import (ABC, RBC, RESET_GRADIENT, RESET_ILLUM, RESET_IMAGE, Engine, FwiError, Net,
                  guarded_step, lib, mfma_probe, parabola_step, ricker_wavelet,
                  stream_copy)

__all__ = ["ABC", "RBC", "RESET_IMAGE", "RESET_GRADIENT", "RESET_ILLUM", "Engine",
           "FwiError", "Net", "guarded_step", "lib", "mfma_probe", "parabola_step", "ricker_wavelet",
           "stream_copy"]
