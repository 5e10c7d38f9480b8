"""dida_letkf -- MI355X-native LETKF analysis engine (host-side Python mirror).

The engine itself is the C-ABI shared library built from ``csrc/``
(HIP kernels for gfx950 plus a C++ runtime); this package holds the
configuration mirror of the reference namelists, the synthetic instance
generator and the ctypes binding.
"""
from .config import (LetkfConfig, PREC_FP64, PREC_FP32, PREC_BF16X3,  # noqa: F401
                     VCOORD_LEVEL, VCOORD_LNP)
