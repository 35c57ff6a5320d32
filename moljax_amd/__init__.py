"""moljax_amd -- MI355X-native (gfx950) kernels behind moljax's stiff-PDE API.

Import paths mirror the reference (``moljax_amd.core.fft_solvers`` for
``moljax.core.fft_solvers`` ...).  Arrays are torch tensors on the ROCm device
(any DLPack producer or NumPy array is accepted and moved there).
"""
from . import core
from ._hip import load_library, MjxError

__version__ = "0.1.0"
