"""odtbrain_amd -- filtered backpropagation for optical diffraction
tomography on AMD Instinct MI355X (gfx950).

Drop-in GPU implementations of ODTbrain's backpropagation functions with the
reference signatures (odtbrain/__init__.py:3-8):

    backpropagate_3d, backpropagate_3d_tilted, backpropagate_2d

The arithmetic runs in hand-written HIP kernels (libodtbrain_hip.so, C ABI
in include/odtbrain_hip.h) called through ctypes; the host side only
validates arguments and plans float64 geometry.
"""
from ._alg2d_bpp import backpropagate_2d
from ._alg3d_bpp import backpropagate_3d
from ._alg3d_bppt import backpropagate_3d_tilted
from . import util
from . import apple
from ._translate_ri import odt_to_ri, opt_to_ri
from ._prepare_sino import sinogram_as_radon, sinogram_as_rytov
from .batch import Batch
from .autograd import backpropagate_3d_adjoint, backpropagate_3d_torch

__version__ = "0.1.0"

__all__ = ["backpropagate_2d", "backpropagate_3d", "backpropagate_3d_tilted",
           "odt_to_ri", "opt_to_ri", "sinogram_as_radon",
           "sinogram_as_rytov", "util", "apple", "Batch", "backpropagate_3d_adjoint",
           "backpropagate_3d_torch"]
