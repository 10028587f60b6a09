"""fastsar_amd: FastSAR's image formation (factorized and exact backprojection) on AMD Instinct MI355X.

The former API of the reference (fastsar/api.py, fastsar/exact.py, fastsar/bp.py) with a 'hip' backend whose
kernels are hand-written HIP for gfx950, reached through the C ABI of include/fsh.h.
"""
__version__ = '0.1.0'

from .api import (form_image, ImageFormer, available_backends, BACKENDS, final_weights)  # noqa: F401
from .exact import ExactFormer, plane_points, backproject  # noqa: F401
from .memory import MemoryWarning  # noqa: F401
from . import sim, plan, autofocus  # noqa: F401
