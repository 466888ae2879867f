"""colb_amd — MI355X-native affine-body-dynamics Newton step with IPC barriers.

The solver class `colb_amd.abd.NewtonAbd` mirrors the reference's `NewtonAbd` (abd.py:583-757)
and drives hand-written HIP kernels (colb_amd/csrc) through the C ABI in include/abd_hip.h.
"""
from .params import Params, reference_compat, north_star  # noqa: F401
from . import scene  # noqa: F401
