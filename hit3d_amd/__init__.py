"""MI355X-native HIT3DP pseudospectral time step.

The compute path is libhit3d_hip.so (hand-written HIP kernels for gfx950)
behind the C ABI in include/hit3d.h; ``hit3d_amd.lib`` binds it.
"""
from .lib import FIELDS, RHS_OLD, WRK, Hit3D, Hit3dError, load, random_draws  # noqa: F401
