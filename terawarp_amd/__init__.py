"""terawarp_amd -- MI355X-native TeraWarp warp path (TPS + affine/rigid).

The compute runs in libterawarp.so (hand-written HIP kernels for gfx950
behind the C ABI in include/terawarp.h).  This package holds the Python
bindings, the seeded synthetic workloads and the build script.
"""
from .warp import (  # noqa: F401
    EXACT, F16, F32, LINEAR, MIXED, NN, PLAIN, ROTATION, SCALING, TRANSLATION,
    Grid, Synth, TeraWarpError, Warper, affine_fit, checksum64, grid_dim, mode_matrix,
)

__all__ = [
    "Warper", "Grid", "Synth", "TeraWarpError", "affine_fit", "mode_matrix", "grid_dim", "checksum64",
    "PLAIN", "MIXED", "EXACT", "LINEAR", "NN", "F32", "F16", "ROTATION", "TRANSLATION", "SCALING",
]
