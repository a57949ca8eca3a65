"""MI355X-native DS2D time-domain hot path.

The compute path is libds2d.so (hand-written HIP kernels for gfx950 behind the
C-ABI in include/ds2d.h); this package only binds it (lib), generates the
seeded synthetic workloads (models), drives the learned-preconditioner
workflow (learned) and builds the library (build).
"""
from .lib import (DS2DError, Solver, ShotSpec, Unet, conv3x3_bf16, load, ricker, laplace_filter,  # noqa: F401
                  laplace_band, device_info)
