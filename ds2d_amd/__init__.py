"""MI355X-native DS2D time-domain hot path.

The compute path is libds2d.so (hand-written HIP kernels for gfx950 behind the
C-ABI in include/ds2d.h); this package only binds it (lib), generates the
seeded synthetic workloads (models) and builds the library (build).
"""
from .lib import DS2DError, Solver, ShotSpec, load, ricker, laplace_filter, laplace_band, device_info  # noqa: F401
