"""MI355X (gfx950) implementation of ECSPRT's per-timestep hybrid-PIC hot path.

The compute path is libecsprt_hip.so (hand-written HIP kernels behind a C ABI,
include/ecsprt_hip.h).  This package is the host-side mirror of the reference's
hot-path interface (src/global.h plugin slots, src/velocities.c, src/positions.c,
src/moments.c, src/potential.c) over that ABI.  There is no CPU fallback: every
call raises if the library is missing or returns an error.
"""
from .pic import Context, HipError, setup_context, default_options  # noqa: F401
