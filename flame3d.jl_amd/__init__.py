"""flame3d_amd: MI355X-native per-RK-stage right-hand side of Flame3D.jl.

The compute path is the C-ABI library ``libflame3d.so`` (hand-written gfx950
HIP kernels, built from csrc/ by build.py).  This package holds the host-side
mirror of the reference's driver API (solver.jl flowAdvance / specAdvance /
RK3 stage, reactions.jl chemistry and surrogate calls) on top of that ABI,
plus the mechanism loader.
"""
from . import mech  # noqa: F401

__all__ = ["mech"]
