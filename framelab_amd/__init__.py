"""framelab_amd -- MI355X-native CT-operator hot path of FrameLab.

Siddon fan/cone forward projectors, exact voxel-driven adjoints, B-spline
tight-framelet transforms and device-resident CGLS / split-Bregman, as
hand-written HIP kernels for gfx950 behind the C ABI in include/framelab_hip.h.
"""
from . import geometry  # noqa: F401

__all__ = ["geometry"]
