"""MI355X-native photon-packet transport for SKIRT7 simulation setups.

The package is a thin host layer over libskirt7_amd.so (hand-written HIP kernels for gfx950
behind the C-ABI in include/skirt7_amd.h). See skirt7_amd.skirt for the reference-shaped
classes and skirt7_amd.workloads for the synthetic configurations.
"""
from .capi import SG_FP64, SG_MIXED, SG_NEIGHBOR, SG_TOPDOWN, SkirtError  # noqa: F401
