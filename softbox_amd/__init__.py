"""softbox_amd -- MI355X-native (gfx950) implementation of SoftBox's per-timestep rigid-rod mechanics.

The compute path is libsoftbox_hip.so (hand-written HIP kernels behind the C ABI of
include/softbox_hip.h).  This package is the host side: ctypes binding, reference-layout stores,
the System-API mirror and the whole-network step.
"""
__version__ = "0.1.0"
