"""MI355X-native (gfx950, FP64) sparse-GP approximate-EP hot path of geepee.

Model classes with the reference's API live in ``geepee_amd.aep_models``; the HIP
kernels are in ``libgpx.so`` (C-ABI in include/gpx.h), bound by ``geepee_amd._lib``.
"""
from .config import JITTER, PROP_MM, PROP_MC, PROP_LIN  # noqa: F401
