"""luxcore_amd — an MI355X (gfx950) ray-intersection device for LuxRays.

The product is the C-ABI shared library ``liblrhip.so`` (include/lrhip.h)
with hand-written HIP traversal kernels; this package holds its sources
(csrc/), a ctypes front end (device.py), the ray-buffer record layouts
(raytypes.py), the seeded synthetic workloads (scenes.py), the batched
launcher and autotuner (batch.py) and the serving loop (serve.py).
"""
from .raytypes import (HIT_DTYPE, NULL_INDEX, QUERY_ANY, QUERY_CLOSEST, RAY_DTYPE, RAY_FLAGS_MASKED,
                       make_rays)

__all__ = ["HIT_DTYPE", "NULL_INDEX", "QUERY_ANY", "QUERY_CLOSEST", "RAY_DTYPE", "RAY_FLAGS_MASKED",
           "make_rays"]
