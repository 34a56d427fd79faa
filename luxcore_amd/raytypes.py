"""Ray-buffer record layouts shared by the host API, the tests and the bench.

``RAY_DTYPE`` is LuxRays' 48-byte ``Ray`` and ``HIT_DTYPE`` its 20-byte
``RayHit`` (include/luxrays/core/geometry/ray.h:35-103,
include/luxrays/core/geometry/ray_types.cl:21-36).  Buffers of these records
are what ``HardwareIntersectionDevice::EnqueueTraceRayBuffer`` moves
(include/luxrays/core/hardwareintersectiondevice.h:41-44).
"""
import numpy as np

RAY_FLAGS_NONE = 0x0
RAY_FLAGS_MASKED = 0x1          # ray_types.cl:21-22
NULL_INDEX = 0xFFFFFFFF         # luxrays_types.cl:21

RAY_DTYPE = np.dtype([
    ("o", "<f4", (3,)),
    ("d", "<f4", (3,)),
    ("mint", "<f4"),
    ("maxt", "<f4"),
    ("time", "<f4"),
    ("flags", "<u4"),
    ("pad", "<f4", (2,)),
])
HIT_DTYPE = np.dtype([
    ("t", "<f4"),
    ("b1", "<f4"),
    ("b2", "<f4"),
    ("meshIndex", "<u4"),
    ("triangleIndex", "<u4"),
])
assert RAY_DTYPE.itemsize == 48 and HIT_DTYPE.itemsize == 20

QUERY_CLOSEST = 0
QUERY_ANY = 1


def make_rays(o, d, mint=1e-4, maxt=np.inf, time=0.0, flags=0):
    """Pack origin/direction arrays (n,3) into a contiguous Ray buffer."""
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    n = o.shape[0]
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["mint"] = mint
    r["maxt"] = maxt
    r["time"] = time
    r["flags"] = flags
    return r


def machine_epsilon(v):
    """``MachineEpsilon::E`` of fp32 values (include/luxrays/core/epsilon.h:50-94):
    |nextfloat_0x80ulps(v) - v| clamped to [1e-5, 1e-1]."""
    v = np.asarray(v, dtype=np.float32)
    bumped = (v.view(np.uint32) + np.uint32(0x80)).view(np.float32)
    e = np.abs(bumped - v).astype(np.float32)
    return np.where(e > np.float32(1e-5), np.where(e < np.float32(1e-1), e, np.float32(1e-1)),
                    np.float32(1e-5)).astype(np.float32)


def machine_epsilon_point(p):
    """``MachineEpsilon::E(Point)``: max over the three coordinates."""
    return machine_epsilon(p).max(axis=-1)
