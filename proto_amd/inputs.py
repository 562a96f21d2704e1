"""Seeded synthetic inputs generated on the device through the C ABI
(pm_fill_separable / pm_fill_uniform).  The arithmetic matches the numpy
restatement used by the tests, so both sides see the same bits."""
import ctypes

import numpy as np

import proto_amd as pa
from proto_amd.device import DeviceArray

_dp = ctypes.POINTER(ctypes.c_double)


def uniform(n, seed, stream_id, dtype=np.float64, stream=None):
    d = DeviceArray(n, dtype)
    pa.call("pm_fill_uniform", pa.PM_F64 if np.dtype(dtype) == np.float64 else pa.PM_F32, d.ptr, n, seed,
            stream_id, 1.0, 0.0, stream)
    return d


def config2_inputs(n, g, dtype=np.float64, stream=None):
    """Config 2: u on [-g, n-1+g]^3 with interior values
    (sin(2 pi x) sin(2 pi y)) sin(2 pi z) + 1e-3 U(seed 1, stream 0), x = (i+1/2)/n,
    and ghost cells the periodic images; f = U(seed 1, stream 1) on [0, n-1]^3.
    Computed in FP64 and stored as dtype."""
    x = (np.arange(n) + 0.5) / n
    s = np.ascontiguousarray(np.sin(2 * np.pi * x))
    dt = pa.PM_F64 if np.dtype(dtype) == np.float64 else pa.PM_F32
    ub = pa.Box.of([-g] * 3, [n - 1 + g] * 3)
    u = DeviceArray(ub.size(), dtype)
    pa.call("pm_fill_separable", dt, u.ptr, ctypes.byref(ub), ctypes.byref(ub), pa.i3([n, n, n]),
            s.ctypes.data_as(_dp), s.ctypes.data_as(_dp), s.ctypes.data_as(_dp), 1e-3, 1, 0, stream)
    f = uniform(n ** 3, 1, 1, dtype, stream)
    return u, f
