"""The allocation-free fused residual (pm_stencil_residual2_noalloc) against
pm_stencil_residual2, through the C ABI: same residual bits and the same
{absMax, sum} pairs, on two calls in a row (the second on a smaller box, so
the fixed scratch holds the previous call's partials), and a clear error for
a box whose partials exceed the fixed scratch."""
import ctypes

import numpy as np
import pytest

import proto_amd as pa
import util
from gpu_util import box, config2_inputs, empty, host, say, vp, zeros

pytestmark = pytest.mark.gpu


def _pair(n, g, dtype, npdt):
    u, f = config2_inputs(n, g, npdt)
    c7, o7, _, _, _ = util.lib_stencil("Laplacian", dtype=npdt)
    c27, o27, _, _, _ = util.lib_stencil("Laplacian_27", dtype=npdt)
    s7, s27 = pa.Stencil(c7, o7, dtype), pa.Stencil(c27, o27, dtype)
    ub, rb = box([-g] * 3, [n - 1 + g] * 3), box([0] * 3, [n - 1] * 3)
    sc = float(n * n)
    out = {}
    for name in ("pm_stencil_residual2", "pm_stencil_residual2_noalloc"):
        r1, r2 = empty(n ** 3, npdt), empty(n ** 3, npdt)
        red = zeros(4, npdt)
        say(name, n, npdt.__name__)
        pa.call(name, s7.h, s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r1), vp(r2), ctypes.byref(rb), sc, sc, vp(red),
                None)
        out[name] = (host(r1).tobytes(), host(r2).tobytes(), host(red).tobytes())
    return out


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_noalloc_residual2_matches_residual2(dtype):
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    for n, g in ((96, 2), (40, 1)):  # the second call is smaller
        out = _pair(n, g, dtype, npdt)
        a, b = out["pm_stencil_residual2"], out["pm_stencil_residual2_noalloc"]
        assert a[0] == b[0] and a[1] == b[1], n
        assert a[2] == b[2], n


def test_noalloc_residual2_rejects_oversize_box():
    # a 16384 x 2048 x 1 index box at the default FP64 tiles (64 x 8) launches
    # 65536 workgroups: 2 MiB of partials, more than the fixed 1 MiB; refused
    # before anything is read or launched
    c7, o7, _, _, _ = util.lib_stencil("Laplacian")
    c27, o27, _, _, _ = util.lib_stencil("Laplacian_27")
    s7, s27 = pa.Stencil(c7, o7), pa.Stencil(c27, o27)
    ub, rb = box([-1, -1, -1], [16384, 2048, 1]), box([0, 0, 0], [16383, 2047, 0])
    dummy = ctypes.c_void_p(16)
    rc = pa.lib().pm_stencil_residual2_noalloc(s7.h, s27.h, dummy, ctypes.byref(ub), dummy, dummy, dummy,
                                               ctypes.byref(rb), ctypes.c_double(1.0), ctypes.c_double(1.0), None,
                                               None)
    assert rc == -1
    assert b"fixed scratch" in pa.lib().pm_last_error()
