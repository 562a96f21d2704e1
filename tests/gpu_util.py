"""GPU test plumbing over the C ABI: device buffers via pm_malloc/pm_memcpy
(proto_amd.device), so the library's own HIP runtime is the only one loaded."""
import sys

import numpy as np

import proto_amd as pa
from proto_amd.device import DeviceArray, gpu_available


def need_gpu():
    assert gpu_available(), "GPU test needs a visible MI355X"


def dev(a):
    need_gpu()
    return DeviceArray.from_numpy(a)


def empty(n, dtype=np.float64):
    need_gpu()
    return DeviceArray(n, dtype)


def zeros(n, dtype=np.float64):
    need_gpu()
    return DeviceArray.zeros(n, dtype)


def vp(d):
    return d.ptr


def host(d):
    return d.numpy()


def box(lo, hi):
    return pa.Box.of(lo, hi)


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def config2_inputs(n, g, dtype=np.float64):
    need_gpu()
    from proto_amd.inputs import config2_inputs as gen
    return gen(n, g, dtype)
