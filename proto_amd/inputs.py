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


def _splitmix_uniform(seed, stream_id, n):
    """The pm_fill_uniform stream on the host (small counts)."""
    M = (1 << 64) - 1
    s0 = (seed + stream_id * 0x632BE59BD9B4E019) & M
    out = np.empty(n)
    for i in range(n):
        z = (s0 + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = ((z >> 11) * 2.0 ** -53) * 2.0 - 1.0
    return out


def gaussian_params(ngauss, seed):
    """Centres in [0,1)^3 and amplitudes +-1 of the Gaussian right-hand sides."""
    u = _splitmix_uniform(seed, 7, 4 * ngauss)
    cen = (u[:3 * ngauss].reshape(ngauss, 3) + 1.0) * 0.5
    amp = np.where(u[3 * ngauss:] >= 0, 1.0, -1.0)
    return cen, amp


def gaussian_level_rhs(layout, n, data, ngauss, seed, sigma=0.05, stream=None):
    """rhs = sum_k a_k exp(-|x - c_k|^2_periodic / (2 sigma^2)) - mean, on the
    valid cells of level data (ghost 0) of a periodic n^3 layout, built on the
    device as a sum of separable products (pm_level_add_separable)."""
    cen, amp = gaussian_params(ngauss, seed)
    x = (np.arange(n) + 0.5) / n
    g0 = pa.i3([0, 0, 0])
    for k in range(ngauss):
        tabs = []
        for a in range(3):
            d = np.minimum(np.abs(x - cen[k, a]), 1.0 - np.abs(x - cen[k, a]))
            tabs.append(np.ascontiguousarray(np.exp(-(d * d) / (2 * sigma * sigma))))
        pa.call("pm_level_add_separable", layout, pa.PM_F64, data.ptr, g0, float(amp[k]),
                tabs[0].ctypes.data_as(_dp), tabs[1].ctypes.data_as(_dp), tabs[2].ctypes.data_as(_dp), stream)
    tot = DeviceArray.zeros(1)
    pa.call("pm_level_reduce", layout, pa.PM_F64, data.ptr, g0, 1, 0, pa.RED_SUM, tot.ptr, 0, stream)
    mean = tot.numpy(stream)[0] / float(n) ** 3
    ones = np.ascontiguousarray(np.ones(n))
    pa.call("pm_level_add_separable", layout, pa.PM_F64, data.ptr, g0, -mean, ones.ctypes.data_as(_dp),
            ones.ctypes.data_as(_dp), ones.ctypes.data_as(_dp), stream)
    return mean


def batch_sizes(nboxes, edges, seed, stream_id=1):
    """Per-axis box extents drawn iid from `edges` (x, y, z of box 0, then box
    1, ...): index floor((u + 1) / 2 * len(edges)) of the pm_fill_uniform
    stream (seed, stream_id)."""
    u = _splitmix_uniform(seed, stream_id, 3 * nboxes)
    idx = np.minimum(((u + 1.0) * 0.5 * len(edges)).astype(np.int64), len(edges) - 1)
    return np.asarray(edges, dtype=np.int32)[idx].reshape(nboxes, 3)


def serve_trace(nreq, lam=200.0, seed=4, sizes=(32, 48, 64, 96, 128, 192, 256), tol=1e-8, keep=()):
    """The served Poisson stream: arrivals at lam requests/s (exponential gaps
    from the SplitMix64 stream (seed, 1)), edge uniform over the index of
    `sizes` (stream 2), K ~ U{4..32} Gaussians (stream 3), request seeds
    1000 + i, tolerance tol; requests whose index is in `keep` keep their
    answer for reading back."""
    u1 = (_splitmix_uniform(seed, 1, nreq) + 1.0) * 0.5
    u2 = (_splitmix_uniform(seed, 2, nreq) + 1.0) * 0.5
    u3 = (_splitmix_uniform(seed, 3, nreq) + 1.0) * 0.5
    t = np.cumsum(-np.log1p(-u1) / lam)
    reqs = []
    for i in range(nreq):
        n = sizes[min(int(u2[i] * len(sizes)), len(sizes) - 1)]
        k = 4 + min(int(u3[i] * 29), 28)
        reqs.append({"arrival_s": float(t[i]), "n": n, "ngauss": k, "seed": 1000 + i, "tol": tol,
                     "keep_result": 1 if i in keep else 0})
    return reqs
