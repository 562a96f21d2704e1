"""Shared test helpers: ctypes bindings of the oracle (oracle/liboracle.so), of
the compiled reference checker (oracle/_ref, present only where it was built),
a numpy restatement of the seeded input generator, and box helpers.

The oracle and the reference checker are test infrastructure; only tests,
__graft_entry__.smoke() and bench.py's cpu_baseline leg load them."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def ptr(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def i3(v):
    v = list(v) + [0] * (3 - len(v))
    return np.ascontiguousarray(np.array(v, dtype=np.int32))


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        lib = ctypes.CDLL(path)
        lib.orc_reduce_f64.restype = ctypes.c_double
        lib.orc_reduce_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_long]
        lib.orc_reduce_f32.restype = ctypes.c_float
        lib.orc_reduce_f32.argtypes = [ctypes.c_int, _fp, ctypes.c_long]
        lib.orc_resnorm.restype = ctypes.c_double
        lib.orc_resnorm.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        lib.orc_mg_solve.restype = ctypes.c_int
        lib.orc_gsrb_relax.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]
        lib.orc_jacobi_relax.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]
        lib.orc_coarse_residual.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        lib.orc_fine_interp.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        _ci = ctypes.c_int
        lib.orc_patch_gsrb_f64.argtypes = [_ci, _ci, _ci, _ci, ctypes.c_double, _dp, _dp]
        lib.orc_patch_gsrb_f32.argtypes = [_ci, _ci, _ci, _ci, ctypes.c_double, _fp, _fp]
        lib.orc_fas_solve.restype = ctypes.c_int
        lib.orc_fas_solve.argtypes = [_ci, _ci, _dp, _dp, _ci, ctypes.c_double, _dp]
        lib.orc_fas_resnorm.restype = ctypes.c_double
        lib.orc_fas_resnorm.argtypes = [_ci, ctypes.c_double, _dp, _dp]
        lib.orc_uniform_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_long, _dp]
        _oracle = lib
    return _oracle


def ref_path(dim):
    return os.path.join(ROOT, "oracle", "_ref", "libprotoref%d.so" % dim)


_refs = {}


def ref(dim=3):
    """Proto's own CPU path (oracle/_ref), or None where it was not built."""
    if dim not in _refs:
        p = ref_path(dim)
        if not os.path.exists(p):
            _refs[dim] = None
        else:
            lib = ctypes.CDLL(p)
            lib.ref_reduce.restype = ctypes.c_double
            lib.ref_reduce.argtypes = [ctypes.c_int, _dp, ctypes.c_long]
            lib.ref_reduce_f.restype = ctypes.c_float
            lib.ref_reduce_f.argtypes = [ctypes.c_int, _fp, ctypes.c_long]
            _refs[dim] = lib
    return _refs[dim]


# ---------------------------------------------------------------- inputs
_G = np.uint64(0x9E3779B97F4A7C15)
_S = np.uint64(0x632BE59BD9B4E019)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, start=0):
    """Values start..start+n-1 of the SplitMix64 stream (seed, stream) mapped to
    [-1, 1): u = ((z >> 11) * 2^-53) * 2 - 1."""
    with np.errstate(over="ignore"):
        s0 = np.uint64(seed) + np.uint64(stream) * _S
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = _mix64(s0 + i * _G)
    return ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) * 2.0 - 1.0


def box_size(lo, hi):
    return int(np.prod([h - l + 1 for l, h in zip(lo, hi)]))


def box_shape(lo, hi):
    """numpy shape (z, y, x) of a box (x fastest in memory)."""
    lo = list(lo) + [0] * (3 - len(lo))
    hi = list(hi) + [0] * (3 - len(hi))
    return (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def lib_stencil(name, arg=0, dim=3, dtype=np.float64):
    """Proto's library stencil as (coefs, offsets[n,3], sr, dr, ds), read from
    the committed golden table (written by Proto itself)."""
    import json
    with open(os.path.join(GOLDEN, "stencil_lib.json")) as f:
        tab = json.load(f)
    key = "%dD/%s/%d/%s" % (dim, name, arg, "f64" if dtype == np.float64 else "f32")
    e = tab[key]
    coefs = np.array([float.fromhex(c) for c in e["coefs"]], dtype=dtype)
    offs = np.array(e["offsets"], dtype=np.int32).reshape(-1, 3)
    return coefs, offs, e["sr"], e["dr"], e["ds"]


def oracle_apply(coefs, offs, src, slo, shi, dst, dlo, dhi, blo, bhi, replace, scale,
                 sr=(1, 1, 1), dr=(1, 1, 1), ds=(0, 0, 0)):
    o = oracle()
    f = o.orc_stencil_apply_f64 if coefs.dtype == np.float64 else o.orc_stencil_apply_f32
    ct = ctypes.c_double if coefs.dtype == np.float64 else ctypes.c_float
    f(ctypes.c_int(len(coefs)), ptr(coefs), ptr(np.ascontiguousarray(offs, dtype=np.int32)),
      ptr(i3(sr)), ptr(i3(dr)), ptr(i3(ds)), ptr(src), ptr(i3(slo)), ptr(i3(shi)),
      ptr(dst), ptr(i3(dlo)), ptr(i3(dhi)), ptr(i3(blo)), ptr(i3(bhi)),
      ctypes.c_int(int(replace)), ct(scale))


def gaussian_rhs(n, ngauss, seed, sigma=0.05):
    """Sum of ngauss periodic Gaussians, centres U[0,1)^3 and amplitudes +-1
    from the seeded stream, minus its mean (x-fastest flat array of n^3)."""
    u = uniform(seed, 7, 4 * ngauss)
    cen = (u[:3 * ngauss].reshape(ngauss, 3) + 1.0) * 0.5
    amp = np.where(u[3 * ngauss:] >= 0, 1.0, -1.0)
    x = (np.arange(n) + 0.5) / n
    out = np.zeros((n, n, n))
    for k in range(ngauss):
        d = [np.minimum(np.abs(x - cen[k, a]), 1.0 - np.abs(x - cen[k, a])) for a in range(3)]
        g = [np.exp(-(dd * dd) / (2 * sigma * sigma)) for dd in d]
        out += amp[k] * (g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :])
    out = out.reshape(-1)
    return out - out.sum() / out.size


def config2_u(n, g):
    """Config-2 input: interior u = (sin(2 pi x) sin(2 pi y)) sin(2 pi z) + 1e-3 U
    (seed 1, stream 0), x = (i+1/2)/n; ghost layers are periodic images."""
    x = (np.arange(n) + 0.5) / n
    s = np.sin(2 * np.pi * x)
    noise = uniform(1, 0, n ** 3).reshape(n, n, n)
    smooth = (s[None, None, :] * s[None, :, None]) * s[:, None, None]
    u = np.pad(smooth + 1e-3 * noise, g, mode="wrap")
    return np.ascontiguousarray(u.reshape(-1))
