"""proto_amd: the MI355X-native Proto hot path.

The product is the native library proto_amd/libprotomi.so (C ABI declared in
include/protomi.h, kernels in proto_amd/csrc) and the C++ API over it in
proto_amd/cpp.  This module only loads that library with ctypes for the
Python tests and the benchmark; it has no compute path of its own and raises
if the library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprotomi.so")

PM_F32, PM_F64 = 0, 1
PM_ERR_CANCELLED = -6
RED_MAX, RED_MIN, RED_ABS, RED_SUM, RED_SUMABS, RED_SUMSQ = range(6)

# Every symbol declared in include/protomi.h with its ctypes signature.
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64
_dpp = ctypes.POINTER(ctypes.c_double)


class Box(ctypes.Structure):
    """pm_box: inclusive [lo, hi] with x fastest (2-D: z = 0)."""
    _fields_ = [("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3)]

    @classmethod
    def of(cls, lo, hi):
        lo = list(lo) + [0] * (3 - len(lo))
        hi = list(hi) + [0] * (3 - len(hi))
        b = cls()
        for d in range(3):
            b.lo[d] = int(lo[d])
            b.hi[d] = int(hi[d])
        return b

    def size(self):
        n = 1
        for d in range(3):
            n *= max(0, self.hi[d] - self.lo[d] + 1)
        return n


_BP = ctypes.POINTER(Box)


class MgOptions(ctypes.Structure):
    """pm_mg_options."""
    _fields_ = [("relax", ctypes.c_int), ("pre", ctypes.c_int), ("post", ctypes.c_int), ("bottom", ctypes.c_int),
                ("interp", ctypes.c_int), ("precision", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("tol", ctypes.c_double), ("use_graph", ctypes.c_int), ("fused", ctypes.c_int),
                ("inner", ctypes.c_int)]


_OP = ctypes.POINTER(MgOptions)


class BatchConfig(ctypes.Structure):
    _fields_ = [("apply_zplanes", ctypes.c_int), ("apply_blocks_per_cu", ctypes.c_int),
                ("gsrb_rows", ctypes.c_int), ("gsrb_blocks_per_cu", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOptions(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int), ("mode", ctypes.c_int),
                ("pre", ctypes.c_int), ("post", ctypes.c_int), ("bottom", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("sigma", ctypes.c_double), ("inner", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("n", ctypes.c_int), ("ngauss", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("tol", ctypes.c_double), ("keep_result", ctypes.c_int)]


class RequestStats(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("iters", ctypes.c_int), ("status", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("bytes", ctypes.c_size_t)]


SERVE_SOLVED, SERVE_NOT_CONVERGED, SERVE_REJECTED = 0, 1, 2


class TileConfig(ctypes.Structure):
    """pm_tile_config."""
    _fields_ = [("rows", ctypes.c_int), ("prefetch", ctypes.c_int), ("nt_store", ctypes.c_int),
                ("zplanes", ctypes.c_int)]

    def as_dict(self):
        return {"rows": self.rows, "prefetch": self.prefetch, "nt_store": self.nt_store, "zplanes": self.zplanes}

SIGNATURES = {
    "pm_last_error": (ctypes.c_char_p, []),
    "pm_version": (ctypes.c_char_p, []),
    "pm_device_count": (_i, [_ip]),
    "pm_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
    "pm_free": (_i, [_vp]),
    "pm_memcpy": (_i, [_vp, _vp, _sz, _i, _vp]),
    "pm_memset_zero": (_i, [_vp, _sz, _vp]),
    "pm_stream_create": (_i, [ctypes.POINTER(_vp)]),
    "pm_stream_destroy": (_i, [_vp]),
    "pm_stream_scratch_info": (_i, [_vp, ctypes.POINTER(_sz), _ip]),
    "pm_stream_sync": (_i, [_vp]),
    "pm_device_sync": (_i, []),
    "pm_stencil_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _vp, _ip, _ip, _ip, _ip]),
    "pm_stencil_destroy": (_i, [_vp]),
    "pm_stencil_kind": (_i, [_vp, _ip]),
    "pm_stencil_set_kind": (_i, [_vp, _i]),
    "pm_stencil_apply": (_i, [_vp, _vp, _BP, _vp, _BP, _BP, _i, _i, _d, _vp]),
    "pm_stencil_apply_reduce": (_i, [_vp, _vp, _BP, _vp, _BP, _BP, _i, _d, _i, _vp, _i, _vp]),
    "pm_stencil_residual": (_i, [_vp, _vp, _BP, _vp, _vp, _BP, _d, _vp, _vp]),
    "pm_stencil_residual2": (_i, [_vp, _vp, _vp, _BP, _vp, _vp, _vp, _BP, _d, _d, _vp, _vp]),
    "pm_stencil_residual2_noalloc": (_i, [_vp, _vp, _vp, _BP, _vp, _vp, _vp, _BP, _d, _d, _vp, _vp]),
    "pm_reduce": (_i, [_i, _i, _vp, _sz, _vp, _i, _vp]),
    "pm_reduce_box": (_i, [_i, _i, _vp, _BP, _BP, _i, _vp, _i, _vp]),
    "pm_reduce_init": (_i, [_i, _i, _vp, _vp]),
    "pm_boxdata_copy": (_i, [_i, _vp, _BP, _vp, _BP, _BP, _ip, _i, _vp]),
    "pm_boxdata_binop": (_i, [_i, _i, _vp, _BP, _vp, _BP, _BP, _i, _vp]),
    "pm_boxdata_scalar": (_i, [_i, _i, _vp, _BP, _BP, _d, _i, _vp]),
    "pm_boxdata_increment": (_i, [_i, _vp, _BP, _vp, _BP, _BP, _d, _i, _vp]),
    "pm_fill_uniform": (_i, [_i, _vp, _sz, _u64, _u64, _d, _d, _vp]),
    "pm_fill_separable": (_i, [_i, _vp, _BP, _BP, _ip, _dpp, _dpp, _dpp, _d, _u64, _u64, _vp]),
    "pm_event_create": (_i, [ctypes.POINTER(_vp)]),
    "pm_event_destroy": (_i, [_vp]),
    "pm_event_record": (_i, [_vp, _vp]),
    "pm_event_sync": (_i, [_vp]),
    "pm_event_elapsed_ms": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "pm_layout_create": (_i, [ctypes.POINTER(_vp), _i, _ip, _ip, _ip, _i, _ip]),
    "pm_layout_destroy": (_i, [_vp]),
    "pm_layout_box_elems": (_i, [_vp, _ip, ctypes.POINTER(_sz)]),
    "pm_level_exchange": (_i, [_vp, _i, _vp, _ip, _i, _vp]),
    "pm_level_gsrb": (_i, [_vp, _i, _vp, _ip, _vp, _ip, _d, _i, _vp]),
    "pm_level_jacobi": (_i, [_vp, _i, _vp, _ip, _vp, _ip, _vp, _d, _vp]),
    "pm_level_resnorm": (_i, [_vp, _i, _vp, _ip, _vp, _ip, _d, _vp, _i, _vp]),
    "pm_level_restrict_residual": (_i, [_vp, _vp, _i, _vp, _ip, _vp, _ip, _d, _vp, _ip, _vp]),
    "pm_level_prolong_add": (_i, [_vp, _vp, _i, _vp, _ip, _vp, _ip, _i, _vp]),
    "pm_level_setval": (_i, [_vp, _i, _vp, _ip, _i, _d, _vp]),
    "pm_level_copy": (_i, [_vp, _vp, _ip, _vp, _vp, _ip, _i, _i, _vp]),
    "pm_level_copy_to": (_i, [_vp, _vp, _ip, _vp, _vp, _ip, _i, _i, _vp]),
    "pm_level_residual": (_i, [_vp, _i, _vp, _ip, _vp, _ip, _d, _i, _vp, _ip, _vp, _vp]),
    "pm_level_add_correction": (_i, [_vp, _i, _vp, _ip, _i, _vp, _ip, _vp]),
    "pm_copy_device": (_i, [_vp, _vp, _sz, _i, _vp]),
    "pm_stream_probe": (_i, [_i, _vp, _vp, _sz, _i, _vp]),
    "pm_stencil_autotune": (_i, [_vp, _vp, _i, _BP, _BP, _i, ctypes.POINTER(TileConfig), ctypes.POINTER(ctypes.c_float), _vp]),
    "pm_stencil_set_tile_config": (_i, [_vp, _vp, _i, _BP, ctypes.POINTER(TileConfig)]),
    "pm_stencil_get_tile_config": (_i, [_vp, _vp, _i, _BP, ctypes.POINTER(TileConfig)]),
    "pm_tune_table_save": (_i, [ctypes.c_char_p]),
    "pm_tune_table_load": (_i, [ctypes.c_char_p]),
    "pm_level_reduce": (_i, [_vp, _i, _vp, _ip, _i, _i, _i, _vp, _i, _vp]),
    "pm_level_add_separable": (_i, [_vp, _i, _vp, _ip, _d, _dpp, _dpp, _dpp, _vp]),
    "pm_batch_create": (_i, [ctypes.POINTER(_vp), _i, _ip, _i]),
    "pm_batch_destroy": (_i, [_vp]),
    "pm_batch_elems": (_i, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "pm_batch_offsets": (_i, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "pm_batch_configure": (_i, [_vp, _vp]),
    "pm_batch_get_config": (_i, [_vp, _vp]),
    "pm_batch_autotune": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _vp, _vp, _vp]),
    "pm_batch_apply": (_i, [_vp, _vp, _vp, _vp, _d, _vp]),
    "pm_batch_gsrb": (_i, [_vp, _i, _vp, _vp, _vp, _d, _vp]),
    "pm_batch_sweep_form": (_i, [_vp, _i, _i]),
    "pm_mg_set_tolerance": (_i, [_vp, _d, _i]),
    "pm_server_default_options": (None, [ctypes.POINTER(ServerOptions)]),
    "pm_server_create": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(ServerOptions)]),
    "pm_server_destroy": (_i, [_vp]),
    "pm_server_request_bytes": (_i, [_vp, _i, ctypes.POINTER(_sz)]),
    "pm_server_resident_bytes": (_i, [_vp, ctypes.POINTER(_sz)]),
    "pm_server_run": (_i, [_vp, _i, ctypes.POINTER(Request), ctypes.POINTER(RequestStats), _vp]),
    "pm_server_result": (_i, [_vp, _i, _vp]),
    "pm_server_rhs": (_i, [_i, _i, _u64, _d, _vp, _vp]),
    "pm_server_set_stop": (_i, [_vp, _i]),
    "pm_server_discard": (_i, [_vp]),
    "pm_server_checkpoint_size": (_i, [_vp, ctypes.POINTER(_sz)]),
    "pm_server_checkpoint": (_i, [_vp, _vp, _sz, _vp]),
    "pm_server_restore": (_i, [_vp, _vp, _sz, _vp]),
    "pm_level_gsrb_sweep": (_i, [_vp, _i, _vp, _ip, _vp, _ip, _vp, _ip, _d, _vp]),
    "pm_level_sweep_form": (_i, [_i]),
    "pm_level_sweep_tune": (_i, [_i, _i]),
    "pm_fas_create": (_i, [ctypes.POINTER(_vp), _vp, _i, _d]),
    "pm_fas_destroy": (_i, [_vp]),
    "pm_fas_bytes": (_i, [_vp, ctypes.POINTER(_sz)]),
    "pm_fas_vcycle": (_i, [_vp, _vp, _vp, _vp]),
    "pm_fas_resnorm": (_i, [_vp, _vp, _vp, ctypes.POINTER(_d), _vp]),
    "pm_fas_residual": (_i, [_vp, _vp, _vp, _vp, ctypes.POINTER(_d), _vp]),
    "pm_fas_solve": (_i, [_vp, _vp, _vp, _i, _d, _vp, ctypes.POINTER(_i), _vp]),
    "pm_level_linear_size": (_i, [_vp, _i, _ip, _i, ctypes.POINTER(_sz)]),
    "pm_level_linear_out": (_i, [_vp, _i, _vp, _ip, _i, _vp, _i, _vp]),
    "pm_level_linear_in": (_i, [_vp, _i, _vp, _ip, _i, _vp, _i, _vp]),
    "pm_graph_begin": (_i, [_vp]),
    "pm_graph_end": (_i, [_vp, ctypes.POINTER(_vp)]),
    "pm_graph_launch": (_i, [_vp, _vp]),
    "pm_graph_destroy": (_i, [_vp]),
    "pm_mg_default_options": (None, [_OP]),
    "pm_mg_create": (_i, [ctypes.POINTER(_vp), _vp, _i, _d, _OP]),
    "pm_mg_destroy": (_i, [_vp]),
    "pm_mg_bytes": (_i, [_vp, ctypes.POINTER(_sz)]),
    "pm_mg_level_layout": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "pm_mg_solve": (_i, [_vp, _vp, _vp, _dpp, _ip, _vp]),
    "pm_mg_vcycle": (_i, [_vp, _vp, _vp, _vp]),
    "pm_mg_resnorm": (_i, [_vp, _vp, _vp, _dpp, _vp]),
}


class NativeError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError("proto_amd native library missing: %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().pm_last_error().decode(errors="replace")
        raise NativeError("%s failed (rc=%d): %s" % (what, rc, msg))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)


def i3(v, fill=0):
    v = list(v) + [fill] * (3 - len(v))
    return (ctypes.c_int * 3)(*[int(x) for x in v])


class Stencil:
    """A device Stencil handle (pm_stencil)."""

    def __init__(self, coefs, offsets, dtype=PM_F64, dim=3, src_ratio=(1, 1, 1), dst_ratio=(1, 1, 1),
                 dst_shift=(0, 0, 0)):
        import numpy as np
        npdt = np.float64 if dtype == PM_F64 else np.float32
        self.coefs = np.ascontiguousarray(np.asarray(coefs, dtype=npdt))
        offs = np.zeros((len(self.coefs), 3), dtype=np.int32)
        o = np.asarray(offsets, dtype=np.int32).reshape(len(self.coefs), -1)
        offs[:, :o.shape[1]] = o
        self.offsets = np.ascontiguousarray(offs)
        self.dtype = dtype
        h = _vp()
        call("pm_stencil_create", ctypes.byref(h), dtype, dim, len(self.coefs),
             self.coefs.ctypes.data_as(_vp), self.offsets.ctypes.data_as(_ip), i3(src_ratio, 1),
             i3(dst_ratio, 1), i3(dst_shift, 0))
        self.h = h

    def kind(self):
        k = ctypes.c_int()
        call("pm_stencil_kind", self.h, ctypes.byref(k))
        return k.value

    def set_kind(self, kind):
        call("pm_stencil_set_kind", self.h, int(kind))
        return self

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pm_stencil_destroy(self.h)
            self.h = None


class Batch:
    """A batch of independent uneven patches (pm_batch): valid extents
    sizes[b] = (nx, ny, nz), `ghost` ghost layers, ghosted and valid storage
    laid out box after box."""

    def __init__(self, sizes, ghost):
        import numpy as np
        sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32).reshape(-1, 3))
        self.sizes = sz
        self.ghost = int(ghost)
        h = _vp()
        call("pm_batch_create", ctypes.byref(h), len(sz), sz.ctypes.data_as(_ip), self.ghost)
        self.h = h
        ng, nv = ctypes.c_size_t(), ctypes.c_size_t()
        call("pm_batch_elems", self.h, ctypes.byref(ng), ctypes.byref(nv))
        self.ghosted_elems, self.valid_elems = ng.value, nv.value
        go = (ctypes.c_size_t * (len(sz) + 1))()
        vo = (ctypes.c_size_t * (len(sz) + 1))()
        call("pm_batch_offsets", self.h, go, vo)
        self.goff = list(go)
        self.voff = list(vo)

    def configure(self, apply_zplanes=0, gsrb_rows=8, blocks_per_cu=4, gsrb_blocks_per_cu=None):
        c = BatchConfig(apply_zplanes, blocks_per_cu, gsrb_rows,
                        blocks_per_cu if gsrb_blocks_per_cu is None else gsrb_blocks_per_cu)
        call("pm_batch_configure", self.h, ctypes.byref(c))

    def config(self):
        c = BatchConfig()
        call("pm_batch_get_config", self.h, ctypes.byref(c))
        return c

    def autotune(self, stencil, src, dst, phi_out, rhs, dx, reps=5, stream=None):
        """Fastest configuration of each operator (kept); returns (config, ms[2])."""
        c = BatchConfig()
        ms = (ctypes.c_float * 2)()
        call("pm_batch_autotune", self.h, stencil.h, src, dst, phi_out, rhs, ctypes.c_double(dx), reps,
             ctypes.byref(c), ms, stream)
        return c, (ms[0], ms[1])

    def apply(self, stencil, src, dst, scale=1.0, stream=None):
        call("pm_batch_apply", self.h, stencil.h, src, dst, ctypes.c_double(scale), stream)

    def gsrb(self, dtype, phi_in, phi_out, rhs, dx, stream=None):
        call("pm_batch_gsrb", self.h, dtype, phi_in, phi_out, rhs, ctypes.c_double(dx), stream)

    def sweep_form(self, form, waves=0):
        """Kernel of gsrb(): 0 automatic, 1 LDS ring, 2 row sweep, 3 row sweep
        with 2-row sub-tiles, 4 row sweep writing its rows' x ghost cells, 5
        pair sweep (two cells per lane), 6 pair sweep writing its rows' ghost
        pairs, 7 quad sweep (FP32, ghost 2, four cells per lane), 8 quad sweep
        with non-temporal stores; waves per workgroup for forms 0, 2, 4 and 5
        (0 automatic)."""
        call("pm_batch_sweep_form", self.h, int(form), int(waves))
        return self

    def ghosted_view(self, flat, b):
        """Box b of a flat ghosted array, shaped (z, y, x)."""
        g = self.ghost
        nx, ny, nz = (int(v) for v in self.sizes[b])
        return flat[self.goff[b]:self.goff[b + 1]].reshape(nz + 2 * g, ny + 2 * g, nx + 2 * g)

    def valid_view(self, flat, b):
        nx, ny, nz = (int(v) for v in self.sizes[b])
        return flat[self.voff[b]:self.voff[b + 1]].reshape(nz, ny, nx)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pm_batch_destroy(self.h)
            self.h = None


class RunStopped(Exception):
    """pm_server_run stopped at its set iteration count (PM_ERR_CANCELLED);
    .stats holds the statistics so far and the server holds the run."""

    def __init__(self, stats):
        super().__init__("server run stopped and held")
        self.stats = stats


class Server:
    """Continuous-batching Poisson solve server (pm_server).  mode 0: batched
    FP32-preconditioned FP64 refinement; mode 1: plain FP64, one at a time."""

    def __init__(self, mode=0, hbm_budget=200 << 30, max_active=64, max_iter=30, sigma=0.05, pre=2, post=2,
                 bottom=10, inner=1):
        o = ServerOptions()
        lib().pm_server_default_options(ctypes.byref(o))
        o.mode, o.hbm_budget, o.max_active, o.max_iter, o.sigma = mode, hbm_budget, max_active, max_iter, sigma
        o.pre, o.post, o.bottom, o.inner = pre, post, bottom, inner
        h = _vp()
        call("pm_server_create", ctypes.byref(h), ctypes.byref(o))
        self.h = h

    def request_bytes(self, n):
        b = ctypes.c_size_t()
        call("pm_server_request_bytes", self.h, n, ctypes.byref(b))
        return b.value

    def resident_bytes(self):
        b = ctypes.c_size_t()
        call("pm_server_resident_bytes", self.h, ctypes.byref(b))
        return b.value

    def run(self, requests, stream=None):
        """requests: list of dicts (arrival_s, n, ngauss, seed, tol, keep_result);
        returns a list of stats dicts in the same order."""
        arr = (Request * len(requests))()
        for i, r in enumerate(requests):
            arr[i] = Request(float(r.get("arrival_s", 0.0)), int(r["n"]), int(r["ngauss"]), int(r["seed"]),
                             float(r.get("tol", 1e-8)), int(r.get("keep_result", 0)))
        st = (RequestStats * len(requests))()
        rc = lib().pm_server_run(self.h, len(requests), arr, st, stream)
        if rc != PM_ERR_CANCELLED:
            check(rc, "pm_server_run")
        out = [{k: getattr(x, k) for k, _ in RequestStats._fields_} for x in st]
        if rc == PM_ERR_CANCELLED:
            raise RunStopped(out)
        return out

    def set_stop(self, iterations):
        """Stop the next run after `iterations` loop iterations (0: never)."""
        call("pm_server_set_stop", self.h, int(iterations))

    def discard(self):
        call("pm_server_discard", self.h)

    def checkpoint(self, stream=None):
        """The held run as bytes (pm_server_checkpoint)."""
        b = ctypes.c_size_t()
        call("pm_server_checkpoint_size", self.h, ctypes.byref(b))
        buf = ctypes.create_string_buffer(b.value)
        call("pm_server_checkpoint", self.h, buf, b.value, stream)
        return buf.raw

    def restore(self, data, stream=None):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        call("pm_server_restore", self.h, buf, len(data), stream)

    def result(self, index, n):
        import numpy as np
        out = np.empty(n ** 3)
        call("pm_server_result", self.h, index, out.ctypes.data_as(_vp))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pm_server_destroy(self.h)
            self.h = None
