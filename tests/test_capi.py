"""CPU checks of the C ABI: the library loads, exports every function that
include/*.h declares, and rejects bad arguments before touching a device."""
import ctypes
import glob
import os
import re

import pytest

import proto_amd as pa
import util


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(util.ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = pa.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and the loader knows the signature of each
    assert names <= set(pa.SIGNATURES), sorted(names - set(pa.SIGNATURES))


def test_version_and_device_count_without_gpu():
    assert b"gfx950" in pa.lib().pm_version()
    n = ctypes.c_int(-1)
    assert pa.lib().pm_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_stencil_create_validation():
    L = pa.lib()
    h = ctypes.c_void_p()
    c = (ctypes.c_double * 1)(1.0)
    o = (ctypes.c_int * 3)(0, 0, 0)
    one = pa.i3([1, 1, 1])
    zero = pa.i3([0, 0, 0])
    assert L.pm_stencil_create(ctypes.byref(h), 7, 3, 1, c, o, one, one, zero) == -1
    assert b"dtype" in L.pm_last_error()
    assert L.pm_stencil_create(ctypes.byref(h), pa.PM_F64, 4, 1, c, o, one, one, zero) == -1
    assert L.pm_stencil_create(ctypes.byref(h), pa.PM_F64, 3, 344, c, o, one, one, zero) == -1
    bad = pa.i3([0, 1, 1])
    assert L.pm_stencil_create(ctypes.byref(h), pa.PM_F64, 3, 1, c, o, bad, one, zero) == -1
    o2 = (ctypes.c_int * 3)(0, 0, 1)
    assert L.pm_stencil_create(ctypes.byref(h), pa.PM_F64, 2, 1, c, o2, one, one, zero) == -1


def test_reduce_and_region_validation():
    L = pa.lib()
    assert L.pm_reduce(9, pa.PM_F64, None, 0, None, 0, None) == -1
    assert L.pm_reduce(pa.RED_SUM, 5, None, 0, None, 0, None) == -1
    b = pa.Box.of([0, 0, 0], [3, 3, 3])
    r = pa.Box.of([0, 0, 0], [4, 3, 3])
    p = ctypes.c_void_p(16)
    assert L.pm_boxdata_scalar(4, pa.PM_F64, p, ctypes.byref(b), ctypes.byref(r), 1.0, 1, None) == -1
    assert b"outside" in L.pm_last_error()
    assert L.pm_boxdata_binop(7, pa.PM_F64, p, ctypes.byref(b), p, ctypes.byref(b), ctypes.byref(b), 1, None) == -1


def test_batch_validation():
    L = pa.lib()
    h = ctypes.c_void_p()
    bad = (ctypes.c_int * 6)(8, 8, 8, 8, 0, 8)
    assert L.pm_batch_create(ctypes.byref(h), 2, bad, 2) == -1
    assert b"box 1" in L.pm_last_error()
    good = (ctypes.c_int * 3)(8, 8, 8)
    assert L.pm_batch_create(ctypes.byref(h), 1, good, 0) == -1
    assert L.pm_batch_create(ctypes.byref(h), 1, good, 5) == -1
    assert L.pm_batch_create(ctypes.byref(h), 0, good, 1) == -1
    assert L.pm_batch_apply(None, None, None, None, 1.0, None) == -1
    assert L.pm_batch_configure(None, None) == -1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(pa, "_lib", None)
    monkeypatch.setattr(pa, "LIB_PATH", "/nonexistent/libprotomi.so")
    with pytest.raises(pa.NativeError):
        pa.lib()
