"""Runs the C++ API programs of tests/cpp (built by __graft_entry__.build())
on the GPU: Proto-style user code over proto_amd/cpp and libprotomi.so, and
Proto's own operator headers compiled against it (tests/cpp/ref)."""
import os
import subprocess

import pytest

import util

pytestmark = pytest.mark.gpu

BIN = os.path.join(util.ROOT, "tests", "cpp", "bin")


ARGS = {"test_api_interp": [os.path.join(util.GOLDEN, "interp_hi.bin"), os.path.join(util.GOLDEN, "interp_bnd.bin")],
        "test_api_boxop": [os.path.join(util.GOLDEN, "boxop.bin")]}


# test_api_boxop compiles Proto's own operator headers (BoxOp_Laplace.H and the
# BoxOp_Test*.H of Proto's tests) unchanged against proto_amd/cpp; it is built
# by __graft_entry__.build() where the reference tree is present.
@pytest.mark.parametrize("name", ["test_api_stencil", "test_api_level", "test_api_interp", "test_api_boxop"])
def test_cpp_program(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "C++ test program not built: run __graft_entry__.build()"
    r = subprocess.run([exe] + ARGS.get(name, []), capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
