"""The Proto-side binding (proto_amd/binding/Proto_protomi.H: what Proto's
DEVICE paths call in place of their kernels) compiled against Proto's own
headers where they lie in the reference tree, and linked with --no-undefined
against libprotomi.so: the binding type-checks against Proto's real Stencil,
BoxData, LevelBoxData and DisjointBoxLayout, and every C-ABI symbol it calls
resolves in the library.  Host compilation only; nothing is run."""
import os
import subprocess
import tempfile

import pytest

import util

REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")), reason="needs the reference headers")
def test_binding_compiles_against_proto_and_links():
    src = os.path.join(util.ROOT, "tests", "cpp", "binding", "proto_binding_check.cpp")
    lib = os.path.join(util.ROOT, "proto_amd", "libprotomi.so")
    assert os.path.exists(lib), "build the library first"
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "libbinding.so")
        cmd = ["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-Wall", "-Werror", "-DDIM=3", "-DNDEBUG",
               "-DPR_DISABLE_INIT_CHECKS", "-DPR_TURN_OFF_TIMERS",
               "-isystem", os.path.join(REF, "include"), "-isystem", os.path.join(REF, "include", "base"),
               "-I", os.path.join(util.ROOT, "include"), "-I", os.path.join(util.ROOT, "proto_amd", "binding"),
               src, "-o", out, "-L", os.path.dirname(lib), "-lprotomi", "-Wl,--no-undefined"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-4000:]
        nm = subprocess.run(["nm", "-D", out], capture_output=True, text=True, check=True).stdout
        used = {l.split()[-1] for l in nm.splitlines() if " U pm_" in l}
        want = {"pm_stencil_create", "pm_stencil_apply", "pm_stencil_destroy", "pm_reduce", "pm_reduce_init",
                "pm_layout_create", "pm_level_exchange", "pm_fas_create", "pm_fas_solve"}
        assert want <= used, sorted(want - used)
