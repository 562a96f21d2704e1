"""GPU tests of the continuous-batching solve server (pm_server).

- A served answer equals, bit for bit, the multigrid driver's FP32-preconditioned
  FP64 solve of the same problem on one periodic box (pm_mg precision 1), with
  the same V-cycle count.  This ties the batched kernels to the level path,
  whose FP64 form is bitwise Proto's.
- A request's answer is the same alone and in a mix of other requests.
- The plain FP64 mode is the multigrid driver's FP64 solve (pm_mg precision 0),
  and the two modes agree to the solve tolerance.
- Admission: over-budget requests are rejected; a budget for one hierarchy
  serialises requests; hierarchies stay resident and within the budget.
- A request that runs out of iterations reports NOT_CONVERGED."""
import ctypes

import numpy as np
import pytest

import proto_amd as pa
from gpu_util import host, say, zeros

pytestmark = pytest.mark.gpu


def levels(n):
    ns = [n]
    while ns[-1] % 2 == 0 and ns[-1] > 8:
        ns.append(ns[-1] // 2)
    return len(ns) - 1


def mg_solve(n, ngauss, seed, tol, precision, sigma=0.05, inner=1):
    """The multigrid driver on a one-box periodic layout; returns (phi n^3, iters)."""
    lay = ctypes.c_void_p()
    pa.call("pm_layout_create", ctypes.byref(lay), 3, pa.i3([n] * 3), pa.i3([n] * 3), pa.i3([1] * 3), 1,
            pa.i3([0, 0, 0]))
    rhs = zeros(n ** 3)
    pa.call("pm_server_rhs", n, ngauss, seed, sigma, rhs.ptr, None)
    opt = pa.MgOptions()
    pa.lib().pm_mg_default_options(ctypes.byref(opt))
    opt.relax, opt.pre, opt.post, opt.bottom = 0, 2, 2, 10
    opt.precision, opt.max_iter, opt.tol, opt.use_graph = precision, 30, tol, 0
    opt.inner = inner
    mg = ctypes.c_void_p()
    pa.call("pm_mg_create", ctypes.byref(mg), lay, levels(n), 1.0 / n, ctypes.byref(opt))
    phi = zeros((n + 2) ** 3)
    it = ctypes.c_int()
    hist = np.zeros(31)
    pa.call("pm_mg_solve", mg, phi.ptr, rhs.ptr, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ctypes.byref(it), None)
    g = host(phi).reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
    pa.call("pm_mg_destroy", mg)
    pa.call("pm_layout_destroy", lay)
    return np.ascontiguousarray(g).reshape(-1), it.value


def req(n, ngauss, seed, t=0.0, tol=1e-8, keep=1):
    return {"arrival_s": t, "n": n, "ngauss": ngauss, "seed": seed, "tol": tol, "keep_result": keep}


@pytest.mark.parametrize("n,ngauss,seed", [(32, 5, 11), (48, 9, 12), (64, 16, 13), (192, 6, 15), (256, 12, 14)])
def test_served_equals_driver_fp32_ir(n, ngauss, seed):
    srv = pa.Server(mode=0)
    st = srv.run([req(n, ngauss, seed)])
    say("served", n, st[0])
    assert st[0]["status"] == pa.SERVE_SOLVED and st[0]["rel_residual"] < 1e-8
    phi_s = srv.result(0, n)
    phi_m, it = mg_solve(n, ngauss, seed, 1e-8, 1)
    assert it == st[0]["iters"]
    assert phi_s.tobytes() == phi_m.tobytes(), np.abs(phi_s - phi_m).max()


@pytest.mark.parametrize("inner", [2, 4])
def test_served_inner_cycles_equal_driver(inner):
    """Several FP32 V-cycles per refinement step: the served answer and its
    step count equal the driver's with the same inner count, alone and in a
    mix of sizes (single-level and multi-level requests)."""
    reqs = [req(64, 9, 51), req(32, 5, 52, 0.0005), req(96, 12, 53, 0.001), req(8, 3, 54, 0.0015)]
    srv = pa.Server(mode=0, inner=inner)
    st = srv.run(reqs)
    for i, r in enumerate(reqs):
        say("inner", inner, "n", r["n"], st[i])
        assert st[i]["status"] == pa.SERVE_SOLVED and st[i]["rel_residual"] < 1e-8
        phi_m, it = mg_solve(r["n"], r["ngauss"], r["seed"], 1e-8, 1, inner=inner)
        assert it == st[i]["iters"]
        assert srv.result(i, r["n"]).tobytes() == phi_m.tobytes()
    one = pa.Server(mode=0).run([reqs[0]])
    assert st[0]["iters"] < one[0]["iters"]


def test_alone_equals_in_mix():
    a, b = req(48, 7, 21), req(32, 30, 22)
    srv = pa.Server(mode=0)
    sa = srv.run([a])
    alone_a = srv.result(0, 48)
    srv.run([b])
    alone_b = srv.result(0, 32)
    mix = [req(64, 4, 30, 0.0, keep=0), dict(b, arrival_s=0.001), req(96, 12, 31, 0.002, keep=0),
           dict(a, arrival_s=0.003), req(32, 8, 32, 0.004, keep=0), req(128, 20, 33, 0.005, keep=0)]
    st = srv.run(mix)
    say("mix", [(s["status"], s["iters"], round(s["finish_s"], 4)) for s in st])
    assert all(s["status"] == pa.SERVE_SOLVED for s in st)
    assert st[3]["iters"] == sa[0]["iters"]
    assert srv.result(3, 48).tobytes() == alone_a.tobytes()
    assert srv.result(1, 32).tobytes() == alone_b.tobytes()
    # continuous batching: the second arrival (1 ms) joined while the first was
    # still in flight, and the last (5 ms) while an earlier one still was
    assert st[1]["admit_s"] < st[0]["finish_s"]
    assert any(st[5]["admit_s"] < st[j]["finish_s"] for j in range(5))


def test_plain_mode_is_driver_fp64_and_modes_agree():
    n, k, seed = 48, 6, 41
    plain = pa.Server(mode=1)
    sp = plain.run([req(n, k, seed)])
    assert sp[0]["status"] == pa.SERVE_SOLVED
    phi_p = plain.result(0, n)
    phi_m, it = mg_solve(n, k, seed, 1e-8, 0)
    assert it == sp[0]["iters"] and phi_p.tobytes() == phi_m.tobytes()
    batched = pa.Server(mode=0)
    batched.run([req(n, k, seed)])
    phi_b = batched.result(0, n)
    scale = np.abs(phi_p).max()
    say("fp32-ir vs fp64 max rel diff", np.abs(phi_b - phi_p).max() / scale)
    assert np.abs(phi_b - phi_p).max() < 1e-6 * scale


def test_admission_budget_and_residency():
    probe = pa.Server(mode=0)
    b32, b64 = probe.request_bytes(32), probe.request_bytes(64)
    # room for one 32^3 hierarchy only: 64^3 is rejected, 32^3 requests take turns
    srv = pa.Server(mode=0, hbm_budget=b32 + b32 // 2)
    st = srv.run([req(32, 4, 51, 0.0), req(64, 4, 52, 0.0), req(32, 5, 53, 0.0), req(32, 6, 54, 0.0)])
    say("budget", [(s["status"], round(s["admit_s"], 4), round(s["finish_s"], 4)) for s in st])
    assert st[1]["status"] == pa.SERVE_REJECTED
    assert [st[i]["status"] for i in (0, 2, 3)] == [pa.SERVE_SOLVED] * 3
    assert st[2]["admit_s"] >= st[0]["finish_s"] and st[3]["admit_s"] >= st[2]["finish_s"]
    assert srv.resident_bytes() == b32  # one hierarchy, kept resident and reused
    assert b64 > b32


def test_not_converged_status():
    srv = pa.Server(mode=0, max_iter=2)
    st = srv.run([req(32, 4, 61, tol=1e-14)])
    assert st[0]["status"] == pa.SERVE_NOT_CONVERGED and st[0]["iters"] == 2
    with pytest.raises(pa.NativeError):
        srv.result(0, 32)
    with pytest.raises(pa.NativeError, match="even"):
        srv.run([req(33, 4, 62)])


def _ck_trace():
    # all present at t = 0 with 3 slots: some in flight, some waiting at the stop
    return [req(32, 5, 41), req(48, 9, 42), req(64, 12, 43), req(32, 20, 44), req(96, 7, 45), req(48, 3, 46)]


def test_checkpoint_resume_same_bits_and_cycles():
    """A run stopped after 3 loop iterations, written to a checkpoint, loaded
    into a new server and resumed gives every request the bits, status and
    V-cycle count of the uninterrupted run.  The in-flight solutions in the
    checkpoint are LevelBoxData<double,1>::linearOut blocks of one-box levels
    (n^3 values, x fastest)."""
    trace = _ck_trace()
    ref = pa.Server(mode=0, max_active=3)
    st_ref = ref.run(trace)
    phis = [ref.result(i, r["n"]) for i, r in enumerate(trace)]
    a = pa.Server(mode=0, max_active=3)
    a.set_stop(3)
    with pytest.raises(pa.RunStopped) as stop:
        a.run(trace)
    held = stop.value.stats
    say("stopped", [(s["status"], s["iters"]) for s in held])
    assert any(s["status"] == -1 for s in held)
    buf = a.checkpoint()
    assert buf[:8] == b"PMSRVCK1"
    del a
    b = pa.Server(mode=0, max_active=3)
    b.restore(buf)
    st = b.run(trace)
    for i, r in enumerate(trace):
        say("resumed", i, st[i]["iters"], st_ref[i]["iters"])
        assert st[i]["status"] == pa.SERVE_SOLVED and st[i]["iters"] == st_ref[i]["iters"], i
        assert st[i]["rel_residual"] == st_ref[i]["rel_residual"], i
        assert b.result(i, r["n"]).tobytes() == phis[i].tobytes(), i


def test_stop_then_continue_in_place_and_bad_checkpoints():
    """A held run continues in the same server with the next pm_server_run;
    a held run refuses another request list; restore refuses a corrupt
    buffer and a server with other sweep counts."""
    trace = _ck_trace()
    ref = pa.Server(mode=0, max_active=3)
    st_ref = ref.run(trace)
    a = pa.Server(mode=0, max_active=3)
    a.set_stop(2)
    with pytest.raises(pa.RunStopped):
        a.run(trace)
    buf = a.checkpoint()
    with pytest.raises(pa.NativeError):
        a.run(trace[:5])  # another list
    a.set_stop(0)
    st = a.run(trace)
    assert [s["iters"] for s in st] == [s["iters"] for s in st_ref]
    assert [s["rel_residual"] for s in st] == [s["rel_residual"] for s in st_ref]
    with pytest.raises(pa.NativeError):
        pa.Server(mode=0, max_active=3).restore(b"XXXXXXXX" + buf[8:])
    with pytest.raises(pa.NativeError):
        pa.Server(mode=0, max_active=3, pre=4).restore(buf)
    with pytest.raises(pa.NativeError):
        pa.Server(mode=0, max_active=3).restore(buf[:200])
