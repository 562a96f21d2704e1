"""The benchmark's batched and served workloads at their full size, under
assertion (the benchmark only times them):

- 2000 uneven patches (per-axis extents iid from {16, 24, 32, 48, 64}, seed 3,
  2 ghost layers, inputs U[-1,1) seed 3), FP64 and FP32: the Laplacian_27 apply
  and one GSRB sweep, each one persistent launch with the autotuned tiles,
  captured together in one graph and replayed; every box of every replay bit
  for bit against the oracle (Proto's hostApply order, Proto_StencilImplem.H:
  725-828; the GSRB sweep of LaplaceOp.H:95-123).
- The served Poisson stream (300 requests at 200/s, edges {32 ... 256}, 4-32
  Gaussians, tolerance 1e-8) through the batched FP32-preconditioned server:
  every request solved to its tolerance, and one request of each size bit for
  bit equal to the same problem solved alone by the multigrid driver."""
import ctypes

import numpy as np
import pytest

import proto_amd as pa
import util
from gpu_util import say
from proto_amd.device import DeviceArray, Stream
from proto_amd.inputs import batch_sizes, serve_trace, uniform
from test_gpu_batch import oracle_apply_batch, oracle_gsrb_batch
from test_gpu_serve import mg_solve

pytestmark = pytest.mark.gpu

NPDT = {pa.PM_F64: np.float64, pa.PM_F32: np.float32}


def _first_bad(B, got, want, valid):
    offs = B.voff if valid else B.goff
    for b in range(len(B.sizes)):
        if got[offs[b]:offs[b + 1]].tobytes() != want[offs[b]:offs[b + 1]].tobytes():
            return b
    return None


@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
def test_batched_2000_patches_every_box(dt):
    t = NPDT[dt]
    s = Stream()
    sizes = batch_sizes(2000, (16, 24, 32, 48, 64), 3)
    B = pa.Batch(sizes, 2)
    say("batch 2000 boxes", dt, "cells", int(np.prod(sizes, axis=1).sum()))
    src = uniform(B.ghosted_elems, 3, 0, t, s.h)
    rhs = uniform(B.valid_elems, 3, 2, t, s.h)
    phi = uniform(B.ghosted_elems, 3, 0, t, s.h)  # the sweep's output holds src's ghosts
    dst = DeviceArray(B.valid_elems, t)
    coefs, offs, *_ = util.lib_stencil("Laplacian_27", 0, 3, t)
    st = pa.Stencil(coefs, offs, dt)
    dx = 1.0 / 64
    cfg, _ = B.autotune(st, src.ptr, dst.ptr, phi.ptr, rhs.ptr, dx, 2, s.h)
    say("tuned", cfg.as_dict())
    g = ctypes.c_void_p()
    pa.call("pm_graph_begin", s.h)
    B.apply(st, src.ptr, dst.ptr, 1.0, s.h)
    B.gsrb(dt, src.ptr, phi.ptr, rhs.ptr, dx, s.h)
    pa.call("pm_graph_end", s.h, ctypes.byref(g))
    src_h, rhs_h = util.uniform(3, 0, B.ghosted_elems).astype(t), util.uniform(3, 2, B.valid_elems).astype(t)
    assert src.numpy(s.h).tobytes() == src_h.tobytes()
    say("oracle apply")
    want_a = oracle_apply_batch(B, coefs, offs, src_h, 1.0)
    say("oracle gsrb")
    want_g = oracle_gsrb_batch(B, src_h, rhs_h, dx)
    try:
        for rep in range(2):
            pa.call("pm_memset_zero", dst.ptr, dst.nbytes, s.h)
            pa.call("pm_graph_launch", g, s.h)
            s.sync()
            say("replay", rep)
            got_a, got_g = dst.numpy(s.h), phi.numpy(s.h)
            assert _first_bad(B, got_a, want_a, True) is None, ("apply", rep, _first_bad(B, got_a, want_a, True))
            assert _first_bad(B, got_g, want_g, False) is None, ("gsrb", rep, _first_bad(B, got_g, want_g, False))
    finally:
        pa.call("pm_graph_destroy", g)


KEEP = (0, 1, 2, 3, 5, 6, 9)  # one request of each edge 32, 64, 192, 128, 96, 256, 48


def test_served_stream_300_requests():
    reqs = serve_trace(300, keep=KEEP)
    assert sorted({reqs[i]["n"] for i in KEEP}) == [32, 48, 64, 96, 128, 192, 256]
    srv = pa.Server(mode=0, hbm_budget=200 << 30, max_active=64)
    st = srv.run(reqs)
    solved = [x for x in st if x["status"] == pa.SERVE_SOLVED]
    say("served", len(solved), "of", len(st), "max rel residual", max(x["rel_residual"] for x in st))
    assert len(solved) == 300
    assert all(x["rel_residual"] < r["tol"] for x, r in zip(st, reqs))
    # continuous batching: some request was admitted while another was in flight
    assert any(st[i + 1]["admit_s"] < st[i]["finish_s"] for i in range(len(st) - 1))
    for i in KEEP:
        r = reqs[i]
        got = srv.result(i, r["n"])
        alone, it = mg_solve(r["n"], r["ngauss"], r["seed"], r["tol"], 1)
        say("request", i, "n", r["n"], "iters", st[i]["iters"], it)
        assert it == st[i]["iters"]
        assert got.tobytes() == alone.tobytes(), (i, np.abs(got - alone).max())
