#!/usr/bin/env python3
"""Benchmark: the config-2 stencil workload on one MI355X.

Workload ("laplacian512_fused_residual"): DIM=3, FP64, one box of 512^3
interior cells with 2 ghost layers; u = sin sin sin + 1e-3 U (seed 1), f = U
(seed 1); one step computes, for the 7-point Laplacian() and the 27-point
Laplacian_27() with scale 1/h^2 (h = 1/512), r = f - L(u) (Proto's forall
axpy) plus absMax(r) and sum(r) -- both operators in one pass over u and f.

value = algorithmic bytes of the step (u read once over its ghosted box, f
read once, r7 and r27 written once: 4 320 330 240 B) / step time, in GB/s.
Inputs are resident in HBM before the timed region.

roofline: the fused kernel's own average duration (HIP events on its stream,
launches without the tiny final-reduction kernel), against 8 TB/s.
cpu_baseline: Proto's own CPU path (oracle/_ref, Release flags, one thread)
running the same program on a bounded z-slab of the same field; its
"all_cores" entry runs 16 such slabs in 16 host processes at once.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the fused residual kernel (rocprofv3 FETCH_SIZE x2.0 +
# WRITE_SIZE x1.0, calibrated on a 1 GiB copy; 4.50-4.53 GB over the z chunks
# 16..256 of the 4-row, 2-plane-prefetch, non-temporal configuration)
TRAFFIC_RESIDUAL2 = 4.52e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-planes", type=int, default=512, help="z-planes of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--only", default="", help="profile helper: run only this kernel family")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default tile configurations)")
    ap.add_argument("--no-mg", action="store_true", help="skip the config-3 multigrid measurement")
    ap.add_argument("--mg-n", type=int, default=512)
    ap.add_argument("--mg-graph", type=int, default=0,
                    help="replay each multigrid cycle as a captured graph (the other form is timed beside it; "
                         "the two measure within 1 %% of each other)")
    ap.add_argument("--serve-n", type=int, default=300, help="requests in the config-5 stream (0 = skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the config-4 batched-patch measurement")
    return ap.parse_args()


def _proto_stencil(name):
    """Coefficients and offsets of one of Proto's library stencils (3-D, FP64),
    from the table Proto itself wrote (tests/golden/stencil_lib.json)."""
    import numpy as np
    with open(os.path.join(ROOT, "tests", "golden", "stencil_lib.json")) as fp:
        e = json.load(fp)["3D/%s/0/f64" % name]
    return (np.array([float.fromhex(c) for c in e["coefs"]]),
            np.array(e["offsets"], dtype=np.int32).reshape(-1, 3))


def stencils(dtype):
    import numpy as np
    import proto_amd as pa
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    # Proto's Laplacian() and Laplacian_27() coefficient lists and orders
    # (Proto_StencilDefs.H:77-87, 130-149), built the way the library does.
    c7 = np.array([-6.0, 1, 1, 1, 1, 1, 1], dtype=npdt)
    o7 = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    c27 = [npdt(-64.0 / 15.0)]
    o27 = [[0, 0, 0]]
    w = {1: 7.0 / 15.0, 2: 1.0 / 10.0, 3: 1.0 / 30.0}
    for z in (-1, 0, 1):
        for y in (-1, 0, 1):
            for x in (-1, 0, 1):
                s = abs(x) + abs(y) + abs(z)
                if s:
                    c27.append(npdt(w[s]))
                    o27.append([x, y, z])
    return pa.Stencil(c7, o7, dtype), pa.Stencil(np.array(c27, dtype=npdt), o27, dtype)


def time_launches(fn, reps, stream):
    """Average duration (ms) of fn() launched reps times, by HIP events on stream."""
    from proto_amd.device import Event
    fn()
    stream.sync()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.sync()
    return e0.elapsed_ms(e1) / reps


def _cpu_slab(n, g, nz, t_go=0.0):
    """Proto's CPU path (oracle/_ref, or the C restatement where the reference
    is not built) on a z-slab of nz planes of the config-2 field: both
    operators, forall residual, absMax, sum.  Returns (wall start, wall end,
    algorithmic bytes, kind); inputs are prepared outside the timed region,
    which starts no earlier than the wall-clock time t_go."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import util
    lib = util.ref(3)
    kind = "reference"
    x = (np.arange(n) + 0.5) / n
    s = np.sin(2 * np.pi * x)
    noise = np.zeros(n * n * nz)
    util.oracle().orc_uniform_f64(1, 0, noise.size, util.ptr(noise))  # same stream as the device inputs
    noise = noise.reshape(nz, n, n)
    smooth = (s[None, None, :] * s[None, :, None]) * s[:nz, None, None]
    ui = smooth + 1e-3 * noise
    u = np.pad(ui, ((g, g), (g, g), (g, g)), mode="wrap")  # periodic ghost fill (of the slab when nz < n)
    u = np.ascontiguousarray(u.reshape(-1))
    f = np.zeros(n * n * nz)
    util.oracle().orc_uniform_f64(1, 1, f.size, util.ptr(f))
    r = np.zeros(n * n * nz)
    ulo, uhi = util.i3([-g, -g, -g]), util.i3([n - 1 + g, n - 1 + g, nz - 1 + g])
    lo, hi = util.i3([0, 0, 0]), util.i3([n - 1, n - 1, nz - 1])
    h = 1.0 / n
    am, sm = ctypes.c_double(), ctypes.c_double()
    time.sleep(max(0.0, t_go - time.time()))
    t0 = time.time()
    if lib is not None:
        for which in (0, 1):
            lib.ref_config2_step(which, util.ptr(u), util.ptr(ulo), util.ptr(uhi), util.ptr(f), util.ptr(r),
                                 util.ptr(lo), util.ptr(hi), ctypes.c_double(1.0 / (h * h)), ctypes.byref(am),
                                 ctypes.byref(sm))
    else:
        kind = "port"
        o = util.oracle()
        for name in ("Laplacian", "Laplacian_27"):
            c, off, _, _, _ = util.lib_stencil(name)
            o.orc_residual_f64(len(c), util.ptr(c), util.ptr(off.reshape(-1).copy()), util.ptr(u), util.ptr(ulo),
                               util.ptr(uhi), util.ptr(f), util.ptr(r), util.ptr(lo), util.ptr(hi),
                               ctypes.c_double(1.0 / (h * h)))
            o.orc_reduce_f64(2, util.ptr(r), r.size)
            o.orc_reduce_f64(3, util.ptr(r), r.size)
    t1 = time.time()
    nbytes = 8 * ((n + 2 * g) * (n + 2 * g) * (nz + 2 * g) + 3 * n * n * nz)
    return t0, t1, nbytes, kind


def _cpu_slab_job(job):
    return _cpu_slab(*job)


def _host_info():
    """CPU model and core counts of this host (the box running the baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count() or 1}


def cpu_baseline(n, g, planes, workers=16, other_workloads=False):
    """Proto's CPU path on a z-slab of `planes` planes of the same config-2
    field, one thread; plus the whole field split into `workers` z-slabs run by
    as many host processes at once (Proto's CPU path is single-threaded, so
    cores are used one process per slab); with other_workloads, also Proto's
    CPU path on bounded samples of the multigrid and batched workloads."""
    import concurrent.futures
    import multiprocessing
    t0, t1, nbytes, kind = _cpu_slab(n, g, planes)
    dt = t1 - t0
    out = {"value": round(nbytes / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind,
           "sample": "config-2 step (7- and 27-point r = f - L(u), absMax, sum) on %dx%dx%d cells, "
                     "%.2f s, Proto's CPU path single-threaded (Release flags)" % (n, n, planes, dt)}
    host = _host_info()
    out["host"] = dict(host, all_cores_workers=min(workers, host["nproc"]),
                       why="one GPU's share of the box's host cores is 16; nproc counts the whole host")
    workers = max(1, min(workers, host["nproc"]))
    if workers > 1:
        nz = max(1, planes // 4)
        ctx = multiprocessing.get_context("spawn")  # fresh interpreters: no GPU state in the children
        t_go = time.time() + 12.0  # every process has its inputs ready by then
        with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
            res = list(ex.map(_cpu_slab_job, [(n, g, nz, t_go)] * workers))
        span = max(r[1] for r in res) - min(min(r[0] for r in res), t_go)
        tot = sum(r[2] for r in res)
        out["all_cores"] = {"value": round(tot / span / 1e9, 4), "unit": "GB/s", "cores": workers,
                            "kind": res[0][3],
                            "sample": "%d host processes, each the same step on its own %dx%dx%d z-slab "
                                      "(%d planes in all), started together, %.2f s to the last end"
                                      % (workers, n, n, nz, workers * nz, span)}
    if other_workloads:
        out["other_workloads"] = cpu_reference()
    return out


def cpu_reference(nbox=200):
    """Part of the CPU baseline: Proto's CPU path (oracle/_ref, Release flags,
    one thread) on bounded samples of the multigrid and batched workloads,
    reported beside the GPU figures: the plain FP64 GSRB V(2,2) correction-scheme cycle at 256^3 in
    64^3 boxes and the library FAS cycle at 128^3 in 64^3 boxes (two cycles
    each; per-solve figures for 512^3 are extrapolated by cell count and cycle
    count and labelled so), and the Laplacian_27 apply on the first `nbox`
    boxes of the batched size list."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import util
    from proto_amd.inputs import batch_sizes
    R3 = util.ref(3)
    if R3 is None:
        return {"note": "oracle/_ref (Proto's CPU build) is not present"}
    out = {"cores": 1, "kind": "reference"}
    n = 256
    rhs = util.gaussian_rhs(n, 16, 2)
    phi, hist = np.zeros(n ** 3), np.zeros(8)
    t0 = time.perf_counter()
    it = R3.ref_mg_gsrb_solve(n, 64, 6, 2, 2, 10, 0, util.ptr(rhs), util.ptr(phi), 2, ctypes.c_double(1e-10),
                              util.ptr(hist))
    dt = (time.perf_counter() - t0) / max(it, 1)
    out["multigrid_gsrb_fp64"] = {
        "s_per_cycle_256": round(dt, 4), "s_per_solve_512_extrapolated": round(dt * 8 * 11, 2),
        "sample": "Proto CPU: 256^3 in 64^3 boxes, %d V(2,2) cycles to 8^3 with residual norms; "
                  "512^3 solve extrapolated as 8x cells x 11 cycles" % it}
    n = 128
    rhs = util.gaussian_rhs(n, 16, 2)
    phi, hist = np.zeros(n ** 3), np.zeros(8)
    t0 = time.perf_counter()
    it = R3.ref_mg_fas_solve(n, 64, 5, util.ptr(rhs), util.ptr(phi), 2, ctypes.c_double(1e-10), util.ptr(hist))
    dt = (time.perf_counter() - t0) / max(it, 1)
    out["library_fas_fp64"] = {
        "s_per_cycle_128": round(dt, 4), "s_per_solve_512_extrapolated": round(dt * 64 * 31, 2),
        "sample": "Proto CPU LevelMultigridSolver<BoxOp_Laplace>: 128^3 in 64^3 boxes, %d FAS cycles; "
                  "512^3 solve extrapolated as 64x cells x 31 cycles" % it}
    c, o, _, _, _ = util.lib_stencil("Laplacian_27")
    offs = np.ascontiguousarray(o.reshape(-1), dtype=np.int32)
    one = util.i3([1, 1, 1])
    zero = util.i3([0, 0, 0])
    ob = np.zeros(6, dtype=np.int32)
    total, tsum = 0, 0.0
    for k, (nx, ny, nz) in enumerate(batch_sizes(2000, (16, 24, 32, 48, 64), 3)[:nbox]):
        src = np.random.default_rng(k).uniform(-1, 1, (nx + 4) * (ny + 4) * (nz + 4))
        dst = np.zeros(nx * ny * nz)
        slo, shi = util.i3([-2, -2, -2]), util.i3([nx + 1, ny + 1, nz + 1])
        lo, hi = util.i3([0, 0, 0]), util.i3([nx - 1, ny - 1, nz - 1])
        t0 = time.perf_counter()
        R3.ref_stencil_apply(len(c), util.ptr(c), util.ptr(offs), util.ptr(one), util.ptr(one), util.ptr(zero),
                             util.ptr(src), util.ptr(slo), util.ptr(shi), util.ptr(dst), util.ptr(lo), util.ptr(hi),
                             util.ptr(lo), util.ptr(hi), 0, 1, ctypes.c_double(1.0), util.ptr(ob))
        tsum += time.perf_counter() - t0
        total += 8 * ((nx + 2) * (ny + 2) * (nz + 2) + nx * ny * nz)
    out["batched_l27_apply_fp64"] = {
        "GB/s": round(total / tsum / 1e9, 4),
        "sample": "Proto CPU Stencil::apply of Laplacian_27, box by box, first %d of the 2000 boxes (%.2f s)"
                  % (nbox, tsum)}
    R2 = util.ref(2)
    if R2 is not None:
        # DIM=2 256^2 Laplacian apply (replace, scale 1/h^2) + absMax
        n = 256
        c, o, _, _, _ = util.lib_stencil("Laplacian", dim=2)
        offs = np.ascontiguousarray(o.reshape(-1), dtype=np.int32)
        src = np.pad(util.uniform(0, 0, n * n).reshape(n, n), 1, mode="wrap").reshape(-1).copy()
        dst = np.zeros(n * n)
        slo, shi = util.i3([-1, -1, 0]), util.i3([n, n, 0])
        lo, hi = util.i3([0, 0, 0]), util.i3([n - 1, n - 1, 0])
        reps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 1.0:
            R2.ref_stencil_apply(len(c), util.ptr(c), util.ptr(offs), util.ptr(one), util.ptr(one), util.ptr(zero),
                                 util.ptr(src), util.ptr(slo), util.ptr(shi), util.ptr(dst), util.ptr(lo),
                                 util.ptr(hi), util.ptr(lo), util.ptr(hi), 0, 1, ctypes.c_double(float(n * n)),
                                 util.ptr(ob))
            R2.ref_reduce(2, util.ptr(dst), n * n)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        out["config1_2d_apply_absmax"] = {
            "us_per_step": round(1e6 * dt, 2),
            "sample": "Proto CPU DIM=2: dst |= Laplacian()(u, 1/h^2) on 256^2 (ghost 1) + absMax, %d steps" % reps}
    return out


def mg_one_solve(pa, stream, which, n=512, bs=64, nlev=6):
    """One config-3 solve for the profiler: which = mg_fp64 / mg_ir (GSRB
    V(2,2), fused sweeps, no graph) or mg_fas (LevelMultigridSolver)."""
    import numpy as np
    from proto_amd.device import DeviceArray
    from proto_amd.inputs import gaussian_level_rhs
    nb = n // bs
    lows = np.array([(x * bs, y * bs, z * bs) for z in range(nb) for y in range(nb) for x in range(nb)],
                    dtype=np.int32).reshape(-1)
    lay = ctypes.c_void_p()
    pa.call("pm_layout_create", ctypes.byref(lay), 3, pa.i3([n] * 3), pa.i3([bs] * 3), pa.i3([1] * 3), nb ** 3,
            lows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    rhs = DeviceArray.zeros(nb ** 3 * bs ** 3)
    gaussian_level_rhs(lay, n, rhs, 16, 2, stream=stream.h)
    hist = np.zeros(41)
    it = ctypes.c_int()
    if which == "mg_fas":
        h = ctypes.c_void_p()
        pa.call("pm_fas_create", ctypes.byref(h), lay, nlev + 1, 1.0 / n)
        st = DeviceArray.zeros(nb ** 3 * (bs + 4) ** 3)
        stream.sync()
        t0 = time.perf_counter()
        pa.call("pm_fas_solve", h, st.ptr, rhs.ptr, 40, 1e-10, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                ctypes.byref(it), stream.h)
        dt = time.perf_counter() - t0
        pa.call("pm_fas_destroy", h)
    else:
        opt = pa.MgOptions()
        pa.lib().pm_mg_default_options(ctypes.byref(opt))
        opt.relax, opt.pre, opt.post, opt.bottom, opt.fused = 0, 2, 2, 10, 1
        opt.precision, opt.max_iter, opt.tol, opt.use_graph = (1 if which == "mg_ir" else 0), 40, 1e-10, 0
        opt.inner = 4  # the refinement as config 3's headline runs it: four FP32 V-cycles per step
        h = ctypes.c_void_p()
        pa.call("pm_mg_create", ctypes.byref(h), lay, nlev, 1.0 / n, ctypes.byref(opt))
        phi = DeviceArray.zeros(nb ** 3 * (bs + 2) ** 3)
        stream.sync()
        t0 = time.perf_counter()
        pa.call("pm_mg_solve", h, phi.ptr, rhs.ptr, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                ctypes.byref(it), stream.h)
        dt = time.perf_counter() - t0
        pa.call("pm_mg_destroy", h)
    pa.call("pm_layout_destroy", lay)
    return {"only": which, "iters": it.value, "solve_s": round(dt, 5), "rel_residual": float(hist[it.value] / hist[0])}


def config1_2d(pa, stream, reps=200):
    """Config 1: DIM=2, 256^2 interior with one periodic ghost layer,
    dst |= Laplacian()(u, 1/h^2), then absMax(dst).  The arrays (1 MB) stay in
    cache, so the figures are times: back-to-back steps on the stream, the same
    two launches replayed as a graph, and the host-observed latency of one step
    (launches + synchronize)."""
    import numpy as np
    from proto_amd.device import DeviceArray
    from proto_amd.inputs import uniform
    n = 256
    h = 1.0 / n
    c = np.array([-4.0, 1, 1, 1, 1])
    o = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]]  # Laplacian() in 2D (Proto_StencilDefs.H:77-87)
    st = pa.Stencil(c, o, pa.PM_F64, 2)
    ui = uniform(n * n, 0, 0, stream=stream.h)  # U[-1,1), seed 0, stream 0, x fastest
    u = DeviceArray.from_numpy(np.pad(ui.numpy().reshape(n, n), 1, mode="wrap").reshape(-1).copy())
    ui.free()
    dst = DeviceArray(n * n)
    out = DeviceArray.zeros(1)
    sb, db = pa.Box.of([-1, -1], [n, n]), pa.Box.of([0, 0], [n - 1, n - 1])
    sp = stream.h

    def step_separate():
        pa.call("pm_stencil_apply", st.h, u.ptr, ctypes.byref(sb), dst.ptr, ctypes.byref(db), ctypes.byref(db), 1, 1,
                1.0 / (h * h), sp)
        pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, dst.ptr, n * n, out.ptr, 0, sp)

    def step():
        # the apply and absMax of the cells it writes in one pass (+ a one-block finish)
        pa.call("pm_stencil_apply_reduce", st.h, u.ptr, ctypes.byref(sb), dst.ptr, ctypes.byref(db), ctypes.byref(db),
                1, 1.0 / (h * h), pa.RED_ABS, out.ptr, 0, sp)

    us_separate = 1e3 * time_launches(step_separate, reps, stream)
    us_stream = 1e3 * time_launches(step, reps, stream)
    gr = ctypes.c_void_p()
    pa.call("pm_graph_begin", sp)
    step()
    pa.call("pm_graph_end", sp, ctypes.byref(gr))
    us_graph = 1e3 * time_launches(lambda: pa.call("pm_graph_launch", gr, sp), reps, stream)
    pa.call("pm_graph_destroy", gr)
    # the runtime's own floor: two one-element kernels per step, launched on the
    # stream and replayed as a two-node graph
    tiny = DeviceArray(1)

    def two_tiny():
        for k in range(2):
            pa.call("pm_fill_uniform", pa.PM_F64, tiny.ptr, 1, 9, k, -1.0, 1.0, sp)
    us_floor_stream = 1e3 * time_launches(two_tiny, reps, stream)
    pa.call("pm_graph_begin", sp)
    two_tiny()
    pa.call("pm_graph_end", sp, ctypes.byref(gr))
    us_floor_graph = 1e3 * time_launches(lambda: pa.call("pm_graph_launch", gr, sp), reps, stream)
    pa.call("pm_graph_destroy", gr)
    tiny.free()
    lat = []
    for _ in range(50):
        t0 = time.perf_counter()
        step()
        stream.sync()
        lat.append(time.perf_counter() - t0)
    amax = float(out.numpy()[0])
    for d in (u, dst, out):
        d.free()
    return {"workload": "laplacian2d_256_absmax", "us_per_step_stream": round(us_stream, 2),
            "us_per_step_graph": round(us_graph, 2), "us_latency_p50": round(1e6 * float(np.median(lat)), 2),
            "us_per_step_apply_then_reduce": round(us_separate, 2),
            "us_two_1elem_kernels_stream": round(us_floor_stream, 2), "us_two_1elem_kernels_graph": round(us_floor_graph, 2),
            "absmax": amax.hex(), "bytes": 8 * (258 * 258 + 2 * 256 * 256)}


def mg_config3(pa, stream, n, bs, nlev, reps=3, use_graph=1):
    """Config 3: periodic n^3 in bs^3 boxes, 16-Gaussian RHS (sigma 0.05, seed 2)
    minus its mean, GSRB V(2,2) to 8^3, relative residual 1e-10.  Plain FP64
    V-cycles vs the FP32-preconditioned FP64 refinement; solves/s from the
    wall time of complete solves (phi = 0 start, hierarchy resident)."""
    import numpy as np
    from proto_amd.device import DeviceArray
    from proto_amd.inputs import gaussian_level_rhs
    nb = n // bs
    lows = np.array([(x * bs, y * bs, z * bs) for z in range(nb) for y in range(nb) for x in range(nb)],
                    dtype=np.int32).reshape(-1)
    lay = ctypes.c_void_p()
    pa.call("pm_layout_create", ctypes.byref(lay), 3, pa.i3([n] * 3), pa.i3([bs] * 3), pa.i3([1] * 3), nb ** 3,
            lows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    rhs = DeviceArray.zeros(nb ** 3 * bs ** 3)
    gaussian_level_rhs(lay, n, rhs, 16, 2, stream=stream.h)
    nphi = nb ** 3 * (bs + 2) ** 3
    out = {"workload": "levelpoisson%d_box%d_gsrb22" % (n, bs), "levels": nlev + 1}
    sols = {}
    # fp32_ir: the refinement with four FP32 V-cycles per step (the preconditioner
    # solves A e = r to about 1e-4, so 3 FP64 residuals and updates instead of 11);
    # fp32_ir_inner1..3: one to three V-cycles per step
    forms = [("fp64_two_colour", 0, 0, use_graph), ("fp64", 0, 1, use_graph), ("fp32_ir", 1, 1, use_graph, 4)]
    # the same two solves launched eagerly (use_graph 0) or replayed as one captured graph per cycle
    forms += [("fp64_graph%d" % (1 - use_graph), 0, 1, 1 - use_graph), ("fp32_ir_graph%d" % (1 - use_graph), 1, 1,
                                                                          1 - use_graph, 4)]
    forms += [("fp32_ir_inner1", 1, 1, use_graph, 1), ("fp32_ir_inner2", 1, 1, use_graph, 2),
              ("fp32_ir_inner3", 1, 1, use_graph, 3)]
    for name, prec, fused, graph, *inner in forms:
        opt = pa.MgOptions()
        pa.lib().pm_mg_default_options(ctypes.byref(opt))
        opt.relax, opt.pre, opt.post, opt.bottom, opt.fused = 0, 2, 2, 10, fused
        opt.precision, opt.max_iter, opt.tol, opt.use_graph = prec, 40, 1e-10, graph
        opt.inner = inner[0] if inner else 1
        mg = ctypes.c_void_p()
        pa.call("pm_mg_create", ctypes.byref(mg), lay, nlev, 1.0 / n, ctypes.byref(opt))
        phi = DeviceArray.zeros(nphi)
        hist = np.zeros(41)
        it = ctypes.c_int()
        times = []
        for r in range(reps + 1):
            pa.call("pm_memset_zero", phi.ptr, phi.nbytes, stream.h)
            stream.sync()
            t0 = time.perf_counter()
            pa.call("pm_mg_solve", mg, phi.ptr, rhs.ptr, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    ctypes.byref(it), stream.h)
            times.append(time.perf_counter() - t0)
        t = sorted(times[1:])[len(times[1:]) // 2]
        nbytes = ctypes.c_size_t()
        pa.call("pm_mg_bytes", mg, ctypes.byref(nbytes))
        out[name] = {"iters": it.value, "solve_s": round(t, 5), "solves_per_s": round(1.0 / t, 3),
                     "rel_residual": float(hist[it.value] / hist[0]), "hierarchy_bytes": nbytes.value}
        out[name]["inner"] = opt.inner if prec == 1 else 0
        if name in ("fp64", "fp32_ir", "fp32_ir_inner1", "fp32_ir_inner2", "fp32_ir_inner3"):
            sols[name] = phi
        pa.call("pm_mg_destroy", mg)
    # Proto's library solver on the same problem: LevelMultigridSolver<BoxOp_Laplace>
    # (FAS, 4th-order flux-form operator, Jacobi), state with 2 ghost layers
    fas = ctypes.c_void_p()
    pa.call("pm_fas_create", ctypes.byref(fas), lay, nlev + 1, 1.0 / n)
    st = DeviceArray.zeros(nb ** 3 * (bs + 4) ** 3)
    hist = np.zeros(41)
    it = ctypes.c_int()
    times = []
    for r in range(2):
        pa.call("pm_memset_zero", st.ptr, st.nbytes, stream.h)
        stream.sync()
        t0 = time.perf_counter()
        pa.call("pm_fas_solve", fas, st.ptr, rhs.ptr, 40, 1e-10, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                ctypes.byref(it), stream.h)
        times.append(time.perf_counter() - t0)
    fb = ctypes.c_size_t()
    pa.call("pm_fas_bytes", fas, ctypes.byref(fb))
    out["library_fas_fp64"] = {"iters": it.value, "solve_s": round(times[-1], 5),
                               "solves_per_s": round(1.0 / times[-1], 3),
                               "rel_residual": float(hist[it.value] / hist[0]), "hierarchy_bytes": fb.value,
                               "operator": "BoxOp_Laplace (4th-order flux form), FAS, Jacobi 6/6/10"}
    pa.call("pm_fas_destroy", fas)
    st.free()
    # one GSRB sweep of the finest level: exchange / red / exchange / black vs one
    # pass, on phi and rhs U[-1,1) (seed 5; all-zero arrays sweep about 3 % faster,
    # so they would overstate the rate)
    g1, g0 = pa.i3([1, 1, 1]), pa.i3([0, 0, 0])
    sweeps = {}
    for tname, dt, t in (("f64", pa.PM_F64, np.float64), ("f32", pa.PM_F32, np.float32)):
        ph = DeviceArray.zeros(nphi, t)
        tm = DeviceArray.zeros(nb ** 3 * bs ** 3, t)
        rh = DeviceArray.zeros(nb ** 3 * bs ** 3, t)
        pa.call("pm_fill_uniform", dt, ph.ptr, nphi, 5, 0, -1.0, 1.0, stream.h)
        pa.call("pm_fill_uniform", dt, rh.ptr, nb ** 3 * bs ** 3, 5, 2, -1.0, 1.0, stream.h)

        def two_colour():
            for color in (0, 1):
                pa.call("pm_level_exchange", lay, dt, ph.ptr, g1, 1, stream.h)
                pa.call("pm_level_gsrb", lay, dt, ph.ptr, g1, rh.ptr, g0, 1.0 / n, color, stream.h)

        def one_pass():
            pa.call("pm_level_gsrb_sweep", lay, dt, ph.ptr, g1, tm.ptr, g0, rh.ptr, g0, 1.0 / n, stream.h)

        def one_pass_back():  # scratch (no ghosts) back into phi (one ghost layer), the solver's other half
            pa.call("pm_level_gsrb_sweep", lay, dt, tm.ptr, g0, ph.ptr, g1, rh.ptr, g0, 1.0 / n, stream.h)
        ms2 = time_launches(two_colour, 5, stream)
        ms1 = time_launches(one_pass, 5, stream)
        alg = np.dtype(t).itemsize * 3 * n ** 3  # phi read, rhs read, phi written
        sweeps[tname] = {"two_colour_ms": round(ms2, 4), "one_pass_ms": round(ms1, 4),
                         "one_pass_GB/s": round(alg / (ms1 * 1e-3) / 1e9, 1),
                         "one_pass_frac": round(alg / (ms1 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        # the one-pass sweep by kernel: LDS tiles, row registers with 4 / 8 rows
        forms = {}
        try:
            for form, fname in ((1, "tile"), (2, "rows4"), (3, "rows8")):
                pa.call("pm_level_sweep_form", form)
                msf = time_launches(one_pass, 5, stream)
                msb = time_launches(one_pass_back, 5, stream)
                forms[fname] = {"ms": round(msf, 4), "frac": round(alg / (msf * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                "back_ms": round(msb, 4)}
        finally:
            pa.call("pm_level_sweep_form", 0)
        sweeps[tname]["by_kernel"] = forms
        for d in (ph, tm, rh):
            d.free()
    out["fine_sweep"] = sweeps
    # error of the mixed-precision answer against the FP64 solve
    # error against the FP64 solve: max |b - a| over max |a|, as it stands and
    # with the difference's mean removed (on the periodic domain the solution is
    # unique up to a constant, which the residual does not see; FP32 rounding
    # gives the FP32 right-hand side a small mean that every sweep carries into
    # the correction's constant part)
    a = sols["fp64"]
    flat = pa.Box.of([0, 0, 0], [nphi - 1, 0, 0])
    m = DeviceArray.zeros(3)
    ncell = n ** 3
    for name in ("fp32_ir", "fp32_ir_inner1", "fp32_ir_inner2", "fp32_ir_inner3"):
        b = sols[name]
        pa.call("pm_boxdata_binop", 1, pa.PM_F64, b.ptr, ctypes.byref(flat), a.ptr, ctypes.byref(flat),
                ctypes.byref(flat), 1, stream.h)
        pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, b.ptr, nphi, m.ptr, 0, stream.h)
        pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, a.ptr, nphi, ctypes.c_void_p(m.ptr.value + 8), 0, stream.h)
        e, amax, _ = m.numpy(stream.h)
        # the mean over valid cells: gather the difference's valid cells on the host
        d = b.numpy(stream.h).reshape(nb, nb, nb, bs + 2, bs + 2, bs + 2)[:, :, :, 1:-1, 1:-1, 1:-1]
        dm = float(d.mean())
        out[name]["rel_error_vs_fp64"] = float(e / amax)
        out[name]["rel_error_vs_fp64_mean_removed"] = float(np.abs(d - dm).max() / amax)
        out[name]["mean_offset_vs_fp64"] = dm
    out["speedup_fp32_ir"] = round(out["fp64"]["solve_s"] / out["fp32_ir"]["solve_s"], 3)
    for k in (1, 2, 3):
        out["speedup_fp32_ir_inner%d" % k] = round(out["fp64"]["solve_s"] / out["fp32_ir_inner%d" % k]["solve_s"], 3)
    out["speedup_fp32_ir_vs_two_colour_fp64"] = round(out["fp64_two_colour"]["solve_s"] / out["fp32_ir"]["solve_s"], 3)
    pa.call("pm_layout_destroy", lay)
    return out


def batch_config4(pa, stream, dtype, reps=10, nboxes=2000, per_box=True):
    """Config 4: nboxes patches, per-axis extents iid from {16,24,32,48,64}
    (seed 3), 2 ghost layers, inputs U[-1,1) over the ghosted boxes (seed 3,
    stream 0; rhs stream 2).  Operators: Laplacian_27 apply (scale 1) and one
    GSRB sweep, each ONE persistent launch over every box, tiles chosen by the
    library's autotuner, then both captured in one graph and replayed.
    Algorithmic bytes per box: the stencil footprint (n+2)^3 read once plus
    n^3 written (apply), plus n^3 of rhs read (sweep)."""
    import numpy as np
    from proto_amd.device import DeviceArray
    from proto_amd.inputs import batch_sizes, uniform
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    esz = np.dtype(npdt).itemsize
    sp = stream.h
    sizes = batch_sizes(nboxes, (16, 24, 32, 48, 64), 3)
    B = pa.Batch(sizes, 2)
    src = uniform(B.ghosted_elems, 3, 0, npdt, sp)
    rhs = uniform(B.valid_elems, 3, 2, npdt, sp)
    dst = DeviceArray(B.valid_elems, npdt)
    phi = uniform(B.ghosted_elems, 3, 0, npdt, sp)  # the sweep's output buffer holds src's ghosts
    st27 = stencils(dtype)[1]
    dx = 1.0 / 64
    valid = int(np.prod(sizes, axis=1).sum())
    foot = int(np.prod(sizes + 2, axis=1).sum())
    apply_bytes = esz * (foot + valid)
    gsrb_bytes = esz * (foot + 2 * valid)
    # the autotuner's sweep candidates are the LDS-ring kernel's tilings: tune it
    # as that kernel, then let the batch pick its kernel again
    B.sweep_form(1)
    cfg, tms = B.autotune(st27, src.ptr, dst.ptr, phi.ptr, rhs.ptr, dx, 5, sp)
    B.sweep_form(0)
    ms_a = time_launches(lambda: B.apply(st27, src.ptr, dst.ptr, 1.0, sp), reps, stream)
    ms_g = time_launches(lambda: B.gsrb(dtype, src.ptr, phi.ptr, rhs.ptr, dx, sp), reps, stream)
    # the sweep by kernel: LDS ring, row sweep (4-row sub-tiles; 2-row; writing the rows' ghost cells), pair sweep
    forms = {}
    try:
        for form, waves, fname in ((1, 0, "lds_ring"), (2, 1, "rows4_w1"), (2, 4, "rows4_w4"), (2, 8, "rows4_w8"),
                                   (3, 0, "rows2"), (4, 1, "rows4_ghostw_w1"), (4, 4, "rows4_ghostw_w4"),
                                   (5, 1, "pairs4_w1"), (5, 2, "pairs4_w2"), (5, 4, "pairs4_w4"), (5, 8, "pairs4_w8"),
                                   (6, 0, "pairs4_ghostw"), (7, 0, "quads2"), (8, 0, "quads2_nt")):
            if form >= 7 and dtype != pa.PM_F32:
                continue  # the quad sweep is FP32
            B.sweep_form(form, waves)
            msf = time_launches(lambda: B.gsrb(dtype, src.ptr, phi.ptr, rhs.ptr, dx, sp), reps, stream)
            forms[fname] = {"ms": round(msf, 4), "frac": round(gsrb_bytes / (msf * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    finally:
        B.sweep_form(0)
    g = ctypes.c_void_p()
    pa.call("pm_graph_begin", sp)
    B.apply(st27, src.ptr, dst.ptr, 1.0, sp)
    B.gsrb(dtype, src.ptr, phi.ptr, rhs.ptr, dx, sp)
    pa.call("pm_graph_end", sp, ctypes.byref(g))
    ms_graph = time_launches(lambda: pa.call("pm_graph_launch", g, sp), reps, stream)
    pa.call("pm_graph_destroy", g)
    # Proto's pattern for comparison: one single-box apply launch per patch
    per_box = []
    for b in range(nboxes):
        nx, ny, nz = (int(v) for v in sizes[b])
        sb = pa.Box.of([-2] * 3, [nx + 1, ny + 1, nz + 1])
        db = pa.Box.of([0] * 3, [nx - 1, ny - 1, nz - 1])
        per_box.append((ctypes.c_void_p(src.ptr.value + esz * B.goff[b]), sb,
                        ctypes.c_void_p(dst.ptr.value + esz * B.voff[b]), db))

    def launch_per_box():
        for sptr, sb, dptr, db in per_box:
            pa.call("pm_stencil_apply", st27.h, sptr, ctypes.byref(sb), dptr, ctypes.byref(db), ctypes.byref(db), 1,
                    1, 1.0, sp)
    ms_pb = time_launches(launch_per_box, 2, stream) if per_box else float("nan")

    def gbs(b, ms):
        return round(b / (ms * 1e-3) / 1e9, 1)
    out = {"tuned": cfg.as_dict(), "autotune_ms": [round(t, 4) for t in tms],
           "apply_L27": {"ms": round(ms_a, 4), "GB/s": gbs(apply_bytes, ms_a),
                         "frac": round(apply_bytes / (ms_a * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "bytes": apply_bytes},
           "gsrb": {"ms": round(ms_g, 4), "GB/s": gbs(gsrb_bytes, ms_g),
                    "frac": round(gsrb_bytes / (ms_g * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "bytes": gsrb_bytes,
                    "by_kernel": forms},
           "graph_apply_plus_gsrb": {"ms": round(ms_graph, 4), "GB/s": gbs(apply_bytes + gsrb_bytes, ms_graph)},
           "apply_L27_launch_per_box": {"ms": round(ms_pb, 4), "GB/s": gbs(apply_bytes, ms_pb),
                                        "launches": nboxes}}
    for d in (src, rhs, dst, phi):
        d.free()
    return {"boxes": nboxes, "valid_cells": valid, "footprint_cells": foot, "result": out}


def serve_trace(nreq, lam=200.0, seed=4, sizes=(32, 48, 64, 96, 128, 192, 256), tol=1e-8):
    """Config 5 request stream (proto_amd.inputs.serve_trace)."""
    from proto_amd.inputs import serve_trace as trace
    return trace(nreq, lam, seed, sizes, tol)


def serve_config5(pa, stream, nreq):
    """Config 5: the served Poisson stream through the batched FP32-IR server
    and through the plain FP64 one-at-a-time solver; the same requests arriving
    at 400 and 800 req/s; plus both at saturation (every request present at
    t = 0) for the capacity in solves/s, at the stream's tolerance 1e-8 and at
    relative residual 1e-10."""
    import numpy as np
    reqs = serve_trace(nreq)

    def summary(st, wall):
        ok = [s for s in st if s["status"] == pa.SERVE_SOLVED]
        lat = np.array([s["finish_s"] - s["arrival_s"] for s in ok])
        span = max(s["finish_s"] for s in st) - min(s["arrival_s"] for s in st)
        return {"solved": len(ok), "rejected": sum(s["status"] == pa.SERVE_REJECTED for s in st),
                "not_converged": sum(s["status"] == pa.SERVE_NOT_CONVERGED for s in st),
                "solves_per_s": round(len(ok) / span, 2), "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
                "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
                "mean_iters": round(float(np.mean([s["iters"] for s in ok])), 2),
                "max_rel_residual": float(max(s["rel_residual"] for s in ok)), "wall_s": round(wall, 3)}
    out = {"workload": "poisson_stream_lambda200", "requests": nreq,
           "arrival_window_s": round(reqs[-1]["arrival_s"], 3), "tol": 1e-8}
    # batched_fp32_ir_inner4: four FP32 V-cycles per refinement step (fewer FP64
    # residuals, updates and norm read-backs per request)
    for name, mode, inner in (("batched_fp32_ir", 0, 1), ("plain_fp64", 1, 1), ("batched_fp32_ir_inner4", 0, 4)):
        srv = pa.Server(mode=mode, hbm_budget=200 << 30, max_active=64, inner=inner)
        srv.run(serve_trace(8), stream.h)  # warm-up: hierarchies and code paths
        t0 = time.perf_counter()
        st = srv.run(reqs, stream.h)
        out[name] = summary(st, time.perf_counter() - t0)
        # the same sizes and seeds arriving faster (Poisson, 400 and 800 req/s):
        # latency under load between the 200/s stream and saturation
        for lam in ((400, 800) if inner == 1 else ()):
            t0 = time.perf_counter()
            st = srv.run(serve_trace(nreq, lam=float(lam)), stream.h)
            out[name]["lambda_%d" % lam] = summary(st, time.perf_counter() - t0)
        sat = [dict(r, arrival_s=0.0) for r in reqs]
        t0 = time.perf_counter()
        st = srv.run(sat, stream.h)
        out[name]["saturated"] = summary(st, time.perf_counter() - t0)
        # the same saturated stream solved to relative residual 1e-10
        sat10 = [dict(r, arrival_s=0.0, tol=1e-10) for r in reqs]
        t0 = time.perf_counter()
        st = srv.run(sat10, stream.h)
        out[name]["saturated_tol_1e-10"] = summary(st, time.perf_counter() - t0)
        out[name]["resident_bytes"] = srv.resident_bytes()
        del srv
    b, p = out["batched_fp32_ir"], out["plain_fp64"]
    out["solves_per_s_ratio"] = round(b["solves_per_s"] / p["solves_per_s"], 3)
    out["saturated_solves_per_s_ratio"] = round(b["saturated"]["solves_per_s"] / p["saturated"]["solves_per_s"], 3)
    out["saturated_tol_1e-10_solves_per_s_ratio"] = round(
        b["saturated_tol_1e-10"]["solves_per_s"] / p["saturated_tol_1e-10"]["solves_per_s"], 3)
    b4 = out["batched_fp32_ir_inner4"]
    out["inner4_solves_per_s_ratio"] = round(b4["solves_per_s"] / p["solves_per_s"], 3)
    out["inner4_saturated_solves_per_s_ratio"] = round(b4["saturated"]["solves_per_s"] / p["saturated"]["solves_per_s"],
                                                       3)
    out["inner4_saturated_tol_1e-10_solves_per_s_ratio"] = round(
        b4["saturated_tol_1e-10"]["solves_per_s"] / p["saturated_tol_1e-10"]["solves_per_s"], 3)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path is a single-GPU kernel); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    # CPU baseline first, before this process touches the GPU (its all-cores
    # leg starts host processes)
    cpu = None if (args.no_cpu or args.only) else cpu_baseline(args.n, 2, args.cpu_planes,
                                                               other_workloads=args.n == 512)
    import numpy as np
    import proto_amd as pa
    from proto_amd.device import DeviceArray, Event, Stream, gpu_available
    from proto_amd.inputs import config2_inputs

    assert gpu_available(), "bench.py needs a GPU"
    n, g = args.n, 2
    h = 1.0 / n
    sc = 1.0 / (h * h)
    stream = Stream()
    sp = stream.h
    s7, s27 = stencils(pa.PM_F64)
    u, f = config2_inputs(n, g, np.float64, sp)
    r1 = DeviceArray(n ** 3)
    r2 = DeviceArray(n ** 3)
    red = DeviceArray.zeros(4)
    ub = pa.Box.of([-g] * 3, [n - 1 + g] * 3)
    rb = pa.Box.of([0] * 3, [n - 1] * 3)
    stream.sync()

    def step():
        pa.call("pm_stencil_residual2", s7.h, s27.h, u.ptr, ctypes.byref(ub), f.ptr, r1.ptr, r2.ptr,
                ctypes.byref(rb), sc, sc, red.ptr, sp)

    def kernel_only():
        pa.call("pm_stencil_residual2", s7.h, s27.h, u.ptr, ctypes.byref(ub), f.ptr, r1.ptr, r2.ptr,
                ctypes.byref(rb), sc, sc, None, sp)

    bytes_u = 8 * (n + 2 * g) ** 3
    bytes_f = 8 * n ** 3
    step_bytes = bytes_u + bytes_f + 2 * bytes_f

    # autotuner (outside the timed region): tile configuration per kernel and shape
    tuned = {}

    def tune(name, st, st2, mode):
        if args.no_tune:
            return
        cfg, ms = pa.TileConfig(), ctypes.c_float()
        pa.call("pm_stencil_autotune", st.h, st2.h if st2 else None, mode, ctypes.byref(ub), ctypes.byref(rb), 3,
                ctypes.byref(cfg), ctypes.byref(ms), sp)
        tuned[name] = dict(cfg.as_dict(), ms=round(ms.value, 4))

    tune("residual2", s7, s27, 3)

    if args.only == "sweep":
        # profiling helper: one-pass level GSRB sweeps of the 512^3 level in 64^3 boxes
        nb, bs = n // 64, 64
        lows = np.array([(x * bs, y * bs, z * bs) for z in range(nb) for y in range(nb) for x in range(nb)],
                        dtype=np.int32).reshape(-1)
        lay = ctypes.c_void_p()
        pa.call("pm_layout_create", ctypes.byref(lay), 3, pa.i3([n] * 3), pa.i3([bs] * 3), pa.i3([1] * 3), nb ** 3,
                lows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        g1, g0 = pa.i3([1, 1, 1]), pa.i3([0, 0, 0])
        for dt, t in ((pa.PM_F64, np.float64), (pa.PM_F32, np.float32)):
            ph = DeviceArray.zeros(nb ** 3 * (bs + 2) ** 3, t)
            tm = DeviceArray.zeros(nb ** 3 * bs ** 3, t)
            rh = DeviceArray.zeros(nb ** 3 * bs ** 3, t)
            pa.call("pm_fill_uniform", dt, ph.ptr, nb ** 3 * (bs + 2) ** 3, 5, 0, -1.0, 1.0, sp)
            pa.call("pm_fill_uniform", dt, rh.ptr, nb ** 3 * bs ** 3, 5, 2, -1.0, 1.0, sp)
            for _ in range(args.steps):  # phi (ghost 1) -> scratch (ghost 0) -> phi, as the solver alternates
                pa.call("pm_level_gsrb_sweep", lay, dt, ph.ptr, g1, tm.ptr, g0, rh.ptr, g0, 1.0 / n, sp)
                pa.call("pm_level_gsrb_sweep", lay, dt, tm.ptr, g0, ph.ptr, g1, rh.ptr, g0, 1.0 / n, sp)
            stream.sync()
            for d in (ph, tm, rh):
                d.free()
        pa.call("pm_layout_destroy", lay)
        cpy = DeviceArray(n ** 3)  # plain 1 GiB device copy to calibrate the counters
        pa.call("pm_memcpy", cpy.ptr, r1.ptr, bytes_f, 2, sp)
        stream.sync()
        print(json.dumps({"only": args.only, "steps": args.steps, "copy_bytes": 2 * bytes_f}))
        return

    if args.only in ("mg_fp64", "mg_ir", "mg_fas"):
        # profiling helper: ONE config-3 solve (512^3 in 64^3 boxes, no graph
        # replay): the fused FP64 V-cycles, the FP32-preconditioned refinement
        # or Proto's library FAS solver; then a plain 1 GiB device copy that
        # calibrates the traffic counters of a --pmc run
        out = mg_one_solve(pa, stream, args.only)
        from proto_amd.device import DeviceArray
        n_cal = 1 << 27
        a_cal, b_cal = DeviceArray.zeros(n_cal), DeviceArray(n_cal)
        pa.call("pm_memcpy", b_cal.ptr, a_cal.ptr, 8 * n_cal, 2, stream.h)
        stream.sync()
        out["copy_bytes"] = 2 * 8 * n_cal
        print(json.dumps(out))
        return

    if args.only in ("serve_batched", "serve_plain"):
        # profiling helper: the config-5 stream (300 requests at 200/s) once
        # through the batched FP32-IR server or the plain FP64 one
        srv = pa.Server(mode=0 if args.only == "serve_batched" else 1, hbm_budget=200 << 30, max_active=64)
        t0 = time.perf_counter()
        st = srv.run(serve_trace(300), stream.h)
        ok = sum(s_["status"] == pa.SERVE_SOLVED for s_ in st)
        print(json.dumps({"only": args.only, "solved": ok, "wall_s": round(time.perf_counter() - t0, 3)}))
        return

    if args.only == "batch":
        # profiling helper: the config-4 batched operators (both precisions)
        # and a plain 1 GiB device copy to calibrate the counters
        for dt in (pa.PM_F64, pa.PM_F32):
            batch_config4(pa, stream, dt, reps=args.steps, per_box=False)
        cpy = DeviceArray(n ** 3)
        pa.call("pm_memcpy", cpy.ptr, r1.ptr, bytes_f, 2, sp)
        stream.sync()
        print(json.dumps({"only": args.only, "steps": args.steps, "copy_bytes": 2 * bytes_f}))
        return

    if args.only:
        # profiling helper: the fused kernel and a plain 1 GiB device copy (the
        # runtime's blit kernel, a known byte count to calibrate the counters)
        cpy = DeviceArray(n ** 3)
        fns = {"residual2": [kernel_only], "step": [step],
               "copy": [lambda: pa.call("pm_memcpy", cpy.ptr, r1.ptr, bytes_f, 2, sp)],
               "profile": [kernel_only, lambda: pa.call("pm_memcpy", cpy.ptr, r1.ptr, bytes_f, 2, sp)]}
        for _ in range(args.warmup + args.steps):
            for fn in fns[args.only]:
                fn()
        stream.sync()
        print(json.dumps({"only": args.only, "steps": args.steps, "copy_bytes": 2 * bytes_f}))
        return

    for _ in range(args.warmup):
        step()
    stream.sync()
    t0 = time.perf_counter()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    e1.sync()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_ms(e1) / args.steps
    value = step_bytes / (ms_step * 1e-3) / 1e9

    k_ms = time_launches(kernel_only, max(5, args.steps), stream)
    cur = pa.TileConfig()
    pa.call("pm_stencil_get_tile_config", s7.h, s27.h, 3, ctypes.byref(rb), ctypes.byref(cur))
    kernel_name = "k_zmarch<double, 7, 3, %d, %d, %d> (7+27-point fused residual, z chunk %d)" % (
        cur.rows, cur.prefetch, cur.nt_store, cur.zplanes)
    achieved = step_bytes / (k_ms * 1e-3) / 1e9
    result = {
        "metric": "stencil apply GB/s (% HBM roofline); MG solves/s + p99 latency vs plain FP64",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "laplacian512_fused_residual", "box": [n, n, n], "ghost": g,
                   "operators": ["Laplacian (7-pt)", "Laplacian_27 (27-pt)"],
                   "fused": "r = f - L(u), absMax(r), sum(r) for both operators, one pass",
                   "step_bytes": step_bytes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": TRAFFIC_RESIDUAL2,
                     "traffic_source": "profiles/zmarch_residual2_laplacian512_f64.txt (FETCH x2.0 + WRITE x1.0)",
                     "kernel": kernel_name, "kernel_ms": round(k_ms, 4)},
        "wall_s": round(wall, 3),
        "tuned": tuned,
    }

    if not args.no_extra:
        extra = {}
        dst = DeviceArray(n ** 3)
        tune("apply_L7_f64", s7, None, 0)
        tune("apply_L27_f64", s27, None, 0)
        tune("residual_L7_f64", s7, None, 2)
        tune("residual_L27_f64", s27, None, 2)
        apply_bytes = bytes_u + bytes_f
        # Stencil<double>::Identity() on the same boxes runs the same z-march kernel
        # with its one centre tap: the same halo'd planes staged, the same stores,
        # so its rate is what the 7-point apply's own access pattern gives
        sid = pa.Stencil(np.array([1.0]), [[0, 0, 0]])
        tune("apply_identity_f64", sid, None, 0)
        for name, st in (("apply_L7_f64", s7), ("apply_L27_f64", s27), ("apply_identity_f64", sid)):
            ms = time_launches(lambda: pa.call("pm_stencil_apply", st.h, u.ptr, ctypes.byref(ub), dst.ptr,
                                               ctypes.byref(rb), ctypes.byref(rb), 1, 1, sc, sp), 10, stream)
            gbs = apply_bytes / (ms * 1e-3) / 1e9
            extra[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "bytes": apply_bytes}
        extra["apply_L7_f64"]["vs_identity"] = round(extra["apply_identity_f64"]["ms"] / extra["apply_L7_f64"]["ms"], 4)
        # stencils outside the two fixed shapes: the generic z-march (taps from an
        # LDS plane ring, any order) against the per-cell kernel it replaced, at
        # the same 512^3 box: Laplacian_19, and the 7-point Laplacian served by
        # each family
        gen = {}
        for sname, kinds in (("Laplacian_19", (1, 0)), ("Laplacian", (7, 1, 0))):
            c_, o_ = _proto_stencil(sname)
            for kind in kinds:
                stg = pa.Stencil(c_, o_).set_kind(kind)
                ms = time_launches(lambda: pa.call("pm_stencil_apply", stg.h, u.ptr, ctypes.byref(ub), dst.ptr,
                                                   ctypes.byref(rb), ctypes.byref(rb), 1, 1, sc, sp), 5, stream)
                gbs = apply_bytes / (ms * 1e-3) / 1e9
                gen["%s_kind%d" % (sname, kind)] = {"ms": round(ms, 4), "GB/s": round(gbs, 1),
                                                    "frac": round(gbs / PEAK_HBM_GBS, 4)}
        extra["generic_apply_f64"] = gen
        for name, st in (("residual_L7_f64", s7), ("residual_L27_f64", s27)):
            ms = time_launches(lambda: pa.call("pm_stencil_residual", st.h, u.ptr, ctypes.byref(ub), f.ptr,
                                               dst.ptr, ctypes.byref(rb), sc, None, sp), 10, stream)
            b = apply_bytes + bytes_f
            gbs = b / (ms * 1e-3) / 1e9
            extra[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "bytes": b}
        dst.free()
        # FP32 variants of the plain applies
        u32, f32 = config2_inputs(n, g, np.float32, sp)
        d32 = DeviceArray(n ** 3, np.float32)
        t7, t27 = stencils(pa.PM_F32)
        tune("apply_L7_f32", t7, None, 0)
        tune("apply_L27_f32", t27, None, 0)
        for name, st in (("apply_L7_f32", t7), ("apply_L27_f32", t27)):
            ms = time_launches(lambda: pa.call("pm_stencil_apply", st.h, u32.ptr, ctypes.byref(ub), d32.ptr,
                                               ctypes.byref(rb), ctypes.byref(rb), 1, 1, sc, sp), 10, stream)
            b = apply_bytes // 2
            gbs = b / (ms * 1e-3) / 1e9
            extra[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "bytes": b}
        u32.free()
        f32.free()
        d32.free()
        # plain device copies of the same size (achievable-bandwidth probes): the
        # runtime's copy kernel and this library's streaming copy (plain / NT stores)
        a = DeviceArray(n ** 3)
        for name, fn in (("copy_runtime", lambda: pa.call("pm_memcpy", a.ptr, r1.ptr, bytes_f, 2, sp)),
                         ("copy_stream", lambda: pa.call("pm_copy_device", a.ptr, r1.ptr, bytes_f, 0, sp)),
                         ("copy_stream_nt", lambda: pa.call("pm_copy_device", a.ptr, r1.ptr, bytes_f, 1, sp))):
            ms = time_launches(fn, 10, stream)
            gbs = 2 * bytes_f / (ms * 1e-3) / 1e9
            extra[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "bytes": 2 * bytes_f}
        # streaming ceilings over 4 GiB (16 workgroups per CU): read-only,
        # write-only (non-temporal), copy
        big = DeviceArray(4 * n ** 3)
        big2 = DeviceArray(4 * n ** 3)
        nb4 = 4 * bytes_f
        for name, kind, moved in (("stream_read", 0, nb4), ("stream_write", 1, nb4), ("stream_copy", 2, 2 * nb4)):
            ms = time_launches(lambda: pa.call("pm_stream_probe", kind, big2.ptr, big.ptr, nb4, 16, sp), 10, stream)
            gbs = moved / (ms * 1e-3) / 1e9
            extra[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "bytes": moved}
        # the 1:1 read:write mix of the stencil (2 GiB read, 2 GiB written) by three
        # copy forms and four grid sizes; the best is the achievable ceiling the
        # 7-point apply is read against
        sweep = {}
        for kind, kname in ((2, "gridstride_nt"), (3, "chunk64k_nt"), (4, "chunk64k")):
            for bpc in (2, 4, 8, 16):
                ms = time_launches(lambda: pa.call("pm_stream_probe", kind, big2.ptr, big.ptr, nb4 // 2, bpc, sp),
                                   10, stream)
                sweep["%s_bpc%d" % (kname, bpc)] = round(nb4 / (ms * 1e-3) / 1e9, 1)
        best = max(sweep, key=sweep.get)
        l7 = extra["apply_L7_f64"]["GB/s"]
        # read-only streams over the same 4 GiB at four grid sizes: the fastest
        # traffic HBM gives, which bounds any mix with writes
        rsweep = {}
        for bpc in (2, 4, 8, 16, 32):
            ms = time_launches(lambda: pa.call("pm_stream_probe", 0, big2.ptr, big.ptr, nb4, bpc, sp), 10, stream)
            rsweep["read_bpc%d" % bpc] = round(nb4 / (ms * 1e-3) / 1e9, 1)
        rbest = max(rsweep, key=rsweep.get)
        extra["copy_sweep"] = {"GB/s": sweep, "best": best, "best_GB/s": sweep[best],
                               "best_frac": round(sweep[best] / PEAK_HBM_GBS, 4),
                               "apply_L7_f64_vs_best_copy": round(l7 / sweep[best], 4),
                               "read_only_GB/s": rsweep, "best_read": rbest, "best_read_GB/s": rsweep[rbest],
                               "best_read_frac": round(rsweep[rbest] / PEAK_HBM_GBS, 4),
                               "apply_L7_f64_vs_best_read": round(l7 / rsweep[rbest], 4)}
        big.free()
        big2.free()
        a.free()
        extra["config1_2d"] = config1_2d(pa, stream)
        result["extra"] = extra

    if not args.no_batch:
        ex = result.setdefault("extra", {})
        for name, dt in (("batched_config4_f64", pa.PM_F64), ("batched_config4_f32", pa.PM_F32)):
            ex[name] = batch_config4(pa, stream, dt)

    if args.serve_n > 0:
        result.setdefault("extra", {})["served_config5"] = serve_config5(pa, stream, args.serve_n)

    if not args.no_mg:
        result.setdefault("extra", {})["multigrid_config3"] = mg_config3(pa, stream, args.mg_n, 64, 6 if args.mg_n == 512 else 3,
                                                                  use_graph=args.mg_graph)

    if cpu is not None:
        result["cpu_baseline"] = cpu
    print(json.dumps(result))


if __name__ == "__main__":
    main()
