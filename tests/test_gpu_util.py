"""GPU tests of the utility entry points: device copy and zero fill, the
seeded separable initialiser, level reductions and separable updates (against
numpy on the same bits), layout queries of the multigrid hierarchy, the FAS
residual norm, and the tune-table round trip."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import proto_amd as pa
import util
from gpu_util import dev, host, say, vp, zeros
from test_gpu_level import G0, G1, Layout

pytestmark = pytest.mark.gpu

_dpp = ctypes.POINTER(ctypes.c_double)


def dptr(a):
    return a.ctypes.data_as(_dpp)


def test_copy_device_and_zero_fill():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2 ** 32, size=(1 << 18) + 12, dtype=np.uint32)  # bytes a multiple of 16
    src = dev(a)
    for nt in (0, 1):
        dst = zeros(a.size, np.uint32)
        pa.call("pm_copy_device", vp(dst), vp(src), a.nbytes, nt, None)
        assert host(dst).tobytes() == a.tobytes()
    assert pa.lib().pm_copy_device(vp(dst), vp(src), a.nbytes - 4, 0, None) == -1
    pa.call("pm_memset_zero", vp(dst), a.nbytes, None)
    assert not host(dst).any()


def test_fill_separable_matches_numpy():
    n, g = 12, 2
    x = (np.arange(n) + 0.5) / n
    tx, ty, tz = np.sin(2 * np.pi * x), np.cos(2 * np.pi * x), np.sin(4 * np.pi * x) + 2
    b = pa.Box.of([-g] * 3, [n - 1 + g] * 3)
    d = zeros(b.size())
    pa.call("pm_fill_separable", pa.PM_F64, vp(d), ctypes.byref(b), ctypes.byref(b), pa.i3([n] * 3), dptr(tx),
            dptr(ty), dptr(tz), 1e-3, 5, 2, None)
    idx = np.arange(-g, n + g) % n
    w = idx[:, None, None], idx[None, :, None], idx[None, None, :]  # z, y, x
    u = util.uniform(5, 2, n ** 3).reshape(n, n, n)
    want = (tx[w[2]] * ty[w[1]]) * tz[w[0]] + 1e-3 * u[w[0], w[1], w[2]]
    assert host(d).tobytes() == np.ascontiguousarray(want).reshape(-1).tobytes()


@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
def test_level_reduce_and_add_separable(dt):
    t = np.float64 if dt == pa.PM_F64 else np.float32
    n, bs = 16, 8
    L = Layout(n, bs)
    glob = util.uniform(91, 0, n ** 3).astype(t)
    data = dev(L.pack(glob, 1))
    out = zeros(1, t)
    say("level reduce", dt)
    for op, f in ((pa.RED_MAX, np.max), (pa.RED_MIN, np.min), (pa.RED_ABS, lambda v: np.abs(v).max())):
        pa.call("pm_level_reduce", L.h, dt, vp(data), G1, 1, 0, op, vp(out), 0, None)
        assert host(out)[0] == f(glob)
    pa.call("pm_level_reduce", L.h, dt, vp(data), G1, 1, 0, pa.RED_SUM, vp(out), 0, None)
    tol = n ** 3 * np.finfo(t).eps * np.abs(glob).sum()
    assert abs(float(host(out)[0]) - float(glob.astype(np.float64).sum())) <= tol
    # data += coef (tx[x] ty[y]) tz[z] on valid cells, ghosts untouched
    tx, ty, tz = (util.uniform(92, k, n) for k in range(3))
    before = host(data)
    pa.call("pm_level_add_separable", L.h, dt, vp(data), G1, 0.25, dptr(tx), dptr(ty), dptr(tz), None)
    prod = (tx[None, None, :] * ty[None, :, None]) * tz[:, None, None]
    want = (glob.reshape(n, n, n).astype(np.float64) + 0.25 * prod).astype(t)  # one rounding to T
    got = L.unpack(host(data), 1).reshape(n, n, n)
    assert got.tobytes() == want.tobytes()
    gb = before.reshape(len(L.lows), bs + 2, bs + 2, bs + 2)
    ga = host(data).reshape(len(L.lows), bs + 2, bs + 2, bs + 2)
    assert ga[:, 0].tobytes() == gb[:, 0].tobytes()  # a ghost plane


def test_multigrid_level_layouts():
    n, bs, nlev = 64, 16, 3
    L = Layout(n, bs)
    opt = pa.MgOptions()
    pa.lib().pm_mg_default_options(ctypes.byref(opt))
    mg = ctypes.c_void_p()
    pa.call("pm_mg_create", ctypes.byref(mg), L.h, nlev, 1.0 / n, ctypes.byref(opt))
    sizes = []
    for l in range(nlev + 1):
        lay = ctypes.c_void_p()
        pa.call("pm_mg_level_layout", mg, l, ctypes.byref(lay))
        e = ctypes.c_size_t()
        pa.call("pm_layout_box_elems", lay, G1, ctypes.byref(e))
        sizes.append(round(e.value ** (1 / 3)) - 2)
    # LevelMultigrid::define: box size kept while it divides the coarse domain, then one box
    assert sizes == [16, 16, 16, 8]
    assert pa.lib().pm_mg_level_layout(mg, nlev + 1, ctypes.byref(ctypes.c_void_p())) == -1
    pa.call("pm_mg_destroy", mg)


def test_fas_resnorm_equals_residual_norm():
    n, bs = 32, 16
    L = Layout(n, bs)
    fas = ctypes.c_void_p()
    pa.call("pm_fas_create", ctypes.byref(fas), L.h, 3, 1.0 / n)
    phi = dev(L.pack(util.uniform(93, 0, n ** 3), 2))
    f = dev(L.pack(util.uniform(93, 1, n ** 3), 0))
    r1, r2 = ctypes.c_double(), ctypes.c_double()
    pa.call("pm_fas_resnorm", fas, vp(phi), vp(f), ctypes.byref(r1), None)
    res = zeros(len(L.lows) * bs ** 3)
    pa.call("pm_fas_residual", fas, vp(phi), vp(f), vp(res), ctypes.byref(r2), None)
    assert r1.value == r2.value == np.abs(host(res)).max()
    # the residual of a constant state is the force
    pa.call("pm_level_setval", L.h, pa.PM_F64, vp(phi), pa.i3([2, 2, 2]), 1, 3.0, None)
    pa.call("pm_fas_residual", fas, vp(phi), vp(f), vp(res), ctypes.byref(r2), None)
    assert np.abs(host(res) - host(f)).max() < 1e-9
    pa.call("pm_fas_destroy", fas)


def test_tune_table_round_trip():
    coefs, offs, *_ = util.lib_stencil("Laplacian", 0, 3, np.float64)
    st = pa.Stencil(coefs, offs, pa.PM_F64)
    n = 64
    ub = pa.Box.of([-1] * 3, [n] * 3)
    rb = pa.Box.of([0] * 3, [n - 1] * 3)
    cfg, ms = pa.TileConfig(), ctypes.c_float()
    pa.call("pm_stencil_autotune", st.h, None, 0, ctypes.byref(ub), ctypes.byref(rb), 1, ctypes.byref(cfg),
            ctypes.byref(ms), None)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.txt").encode()
        pa.call("pm_tune_table_save", path)
        other = pa.TileConfig(1, 2, 1 - cfg.nt_store, 16)
        pa.call("pm_stencil_set_tile_config", st.h, None, 0, ctypes.byref(rb), ctypes.byref(other))
        pa.call("pm_tune_table_load", path)
        got = pa.TileConfig()
        pa.call("pm_stencil_get_tile_config", st.h, None, 0, ctypes.byref(rb), ctypes.byref(got))
        assert got.as_dict() == cfg.as_dict()
        assert len(open(path).read().strip().splitlines()) >= 1


@pytest.mark.parametrize("n,bs,g", [(16, 8, 1), (8, 4, 2)])
def test_level_linear_out_in(n, bs, g):
    """LevelBoxData::linearOut / linearIn in Proto's buffer format (golden
    buffers written by Proto), to host and device buffers."""
    z = np.load(os.path.join(util.GOLDEN, "linear.npz"), allow_pickle=False)
    key = "n%d_bs%d_g%d" % (n, bs, g)
    gold, lows = z[key + ":buffer"], np.ascontiguousarray(z[key + ":lows"].astype(np.int32))
    lay = ctypes.c_void_p()
    pa.call("pm_layout_create", ctypes.byref(lay), 3, pa.i3([n] * 3), pa.i3([bs] * 3), pa.i3([1] * 3), len(lows),
            lows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    gg = pa.i3([g] * 3)
    size = ctypes.c_size_t()
    pa.call("pm_level_linear_size", lay, pa.PM_F64, gg, 2, ctypes.byref(size))
    assert size.value == gold.nbytes
    data = zeros(gold.size)
    pa.call("pm_level_linear_in", lay, pa.PM_F64, vp(data), gg, 2, gold.ctypes.data_as(ctypes.c_void_p), 1, None)
    out = np.zeros_like(gold)
    pa.call("pm_level_linear_out", lay, pa.PM_F64, vp(data), gg, 2, out.ctypes.data_as(ctypes.c_void_p), 1, None)
    assert out.tobytes() == gold.tobytes()
    dbuf = zeros(gold.size)
    pa.call("pm_level_linear_out", lay, pa.PM_F64, vp(data), gg, 2, vp(dbuf), 0, None)
    back = zeros(gold.size)
    pa.call("pm_level_linear_in", lay, pa.PM_F64, vp(back), gg, 2, vp(dbuf), 0, None)
    assert host(back).tobytes() == gold.tobytes()
    pa.call("pm_layout_destroy", lay)


def test_stream_probe_kinds():
    """The bandwidth probe: copies (grid-stride, chunked) reproduce the source, write fills the
    word pattern (w, w+1, w+2, w+3) of 16-byte vector w, read leaves memory
    alone; bad arguments are refused."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2 ** 32, size=4 * 4099, dtype=np.uint32)
    src = dev(a)
    dst = zeros(a.size, np.uint32)
    pa.call("pm_stream_probe", 2, vp(dst), vp(src), a.nbytes, 4, None)
    assert host(dst).tobytes() == a.tobytes()
    pa.call("pm_stream_probe", 1, vp(dst), None, a.nbytes, 4, None)
    w = np.arange(a.size // 4, dtype=np.uint32)[:, None] + np.arange(4, dtype=np.uint32)[None, :]
    assert host(dst).tobytes() == w.reshape(-1).tobytes()
    sink = zeros(4, np.uint32)
    pa.call("pm_stream_probe", 0, vp(sink), vp(src), a.nbytes, 4, None)
    assert not host(sink).any() and host(src).tobytes() == a.tobytes()
    for kind in (3, 4):  # chunked copies: one 64 KB chunk and a 3-word tail
        dst2 = zeros(a.size, np.uint32)
        pa.call("pm_stream_probe", kind, vp(dst2), vp(src), a.nbytes, 4, None)
        assert host(dst2).tobytes() == a.tobytes(), kind
    assert pa.lib().pm_stream_probe(5, vp(dst), vp(src), a.nbytes, 4, None) == -1
    assert pa.lib().pm_stream_probe(2, vp(dst), vp(src), a.nbytes - 4, 4, None) == -1
    assert pa.lib().pm_stream_probe(2, vp(dst), vp(src), a.nbytes, 65, None) == -1


def test_level_copy_to_regions():
    """LevelBoxData::copyTo regions (Proto_LevelBoxDataImplem.H:346-386,
    795-870) against a numpy model.  Different layouts: every ghosted
    destination cell covered by a source patch (periodic images included)
    takes the source's valid value, uncovered cells keep theirs.  Compatible
    layouts: box-by-box copy of the common ghosted region, then exchange."""
    n = 16
    glob = util.uniform(94, 0, n ** 3)
    G = glob.reshape(n, n, n)
    # source: 8^3 boxes, only the patches with x-patch index 0 (half the domain)
    S = Layout(n, 8, lows=[(0, y, z) for z in (0, 8) for y in (0, 8)])
    src = dev(S.pack(glob, 1))
    D = Layout(n, 4)
    gd = 2
    e = 4 + 2 * gd
    old = np.full(len(D.lows) * e ** 3, -5.0)
    dst = dev(old)
    pa.call("pm_level_copy_to", S.h, vp(src), G1, D.h, vp(dst), pa.i3([gd] * 3), pa.PM_F64, 1, None)
    got = host(dst).reshape(len(D.lows), e, e, e)
    want = old.reshape(len(D.lows), e, e, e).copy()
    for b, lo in enumerate(D.lows):
        for k in range(e):
            for j in range(e):
                for i in range(e):
                    x, y, z = ((lo[0] - gd + i) % n, (lo[1] - gd + j) % n, (lo[2] - gd + k) % n)
                    if x < 8:
                        want[b, k, j, i] = G[z, y, x]
    assert got.tobytes() == want.tobytes()
    # compatible layouts: ghost 1 source (ghosts hold a marker), ghost 2 destination
    A = Layout(n, 8)
    srcA = A.pack(glob, 1).reshape(len(A.lows), 10, 10, 10)
    srcA[:, 0] = 3.0  # marker on one ghost face; exchange must overwrite it in the destination
    srcA_d = dev(srcA.reshape(-1))
    dstA = dev(np.full(len(A.lows) * 12 ** 3, -5.0))
    pa.call("pm_level_copy_to", A.h, vp(srcA_d), G1, A.h, vp(dstA), pa.i3([2] * 3), pa.PM_F64, 1, None)
    assert host(dstA).tobytes() == A.pack(glob, 2).tobytes()
    assert pa.lib().pm_level_copy_to(A.h, vp(srcA_d), G1, A.h, vp(srcA_d), G1, pa.PM_F64, 1, None) == -1


def test_graph_scratch_survives_growth():
    """A graph captured around a reduction keeps the stream scratch it
    recorded: a later, larger eager call on the same stream grows the scratch
    without freeing the captured buffer (it is retired until the stream is
    destroyed), so the replay writes its partials into live memory and gives
    the eager answer.  A buffer allocated after the growth is not touched by
    the replay."""
    from proto_amd.device import DeviceArray, Stream
    s = Stream()
    a = util.uniform(21, 0, 5000)
    ad, out, ref = dev(a), zeros(1), zeros(1)
    pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, vp(ad), a.size, vp(ref), 0, s.h)  # sizes the scratch eagerly
    s.sync()
    g = ctypes.c_void_p()
    pa.call("pm_graph_begin", s.h)
    pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, vp(ad), a.size, vp(out), 0, s.h)
    pa.call("pm_graph_end", s.h, ctypes.byref(g))
    nb, nret = ctypes.c_size_t(), ctypes.c_int()
    pa.call("pm_stream_scratch_info", s.h, ctypes.byref(nb), ctypes.byref(nret))
    say("scratch before growth", nb.value, nret.value)
    assert nret.value == 0
    small = nb.value
    # 2048 x 2048 x 1 residual: 32 x 256 workgroups x 2 FP64 partials = 128 KiB of scratch
    n = 2048
    c7, o7, _, _, _ = util.lib_stencil("Laplacian")
    st = pa.Stencil(c7, o7)
    ub, rb = pa.Box.of([-1, -1, -1], [n, n, 1]), pa.Box.of([0, 0, 0], [n - 1, n - 1, 0])
    u = zeros((n + 2) * (n + 2) * 3)
    f, r, red = zeros(n * n), zeros(n * n), zeros(2)
    pa.call("pm_stencil_residual", st.h, vp(u), ctypes.byref(ub), vp(f), vp(r), ctypes.byref(rb), 1.0, vp(red), s.h)
    s.sync()
    pa.call("pm_stream_scratch_info", s.h, ctypes.byref(nb), ctypes.byref(nret))
    say("scratch after growth", nb.value, nret.value)
    assert nb.value > small and nret.value == 1
    filler = DeviceArray(small // 8)
    pa.call("pm_fill_uniform", pa.PM_F64, vp(filler), filler.n, 4, 0, 0.0, 7.0, s.h)
    try:
        for rep in range(3):
            pa.call("pm_memset_zero", vp(out), 8, s.h)
            pa.call("pm_graph_launch", g, s.h)
            s.sync()
            assert host(out)[0] == host(ref)[0] == np.abs(a).max(), rep
        assert (host(filler) == 7.0).all()
    finally:
        pa.call("pm_graph_destroy", g)


def test_scratch_growth_under_capture_is_refused():
    """A call that needs more stream scratch than exists while the stream is
    capturing fails with PM_ERR_UNSUPPORTED before recording anything (no
    drain or allocation inside a capture)."""
    from proto_amd.device import Stream
    s = Stream()
    n = 2048
    c7, o7, _, _, _ = util.lib_stencil("Laplacian")
    st = pa.Stencil(c7, o7)
    ub, rb = pa.Box.of([-1, -1, -1], [n, n, 1]), pa.Box.of([0, 0, 0], [n - 1, n - 1, 0])
    u = zeros((n + 2) * (n + 2) * 3)
    f, r, red = zeros(n * n), zeros(n * n), zeros(2)
    g = ctypes.c_void_p()
    pa.call("pm_graph_begin", s.h)
    rc = pa.lib().pm_stencil_residual(st.h, vp(u), ctypes.byref(ub), vp(f), vp(r), ctypes.byref(rb), 1.0, vp(red),
                                      s.h)
    pa.call("pm_graph_end", s.h, ctypes.byref(g))
    pa.call("pm_graph_destroy", g)
    assert rc == -4
    assert b"under graph capture" in pa.lib().pm_last_error()
