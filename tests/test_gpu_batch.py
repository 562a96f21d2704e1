"""GPU parity of the batched-patch operators (pm_batch_apply, pm_batch_gsrb)
against the oracle, box by box, bitwise: the 7- and 27-point Laplacian apply
of Proto's Stencil::apply (Proto_StencilImplem.H hostApply) and one GSRB sweep
with fixed ghosts (LaplaceOp.H:95-123), FP64 and FP32.  Also: the work queue
resets itself (repeated launches and every tiling knob give the same bits),
and a captured graph replays correctly with eager work between replays."""
import ctypes

import numpy as np
import pytest

import proto_amd as pa
import util
from gpu_util import dev, host, say, vp, zeros

pytestmark = pytest.mark.gpu

# uneven patches: odd and even extents, 1-cell boxes, several 64-wide x tiles
SIZES = [(16, 16, 16), (24, 24, 24), (33, 7, 5), (1, 1, 1), (2, 3, 4), (64, 64, 8), (65, 9, 11),
         (130, 5, 6), (5, 130, 3), (3, 4, 130), (48, 32, 16), (7, 7, 7), (100, 1, 2), (1, 50, 1),
         (17, 19, 23), (32, 32, 32)]
GHOST = 2
NPDT = {pa.PM_F64: np.float64, pa.PM_F32: np.float32}


def make_batch(sizes=SIZES, ghost=GHOST):
    return pa.Batch(sizes, ghost)


def inputs(B, dt, seed):
    t = NPDT[dt]
    src = util.uniform(seed, 0, B.ghosted_elems).astype(t)
    rhs = util.uniform(seed, 1, B.valid_elems).astype(t)
    return src, rhs


def oracle_apply_batch(B, coefs, offs, src, scale):
    out = np.zeros(B.valid_elems, dtype=src.dtype)
    g = B.ghost
    for b in range(len(B.sizes)):
        n = [int(v) for v in B.sizes[b]]
        s = np.ascontiguousarray(src[B.goff[b]:B.goff[b + 1]])
        d = np.zeros(B.voff[b + 1] - B.voff[b], dtype=src.dtype)
        hi = [v - 1 for v in n]
        util.oracle_apply(coefs, offs, s, [-g] * 3, [v + g - 1 for v in n], d, [0, 0, 0], hi, [0, 0, 0], hi,
                          True, scale)
        out[B.voff[b]:B.voff[b + 1]] = d
    return out


def oracle_gsrb_batch(B, phi, rhs, dx):
    out = phi.copy()
    f = util.oracle().orc_patch_gsrb_f64 if phi.dtype == np.float64 else util.oracle().orc_patch_gsrb_f32
    for b in range(len(B.sizes)):
        nx, ny, nz = (int(v) for v in B.sizes[b])
        p = np.ascontiguousarray(out[B.goff[b]:B.goff[b + 1]])
        r = np.ascontiguousarray(rhs[B.voff[b]:B.voff[b + 1]])
        f(nx, ny, nz, B.ghost, dx, util.ptr(r), util.ptr(p))
        out[B.goff[b]:B.goff[b + 1]] = p
    return out


def first_diff(B, got, want, valid):
    for b in range(len(B.sizes)):
        lo, hi = (B.voff[b], B.voff[b + 1]) if valid else (B.goff[b], B.goff[b + 1])
        if got[lo:hi].tobytes() != want[lo:hi].tobytes():
            return "box %d %s" % (b, tuple(int(v) for v in B.sizes[b]))
    return None


@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
@pytest.mark.parametrize("name", ["Laplacian", "Laplacian_27"])
def test_batch_apply_vs_oracle(dt, name):
    B = make_batch()
    t = NPDT[dt]
    coefs, offs, *_ = util.lib_stencil(name, 0, 3, t)
    st = pa.Stencil(coefs, offs, dt)
    src, _ = inputs(B, dt, 11)
    s_d, d_d = dev(src), zeros(B.valid_elems, t)
    for scale in (1.0, 0.37):
        want = oracle_apply_batch(B, coefs, offs, src, scale)
        for cfg in ((0, 8, 4), (4, 8, 1), (1, 8, 16), (0, 8, 2)):
            B.configure(*cfg)
            for rep in range(2):
                say("batch apply", name, dt, "scale", scale, "cfg", cfg, "rep", rep)
                B.apply(st, vp(s_d), vp(d_d), scale)
                got = host(d_d)
                assert first_diff(B, got, want, True) is None, (cfg, rep, first_diff(B, got, want, True))


@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
def test_batch_gsrb_vs_oracle(dt):
    B = make_batch()
    t = NPDT[dt]
    phi, rhs = inputs(B, dt, 12)
    dx = 1.0 / 64
    want1 = oracle_gsrb_batch(B, phi, rhs, dx)
    want2 = oracle_gsrb_batch(B, want1, rhs, dx)
    p_in, r_d = dev(phi), dev(rhs)
    for rows in (1, 3, 8):
        B.configure(0, rows, 4)
        for rep in range(2):
            say("batch gsrb", dt, "rows", rows, "rep", rep)
            p_out = dev(phi)  # ghosts of phi_out = ghosts of phi_in
            B.gsrb(dt, vp(p_in), vp(p_out), vp(r_d), dx)
            got = host(p_out)
            assert first_diff(B, got, want1, False) is None, (rows, rep, first_diff(B, got, want1, False))
            # ping-pong: a second sweep back into a buffer holding phi's ghosts
            p_back = dev(phi)
            B.gsrb(dt, vp(p_out), vp(p_back), vp(r_d), dx)
            got2 = host(p_back)
            assert first_diff(B, got2, want2, False) is None, (rows, rep, first_diff(B, got2, want2, False))


# every box 2..64 wide: the row-register sweep serves these (odd extents,
# rows not a multiple of the sub-tile height, one- and two-plane boxes, one
# to four sub-tiles per wave)
SIZES_ROWS = [(16, 16, 16), (24, 24, 24), (33, 7, 5), (2, 3, 4), (64, 64, 8), (48, 32, 16), (7, 7, 7),
              (17, 19, 23), (32, 32, 32), (64, 1, 2), (2, 50, 1), (63, 5, 3), (21, 4, 9), (40, 13, 2), (3, 2, 1)]


@pytest.mark.parametrize("ghost", [1, 2])
@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
def test_batch_row_sweep_vs_oracle(dt, ghost):
    """The row sweep (forms 0 = automatic, 2, 3, 4) and the LDS-ring kernel
    (form 1) against the oracle, box by box, bitwise, and a second sweep back."""
    B = pa.Batch(SIZES_ROWS, ghost)
    t = NPDT[dt]
    phi, rhs = inputs(B, dt, 14 + ghost)
    dx = 1.0 / 32
    want1 = oracle_gsrb_batch(B, phi, rhs, dx)
    want2 = oracle_gsrb_batch(B, want1, rhs, dx)
    p_in, r_d = dev(phi), dev(rhs)
    for form, waves in ((0, 0), (2, 1), (2, 2), (2, 8), (3, 0), (4, 1), (4, 4), (1, 0)):
        B.sweep_form(form, waves)
        say("batch row sweep", dt, "ghost", ghost, "form", form, "waves", waves)
        p_out = dev(phi)
        B.gsrb(dt, vp(p_in), vp(p_out), vp(r_d), dx)
        got = host(p_out)
        assert first_diff(B, got, want1, False) is None, (form, first_diff(B, got, want1, False))
        p_back = dev(phi)
        B.gsrb(dt, vp(p_out), vp(p_back), vp(r_d), dx)
        assert first_diff(B, host(p_back), want2, False) is None, (form, first_diff(B, host(p_back), want2, False))
    # a batch with a box wider than 64 cells has no row sweep
    with pytest.raises(pa.NativeError, match="2..64"):
        make_batch().sweep_form(2)


# every box even and 4..128 wide: the pair sweep (two cells per lane) serves these
SIZES_PAIRS = [(16, 16, 16), (24, 24, 24), (34, 7, 5), (4, 3, 4), (64, 64, 8), (48, 32, 16), (6, 7, 7),
               (18, 19, 23), (32, 32, 32), (64, 1, 2), (4, 50, 1), (62, 5, 3), (22, 4, 9), (40, 13, 2), (4, 2, 1),
               (128, 6, 3), (96, 9, 4)]


@pytest.mark.parametrize("ghost", [2, 4])
@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
def test_batch_pair_sweep_vs_oracle(dt, ghost):
    """The pair sweep (forms 0 = automatic for FP32, 5, 6) and the LDS ring
    against the oracle, box by box, bitwise, and a second sweep back."""
    B = pa.Batch(SIZES_PAIRS, ghost)
    t = NPDT[dt]
    phi, rhs = inputs(B, dt, 30 + ghost)
    dx = 1.0 / 48
    want1 = oracle_gsrb_batch(B, phi, rhs, dx)
    want2 = oracle_gsrb_batch(B, want1, rhs, dx)
    p_in, r_d = dev(phi), dev(rhs)
    for form, waves in ((0, 0), (5, 1), (5, 2), (5, 8), (6, 0), (1, 0)):
        B.sweep_form(form, waves)
        say("batch pair sweep", dt, "ghost", ghost, "form", form, "waves", waves)
        p_out = dev(phi)
        B.gsrb(dt, vp(p_in), vp(p_out), vp(r_d), dx)
        got = host(p_out)
        assert first_diff(B, got, want1, False) is None, (form, first_diff(B, got, want1, False))
        p_back = dev(phi)
        B.gsrb(dt, vp(p_out), vp(p_back), vp(r_d), dx)
        assert first_diff(B, host(p_back), want2, False) is None, (form, first_diff(B, host(p_back), want2, False))
    with pytest.raises(pa.NativeError, match="one wave"):
        B.sweep_form(6, 4)
    with pytest.raises(pa.NativeError, match="waves per workgroup"):
        B.sweep_form(5, 3)
    # odd widths or an odd ghost have no pair sweep
    with pytest.raises(pa.NativeError, match="pair sweep"):
        pa.Batch(SIZES_ROWS, 2).sweep_form(5)
    with pytest.raises(pa.NativeError, match="pair sweep"):
        pa.Batch(SIZES_PAIRS, 1).sweep_form(5)


# every box 4..252 wide in multiples of 4: the quad sweep (four cells per lane,
# FP32, two ghost layers) serves these
SIZES_QUADS = [(16, 16, 16), (24, 24, 24), (4, 3, 4), (64, 64, 8), (48, 32, 16), (20, 19, 23), (32, 32, 32), (64, 1, 2),
               (4, 50, 1), (60, 5, 3), (8, 4, 9), (40, 13, 2), (4, 2, 1), (128, 6, 3), (96, 9, 4), (252, 3, 2)]


def test_batch_quad_sweep_vs_oracle():
    """The quad sweep (form 0 = automatic for FP32 with two ghost layers, 7 and 8)
    against the oracle, box by box, bitwise, and a second sweep back; the rows'
    ghost columns keep their input values.  FP64, odd multiples of 2 and other
    ghost widths have no quad sweep."""
    dt = pa.PM_F32
    B = pa.Batch(SIZES_QUADS, 2)
    phi, rhs = inputs(B, dt, 52)
    dx = 1.0 / 40
    want1 = oracle_gsrb_batch(B, phi, rhs, dx)
    want2 = oracle_gsrb_batch(B, want1, rhs, dx)
    p_in, r_d = dev(phi), dev(rhs)
    for form in (0, 7, 8):
        B.sweep_form(form)
        say("batch quad sweep form", form)
        p_out = dev(phi)
        B.gsrb(dt, vp(p_in), vp(p_out), vp(r_d), dx)
        got = host(p_out)
        assert first_diff(B, got, want1, False) is None, (form, first_diff(B, got, want1, False))
        p_back = dev(phi)
        B.gsrb(dt, vp(p_out), vp(p_back), vp(r_d), dx)
        assert first_diff(B, host(p_back), want2, False) is None, (form, first_diff(B, host(p_back), want2, False))
    B.sweep_form(7)
    f64 = pa.Batch(SIZES_QUADS, 2)
    f64.sweep_form(7)
    phi64, rhs64 = inputs(f64, pa.PM_F64, 53)
    a64, b64, r64 = dev(phi64), dev(phi64), dev(rhs64)
    with pytest.raises(pa.NativeError, match="quad sweep is FP32"):
        f64.gsrb(pa.PM_F64, vp(a64), vp(b64), vp(r64), dx)
    with pytest.raises(pa.NativeError, match="quad sweep"):
        pa.Batch(SIZES_PAIRS, 2).sweep_form(7)
    with pytest.raises(pa.NativeError, match="quad sweep"):
        pa.Batch(SIZES_QUADS, 4).sweep_form(7)


def test_batch_gsrb_limits():
    B = make_batch()
    phi, rhs = inputs(B, pa.PM_F64, 13)
    p_in, r_d = dev(phi), dev(rhs)
    with pytest.raises(pa.NativeError, match="out of place"):
        B.gsrb(pa.PM_F64, vp(p_in), vp(p_in), vp(r_d), 0.1)
    # 32 rows of 132-wide FP64 planes do not fit a workgroup's LDS: the tiles
    # are cut to fewer rows and the bits do not change
    B.configure(0, 32, 4)
    p_out = dev(phi)
    B.gsrb(pa.PM_F64, vp(p_in), vp(p_out), vp(r_d), 0.1)
    assert first_diff(B, host(p_out), oracle_gsrb_batch(B, phi, rhs, 0.1), False) is None
    wide = pa.Batch([(300, 4, 4)], 1)
    src = dev(np.zeros(wide.ghosted_elems))
    with pytest.raises(pa.NativeError, match="wider than 256"):
        wide.gsrb(pa.PM_F64, vp(src), vp(dev(np.zeros(wide.ghosted_elems))), vp(dev(np.zeros(wide.valid_elems))), 0.1)


def test_batch_many_small_boxes_queue():
    """Far more tiles than persistent workgroups, one workgroup per CU."""
    rng = np.random.default_rng(5)
    sizes = rng.choice([1, 2, 4, 5, 8], size=(3000, 3))
    B = pa.Batch(sizes, 1)
    B.configure(0, 2, 1)
    coefs, offs, *_ = util.lib_stencil("Laplacian", 0, 3, np.float64)
    st = pa.Stencil(coefs, offs, pa.PM_F64)
    src, rhs = inputs(B, pa.PM_F64, 14)
    want = oracle_apply_batch(B, coefs, offs, src, 1.0)
    wantg = oracle_gsrb_batch(B, src, rhs, 0.25)
    s_d, d_d, r_d = dev(src), zeros(B.valid_elems), dev(rhs)
    for rep in range(3):
        say("small boxes rep", rep)
        B.apply(st, vp(s_d), vp(d_d))
        assert first_diff(B, host(d_d), want, True) is None
        p_out = dev(src)
        B.gsrb(pa.PM_F64, vp(s_d), vp(p_out), vp(r_d), 0.25)
        assert first_diff(B, host(p_out), wantg, False) is None


def test_batch_graph_replay():
    """apply + GSRB captured once, replayed three times with eager refills of
    the inputs between replays; every replay matches the oracle."""
    from proto_amd.device import Stream
    B = make_batch()
    coefs, offs, *_ = util.lib_stencil("Laplacian_27", 0, 3, np.float64)
    st = pa.Stencil(coefs, offs, pa.PM_F64)
    s = Stream()
    src_d, dst_d = zeros(B.ghosted_elems), zeros(B.valid_elems)
    rhs_d, out_d = zeros(B.valid_elems), zeros(B.ghosted_elems)
    dx = 1.0 / 32
    g = ctypes.c_void_p()
    pa.call("pm_graph_begin", s.h)
    B.apply(st, vp(src_d), vp(dst_d), 1.0, s.h)
    B.gsrb(pa.PM_F64, vp(src_d), vp(out_d), vp(rhs_d), dx, s.h)
    pa.call("pm_graph_end", s.h, ctypes.byref(g))
    try:
        for rep in range(3):
            say("graph replay", rep)
            pa.call("pm_fill_uniform", pa.PM_F64, vp(src_d), B.ghosted_elems, 100 + rep, 0,
                    ctypes.c_double(1.0), ctypes.c_double(0.0), s.h)
            pa.call("pm_fill_uniform", pa.PM_F64, vp(rhs_d), B.valid_elems, 100 + rep, 1,
                    ctypes.c_double(1.0), ctypes.c_double(0.0), s.h)
            pa.call("pm_memcpy", vp(out_d), vp(src_d), B.ghosted_elems * 8, 2, s.h)
            s.sync()
            src, rhs = host(src_d), host(rhs_d)
            pa.call("pm_graph_launch", g, s.h)
            s.sync()
            want = oracle_apply_batch(B, coefs, offs, src, 1.0)
            assert first_diff(B, host(dst_d), want, True) is None, rep
            assert first_diff(B, host(out_d), oracle_gsrb_batch(B, src, rhs, dx), False) is None, rep
    finally:
        pa.call("pm_graph_destroy", g)


def test_batch_autotune_keeps_bits():
    B = make_batch()
    coefs, offs, *_ = util.lib_stencil("Laplacian_27", 0, 3, np.float32)
    st = pa.Stencil(coefs, offs, pa.PM_F32)
    src, rhs = inputs(B, pa.PM_F32, 15)
    s_d, d_d, r_d, p_d = dev(src), zeros(B.valid_elems, np.float32), dev(rhs), dev(src)
    cfg, ms = B.autotune(st, vp(s_d), vp(d_d), vp(p_d), vp(r_d), 0.5, 2)
    say("autotuned", cfg.as_dict(), ms)
    assert B.config().as_dict() == cfg.as_dict() and ms[0] > 0 and ms[1] > 0
    B.apply(st, vp(s_d), vp(d_d))
    assert first_diff(B, host(d_d), oracle_apply_batch(B, coefs, offs, src, 1.0), True) is None
    p_d = dev(src)
    B.gsrb(pa.PM_F32, vp(s_d), vp(p_d), vp(r_d), 0.5)
    assert first_diff(B, host(p_d), oracle_gsrb_batch(B, src, rhs, 0.5), False) is None


@pytest.mark.parametrize("dt", [pa.PM_F64, pa.PM_F32])
@pytest.mark.parametrize("ghost", [1, 2])
def test_batch_gsrb_even_widths_vs_oracle(dt, ghost):
    """Batches whose boxes all have even widths take the colour-by-colour sweep."""
    sizes = [(16, 16, 16), (24, 8, 6), (2, 2, 2), (64, 10, 4), (4, 30, 7), (8, 2, 20), (48, 16, 16), (6, 6, 6)]
    B = pa.Batch(sizes, ghost)
    t = NPDT[dt]
    phi, rhs = inputs(B, dt, 31 + ghost)
    dx = 1.0 / 48
    want = oracle_gsrb_batch(B, phi, rhs, dx)
    p_in, r_d = dev(phi), dev(rhs)
    for rows in (1, 3, 8, 16):
        B.configure(0, rows, 4)
        say("even batch gsrb", dt, ghost, rows)
        p_out = dev(phi)
        B.gsrb(dt, vp(p_in), vp(p_out), vp(r_d), dx)
        assert first_diff(B, host(p_out), want, False) is None, (rows, first_diff(B, host(p_out), want, False))
