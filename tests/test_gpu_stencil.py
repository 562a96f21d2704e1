"""GPU parity of Stencil apply, the fused residual and Reduction, through the
C ABI, against Proto's golden vectors and the oracle.  Integer-exact semantics
(bitwise) for the stencil arithmetic; Sum within n*eps*sum|x| of the
sequential order, as the tolerance states below."""
import ctypes
import os

import numpy as np
import pytest

import proto_amd as pa
import util
from gpu_util import box, config2_inputs, dev, empty, host, say, vp, zeros
from test_oracle import merge

pytestmark = pytest.mark.gpu


def apply_gpu(st, src, slo, shi, dst, dlo, dhi, blo, bhi, replace, scale, ncomp=1):
    s_d, d_d = dev(src), dev(dst)
    pa.call("pm_stencil_apply", st.h, vp(s_d), ctypes.byref(box(slo, shi)), vp(d_d),
            ctypes.byref(box(dlo, dhi)), ctypes.byref(box(blo, bhi)), ncomp, int(replace), float(scale), None)
    return host(d_d)


def test_golden_apply_cases():
    z = np.load(os.path.join(util.GOLDEN, "stencil_apply.npz"), allow_pickle=False)
    keys = sorted({k.split(":")[0] for k in z.files})
    for k in keys:
        g = {f.split(":")[1]: z[f] for f in z.files if f.startswith(k + ":")}
        say("golden case", k)
        dim = int(g["dim"])
        c, o = merge(g["coefs"], g["offs"])
        dt = pa.PM_F64 if c.dtype == np.float64 else pa.PM_F32
        st = pa.Stencil(c, o, dt, dim, g["sr"], g["dr"], g["ds"])
        b = g["box"]
        out = apply_gpu(st, g["src"], g["slo"], g["shi"], g["dst0"].copy(), g["dlo"], g["dhi"], b[:3], b[3:],
                        bool(g["replace"]), g["scale"].item())
        assert out.tobytes() == g["dst"].tobytes(), k


def test_laplacian_kinds():
    for name, kind in (("Laplacian", 7), ("Laplacian_27", 27)):
        c, o, _, _, _ = util.lib_stencil(name)
        assert pa.Stencil(c, o).kind() == kind
    # any other unit-ratio stencil within [-1,1]^3: the generic z-march (taps from LDS)
    c, o, _, _, _ = util.lib_stencil("Laplacian_19")
    assert pa.Stencil(c, o).kind() == 1
    c, o, _, _, _ = util.lib_stencil("Laplacian")
    assert pa.Stencil(c[::-1].copy(), o[::-1].copy()).kind() == 1  # other order
    # refinement ratios or taps further out: the per-cell kernel
    c, o, sr, dr, ds = util.lib_stencil("AvgDown", 2)
    assert pa.Stencil(c, o, src_ratio=sr, dst_ratio=dr, dst_shift=ds).kind() == 0


@pytest.mark.parametrize("name,dtype", [("Laplacian", pa.PM_F64), ("Laplacian_27", pa.PM_F64),
                                        ("Laplacian", pa.PM_F32), ("Laplacian_27", pa.PM_F32),
                                        ("Laplacian_19", pa.PM_F64)])
def test_random_boxes_vs_oracle(name, dtype):
    rng = np.random.default_rng(11)
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    c, o, _, _, _ = util.lib_stencil(name, dtype=npdt)
    st = pa.Stencil(c, o, dtype)
    for trial in range(8):
        g = int(rng.integers(1, 3))
        lo = rng.integers(-20, 20, size=3)
        n = rng.integers(1, 90, size=3) if trial else np.array([130, 67, 33])
        hi = lo + n - 1
        slo, shi = lo - g, hi + g
        dlo, dhi = lo - rng.integers(0, 2, size=3), hi + rng.integers(0, 2, size=3)
        replace = trial % 3 != 2
        scale = [1.0, 4096.0, -0.75][trial % 3]
        say(name, dtype, "trial", trial, "box", lo.tolist(), hi.tolist())
        src = util.uniform(300 + trial, 0, util.box_size(slo, shi)).astype(npdt)
        dst0 = util.uniform(300 + trial, 1, util.box_size(dlo, dhi)).astype(npdt)
        ref = dst0.copy()
        util.oracle_apply(c, o, src, slo, shi, ref, dlo, dhi, lo, hi, replace, scale)
        out = apply_gpu(st, src, slo, shi, dst0.copy(), dlo, dhi, lo, hi, replace, scale)
        assert out.tobytes() == ref.tobytes(), (name, trial)


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_generic_zmarch_random_stencils(dtype):
    """Kind-1 stencils (any taps in [-1,1]^DIM, any order, random coefficients):
    3-D with and without z taps (the one-plane form), 2-D, two components,
    replace and accumulate, scale != 1, boxes that are not tile multiples and
    destinations larger than the index box; bitwise against the oracle."""
    rng = np.random.default_rng(23)
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    allo = np.array([(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)], dtype=np.int32)
    for trial in range(10):
        dim = 2 if trial % 4 == 3 else 3
        pool = allo[allo[:, 2] == 0] if (dim == 2 or trial % 4 == 2) else allo
        k = int(rng.integers(1, len(pool) + 1))
        o = pool[rng.permutation(len(pool))[:k]].copy()
        c = rng.uniform(-2, 2, size=k).astype(npdt)
        st = pa.Stencil(c, o, dtype, dim)
        assert st.kind() == 1
        nc = 2 if trial % 5 == 4 else 1
        lo = rng.integers(-9, 9, size=3)
        n = rng.integers(1, 80, size=3) if trial else np.array([131, 70, 21])
        if dim == 2:
            lo[2], n[2] = 0, 1
        hi = lo + n - 1
        g = np.array([1, 1, 1 if dim == 3 else 0])
        slo, shi = lo - g, hi + g
        dlo, dhi = lo - rng.integers(0, 2, size=3) * (g > 0), hi + rng.integers(0, 2, size=3) * (g > 0)
        replace = trial % 3 != 1
        scale = [1.0, 0.5, -3.0][trial % 3]
        say("generic", dtype, "trial", trial, "dim", dim, "taps", k, "box", lo.tolist(), hi.tolist())
        ns, nd = util.box_size(slo, shi), util.box_size(dlo, dhi)
        src = util.uniform(700 + trial, 0, nc * ns).astype(npdt)
        dst0 = util.uniform(700 + trial, 1, nc * nd).astype(npdt)
        ref = dst0.copy()
        for q in range(nc):
            r = ref[q * nd:(q + 1) * nd].copy()
            util.oracle_apply(c, o, src[q * ns:(q + 1) * ns].copy(), slo, shi, r, dlo, dhi, lo, hi, replace, scale)
            ref[q * nd:(q + 1) * nd] = r
        out = apply_gpu(st, src, slo, shi, dst0.copy(), dlo, dhi, lo, hi, replace, scale, ncomp=nc)
        assert out.tobytes() == ref.tobytes(), trial


@pytest.mark.parametrize("name", ["Laplacian", "Laplacian_27"])
def test_kernel_families_agree(name):
    """The 7/27-point z-march, the generic z-march and the per-cell kernel
    give the same bits for the same stencil (the set_kind override)."""
    c, o, _, _, _ = util.lib_stencil(name)
    lo, hi = np.array([-3, 2, 5]), np.array([70, 41, 38])
    slo, shi = lo - 1, hi + 1
    src = util.uniform(91, 0, util.box_size(slo, shi))
    dst0 = util.uniform(91, 1, util.box_size(lo, hi))
    outs = {}
    for kind in (7 if name == "Laplacian" else 27, 1, 0):
        st = pa.Stencil(c, o).set_kind(kind)
        assert st.kind() == kind
        outs[kind] = apply_gpu(st, src, slo, shi, dst0.copy(), lo, hi, lo, hi, False, 4096.0).tobytes()
    assert len(set(outs.values())) == 1
    with pytest.raises(pa.NativeError):
        pa.Stencil(c, o).set_kind(27 if name == "Laplacian" else 7)


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_one_tap_stencil(dtype):
    """Stencil<T>::Identity() and its multiples (kind 2: the z-march kernel with
    its centre tap) against the oracle and the closed form dst = (replace ? 0 :
    dst) + c * src, on boxes with edge tiles, replace and accumulate; the
    generic z-march and the per-cell kernel give the same bits.  The last box
    has no halo (g = 0): the z-march stages planes z - 1 and z + 1 unchecked, so
    kind 2 must take the bounds-checked generic z-march there."""
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    o = [[0, 0, 0]]
    for ci, coef in enumerate((1.0, -0.75)):
        c = np.array([coef], dtype=npdt)
        assert pa.Stencil(c, o, dtype).kind() == 2
        for trial, (lo, hi, g) in enumerate((([0, 0, 0], [511, 63, 40], 2), ([-3, 5, 1], [70, 44, 29], 1),
                                             ([2, -7, 0], [2, 9, 17], 0))):
            lo, hi = np.array(lo), np.array(hi)
            slo, shi = lo - g, hi + g
            src = util.uniform(70 + trial, ci, util.box_size(slo, shi)).astype(npdt)
            dst0 = util.uniform(80 + trial, ci, util.box_size(lo, hi)).astype(npdt)
            for replace in (True, False):
                say("one-tap", dtype, coef, "box", lo.tolist(), hi.tolist(), "replace", replace)
                ref = dst0.copy()
                util.oracle_apply(c, o, src, slo, shi, ref, lo, hi, lo, hi, replace, 1.0)
                s3 = src.reshape(util.box_shape(slo, shi))
                inner = s3[g:s3.shape[0] - g, g:s3.shape[1] - g, g:s3.shape[2] - g].reshape(-1)
                closed = ((np.zeros_like(dst0) if replace else dst0) + c[0] * inner).astype(npdt)
                assert ref.tobytes() == closed.tobytes()
                outs = set()
                for kind in (2, 1, 0):
                    st = pa.Stencil(c, o, dtype).set_kind(kind)
                    out = apply_gpu(st, src, slo, shi, dst0.copy(), lo, hi, lo, hi, replace, 1.0)
                    assert out.tobytes() == ref.tobytes(), (coef, trial, replace, kind)
                    outs.add(out.tobytes())
                assert len(outs) == 1


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_apply_reduce(dtype):
    """pm_stencil_apply_reduce: dst as pm_stencil_apply writes it; Max, Min
    and Abs of the written cells exact, Sum within the stated bound; 3-D and
    2-D generic stencils (one pass), a 7-point one (apply, then reduce)."""
    rng = np.random.default_rng(5)
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    c2 = np.array([-4.0, 1, 1, 1, 1], dtype=npdt)
    o2 = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]]
    c19, o19, _, _, _ = util.lib_stencil("Laplacian_19", dtype=npdt)
    c7, o7, _, _, _ = util.lib_stencil("Laplacian", dtype=npdt)
    cases = [(c2, o2, 2, [3, -4, 0], [130, 70, 0]), (c19, o19, 3, [-5, 2, 1], [60, 41, 37]),
             (c7, o7, 3, [0, 0, 0], [47, 33, 29])]
    for ci, (c, o, dim, lo, hi) in enumerate(cases):
        st = pa.Stencil(c, o, dtype, dim)
        lo, hi = np.array(lo), np.array(hi)
        g = np.array([1, 1, 1 if dim == 3 else 0])
        slo, shi = lo - g, hi + g
        src = util.uniform(60 + ci, 0, util.box_size(slo, shi)).astype(npdt)
        ref = np.zeros(util.box_size(lo, hi), dtype=npdt)
        util.oracle_apply(c, o, src, slo, shi, ref, lo, hi, lo, hi, True, 256.0)
        for op in (pa.RED_MAX, pa.RED_MIN, pa.RED_ABS, pa.RED_SUM):
            s_d, d_d, out = dev(src), zeros(ref.size, npdt), zeros(1, npdt)
            say("apply_reduce", dtype, ci, op)
            pa.call("pm_stencil_apply_reduce", st.h, vp(s_d), ctypes.byref(box(slo, shi)), vp(d_d),
                    ctypes.byref(box(lo, hi)), ctypes.byref(box(lo, hi)), 1, 256.0, op, vp(out), 0, None)
            assert host(d_d).tobytes() == ref.tobytes(), (ci, op)
            got = host(out)[0]
            if op == pa.RED_MAX:
                assert got == ref.max()
            elif op == pa.RED_MIN:
                assert got == ref.min()
            elif op == pa.RED_ABS:
                assert got == np.abs(ref).max()
            else:
                assert abs(float(got) - float(ref.astype(np.float64).sum())) <= _sum_bound(ref)


def test_multicomponent_and_generic_ratios():
    # 3 components, AvgDown (srcRatio 2) and a destination-refined stencil
    c, o, sr, dr, ds = util.lib_stencil("AvgDown", 2)
    st = pa.Stencil(c, o, pa.PM_F64, 3, sr, dr, ds)
    slo, shi = [0, 0, 0], [15, 13, 9]
    dlo, dhi = [0, 0, 0], [7, 6, 4]
    nc = 3
    src = util.uniform(5, 0, nc * util.box_size(slo, shi))
    dst = np.zeros(nc * util.box_size(dlo, dhi))
    ref = dst.copy()
    n_s, n_d = util.box_size(slo, shi), util.box_size(dlo, dhi)
    for k in range(nc):
        r = ref[k * n_d:(k + 1) * n_d]
        util.oracle_apply(c, o, src[k * n_s:(k + 1) * n_s].copy(), slo, shi, r, dlo, dhi, dlo, dhi, True, 1.0,
                          sr, dr, ds)
        ref[k * n_d:(k + 1) * n_d] = r
    out = apply_gpu(st, src, slo, shi, dst, dlo, dhi, dlo, dhi, True, 1.0, ncomp=nc)
    assert out.tobytes() == ref.tobytes()


def test_out_of_bounds_request_is_rejected():
    c, o, _, _, _ = util.lib_stencil("Laplacian")
    st = pa.Stencil(c, o)
    src = dev(np.zeros(10 ** 3))
    dst = dev(np.zeros(8 ** 3))
    rc = pa.lib().pm_stencil_apply(st.h, vp(src), ctypes.byref(box([0] * 3, [9] * 3)), vp(dst),
                                   ctypes.byref(box([0] * 3, [7] * 3)), ctypes.byref(box([0] * 3, [7] * 3)),
                                   1, 1, 1.0, None)
    assert rc == -1 and b"outside" in pa.lib().pm_last_error()


def _sum_bound(r):
    n = r.size
    eps = np.finfo(r.dtype).eps / 2
    return 2 * n * eps * float(np.abs(r.astype(np.float64)).sum())


def test_residual_config2_small_golden():
    z = np.load(os.path.join(util.GOLDEN, "config2_small.npz"), allow_pickle=False)
    n, g = 32, 2
    u, f = config2_inputs(n, g)
    # the device generator reproduces the numpy generator bit for bit
    assert host(u).tobytes() == util.config2_u(n, g).tobytes()
    assert host(f).tobytes() == util.uniform(1, 1, n ** 3).tobytes()
    h = 1.0 / n
    for name in ("Laplacian", "Laplacian_27"):
        say("residual", name)
        c, o, _, _, _ = util.lib_stencil(name)
        st = pa.Stencil(c, o)
        r = empty(n ** 3)
        red = zeros(2)
        pa.call("pm_stencil_residual", st.h, vp(u), ctypes.byref(box([-g] * 3, [n - 1 + g] * 3)), vp(f), vp(r),
                ctypes.byref(box([0] * 3, [n - 1] * 3)), 1.0 / (h * h), vp(red), None)
        rr = host(r)
        assert rr.tobytes() == z[name + ":r"].tobytes()
        amax, s = host(red)
        assert amax == z[name + ":absmax"]
        # Sum tolerance: |S_gpu - S_seq| <= 2 n (eps/2) sum|r|
        assert abs(s - z[name + ":sum"]) <= _sum_bound(rr)


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_residual2_one_pass_equals_two(dtype):
    n, g = 40, 1
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    u, f = config2_inputs(n, g, npdt)
    h = 1.0 / n
    c7, o7, _, _, _ = util.lib_stencil("Laplacian", dtype=npdt)
    c27, o27, _, _, _ = util.lib_stencil("Laplacian_27", dtype=npdt)
    s7, s27 = pa.Stencil(c7, o7, dtype), pa.Stencil(c27, o27, dtype)
    ub, rb = box([-g] * 3, [n - 1 + g] * 3), box([0] * 3, [n - 1] * 3)
    r1, r2, a1, a2 = (empty(n ** 3, npdt) for _ in range(4))
    red = zeros(4, npdt)
    red1 = zeros(2, npdt)
    red2 = zeros(2, npdt)
    sc = 1.0 / (h * h)
    pa.call("pm_stencil_residual2", s7.h, s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r1), vp(r2), ctypes.byref(rb),
            sc, sc, vp(red), None)
    pa.call("pm_stencil_residual", s7.h, vp(u), ctypes.byref(ub), vp(f), vp(a1), ctypes.byref(rb), sc, vp(red1), None)
    pa.call("pm_stencil_residual", s27.h, vp(u), ctypes.byref(ub), vp(f), vp(a2), ctypes.byref(rb), sc, vp(red2), None)
    assert host(r1).tobytes() == host(a1).tobytes()
    assert host(r2).tobytes() == host(a2).tobytes()
    R, R1, R2 = host(red), host(red1), host(red2)
    assert R[0] == R1[0] and R[2] == R2[0]
    # oracle check of r (FP32 included: same order, one rounding per op)
    uh, fh = host(u), host(f)
    for cc, oo, rr in ((c7, o7, r1), (c27, o27, r2)):
        ref = np.zeros(n ** 3, dtype=npdt)
        fn = util.oracle().orc_residual_f64 if dtype == pa.PM_F64 else util.oracle().orc_residual_f32
        ct = ctypes.c_double if dtype == pa.PM_F64 else ctypes.c_float
        fn(len(cc), util.ptr(cc), util.ptr(oo.reshape(-1).copy()), util.ptr(uh), util.ptr(util.i3([-g] * 3)),
           util.ptr(util.i3([n - 1 + g] * 3)), util.ptr(fh), util.ptr(ref), util.ptr(util.i3([0] * 3)),
           util.ptr(util.i3([n - 1] * 3)), ct(sc))
        assert host(rr).tobytes() == ref.tobytes()


def test_config2_full_size_512():
    """Config 2 at its real size: 512^3 interior, 2 ghosts, both operators fused
    with r = f - L(u); every cell bitwise against the oracle, absMax exact,
    sum within the stated bound, and a second run bit-identical."""
    n, g = 512, 2
    say("config2 512^3 inputs")
    u, f = config2_inputs(n, g)
    h = 1.0 / n
    c7, o7, _, _, _ = util.lib_stencil("Laplacian")
    c27, o27, _, _, _ = util.lib_stencil("Laplacian_27")
    s7, s27 = pa.Stencil(c7, o7), pa.Stencil(c27, o27)
    ub, rb = box([-g] * 3, [n - 1 + g] * 3), box([0] * 3, [n - 1] * 3)
    r1 = empty(n ** 3)
    r2 = empty(n ** 3)
    red = zeros(4)
    sc = 1.0 / (h * h)
    say("config2 512^3 fused residual2")
    pa.call("pm_stencil_residual2", s7.h, s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r1), vp(r2), ctypes.byref(rb),
            sc, sc, vp(red), None)
    R = host(red).copy()
    uh, fh = host(u), host(f)
    for k, (cc, oo, rr) in enumerate(((c7, o7, r1), (c27, o27, r2))):
        say("config2 512^3 oracle operator", k)
        ref = np.zeros(n ** 3)
        util.oracle().orc_residual_f64(len(cc), util.ptr(cc), util.ptr(oo.reshape(-1).copy()), util.ptr(uh),
                                       util.ptr(util.i3([-g] * 3)), util.ptr(util.i3([n - 1 + g] * 3)),
                                       util.ptr(fh), util.ptr(ref), util.ptr(util.i3([0] * 3)),
                                       util.ptr(util.i3([n - 1] * 3)), ctypes.c_double(sc))
        got = host(rr)
        assert got.tobytes() == ref.tobytes()
        assert R[2 * k] == util.oracle().orc_reduce_f64(2, util.ptr(ref), ref.size)
        seq = util.oracle().orc_reduce_f64(3, util.ptr(ref), ref.size)
        assert abs(R[2 * k + 1] - seq) <= _sum_bound(ref)
        del ref, got
    red2 = zeros(4)
    pa.call("pm_stencil_residual2", s7.h, s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r1), vp(r2), ctypes.byref(rb),
            sc, sc, vp(red2), None)
    assert host(red2).tobytes() == R.tobytes()


def test_config1_full_size_256_2d():
    """Config 1 at its real size: DIM=2, 256^2 interior, one periodic ghost layer,
    dst |= Laplacian()(u, 1/h^2), then absMax and sum of dst.  Every cell and the
    absMax bitwise against the values Proto's CPU path wrote (config1.json); the sum
    within the bound stated in _sum_bound of Proto's sequential sum."""
    import hashlib
    import json
    meta = json.load(open(os.path.join(util.GOLDEN, "config1.json")))
    n = 256
    h = 1.0 / n
    c, o, _, _, _ = util.lib_stencil("Laplacian", dim=2)
    st = pa.Stencil(c, o, pa.PM_F64, 2)
    u = np.pad(util.uniform(0, 0, n * n).reshape(n, n), 1, mode="wrap").reshape(-1).copy()
    say("config1 256^2 apply")
    dst = apply_gpu(st, u, [-1, -1], [n, n], np.full(n * n, 7.77e77), [0, 0], [n - 1, n - 1], [0, 0],
                    [n - 1, n - 1], True, 1.0 / (h * h))
    assert hashlib.sha256(dst.tobytes()).hexdigest() == meta["sha256"]
    dd = dev(dst)
    out = zeros(1)
    pa.call("pm_reduce", pa.RED_ABS, pa.PM_F64, vp(dd), n * n, vp(out), 0, None)
    assert host(out)[0] == float.fromhex(meta["absmax"])
    pa.call("pm_reduce", pa.RED_SUM, pa.PM_F64, vp(dd), n * n, vp(out), 0, None)
    assert abs(host(out)[0] - float.fromhex(meta["sum"])) <= _sum_bound(dst)


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_reductions_vs_oracle(dtype):
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    for n in (1, 7, 1000, 65536, 3000017):
        x = util.uniform(900 + n, 0, n).astype(npdt)
        xd = dev(x)
        out = zeros(1, npdt)
        for op in range(6):
            say("reduce", dtype, n, op)
            pa.call("pm_reduce", op, dtype, vp(xd), n, vp(out), 0, None)
            got = host(out)[0]
            f = util.oracle().orc_reduce_f64 if dtype == pa.PM_F64 else util.oracle().orc_reduce_f32
            want = f(op, util.ptr(x), n)
            if op in (0, 1, 2):
                assert got == want, (n, op)
            else:
                terms = np.abs(x.astype(np.float64)) if op != 5 else x.astype(np.float64) ** 2
                bound = 2 * n * np.finfo(npdt).eps * float(terms.sum()) + 1e-300
                assert abs(float(got) - float(want)) <= bound, (n, op, got, want)
            # deterministic: same bits twice; combine folds into the old value
            pa.call("pm_reduce", op, dtype, vp(xd), n, vp(out), 0, None)
            assert host(out)[0] == got
    # Max starts at numeric_limits<T>::min() (Proto_ReductionImplem.H:68)
    neg = dev(-np.ones(100, dtype=npdt))
    out = zeros(1, npdt)
    pa.call("pm_reduce", pa.RED_MAX, dtype, vp(neg), 100, vp(out), 0, None)
    assert host(out)[0] == np.finfo(npdt).tiny
    pa.call("pm_reduce_init", pa.RED_ABS, dtype, vp(out), None)
    pa.call("pm_reduce", pa.RED_ABS, dtype, vp(neg), 100, vp(out), 1, None)
    assert host(out)[0] == 1.0


def test_reduce_box_region():
    lo, hi = [-2, -2, -2], [9, 8, 7]
    nc = 2
    x = util.uniform(77, 0, nc * util.box_size(lo, hi))
    xd = dev(x)
    out = zeros(1)
    rlo, rhi = [0, 0, 0], [7, 6, 5]
    pa.call("pm_reduce_box", pa.RED_ABS, pa.PM_F64, vp(xd), ctypes.byref(box(lo, hi)), ctypes.byref(box(rlo, rhi)),
            1, vp(out), 0, None)
    a = x[util.box_size(lo, hi):].reshape(util.box_shape(lo, hi))[2:8, 2:9, 2:10]
    assert host(out)[0] == np.abs(a).max()


def test_boxdata_ops():
    slo, shi = [0, 0, 0], [9, 9, 9]
    dlo, dhi = [-3, 1, 2], [12, 14, 16]
    src = util.uniform(1, 0, util.box_size(slo, shi))
    dst0 = util.uniform(1, 1, util.box_size(dlo, dhi))
    S = src.reshape(util.box_shape(slo, shi))
    D = dst0.reshape(util.box_shape(dlo, dhi)).copy()
    sd, dd = dev(src), dev(dst0)
    region = box([1, 2, 3], [8, 7, 6])
    shift = pa.i3([2, 3, 4])
    pa.call("pm_boxdata_copy", pa.PM_F64, vp(sd), ctypes.byref(box(slo, shi)), vp(dd), ctypes.byref(box(dlo, dhi)),
            ctypes.byref(region), shift, 1, None)
    # dst[p + shift] = src[p], p in region; numpy index = p - lo
    D[3 + 4 - 2:6 + 4 - 2 + 1, 2 + 3 - 1:7 + 3 - 1 + 1, 1 + 2 + 3:8 + 2 + 3 + 1] = S[3:7, 2:8, 1:9]
    assert host(dd).tobytes() == D.reshape(-1).tobytes()
    # increment, binop, scalar on a common box
    a0 = util.uniform(2, 0, 1000)
    b0 = util.uniform(2, 1, 1000) + 3.0
    ad, bd = dev(a0), dev(b0)
    B = box([0, 0, 0], [9, 9, 9])
    pa.call("pm_boxdata_increment", pa.PM_F64, vp(ad), ctypes.byref(B), vp(bd), ctypes.byref(B), ctypes.byref(B),
            0.5, 1, None)
    ref = a0 + b0 * 0.5
    assert host(ad).tobytes() == ref.tobytes()
    for op, fn in ((0, np.add), (1, np.subtract), (2, np.multiply), (3, np.divide)):
        pa.call("pm_boxdata_binop", op, pa.PM_F64, vp(ad), ctypes.byref(B), vp(bd), ctypes.byref(B),
                ctypes.byref(B), 1, None)
        ref = fn(ref, b0)
        assert host(ad).tobytes() == ref.tobytes()
        pa.call("pm_boxdata_scalar", op, pa.PM_F64, vp(ad), ctypes.byref(B), ctypes.byref(B), 1.25, 1, None)
        ref = fn(ref, 1.25)
        assert host(ad).tobytes() == ref.tobytes()
    pa.call("pm_boxdata_scalar", 4, pa.PM_F64, vp(ad), ctypes.byref(B), ctypes.byref(B), -2.5, 1, None)
    assert np.all(host(ad) == -2.5)


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_every_tile_configuration_is_bit_identical(dtype):
    """The autotuner may pick any tile configuration: each one must give the
    oracle's bits (ragged box not a multiple of any tile, odd z extent)."""
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    lo, hi = [-3, 2, 1], [70, 44, 37]
    slo, shi = [x - 1 for x in lo], [x + 1 for x in hi]
    src = util.uniform(61, 0, util.box_size(slo, shi)).astype(npdt)
    f = util.uniform(61, 1, util.box_size(lo, hi)).astype(npdt)
    s_d, f_d = dev(src), dev(f)
    ib, sb = box(lo, hi), box(slo, shi)
    refs = {}
    for name in ("Laplacian", "Laplacian_27"):
        c, o, _, _, _ = util.lib_stencil(name, dtype=npdt)
        ref = np.zeros(util.box_size(lo, hi), dtype=npdt)
        util.oracle_apply(c, o, src, slo, shi, ref, lo, hi, lo, hi, True, 3.0)
        refs[name] = (pa.Stencil(c, o, dtype), ref)
    s7, r7 = refs["Laplacian"]
    s27, r27 = refs["Laplacian_27"]
    out, out2, red = empty(util.box_size(lo, hi), npdt), empty(util.box_size(lo, hi), npdt), zeros(4, npdt)
    for rows in (1, 2, 4):
        for pf in (1, 2):
            for nt in (0, 1):
                for zp in (0, 5, 16):
                    cfg = pa.TileConfig(rows, pf, nt, zp)
                    say("tile config", dtype, rows, pf, nt, zp)
                    for st, ref in ((s7, r7), (s27, r27)):
                        pa.call("pm_stencil_set_tile_config", st.h, None, 0, ctypes.byref(ib), ctypes.byref(cfg))
                        pa.call("pm_stencil_apply", st.h, vp(s_d), ctypes.byref(sb), vp(out), ctypes.byref(ib),
                                ctypes.byref(ib), 1, 1, 3.0, None)
                        assert host(out).tobytes() == ref.tobytes()
                    pa.call("pm_stencil_set_tile_config", s7.h, s27.h, 3, ctypes.byref(ib), ctypes.byref(cfg))
                    pa.call("pm_stencil_residual2", s7.h, s27.h, vp(s_d), ctypes.byref(sb), vp(f_d), vp(out),
                            vp(out2), ctypes.byref(ib), 3.0, 3.0, vp(red), None)
                    assert host(out).tobytes() == (f - r7).tobytes()
                    assert host(out2).tobytes() == (f - r27).tobytes()
    # the autotuner itself runs and records a valid configuration
    best, ms = pa.TileConfig(), ctypes.c_float()
    pa.call("pm_stencil_autotune", s7.h, s27.h, 3, ctypes.byref(sb), ctypes.byref(ib), 2, ctypes.byref(best),
            ctypes.byref(ms), None)
    got = pa.TileConfig()
    pa.call("pm_stencil_get_tile_config", s7.h, s27.h, 3, ctypes.byref(ib), ctypes.byref(got))
    assert got.as_dict() == best.as_dict() and ms.value > 0


@pytest.mark.parametrize("dtype", [pa.PM_F64, pa.PM_F32])
def test_residual_reductions_every_call(dtype):
    """Back-to-back residual calls (the reduction output cleared in between)
    each rewrite absMax and sum, agree with a host reduction of the residual
    arrays they wrote, and give the same bits every time."""
    n, g = 48, 2
    npdt = np.float64 if dtype == pa.PM_F64 else np.float32
    u, f = config2_inputs(n, g, npdt)
    h = 1.0 / n
    c7, o7, _, _, _ = util.lib_stencil("Laplacian", dtype=npdt)
    c27, o27, _, _, _ = util.lib_stencil("Laplacian_27", dtype=npdt)
    s7, s27 = pa.Stencil(c7, o7, dtype), pa.Stencil(c27, o27, dtype)
    ub, rb = box([-g] * 3, [n - 1 + g] * 3), box([0] * 3, [n - 1] * 3)
    r1, r2 = empty(n ** 3, npdt), empty(n ** 3, npdt)
    red, red1 = zeros(4, npdt), zeros(2, npdt)
    sc = 1.0 / (h * h)
    first = None
    for it in range(3):
        pa.call("pm_memset_zero", vp(red), 4 * np.dtype(npdt).itemsize, None)
        pa.call("pm_memset_zero", vp(red1), 2 * np.dtype(npdt).itemsize, None)
        pa.call("pm_stencil_residual2", s7.h, s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r1), vp(r2),
                ctypes.byref(rb), sc, sc, vp(red), None)
        pa.call("pm_stencil_residual", s27.h, vp(u), ctypes.byref(ub), vp(f), vp(r2), ctypes.byref(rb), sc,
                vp(red1), None)
        R, R1 = host(red), host(red1)
        a, b = host(r1).astype(np.float64), host(r2).astype(np.float64)
        assert R[0] == np.abs(a).max() and R[2] == np.abs(b).max() and R1[0] == R[2]
        tol = n ** 3 * np.finfo(npdt).eps
        assert abs(R[1] - a.sum()) <= tol * np.abs(a).sum() and abs(R[3] - b.sum()) <= tol * np.abs(b).sum()
        if first is None:
            first = R.tobytes()
        assert R.tobytes() == first  # deterministic
