"""Pinning the oracle (oracle/cpu_ref.c) on the CPU: Proto's golden vectors,
the compiled reference where it is built, and closed forms chosen so that a
dropped term, a wrong sign or a swapped index fails at least one of them."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import util
from util import i3, ptr, uniform

G = util.GOLDEN


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def cases():
    z = load("stencil_apply.npz")
    keys = sorted({k.split(":")[0] for k in z.files})
    return z, keys


def test_stencil_library_order_and_values():
    # Laplacian: centre -2*DIM first, then +e_d, -e_d (Proto_StencilDefs.H:77-87)
    c, o, sr, dr, ds = util.lib_stencil("Laplacian")
    assert c.tolist() == [-6.0, 1, 1, 1, 1, 1, 1]
    assert o.tolist() == [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    # Laplacian_27: centre -64/15, then Box::Cube(3)-1 x-fastest minus the centre
    c, o, _, _, _ = util.lib_stencil("Laplacian_27")
    assert len(c) == 27 and c[0] == -64.0 / 15.0 and o[0].tolist() == [0, 0, 0]
    cube = [[x, y, z] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1) if (x, y, z) != (0, 0, 0)]
    assert o[1:].tolist() == cube
    w = {1: 7.0 / 15.0, 2: 1.0 / 10.0, 3: 1.0 / 30.0}
    assert all(c[j] == w[int(np.abs(o[j]).sum())] for j in range(1, 27))
    # coefficients do not sum to 0 in FP64 (sequential sum)
    s = 0.0
    for x in c:
        s += x
    assert s != 0.0 and abs(s) < 1e-15
    # AvgDown(2): 1/8 over Box(2) x-fastest with srcRatio 2
    c, o, sr, dr, ds = util.lib_stencil("AvgDown", 2)
    assert c.tolist() == [0.125] * 8 and sr == [2, 2, 2]
    assert o.tolist() == [[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    # float library values are the double constants rounded to float
    cf, _, _, _, _ = util.lib_stencil("Laplacian_27", dtype=np.float32)
    assert np.array_equal(cf, util.lib_stencil("Laplacian_27")[0].astype(np.float32))


@pytest.mark.skipif(util.ref(3) is None, reason="reference checker not built here")
def test_stencil_library_table_is_protos():
    tab = json.load(open(os.path.join(G, "stencil_lib.json")))
    R = util.ref(3)
    coefs = np.zeros(64)
    offs = np.zeros(192, dtype=np.int32)
    sr, dr, ds = i3([0] * 3), i3([0] * 3), i3([0] * 3)
    n = R.ref_stencil_lib(b"Laplacian_27", 0, ptr(coefs), ptr(offs), ptr(sr), ptr(dr), ptr(ds), 64)
    e = tab["3D/Laplacian_27/0/f64"]
    assert [float(x).hex() for x in coefs[:n]] == e["coefs"]
    assert offs[:3 * n].tolist() == e["offsets"]


def merge(c, o):
    """Build the stencil as Proto's addCoef does (merge repeated offsets)."""
    if c.dtype != np.float64:
        return c, o
    mc = np.zeros(len(c))
    mo = np.zeros(3 * len(c), dtype=np.int32)
    m = util.oracle().orc_stencil_merge(len(c), ptr(c), ptr(np.ascontiguousarray(o.reshape(-1), dtype=np.int32)),
                                        ptr(mc), ptr(mo))
    return mc[:m], mo[:3 * m].reshape(m, 3)


def test_stencil_merge_duplicates():
    c, o = merge(np.array([1.0, 2.0, 0.5, -3.0]), np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [0, -1, 0]]))
    assert c.tolist() == [1.5, 2.0, -3.0]
    assert o.tolist() == [[0, 0, 0], [1, 0, 0], [0, -1, 0]]


def _infer(c, o, sr, dr, ds, slo, shi, dlo, dhi, dim):
    blo, bhi = i3([0] * 3), i3([0] * 3)
    span_lo = i3(o.min(axis=0).tolist())
    span_hi = i3(o.max(axis=0).tolist())
    ok = util.oracle().orc_infer_box(dim, ptr(span_lo), ptr(span_hi), ptr(i3(sr)), ptr(i3(dr)),
                                     ptr(i3(ds)), ptr(i3(slo)), ptr(i3(shi)), ptr(i3(dlo)),
                                     ptr(i3(dhi)), ptr(blo), ptr(bhi))
    return ok, blo, bhi


def test_apply_golden_bitwise():
    z, keys = cases()
    assert len(keys) >= 20
    for k in keys:
        g = {f.split(":")[1]: z[f] for f in z.files if f.startswith(k + ":")}
        dim = int(g["dim"])
        c, o = merge(g["coefs"], g["offs"])
        ok, blo, bhi = _infer(c, o, g["sr"], g["dr"], g["ds"], g["slo"], g["shi"], g["dlo"], g["dhi"], dim)
        assert ok
        assert blo.tolist() == g["box"][:3].tolist() and bhi.tolist() == g["box"][3:].tolist(), k
        dst = g["dst0"].copy()
        util.oracle_apply(c, o, g["src"], g["slo"], g["shi"], dst, g["dlo"], g["dhi"], blo, bhi,
                          bool(g["replace"]), g["scale"].item(), g["sr"], g["dr"], g["ds"])
        assert dst.tobytes() == g["dst"].tobytes(), k


@pytest.mark.skipif(util.ref(3) is None, reason="reference checker not built here")
def test_apply_random_vs_reference():
    rng = np.random.default_rng(5)
    R = util.ref(3)
    for trial in range(25):
        n = int(rng.integers(1, 12))
        offs = rng.integers(-2, 3, size=(n, 3)).astype(np.int32)
        coefs = uniform(100 + trial, 3, n)
        slo = rng.integers(-4, 2, size=3)
        shi = slo + rng.integers(5, 12, size=3)
        dlo = slo + rng.integers(0, 3, size=3)
        dhi = dlo + rng.integers(2, 8, size=3)
        replace = bool(trial % 2)
        scale = [1.0, 0.0, 3.5, -2.0][trial % 4]
        src = uniform(200 + trial, 0, util.box_size(slo, shi))
        dst0 = uniform(200 + trial, 1, util.box_size(dlo, dhi))
        a = dst0.copy()
        b = dst0.copy()
        box = np.zeros(6, dtype=np.int32)
        R.ref_stencil_apply(n, ptr(coefs), ptr(offs.reshape(-1).copy()), ptr(i3([1] * 3)), ptr(i3([1] * 3)),
                            ptr(i3([0] * 3)), ptr(src), ptr(i3(slo)), ptr(i3(shi)), ptr(a), ptr(i3(dlo)),
                            ptr(i3(dhi)), ptr(i3([0] * 3)), ptr(i3([-1] * 3)), 1, int(replace),
                            ctypes.c_double(scale), ptr(box))
        mc, mo = merge(coefs, offs)
        ok, blo, bhi = _infer(mc, mo, [1] * 3, [1] * 3, [0] * 3, slo, shi, dlo, dhi, 3)
        if not ok:
            continue
        assert blo.tolist() == box[:3].tolist() and bhi.tolist() == box[3:].tolist()
        util.oracle_apply(mc, mo, src, slo, shi, b, dlo, dhi, blo, bhi, replace, scale)
        assert a.tobytes() == b.tobytes(), trial


def test_box_inference_stencil_tests_cases():
    # tests/StencilTests.cpp:80-190 (Domain_Range / BoxInference): a Laplacian on a
    # source box B yields the index box B.grow(-1); AvgDown(2) on [0,15]^3 -> [0,7]^3;
    # a destination-refined stencil (dr=2) into [0,15] from [0,7] -> [0,7]
    c, o, sr, dr, ds = util.lib_stencil("Laplacian")
    ok, blo, bhi = _infer(c, o, sr, dr, ds, (-1, -1, -1), (8, 8, 8), (-5, -5, -5), (20, 20, 20), 3)
    assert ok and blo.tolist() == [0, 0, 0] and bhi.tolist() == [7, 7, 7]
    c, o, sr, dr, ds = util.lib_stencil("AvgDown", 2)
    ok, blo, bhi = _infer(c, o, sr, dr, ds, (0, 0, 0), (15, 15, 15), (-9, -9, -9), (99, 99, 99), 3)
    assert ok and blo.tolist() == [0, 0, 0] and bhi.tolist() == [7, 7, 7]
    ok, blo, bhi = _infer(np.ones(1), np.zeros((1, 3), dtype=np.int32), (1, 1, 1), (2, 2, 2), (1, 1, 1),
                          (0, 0, 0), (7, 7, 7), (0, 0, 0), (15, 15, 15), 3)
    assert ok and blo.tolist() == [0, 0, 0] and bhi.tolist() == [7, 7, 7]
    # destShift 1 with odd upper bound: taperCoarsen truncates (15-1)/2 = 7
    ok, blo, bhi = _infer(np.ones(1), np.zeros((1, 3), dtype=np.int32), (1, 1, 1), (2, 2, 2), (1, 1, 1),
                          (0, 0, 0), (9, 9, 9), (0, 0, 0), (14, 14, 14), 3)
    assert ok and bhi.tolist() == [6, 6, 6]


def test_laplacian_of_paraboloid_is_2dim():
    # examples/_archived/Multigrid/exec/Multigrid.cpp:233-240: L(sum x_d^2) = 2*DIM
    for dim, name in ((3, "Laplacian"), (3, "Laplacian_27"), (3, "Laplacian_19"), (2, "Laplacian")):
        c, o, _, _, _ = util.lib_stencil(name, dim=dim)
        n = 6
        lo, hi = [-1] * dim, [n] * dim
        ax = np.arange(-1, n + 1, dtype=np.float64)
        if dim == 3:
            src = (ax[None, None, :] ** 2 + ax[None, :, None] ** 2 + ax[:, None, None] ** 2).reshape(-1)
        else:
            src = (ax[None, :] ** 2 + ax[:, None] ** 2).reshape(-1)
        dst = np.zeros(n ** dim)
        util.oracle_apply(c, o, src.copy(), lo, hi, dst, [0] * dim, [n - 1] * dim, [0] * dim,
                          [n - 1] * dim, True, 1.0)
        assert np.allclose(dst, 2 * dim, rtol=0, atol=1e-12), name
        if name == "Laplacian":
            assert np.all(dst == 2 * dim)


def test_constant_input():
    c, o, _, _, _ = util.lib_stencil("Laplacian")
    src = np.full(10 ** 3, 3.25)
    dst = np.ones(8 ** 3)
    util.oracle_apply(c, o, src, [-1] * 3, [8] * 3, dst, [0] * 3, [7] * 3, [0] * 3, [7] * 3, True, 2.0)
    assert np.all(dst == 0.0)
    c, o, _, _, _ = util.lib_stencil("Laplacian_27")
    src = np.ones(10 ** 3)
    util.oracle_apply(c, o, src, [-1] * 3, [8] * 3, dst, [0] * 3, [7] * 3, [0] * 3, [7] * 3, True, 1.0)
    s = 0.0
    for x in c:
        s += x
    assert np.all(dst == s) and s != 0  # the ~1e-17 remainder of Proto's FP64 weights


def test_scale_zero_and_accumulate_semantics():
    c, o, _, _, _ = util.lib_stencil("Laplacian")
    src = uniform(3, 0, 10 ** 3)
    d0 = uniform(3, 1, 8 ** 3)
    d = d0.copy()
    util.oracle_apply(c, o, src, [-1] * 3, [8] * 3, d, [0] * 3, [7] * 3, [0] * 3, [7] * 3, False, 0.0)
    assert d.tobytes() == d0.tobytes()  # += with scale 0 leaves dst alone
    util.oracle_apply(c, o, src, [-1] * 3, [8] * 3, d, [0] * 3, [7] * 3, [0] * 3, [7] * 3, True, 0.0)
    assert np.all(d == 0)  # |= with scale 0 zeroes (hostApply 750-768)


def test_reductions():
    z = load("reductions.npz")
    o = util.oracle()
    for name in ("rand", "neg", "ints"):
        d = z[name + ":data"]
        got = [o.orc_reduce_f64(op, ptr(d), len(d)) for op in range(6)]
        assert np.array(got).tobytes() == z[name + ":f64"].tobytes(), name
    d = z["rand:data_f32"]
    got = np.array([o.orc_reduce_f32(op, ptr(d), len(d)) for op in range(6)], dtype=np.float32)
    assert got.tobytes() == z["rand:f32"].tobytes()
    # Max starts at numeric_limits<T>::min(): all-negative data gives 2.2e-308
    assert z["neg:f64"][0] == np.finfo(np.float64).tiny
    # tests/ReductionTests.cpp:17-31 closed forms for +-ii data
    ints = z["ints:data"]
    n = len(ints)
    assert z["ints:f64"][2] == n - 1 if (n - 1) % 2 else True
    assert z["ints:f64"][4] == sum(range(n))
    assert z["ints:f64"][3] == sum(i if i % 2 == 0 else -i for i in range(n))
    assert z["ints:f64"][5] == sum(i * i for i in range(n))


def test_interp_golden_and_invariants():
    z = load("interp.npz")
    for linear in (0, 1):
        for rep in (0, 1):
            k = "lin%d_rep%d" % (linear, rep)
            clo, chi, flo, fhi = z[k + ":boxes"]
            fine = z[k + ":fine0"].copy()
            util.oracle().orc_interp_apply(3, linear, 2, ptr(z[k + ":crse"]), ptr(i3(clo)), ptr(i3(chi)),
                                           ptr(fine), ptr(i3(flo)), ptr(i3(fhi)), rep)
            assert fine.tobytes() == z[k + ":fine"].tobytes(), k
    # prolongation of 1 is 1 (Multigrid.cpp:226-231); Constant is injection
    # (tests/InterpTests.cpp:42-113); Linear reproduces a linear function
    clo, chi, flo, fhi = i3([-1] * 3), i3([4] * 3), i3([0] * 3), i3([7] * 3)
    for linear in (0, 1):
        fine = np.zeros(512)
        util.oracle().orc_interp_apply(3, linear, 2, ptr(np.ones(216)), ptr(clo), ptr(chi), ptr(fine),
                                       ptr(flo), ptr(fhi), 1)
        assert np.all(fine == 1.0)
    a = np.arange(-1, 5, dtype=np.float64)
    crse = np.broadcast_to(a[None, None, :] + 2 * a[None, :, None] - a[:, None, None], (6, 6, 6)).copy()
    fine = np.zeros(512)
    util.oracle().orc_interp_apply(3, 0, 2, ptr(crse.reshape(-1)), ptr(clo), ptr(chi), ptr(fine),
                                   ptr(flo), ptr(fhi), 1)
    f = fine.reshape(8, 8, 8)
    idx = np.arange(8) // 2
    assert np.array_equal(f, crse[1:, 1:, 1:][idx][:, idx][:, :, idx])
    fine = np.zeros(512)
    util.oracle().orc_interp_apply(3, 1, 2, ptr(crse.reshape(-1)), ptr(clo), ptr(chi), ptr(fine),
                                   ptr(flo), ptr(fhi), 1)
    f = fine.reshape(8, 8, 8)
    t = np.arange(8) / 2.0  # vertex-style: fine 2i+s at coarse coordinate i + s/2
    expect = t[None, None, :] + 2 * t[None, :, None] - t[:, None, None]
    assert np.array_equal(f, expect)


def test_exchange_golden_and_brute_force():
    z = load("exchange.npz")
    for key in ("n8_bs4_g1", "n8_bs4_g2", "n12_bs4_g1"):
        n, bs, g = [int(t[1:].replace("s", "")) if False else None for t in key.split("_")][0:3]
        parts = key.split("_")
        n, bs, g = int(parts[0][1:]), int(parts[1][2:]), int(parts[2][1:])
        glob = z[key + ":global"]
        nb = n // bs
        e = bs + 2 * g
        out = np.zeros(nb ** 3 * e ** 3)
        util.oracle().orc_level_exchange(n, bs, g, ptr(glob), ptr(out))
        out = out.reshape(nb, nb, nb, e, e, e)
        lows = z[key + ":lows"]
        gold = z[key + ":boxes"]
        for ib in range(len(lows)):
            bx, by, bz = (lows[ib] // bs).tolist()
            assert out[bz, by, bx].tobytes() == gold[ib].tobytes(), (key, ib)
        # brute force periodic image for one box
        G3 = glob.reshape(n, n, n)
        idx = (np.arange(-g, bs + g)) % n
        assert np.array_equal(out[0, 0, 0], G3[idx][:, idx][:, :, idx])


def _mg(relax, linear, n=32, maxiter=20):
    rhs = util.gaussian_rhs(n, 6, 2)
    phi = np.zeros(n ** 3)
    hist = np.zeros(maxiter + 2)
    it = util.oracle().orc_mg_solve(n, 4, relax, 2, 2, 10, linear, ptr(rhs), ptr(phi), maxiter,
                                    ctypes.c_double(1e-10), ptr(hist))
    return it, hist[:it + 1], phi


@pytest.mark.parametrize("key,relax,linear", [("gsrb_const", 0, 0), ("gsrb_linear", 0, 1), ("jacobi", 1, 0)])
def test_multigrid_residual_history_golden(key, relax, linear):
    z = load("multigrid.npz")
    meta = json.load(open(os.path.join(G, "multigrid.json")))
    it, hist, phi = _mg(relax, linear)
    assert it == meta[key]["iters"]
    assert hist.tobytes() == z[key + ":hist"].tobytes()
    assert hashlib.sha256(phi.tobytes()).hexdigest() == meta[key]["phi_sha256"]
    # the V-cycle converges (Proto's vertex-style Linear prolongation is
    # mis-centred for cell data and only contracts by about 1/2 per cycle)
    if linear:
        assert np.all(np.diff(hist) < 0) and hist[-1] < 1e-7 * hist[0]
    else:
        assert hist[-1] < 1e-10 * hist[0]


def test_gsrb_relax_golden_and_halfsweep_invariant():
    z = load("multigrid.npz")
    n = 32
    phi = uniform(70, 0, n ** 3)
    rhs = uniform(70, 1, n ** 3)
    util.oracle().orc_gsrb_relax(n, 1.0 / n, 3, ptr(rhs), ptr(phi))
    assert phi.tobytes() == z["relax:phi"].tobytes()
    # after a full GSRB iteration the residual vanishes (to rounding) at the
    # colour updated last (black, odd parity) and not at the other one
    P = phi.reshape(n, n, n)
    h2 = (1.0 / n) ** 2
    lap = (sum(np.roll(P, s, a) for a in range(3) for s in (1, -1)) - 6 * P) / h2
    res = (lap - rhs.reshape(n, n, n))
    par = (np.arange(n)[None, None, :] + np.arange(n)[None, :, None] + np.arange(n)[:, None, None]) % 2
    scale = np.abs(rhs).max() + np.abs(lap).max()
    assert np.abs(res[par == 1]).max() < 1e-12 * scale
    assert np.abs(res[par == 0]).max() > 1e-3


def test_config1_full_size_golden():
    meta = json.load(open(os.path.join(G, "config1.json")))
    n = 256
    c, o, _, _, _ = util.lib_stencil("Laplacian", dim=2)
    u = np.pad(uniform(0, 0, n * n).reshape(n, n), 1, mode="wrap").reshape(-1).copy()
    dst = np.zeros(n * n)
    h = 1.0 / 256
    util.oracle_apply(c, o, u, [-1, -1], [n, n], dst, [0, 0], [n - 1, n - 1], [0, 0], [n - 1, n - 1],
                      True, 1.0 / (h * h))
    assert hashlib.sha256(dst.tobytes()).hexdigest() == meta["sha256"]
    assert util.oracle().orc_reduce_f64(2, ptr(dst), dst.size).hex() == meta["absmax"]
    assert util.oracle().orc_reduce_f64(3, ptr(dst), dst.size).hex() == meta["sum"]


def test_config2_small_golden_and_analytic():
    z = load("config2_small.npz")
    n, g = 32, 2
    u = util.config2_u(n, g)
    f = uniform(1, 1, n ** 3)
    h = 1.0 / n
    for name in ("Laplacian", "Laplacian_27"):
        c, o, _, _, _ = util.lib_stencil(name)
        r = np.zeros(n ** 3)
        util.oracle().orc_residual_f64(len(c), ptr(c), ptr(o.reshape(-1).copy()), ptr(u), ptr(i3([-g] * 3)),
                                       ptr(i3([n - 1 + g] * 3)), ptr(f), ptr(r), ptr(i3([0] * 3)),
                                       ptr(i3([n - 1] * 3)), ctypes.c_double(1.0 / (h * h)))
        assert r.tobytes() == z[name + ":r"].tobytes()
        assert util.oracle().orc_reduce_f64(2, ptr(r), r.size) == z[name + ":absmax"]
        assert util.oracle().orc_reduce_f64(3, ptr(r), r.size) == z[name + ":sum"]
    # analytic check on the smooth part (linearity removes the noise):
    # L7 s/h^2 = (-12 pi^2 + 4 pi^4 h^2) s + O(h^4), L27: (-12 pi^2 + 12 pi^4 h^2) s
    x = (np.arange(-g, n + g) + 0.5) / n
    s1 = np.sin(2 * np.pi * x)
    s = ((s1[None, None, :] * s1[None, :, None]) * s1[:, None, None]).reshape(-1).copy()
    for name, k in (("Laplacian", 4), ("Laplacian_27", 12)):
        c, o, _, _, _ = util.lib_stencil(name)
        Ls = np.zeros(n ** 3)
        util.oracle_apply(c, o, s, [-g] * 3, [n - 1 + g] * 3, Ls, [0] * 3, [n - 1] * 3, [0] * 3,
                          [n - 1] * 3, True, 1.0 / (h * h))
        exact = (-12 * np.pi ** 2 + k * np.pi ** 4 * h * h) * s.reshape(n + 2 * g, n + 2 * g, n + 2 * g)[
            g:-g, g:-g, g:-g].reshape(-1)
        assert np.abs(Ls - exact).max() < 60 * (2 * np.pi * h) ** 4 * 12 * np.pi ** 2


def test_uniform_generator():
    a = np.zeros(1000)
    util.oracle().orc_uniform_f64(7, 3, 1000, ptr(a))
    assert a.tobytes() == uniform(7, 3, 1000).tobytes()
    assert np.array_equal(uniform(7, 3, 10, start=990), a[990:])
    assert a.min() >= -1.0 and a.max() < 1.0 and abs(a.mean()) < 0.1
    # known first value of stream (0, 0): SplitMix64's first output from state 0
    z = 0x9E3779B97F4A7C15
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    z ^= z >> 31
    assert z == 0xE220A8397B1DCDAF
    assert uniform(0, 0, 1)[0] == ((z >> 11) * 2.0 ** -53) * 2 - 1


@pytest.mark.parametrize("nx,ny,nz,g", [(7, 5, 6, 2), (8, 8, 8, 1), (13, 4, 9, 2)])
def test_patch_gsrb_golden_and_invariant(nx, ny, nz, g):
    z = load("multigrid.npz")
    ph = uniform(80 + nx, 0, (nx + 2 * g) * (ny + 2 * g) * (nz + 2 * g))
    rh = uniform(80 + nx, 1, nx * ny * nz)
    dx = 1.0 / 32
    util.oracle().orc_patch_gsrb_f64(nx, ny, nz, g, ctypes.c_double(dx), ptr(rh), ptr(ph))
    assert ph.tobytes() == z["patch_%d_%d_%d_%d" % (nx, ny, nz, g)].tobytes()
    # black cells (updated last, from final red neighbours and fixed ghosts)
    # satisfy the discrete equation to rounding; red cells do not
    P = ph.reshape(nz + 2 * g, ny + 2 * g, nx + 2 * g)
    c = P[g:-g, g:-g, g:-g]
    lap = (P[g:-g, g:-g, g + 1:nx + g + 1] + P[g:-g, g:-g, g - 1:nx + g - 1] + P[g:-g, g + 1:ny + g + 1, g:-g]
           + P[g:-g, g - 1:ny + g - 1, g:-g] + P[g + 1:nz + g + 1, g:-g, g:-g] + P[g - 1:nz + g - 1, g:-g, g:-g]
           - 6 * c) / (dx * dx)
    res = lap - rh.reshape(nz, ny, nx)
    par = (np.arange(nx)[None, None, :] + np.arange(ny)[None, :, None] + np.arange(nz)[:, None, None]) % 2
    assert np.abs(res[par == 1]).max() < 1e-9 * (np.abs(lap).max() + 1)
    assert np.abs(res[par == 0]).max() > 1e-3
    # the f32 form stays within float rounding of the f64 sweep
    ph32 = uniform(80 + nx, 0, ph.size).astype(np.float32)
    util.oracle().orc_patch_gsrb_f32(nx, ny, nz, g, ctypes.c_double(dx), ptr(rh.astype(np.float32)), ptr(ph32))
    assert np.abs(ph32 - ph).max() < 1e-4 * (np.abs(ph).max() + 1)


def test_fas_matches_proto_history():
    """Library FAS (LevelMultigridSolver<BoxOp_Laplace>) restatement: Proto's
    residual history and solution bits at 32^3, 4 levels."""
    z = load("multigrid.npz")
    meta = json.load(open(os.path.join(util.GOLDEN, "multigrid.json")))
    n = 32
    rhs = util.gaussian_rhs(n, 6, 2)
    phi = np.zeros(n ** 3)
    hist = np.zeros(32)
    it = util.oracle().orc_fas_solve(n, 4, ptr(rhs), ptr(phi), 20, 1e-10, ptr(hist))
    assert it == meta["fas"]["iters"]
    assert hist[:it + 1].tobytes() == z["fas:hist"].tobytes()
    assert hashlib.sha256(phi.tobytes()).hexdigest() == meta["fas"]["phi_sha256"]
    # the last residual meets the tolerance and the history falls monotonically
    assert hist[it] < 1e-10 * hist[0] and np.all(np.diff(hist[:it + 1]) < 0)


@pytest.mark.parametrize("n,bs,g", [(16, 8, 1), (8, 4, 2)])
def test_linear_out_format_is_the_level_storage(n, bs, g):
    """Proto's LevelBoxData::linearOut buffer (written by Proto) is box after
    box in layout order, component after component, each over the ghosted box
    x fastest -- the one-allocation level storage of this library."""
    z = load("linear.npz")
    key = "n%d_bs%d_g%d" % (n, bs, g)
    parts = []
    for lo in z[key + ":lows"]:
        r = [np.arange(lo[d] - g, lo[d] + bs + g) for d in range(3)]
        pz, py, px = np.meshgrid(r[2], r[1], r[0], indexing="ij")
        for c in range(2):
            parts.append((px + 1000.0 * py + 1e6 * pz + 0.5 * c).reshape(-1))
    assert np.concatenate(parts).tobytes() == z[key + ":buffer"].tobytes()
