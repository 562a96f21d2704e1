"""Write the golden fixtures in this folder from Proto's own CPU code.

Run in a container where the reference tree exists and oracle/_ref is built
(`make -C oracle ref`):  python tests/golden/make_golden.py

Every expected value below is produced by the compiled reference checker
(oracle/ref_driver.cpp driving Proto's headers); inputs come from the seeded
SplitMix64 generator of tests/util.py.  The fixtures are data only: inputs,
boxes, and Proto's outputs.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import util  # noqa: E402
from util import i3, ptr, uniform  # noqa: E402


def ref_lib(R, name, arg, f32):
    coefs = np.zeros(400, dtype=np.float32 if f32 else np.float64)
    offs = np.zeros(1200, dtype=np.int32)
    sr, dr, ds = i3([0, 0, 0]), i3([0, 0, 0]), i3([0, 0, 0])
    fn = R.ref_stencil_lib_f if f32 else R.ref_stencil_lib
    n = fn(name.encode(), arg, ptr(coefs), ptr(offs), ptr(sr), ptr(dr), ptr(ds), 400)
    return n, coefs[:n], offs[:3 * n].reshape(n, 3), sr, dr, ds


def stencil_table():
    tab = {}
    for dim in (2, 3):
        R = util.ref(dim)
        names = [("Laplacian", 0), ("AvgDown", 2), ("Sum", 2)]
        names += [("FluxDivergence", d) for d in range(dim)]
        names += [("DiffCellToFaceLo", d) for d in range(dim)]
        names += [("DiffCellToFaceHi", d) for d in range(dim)]
        names += [("Laplacian_27", 0), ("Laplacian_19", 0)] if dim == 3 else [("Laplacian_9", 0)]
        for name, arg in names:
            for f32 in (False, True):
                if f32 and name.startswith("DiffCellToFace"):
                    continue
                n, c, o, sr, dr, ds = ref_lib(R, name, arg, f32)
                key = "%dD/%s/%d/%s" % (dim, name, arg, "f32" if f32 else "f64")
                tab[key] = {
                    "coefs": [float(x).hex() for x in c],
                    "offsets": o.reshape(-1).tolist(),
                    "sr": sr.tolist(), "dr": dr.tolist(), "ds": ds.tolist(),
                }
    with open(os.path.join(HERE, "stencil_lib.json"), "w") as f:
        json.dump(tab, f, indent=0, sort_keys=True)
    return tab


def ext_cases(dim):
    """(name, a, b, c) argument sets of the rest of Proto's stencil library."""
    cases = [("Identity", 0, 0, 0)]
    for d in range(dim):
        cases.append(("LaplacianFace", d, 2, 0))
        for side in (0, 1):
            cases += [("CellToFace", d, side, o) for o in (2, 4, 5)]
            cases += [("DiffCellToFace", d, side, o) for o in (2, 4)]
            cases += [("AvgDownFace", d, side, r) for r in (2, 4)]
        cases += [("CellToFaceL", d, o, 0) for o in (2, 4, 5)]
        cases += [("CellToFaceH", d, o, 0) for o in (2, 4, 5)]
        cases += [("faceToCell", d, o, 0) for o in (2, 4)]
        cases += [("cornersToFaces", d, o, 0) for o in (2, 4)]
        for n_, o in ((1, 1), (1, 2), (1, 4), (1, 6), (1, 12), (2, 2), (2, 4), (2, 6), (2, 12), (3, 2), (3, 4),
                      (4, 2), (4, 4), (5, 2), (6, 4), (7, 2), (12, 2)):
            cases.append(("Derivative", n_, d, o))
    cases += [("CornersToCells", o, 0, 0) for o in (2, 4)]
    return cases


def stencil_table_ext():
    """Coefficient lists of LaplacianFace, CellToFace(L/H), DiffCellToFace
    (orders 2 and 4), Derivative, AvgDownFace, faceToCell, cornersToFaces,
    CornersToCells and Identity, FP64, as Proto builds them."""
    tab = {}
    for dim in (2, 3):
        R = util.ref(dim)
        for name, a, b, c in ext_cases(dim):
            coefs = np.zeros(400)
            offs = np.zeros(1200, dtype=np.int32)
            sr, dr, ds = i3([0, 0, 0]), i3([0, 0, 0]), i3([0, 0, 0])
            n = R.ref_stencil_lib_ext(name.encode(), a, b, c, ptr(coefs), ptr(offs), ptr(sr), ptr(dr), ptr(ds), 400)
            assert n > 0, (dim, name, a, b, c, n)
            tab["%dD/%s/%d_%d_%d" % (dim, name, a, b, c)] = {
                "coefs": [float(x).hex() for x in coefs[:n]],
                "offsets": offs[:3 * n].tolist(),
                "sr": sr.tolist(), "dr": dr.tolist(), "ds": ds.tolist(),
            }
    with open(os.path.join(HERE, "stencil_lib_ext.json"), "w") as f:
        json.dump(tab, f, indent=0, sort_keys=True)
    return tab


def ref_apply(dim, coefs, offs, sr, dr, ds, src, slo, shi, dst, dlo, dhi, infer, replace, scale,
              blo=(0, 0, 0), bhi=(-1, -1, -1)):
    R = util.ref(dim)
    out = np.zeros(6, dtype=np.int32)
    f32 = coefs.dtype == np.float32
    fn = R.ref_stencil_apply_f if f32 else R.ref_stencil_apply
    ct = ctypes.c_float if f32 else ctypes.c_double
    fn(len(coefs), ptr(coefs), ptr(np.ascontiguousarray(offs.reshape(-1), dtype=np.int32)),
       ptr(i3(sr)), ptr(i3(dr)), ptr(i3(ds)), ptr(src), ptr(i3(slo)), ptr(i3(shi)),
       ptr(dst), ptr(i3(dlo)), ptr(i3(dhi)), ptr(i3(blo)), ptr(i3(bhi)),
       int(infer), int(replace), ct(scale), ptr(out))
    return out


def apply_cases(tab):
    """Each case: stencil, src/dst boxes, replace, scale, inputs, Proto's output."""
    cases = {}

    def lib(dim, name, arg, f32):
        e = tab["%dD/%s/%d/%s" % (dim, name, arg, "f32" if f32 else "f64")]
        dt = np.float32 if f32 else np.float64
        return (np.array([float.fromhex(c) for c in e["coefs"]], dtype=dt),
                np.array(e["offsets"], dtype=np.int32).reshape(-1, 3), e["sr"], e["dr"], e["ds"])

    def add(key, dim, st, slo, shi, dlo, dhi, replace, scale, seed, f32=False, init_dst=True):
        coefs, offs, sr, dr, ds = st
        dt = np.float32 if f32 else np.float64
        src = uniform(seed, 0, util.box_size(slo, shi)).astype(dt)
        dst0 = uniform(seed, 1, util.box_size(dlo, dhi)).astype(dt) if init_dst else \
            np.zeros(util.box_size(dlo, dhi), dtype=dt)
        dst = dst0.copy()
        box = ref_apply(dim, coefs, offs, sr, dr, ds, src, slo, shi, dst, dlo, dhi, True, replace, scale)
        cases[key] = dict(dim=dim, coefs=coefs, offs=offs, sr=np.array(sr), dr=np.array(dr),
                          ds=np.array(ds), slo=np.array(slo), shi=np.array(shi), dlo=np.array(dlo),
                          dhi=np.array(dhi), replace=replace, scale=np.array(scale, dtype=dt),
                          src=src, dst0=dst0, dst=dst, box=box)

    L7 = lib(3, "Laplacian", 0, False)
    L27 = lib(3, "Laplacian_27", 0, False)
    L5 = lib(2, "Laplacian", 0, False)
    add("L7_replace", 3, L7, (-1, -1, -1), (12, 10, 9), (0, 0, 0), (11, 9, 8), True, 144.0, 11)
    add("L7_accum", 3, L7, (-1, -1, -1), (12, 10, 9), (0, 0, 0), (11, 9, 8), False, 3.0, 12)
    add("L7_scale1_g2", 3, L7, (-2, -2, -2), (13, 13, 13), (0, 0, 0), (11, 11, 11), True, 1.0, 13)
    add("L7_scale0", 3, L7, (-1, -1, -1), (8, 8, 8), (0, 0, 0), (7, 7, 7), True, 0.0, 14)
    add("L7_offset_boxes", 3, L7, (-5, 3, -2), (9, 14, 7), (-4, 4, -1), (8, 13, 6), True, 262144.0, 15)
    add("L27_replace", 3, L27, (-1, -1, -1), (12, 10, 9), (0, 0, 0), (11, 9, 8), True, 262144.0, 16)
    add("L27_accum", 3, L27, (-2, -2, -2), (13, 11, 10), (0, 0, 0), (11, 9, 8), False, 0.3, 17)
    add("L27_f32", 3, lib(3, "Laplacian_27", 0, True), (-1, -1, -1), (12, 10, 9), (0, 0, 0),
        (11, 9, 8), True, 65536.0, 18, f32=True)
    add("L7_f32", 3, lib(3, "Laplacian", 0, True), (-1, -1, -1), (12, 10, 9), (0, 0, 0),
        (11, 9, 8), True, 3.0, 19, f32=True)
    add("L5_2d", 2, L5, (-1, -1), (16, 16), (0, 0), (15, 15), True, 256.0 * 256.0, 20)
    add("L9_2d", 2, lib(2, "Laplacian_9", 0, False), (-1, -1), (16, 12), (0, 0), (15, 11), True, 7.0, 21)
    add("L19", 3, lib(3, "Laplacian_19", 0, False), (-1, -1, -1), (9, 9, 9), (0, 0, 0), (8, 8, 8),
        True, 5.0, 22)
    add("AvgDown2", 3, lib(3, "AvgDown", 2, False), (0, 0, 0), (15, 11, 7), (0, 0, 0), (7, 5, 3),
        True, 1.0, 23)
    add("AvgDown2_2d", 2, lib(2, "AvgDown", 2, False), (0, 0), (15, 11), (0, 0), (7, 5), True, 1.0, 24)
    # GSRB sub-lattice stencil of LaplaceOp.H:56-73: Laplacian * (-1/h^2) Shift(k), ratios 2, shift k
    h = 1.0 / 16
    k = (1, 0, 1)
    c, o, _, _, _ = L7
    gs = ((c * (-1.0 / (h * h))).astype(np.float64), o + np.array(k, dtype=np.int32),
          [2, 2, 2], [2, 2, 2], list(k))
    add("GSRB_sublattice", 3, gs, (-1, -1, -1), (16, 16, 16), (0, 0, 0), (15, 15, 15), True, 1.0, 25)
    add("GSRB_sublattice_acc", 3, gs, (-1, -1, -1), (16, 16, 16), (0, 0, 0), (15, 15, 15), False, 0.5, 26)
    # a stencil whose span excludes 0 (hostApply's shift-and-recompose path, 737-746)
    pos = (np.array([0.25, -1.5, 3.0], dtype=np.float64),
           np.array([[1, 1, 0], [2, 1, 1], [1, 2, 2]], dtype=np.int32), [1, 1, 1], [1, 1, 1], [0, 0, 0])
    add("span_excludes_zero", 3, pos, (0, 0, 0), (11, 11, 11), (0, 0, 0), (9, 9, 9), True, 2.0, 27)
    # duplicate offsets merge into the first occurrence (addCoef, StencilImplem.H:237-276)
    dup = (np.array([1.0, 2.0, 0.5, -3.0], dtype=np.float64),
           np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [0, -1, 0]], dtype=np.int32), [1, 1, 1], [1, 1, 1],
           [0, 0, 0])
    add("dup_offsets", 3, dup, (0, -1, 0), (9, 9, 9), (0, 0, 0), (9, 9, 9), True, 1.0, 28)
    # FluxDivergence: 2 coefficients, one-sided
    add("FluxDiv_z", 3, lib(3, "FluxDivergence", 2, False), (0, 0, 0), (7, 8, 9), (0, 0, 0), (7, 8, 8),
        True, 1.0, 29)
    # source refinement (coarse <- fine sum) and destination refinement with shift
    add("Sum2", 3, lib(3, "Sum", 2, False), (0, 0, 0), (9, 9, 9), (0, 0, 0), (4, 4, 4), True, 1.0, 30)
    dref = (np.array([1.0, 0.5], dtype=np.float64), np.array([[0, 0, 0], [1, 0, 0]], dtype=np.int32),
            [1, 1, 1], [2, 1, 1], [1, 0, 0])
    add("DestRefine_shift", 3, dref, (0, 0, 0), (7, 5, 4), (0, 0, 0), (15, 5, 4), True, 1.0, 31)
    # config-1 at full size (DIM=2, 256^2, h = 1/256): only the hash of the output
    # and its absMax are kept, the oracle recomputes the rest
    return cases


def save_cases(cases):
    flat = {}
    for key, c in cases.items():
        for k, v in c.items():
            flat["%s:%s" % (key, k)] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "stencil_apply.npz"), **flat)


def reductions():
    R3 = util.ref(3)
    out = {}
    x = uniform(40, 0, 1000)
    xf = x.astype(np.float32)
    neg = -np.abs(uniform(41, 0, 300)) - 1.0
    ints = np.array([(ii if ii % 2 == 0 else -ii) for ii in range(64)], dtype=np.float64)
    for name, arr in (("rand", x), ("neg", neg), ("ints", ints)):
        out["%s:data" % name] = arr
        out["%s:f64" % name] = np.array([R3.ref_reduce(op, ptr(arr), len(arr)) for op in range(6)])
    out["rand:data_f32"] = xf
    out["rand:f32"] = np.array([R3.ref_reduce_f(op, ptr(xf), len(xf)) for op in range(6)],
                               dtype=np.float32)
    np.savez_compressed(os.path.join(HERE, "reductions.npz"), **out)


def interp():
    R3 = util.ref(3)
    out = {}
    for linear in (0, 1):
        clo, chi = (-1, -1, -1), (4, 3, 5)
        flo, fhi = (0, 0, 0), (7, 5, 9)
        crse = uniform(50 + linear, 0, util.box_size(clo, chi))
        fine0 = uniform(52 + linear, 1, util.box_size(flo, fhi))
        for replace in (0, 1):
            fine = fine0.copy()
            R3.ref_interp_apply(linear, 2, ptr(crse), ptr(i3(clo)), ptr(i3(chi)), ptr(fine),
                                ptr(i3(flo)), ptr(i3(fhi)), replace)
            key = "lin%d_rep%d" % (linear, replace)
            out[key + ":crse"] = crse
            out[key + ":fine0"] = fine0
            out[key + ":fine"] = fine
            out[key + ":boxes"] = np.array([clo, chi, flo, fhi])
    np.savez_compressed(os.path.join(HERE, "interp.npz"), **out)


def interp_hi():
    """Quadratic and FiniteVolume InterpStencils applied by Proto, as a binary
    fixture the C++ API test reads: per case int32 [kind, order, ratio,
    replace, clo[3], chi[3], flo[3], fhi[3]], then float64 crse, fine0, fine."""
    R3 = util.ref(3)
    cases = [(2, 0, 2), (2, 0, 4), (3, 2, 2), (3, 4, 2), (3, 5, 2), (3, 3, 4), (3, 4, 4)]
    clo, chi = (-2, -2, -2), (5, 4, 6)
    buf = bytearray()
    for k, (kind, order, ratio) in enumerate(cases):
        flo = (0, 0, 0)
        fhi = (4 * ratio - 1, 3 * ratio - 1, 5 * ratio - 1)
        crse = uniform(60 + k, 0, util.box_size(clo, chi))
        fine0 = uniform(60 + k, 1, util.box_size(flo, fhi))
        for replace in (0, 1):
            fine = fine0.copy()
            R3.ref_interp_apply_kind(kind, order, ratio, ptr(crse), ptr(i3(clo)), ptr(i3(chi)), ptr(fine),
                                     ptr(i3(flo)), ptr(i3(fhi)), replace)
            head = np.array([kind, order, ratio, replace, *clo, *chi, *flo, *fhi], dtype=np.int32)
            buf += head.tobytes() + crse.tobytes() + fine0.tobytes() + fine.tobytes()
    with open(os.path.join(HERE, "interp_hi.bin"), "wb") as f:
        f.write(bytes(buf))


def exchange():
    R3 = util.ref(3)
    out = {}
    for n, bs, g in ((8, 4, 1), (8, 4, 2), (12, 4, 1)):
        glob = uniform(60, 0, n ** 3)
        nb = (n // bs) ** 3
        e = bs + 2 * g
        lows = np.zeros(3 * nb, dtype=np.int32)
        boxes = np.zeros(nb * e ** 3)
        R3.ref_level_exchange(n, bs, g, ptr(glob), ctypes.c_double(-99.0), ptr(lows), ptr(boxes))
        key = "n%d_bs%d_g%d" % (n, bs, g)
        out[key + ":global"] = glob
        out[key + ":lows"] = lows.reshape(nb, 3)
        out[key + ":boxes"] = boxes.reshape(nb, e, e, e)
    np.savez_compressed(os.path.join(HERE, "exchange.npz"), **out)


def multigrid():
    R3 = util.ref(3)
    out = {}
    meta = {}
    n = 32
    rhs = util.gaussian_rhs(n, 6, 2)
    runs = {
        "gsrb_const": lambda p, h: R3.ref_mg_gsrb_solve(n, 16, 4, 2, 2, 10, 0, ptr(rhs), ptr(p), 20,
                                                       ctypes.c_double(1e-10), ptr(h)),
        "gsrb_linear": lambda p, h: R3.ref_mg_gsrb_solve(n, 16, 4, 2, 2, 10, 1, ptr(rhs), ptr(p), 20,
                                                        ctypes.c_double(1e-10), ptr(h)),
        "jacobi": lambda p, h: R3.ref_mg_jacobi_solve(n, 16, 4, ptr(rhs), ptr(p), 20,
                                                      ctypes.c_double(1e-10), ptr(h)),
        "fas": lambda p, h: R3.ref_mg_fas_solve(n, 16, 4, ptr(rhs), ptr(p), 20,
                                                ctypes.c_double(1e-10), ptr(h)),
    }
    for key, fn in runs.items():
        phi = np.zeros(n ** 3)
        hist = np.zeros(32)
        it = fn(phi, hist)
        out[key + ":hist"] = hist[:it + 1]
        meta[key] = {"iters": it, "phi_sha256": hashlib.sha256(phi.tobytes()).hexdigest(),
                     "phi_absmax": float(np.abs(phi).max()).hex()}
    # one GSRB relaxation pass on its own
    phi = uniform(70, 0, n ** 3)
    rhs2 = uniform(70, 1, n ** 3)
    R3.ref_gsrb_relax(n, 16, 3, ptr(rhs2), ptr(phi))
    out["relax:phi"] = phi
    # one GSRB sweep on isolated patches with fixed ghosts (odd and even extents)
    for (nx, ny, nz, g) in ((7, 5, 6, 2), (8, 8, 8, 1), (13, 4, 9, 2)):
        key = "patch_%d_%d_%d_%d" % (nx, ny, nz, g)
        ph = uniform(80 + nx, 0, (nx + 2 * g) * (ny + 2 * g) * (nz + 2 * g))
        rh = uniform(80 + nx, 1, nx * ny * nz)
        R3.ref_patch_gsrb(nx, ny, nz, g, ctypes.c_double(1.0 / 32), ptr(rh), ptr(ph))
        out[key] = ph
    np.savez_compressed(os.path.join(HERE, "multigrid.npz"), **out)
    with open(os.path.join(HERE, "multigrid.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def config1():
    """DIM=2, 256^2, 1 ghost, u = U[-1,1) seed 0 on the interior, periodic ghost
    fill, dst |= Laplacian()(u, 1/h^2), h = 1/256, then absMax."""
    R2 = util.ref(2)
    n = 256
    tab = json.load(open(os.path.join(HERE, "stencil_lib.json")))
    e = tab["2D/Laplacian/0/f64"]
    coefs = np.array([float.fromhex(c) for c in e["coefs"]])
    offs = np.array(e["offsets"], dtype=np.int32)
    interior = uniform(0, 0, n * n).reshape(n, n)
    u = np.pad(interior, 1, mode="wrap")
    dst = np.zeros(n * n)
    h = 1.0 / 256
    ref_apply(2, coefs, offs, [1, 1, 1], [1, 1, 1], [0, 0, 0], np.ascontiguousarray(u.reshape(-1)),
              (-1, -1), (n, n), dst, (0, 0), (n - 1, n - 1), True, True, 1.0 / (h * h))
    amax = R2.ref_reduce(2, ptr(dst), n * n)
    meta = {"sha256": hashlib.sha256(dst.tobytes()).hexdigest(), "absmax": amax.hex(),
            "sum": R2.ref_reduce(3, ptr(dst), n * n).hex()}
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def config2_small():
    """Config-2 recipe at 32^3 (2 ghosts): u = sin sin sin + 1e-3 U (seed 1,
    stream 0), f = U (seed 1, stream 1), r = f - L(u, 1/h^2) for L7 and L27,
    absMax and sum (Proto's sequential order)."""
    R3 = util.ref(3)
    tab = json.load(open(os.path.join(HERE, "stencil_lib.json")))
    out = {}
    n, g = 32, 2
    u = util.config2_u(n, g)
    f = uniform(1, 1, n ** 3)
    for name in ("Laplacian", "Laplacian_27"):
        e = tab["3D/%s/0/f64" % name]
        coefs = np.array([float.fromhex(c) for c in e["coefs"]])
        offs = np.array(e["offsets"], dtype=np.int32)
        Lu = np.zeros(n ** 3)
        h = 1.0 / n
        ref_apply(3, coefs, offs, [1, 1, 1], [1, 1, 1], [0, 0, 0], u, (-g,) * 3, (n - 1 + g,) * 3, Lu,
                  (0, 0, 0), (n - 1,) * 3, True, True, 1.0 / (h * h))
        r = f - Lu
        out[name + ":r"] = r
        out[name + ":absmax"] = np.array(R3.ref_reduce(2, ptr(r), r.size))
        out[name + ":sum"] = np.array(R3.ref_reduce(3, ptr(r), r.size))
    np.savez_compressed(os.path.join(HERE, "config2_small.npz"), **out)


def linear():
    """LevelBoxData::linearOut buffers written by Proto (2 components, ghosts
    included) with the box lows in Proto's layout order."""
    R3 = util.ref(3)
    R3.ref_level_linear_out.restype = ctypes.c_long
    out = {}
    for (n, bs, g) in ((16, 8, 1), (8, 4, 2)):
        nb = (n // bs) ** 3
        buf = np.zeros(nb * 2 * (bs + 2 * g) ** 3)
        lows = np.zeros(3 * nb, dtype=np.int32)
        nbytes = R3.ref_level_linear_out(n, bs, g, ptr(buf), ptr(lows))
        assert nbytes == buf.nbytes
        key = "n%d_bs%d_g%d" % (n, bs, g)
        out[key + ":buffer"] = buf
        out[key + ":lows"] = lows.reshape(nb, 3)
    np.savez_compressed(os.path.join(HERE, "linear.npz"), **out)


def interp_bnd():
    """interpBoundaries run by Proto (coarse periodic 16^3 in 8^3 boxes, ghost
    1; ratio-2 fine level over patches {1,2,3}^3 of the 32^3 domain, ghost 2;
    ghosts start at -7) with Constant, Linear and Quadratic InterpStencils, as
    a binary fixture the C++ API test reads: int32 ncase, float64 crse (16^3)
    and fine (32^3) valid inputs, then per case int32 [kind, nboxes], int32
    lows[nboxes][3], float64 ghosted data of every fine box in layout order."""
    R3 = util.ref(3)
    crse = uniform(70, 0, 16 ** 3)
    fine = uniform(70, 1, 32 ** 3)
    buf = bytearray(np.array([3], dtype=np.int32).tobytes() + crse.tobytes() + fine.tobytes())
    for kind in (0, 1, 2):
        lows = np.zeros(3 * 27, dtype=np.int32)
        out = np.zeros(27 * 12 ** 3)
        nb = R3.ref_interp_boundaries(kind, ptr(crse), ptr(fine), ctypes.c_double(-7.0), ptr(lows), ptr(out))
        assert nb == 27
        buf += np.array([kind, nb], dtype=np.int32).tobytes() + lows.tobytes() + out.tobytes()
    with open(os.path.join(HERE, "interp_bnd.bin"), "wb") as f:
        f.write(bytes(buf))


def _records(path, recs):
    """Binary fixture read by the C++ API tests: int32 count, then per record
    int32 name length, the name, int64 value count, float64 values."""
    buf = bytearray(np.array([len(recs)], dtype=np.int32).tobytes())
    for name, a in recs:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        nb = name.encode()
        buf += np.array([len(nb)], dtype=np.int32).tobytes() + nb
        buf += np.array([a.size], dtype=np.int64).tobytes() + a.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(buf))


def boxop():
    """Proto's BoxOp / LevelOp / LevelMultigridSolver on seeded inputs, for the
    C++ API test that compiles Proto's own operator headers against this
    library: the FAS solve of LevelMultigridSolver<BoxOp_Laplace> (32^3 in
    16^3 boxes, 4 levels: right-hand side, residual history, solution),
    LevelOp<BoxOp_TestLaplace> on a 16^3 periodic level in 8^3 boxes, and
    BoxOp_TestFlux / _TestSource / _TestDiag on one 16^3 box."""
    R3 = util.ref(3)
    recs = []
    n = 32
    rhs = util.gaussian_rhs(n, 6, 2)
    phi = np.zeros(n ** 3)
    hist = np.zeros(32)
    it = R3.ref_mg_fas_solve(n, 16, 4, ptr(rhs), ptr(phi), 20, ctypes.c_double(1e-10), ptr(hist))
    recs += [("fas:rhs", rhs), ("fas:hist", hist[:it + 1]), ("fas:phi", phi)]
    m = 16
    lp = uniform(90, 0, m ** 3)
    lo = np.zeros(m ** 3)
    R3.ref_levelop_apply(m, 8, ptr(lp), ptr(lo))
    recs += [("levelop:phi", lp), ("levelop:out", lo)]
    src = uniform(91, 0, (m + 2) ** 3)
    aux = uniform(91, 1, (m + 2) ** 3)
    recs += [("boxop:src", src), ("boxop:aux", aux)]
    for which in (0, 1, 2):
        dst = np.zeros(m ** 3)
        flux = np.zeros(3 * (m + 1) * m * m)
        R3.ref_boxop_apply(which, m, ptr(src), ptr(aux), ptr(dst), ptr(flux))
        recs.append(("boxop%d:dst" % which, dst))
        if which == 0:
            recs.append(("boxop0:flux", flux))
    _records(os.path.join(HERE, "boxop.bin"), recs)


if __name__ == "__main__":
    if util.ref(3) is None or util.ref(2) is None:
        sys.exit("build the reference checker first: make -C oracle ref")
    tab = stencil_table()
    stencil_table_ext()
    save_cases(apply_cases(tab))
    reductions()
    interp()
    exchange()
    multigrid()
    linear()
    interp_hi()
    interp_bnd()
    config1()
    config2_small()
    boxop()
    print("golden fixtures written to", HERE)
