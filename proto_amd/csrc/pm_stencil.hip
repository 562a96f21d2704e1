// Stencil<T> on the device: handle, generic kernel, and the z-marching
// LDS-tiled kernels for the 7-point and 27-point Laplacian shapes.
//
// Parity: every kernel accumulates in Proto's coefficient order with one
// rounding per multiply and per add (the library is compiled with
// -ffp-contract=off), starting from 0 (replace) or the old dst (accumulate),
// with coefficients pre-multiplied by scale when scale != 1 -- the arithmetic
// of Stencil::hostApply (Proto_StencilImplem.H:725-828), so results are
// bit-identical to Proto's CPU path.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "pm_internal.hpp"

struct pm_stencil {
    int dtype, dim, ncoef, kind;
    std::vector<double> coefs;  // exact T values held as double
    std::vector<int> offs;      // ncoef x 3
    int sr[3], dr[3], ds[3];
    int span_lo[3], span_hi[3];
    void* d_coefs;  // T[ncoef]
    int* d_offs;    // int[3 ncoef]
};

namespace pm {

// ------------------------------------------------------------ shapes
static int detect_kind(const pm_stencil* s)
{
    bool unit = true;
    for (int d = 0; d < 3; d++) {
        if (s->sr[d] != 1 || s->dr[d] != 1 || s->ds[d] != 0) { unit = false; }
    }
    if (!unit) { return 0; }
    // kind 1: taps read from LDS by the generic z-march (k_zgen)
    bool near = s->ncoef >= 1 && s->ncoef <= 27;
    for (int j = 0; j < s->ncoef && near; j++)
        for (int d = 0; d < 3; d++)
            if (s->offs[3 * j + d] < -1 || s->offs[3 * j + d] > 1) { near = false; }
    const int generic = near ? 1 : 0;
    if (s->dim != 3) { return generic; }
    // kind 2: one tap at the origin, on the z-march kernel with SHAPE 1
    if (s->ncoef == 1 && s->offs[0] == 0 && s->offs[1] == 0 && s->offs[2] == 0) { return 2; }
    if (s->ncoef == 7) {
        for (int j = 0; j < 7; j++)
            for (int d = 0; d < 3; d++)
                if (s->offs[3 * j + d] != lap7_off(j, d)) { return generic; }
        return 7;
    }
    if (s->ncoef == 27) {
        for (int j = 0; j < 27; j++)
            for (int d = 0; d < 3; d++)
                if (s->offs[3 * j + d] != lap27_off(j, d)) { return generic; }
        return 27;
    }
    return generic;
}

// ------------------------------------------------------------ generic kernel
template<typename T>
struct GenArgs {
    const T* src;
    T* dst;
    Box3 sbox, dbox, ibox;
    const T* coefs;
    const int* offs;
    int ncoef, ncomp, replace, apply_scale;
    T scale;
    int sr[3], dr[3], ds[3];
};

// One thread per index point; the destination pencil start must lie in dst
// (hostApply skips the pencil otherwise, Proto_StencilImplem.H:811).
template<typename T>
__global__ __launch_bounds__(256) void k_stencil_generic(GenArgs<T> a)
{
    const long nx = box_nx(a.ibox), ny = box_ny(a.ibox);
    const long n = box_size(a.ibox);
    const long sx = box_nx(a.sbox), sxy = sx * box_ny(a.sbox);
    const long ssz = box_size(a.sbox), dsz = box_size(a.dbox);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = a.ibox.lo[0] + (int)(t % nx);
        const long r = t / nx;
        const int j = a.ibox.lo[1] + (int)(r % ny);
        const int k = a.ibox.lo[2] + (int)(r / ny);
        const int px = i * a.dr[0] + a.ds[0], py = j * a.dr[1] + a.ds[1], pz = k * a.dr[2] + a.ds[2];
        if (!box_contains(a.dbox, a.ibox.lo[0] * a.dr[0] + a.ds[0], py, pz)) { continue; }
        if (!box_contains(a.dbox, px, py, pz)) { continue; }
        const long di = box_index(a.dbox, px, py, pz);
        const long si = box_index(a.sbox, i * a.sr[0], j * a.sr[1], k * a.sr[2]);
        for (int c = 0; c < a.ncomp; c++) {
            T acc = a.replace ? (T)0 : a.dst[di + c * dsz];
            for (int jj = 0; jj < a.ncoef; jj++) {
                T cc = a.coefs[jj];
                if (a.apply_scale) { cc = cc * a.scale; }
                const long o = a.offs[3 * jj] + a.offs[3 * jj + 1] * sx + a.offs[3 * jj + 2] * sxy;
                acc = acc + cc * a.src[si + c * ssz + o];
            }
            a.dst[di + c * dsz] = acc;
        }
    }
}

// ------------------------------------------------------------ z-march kernels
enum { ZM_REPLACE = 0, ZM_ACCUM = 1, ZM_RESIDUAL = 2, ZM_RESIDUAL2 = 3 };

template<typename T>
struct ZArgs {
    const T* src;   // u on sbox
    T* dst;         // dst on dbox (or r on ibox for residual modes)
    T* dst2;        // second residual (ZM_RESIDUAL2), on ibox
    const T* f;     // f on ibox (residual modes)
    T* partials;    // per block {amax, sum} (x2 for ZM_RESIDUAL2)
    Box3 sbox, dbox, ibox;
    int ntx, nty, tz;
    T c[27];        // coefficients, pre-scaled (first operator)
    T c2[27];       // second operator (ZM_RESIDUAL2, always the 27-point shape)
    int ntap;       // k_zgen: taps in the stencil's order, (dx+1) | (dy+1) << 8 | (dz+1) << 16
    int tap[27];
    int op;         // k_zgen with RED: reduction operator of the written values (partials per block)
};

template<int NT, typename T>
__device__ inline void put(T* p, T v)
{
    if (NT) { __builtin_nontemporal_store(v, p); } else { *p = v; }
}

template<int RY>
struct ZGeom {
    static constexpr int TX = 64, TY = 4 * RY, LX = TX + 2, LY = TY + 2;
    static constexpr int NE = LX * LY, NL = (NE + 255) / 256, NP = NL * 256;  // NP: plane padded to whole passes
};

// Interior tiles (tile, halo and outputs all inside their boxes).  The loop
// holds no divergent branch, so the compiler can count outstanding memory
// operations exactly and each wait covers only the load it needs: loads of
// plane zz + PF stay in flight while plane zz - 1 is computed.  Three LDS
// planes, three register planes and three staging sets rotate with the
// 3-way unrolled loop, so no register is copied from plane to plane.  Loads
// past the end of the z range are clamped to plane ze (in the source box)
// and land in sets that are never stored.
template<typename T, int SHAPE, int MODE, int RY, int PF, int NT>
__device__ __forceinline__ void zm_interior(const ZArgs<T>& a, T (*sm)[ZGeom<RY>::NP], int x0, int y0, int zb, int ze,
                                            T& amax, T& sum, T& amax2, T& sum2)
{
    using G = ZGeom<RY>;
    constexpr int NL = G::NL, LX = G::LX;
    constexpr int NC = SHAPE;
    constexpr bool RES = (MODE == ZM_RESIDUAL || MODE == ZM_RESIDUAL2);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long sx = box_nx(a.sbox), sxy = sx * box_ny(a.sbox);
    long eoff[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = min(tid + 256 * k, G::NE - 1);
        const int row = e / LX, col = e - (e / LX) * LX;
        eoff[k] = (long)(x0 - 1 + col - a.sbox.lo[0]) + sx * (long)(y0 - 1 + row - a.sbox.lo[1]);
    }
    T P[3][NL];
    auto gload = [&](T (&p)[NL], int z) {
        const T* base = a.src + sxy * (long)(z - a.sbox.lo[2]);
#pragma unroll
        for (int k = 0; k < NL; k++) { p[k] = base[eoff[k]]; }
    };
    auto lstore = [&](const T (&p)[NL], T* s) {
#pragma unroll
        for (int k = 0; k < NL; k++) { s[tid + 256 * k] = p[k]; }
    };
    T q[3][RY + 2][3];
    auto lread = [&](T (&qq)[RY + 2][3], const T* s) {
#pragma unroll
        for (int r = 0; r < RY + 2; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) { qq[r][c] = s[(w * RY + r) * LX + lane + c]; }
    };
    // outputs (and f, or the old dst when accumulating) of rows y0 + w RY + o
    const Box3& ob = RES ? a.ibox : a.dbox;
    const long onx = box_nx(ob), oxy = onx * box_ny(ob);
    const long o0 = box_index(ob, x0 + lane, y0 + w * RY, zb);
    const T* pv = (MODE == ZM_ACCUM) ? a.dst : a.f;
    T F[3][RY];
    auto fload = [&](T (&f)[RY], int z) {
        if (MODE == ZM_REPLACE) { return; }
#pragma unroll
        for (int o = 0; o < RY; o++) { f[o] = pv[o0 + o * onx + oxy * (long)(z - zb)]; }
    };

    // one z step: stage plane zz (slot S), then emit plane zz - 1 from the
    // register planes in slots S+1 (zz - 2), S+2 (zz - 1) and S (zz)
    auto step = [&](auto sc, int zz) {
        constexpr int S = decltype(sc)::value, SB = (S + 1) % 3, SC = (S + 2) % 3;
        lstore(P[S], sm[S]);
        const int z = zz - 1;
        fload(F[(S + 1) % 3], min(z + 1, ze - 1));
        gload(P[(S + PF) % 3], min(zz + PF, ze));
        __syncthreads();
        lread(q[S], sm[S]);
#pragma unroll
        for (int o = 0; o < RY; o++) {
            auto tap = [&](int dx, int dy, int dz) -> T {
                return dz < 0 ? q[SB][o + 1 + dy][1 + dx] : (dz == 0 ? q[SC][o + 1 + dy][1 + dx] : q[S][o + 1 + dy][1 + dx]);
            };
            const long oi = o0 + o * onx + oxy * (long)(z - zb);
            T acc = (MODE == ZM_ACCUM) ? F[S][o] : (T)0;
#pragma unroll
            for (int j = 0; j < NC; j++) {
                acc = acc + a.c[j] * tap(shape_off<SHAPE>(j, 0), shape_off<SHAPE>(j, 1), shape_off<SHAPE>(j, 2));
            }
            if (!RES) {
                put<NT>(&a.dst[oi], acc);
            } else {
                const T fv = F[S][o];
                const T rv = fv - acc;
                put<NT>(&a.dst[oi], rv);
                const T av = rv > 0 ? rv : -rv;
                amax = amax < av ? av : amax;
                sum = sum + rv;
                if (MODE == ZM_RESIDUAL2) {
                    T acc2 = 0;
#pragma unroll
                    for (int j = 0; j < 27; j++) {
                        acc2 = acc2 + a.c2[j] * tap(lap27_off(j, 0), lap27_off(j, 1), lap27_off(j, 2));
                    }
                    const T rv2 = fv - acc2;
                    put<NT>(&a.dst2[oi], rv2);
                    const T av2 = rv2 > 0 ? rv2 : -rv2;
                    amax2 = amax2 < av2 ? av2 : amax2;
                    sum2 = sum2 + rv2;
                }
            }
        }
    };

    gload(P[0], zb - 1);
    if (PF == 2) { gload(P[1], zb); }
    fload(F[2], zb);
    // planes zb - 1 and zb: staged, nothing to emit yet
    lstore(P[0], sm[0]);
    gload(P[PF % 3], zb - 1 + PF);
    __syncthreads();
    lread(q[0], sm[0]);
    lstore(P[1], sm[1]);
    gload(P[(1 + PF) % 3], min(zb + PF, ze));
    __syncthreads();
    lread(q[1], sm[1]);
    int zz = zb + 1;
    for (;;) {
        step(std::integral_constant<int, 2>{}, zz);
        if (++zz > ze) { break; }
        step(std::integral_constant<int, 0>{}, zz);
        if (++zz > ze) { break; }
        step(std::integral_constant<int, 1>{}, zz);
        if (++zz > ze) { break; }
    }
}

// Tiles that touch a box edge: every load and store is bounds-checked (zero
// outside the source box), two LDS planes, register planes shifted per step.
template<typename T, int SHAPE, int MODE, int RY, int PF, int NT>
__device__ __forceinline__ void zm_edge(const ZArgs<T>& a, T (*sm)[ZGeom<RY>::NP], int x0, int y0, int zb, int ze,
                                        T& amax, T& sum, T& amax2, T& sum2)
{
    using G = ZGeom<RY>;
    constexpr int NL = G::NL, LX = G::LX, NE = G::NE;
    constexpr int NC = SHAPE;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long sx = box_nx(a.sbox), sxy = sx * box_ny(a.sbox);
    long eoff[NL];
    bool eok[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = tid + 256 * k;
        const int row = e / LX, col = e - (e / LX) * LX;
        const int gx = x0 - 1 + col, gy = y0 - 1 + row;
        eok[k] = (e < NE) && gx >= a.sbox.lo[0] && gx <= a.sbox.hi[0] && gy >= a.sbox.lo[1] && gy <= a.sbox.hi[1];
        eoff[k] = (long)(gx - a.sbox.lo[0]) + sx * (long)(gy - a.sbox.lo[1]);
    }
    auto load = [&](T (&pre)[NL], int z) {
        const T* base = a.src + sxy * (long)(z - a.sbox.lo[2]);
#pragma unroll
        for (int k = 0; k < NL; k++) { pre[k] = eok[k] ? base[eoff[k]] : (T)0; }
    };
    auto store = [&](const T (&pre)[NL], int buf) {
#pragma unroll
        for (int k = 0; k < NL; k++) { sm[buf][tid + 256 * k] = pre[k]; }
    };

    T q[3][RY + 2][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int r = 0; r < RY + 2; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) q[p][r][c] = (T)0;

    const int gx = x0 + lane;
    const bool xok = gx <= a.ibox.hi[0];
    // residual modes: f of the next output plane, loaded one plane ahead
    const long ixy = (long)box_nx(a.ibox) * box_ny(a.ibox);
    T fpre[RY];
#pragma unroll
    for (int o = 0; o < RY; o++) {
        const int gy = y0 + w * RY + o;
        fpre[o] = (T)0;
        if ((MODE == ZM_RESIDUAL || MODE == ZM_RESIDUAL2) && xok && gy <= a.ibox.hi[1] && zb < ze) {
            fpre[o] = a.f[box_index(a.ibox, gx, gy, zb)];
        }
    }

    // shift the register queue, read plane zz from LDS, emit plane zz-1
    auto compute = [&](int zz, int buf) {
#pragma unroll
        for (int r = 0; r < RY + 2; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                q[0][r][c] = q[1][r][c];
                q[1][r][c] = q[2][r][c];
                q[2][r][c] = sm[buf][(w * RY + r) * LX + lane + c];
            }
        if (zz < zb + 1) { return; }
        const int z = zz - 1;
#pragma unroll
        for (int o = 0; o < RY; o++) {
            const int gy = y0 + w * RY + o;
            if (!(xok && gy <= a.ibox.hi[1])) { continue; }
            T acc = 0;
            if (MODE == ZM_ACCUM) { acc = a.dst[box_index(a.dbox, gx, gy, z)]; }
#pragma unroll
            for (int j = 0; j < NC; j++) {
                const int dx = shape_off<SHAPE>(j, 0), dy = shape_off<SHAPE>(j, 1), dz = shape_off<SHAPE>(j, 2);
                acc = acc + a.c[j] * q[1 + dz][o + 1 + dy][1 + dx];
            }
            if (MODE == ZM_REPLACE || MODE == ZM_ACCUM) {
                put<NT>(&a.dst[box_index(a.dbox, gx, gy, z)], acc);
            } else {
                const long ii = box_index(a.ibox, gx, gy, z);
                const T fv = fpre[o];
                if (z + 1 < ze) { fpre[o] = a.f[ii + ixy]; }
                const T rv = fv - acc;
                put<NT>(&a.dst[ii], rv);
                const T av = rv > 0 ? rv : -rv;
                amax = amax < av ? av : amax;
                sum = sum + rv;
                if (MODE == ZM_RESIDUAL2) {
                    T acc2 = 0;
#pragma unroll
                    for (int j = 0; j < 27; j++) {
                        const int dx = lap27_off(j, 0), dy = lap27_off(j, 1), dz = lap27_off(j, 2);
                        acc2 = acc2 + a.c2[j] * q[1 + dz][o + 1 + dy][1 + dx];
                    }
                    const T rv2 = fv - acc2;
                    put<NT>(&a.dst2[ii], rv2);
                    const T av2 = rv2 > 0 ? rv2 : -rv2;
                    amax2 = amax2 < av2 ? av2 : amax2;
                    sum2 = sum2 + rv2;
                }
            }
        }
    };

    T preA[NL], preB[NL];
    if (PF == 1) {
        int buf = 0;
        load(preA, zb - 1);
        for (int zz = zb - 1; zz <= ze; ++zz) {
            store(preA, buf);
            if (zz + 1 <= ze) { load(preA, zz + 1); }
            __syncthreads();
            compute(zz, buf);
            buf ^= 1;
        }
    } else {
        // two planes in flight: plane zz in preA (even steps) / preB (odd steps)
        load(preA, zb - 1);
        load(preB, zb);
        int zz = zb - 1;
        for (;;) {
            store(preA, 0);
            if (zz + 2 <= ze) { load(preA, zz + 2); }
            __syncthreads();
            compute(zz, 0);
            if (++zz > ze) { break; }
            store(preB, 1);
            if (zz + 2 <= ze) { load(preB, zz + 2); }
            __syncthreads();
            compute(zz, 1);
            if (++zz > ze) { break; }
        }
    }
}

// Block = 4 wave64s over a 64 x (4 RY) tile of one (x, y) column range; the
// block walks its z range plane by plane.  Each plane (with a 1-cell halo in
// x and y) is staged once from HBM into LDS; each thread keeps the 3 x 3
// neighbourhoods of its RY rows for planes z-1, z, z+1 in registers, so one
// LDS pass per plane feeds all 27 taps.  Global loads of the next PF planes
// are in flight while a plane is computed.  NT = 1 writes results with
// non-temporal stores (they are streamed out, never re-read by this kernel).
// Tile rows (RY), prefetch depth (PF), NT and the z extent per block are the
// autotuner's knobs.
template<typename T, int SHAPE, int MODE, int RY, int PF, int NT>
__global__ __launch_bounds__(256) void k_zmarch(ZArgs<T> a)
{
    using G = ZGeom<RY>;
    __shared__ T sm[3][G::NP];
    __shared__ T red[16];

    const int tid = threadIdx.x;
    unsigned b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = b % a.ntx;
    b /= a.ntx;
    const int ty = b % a.nty;
    const int tzi = b / a.nty;
    const int x0 = a.ibox.lo[0] + tx * G::TX, y0 = a.ibox.lo[1] + ty * G::TY;
    const int zb = a.ibox.lo[2] + tzi * a.tz;
    const int ze = min(zb + a.tz, a.ibox.hi[2] + 1);  // outputs z in [zb, ze)
    const bool interior = x0 - 1 >= a.sbox.lo[0] && x0 + G::TX <= a.sbox.hi[0] && y0 - 1 >= a.sbox.lo[1] &&
                          y0 + G::TY <= a.sbox.hi[1] && x0 + G::TX - 1 <= a.ibox.hi[0] &&
                          y0 + G::TY - 1 <= a.ibox.hi[1];

    T amax = 0, sum = 0, amax2 = 0, sum2 = 0;
    if (interior) {
        zm_interior<T, SHAPE, MODE, RY, PF, NT>(a, sm, x0, y0, zb, ze, amax, sum, amax2, sum2);
    } else {
        zm_edge<T, SHAPE, MODE, RY, PF, NT>(a, sm, x0, y0, zb, ze, amax, sum, amax2, sum2);
    }
    if (MODE == ZM_RESIDUAL || MODE == ZM_RESIDUAL2) {
        const unsigned pb = blockIdx.x;
        const int per = (MODE == ZM_RESIDUAL2) ? 4 : 2;
        T v = block_reduce(PM_RED_ABS, amax, red);
        if (tid == 0) { a.partials[per * pb + 0] = v; }
        v = block_reduce(PM_RED_SUM, sum, red);
        if (tid == 0) { a.partials[per * pb + 1] = v; }
        if (MODE == ZM_RESIDUAL2) {
            v = block_reduce(PM_RED_ABS, amax2, red);
            if (tid == 0) { a.partials[per * pb + 2] = v; }
            v = block_reduce(PM_RED_SUM, sum2, red);
            if (tid == 0) { a.partials[per * pb + 3] = v; }
        }
    }
}

// Any unit-ratio stencil whose offsets lie in [-1,1]^3 (or [-1,1]^2 in 2-D),
// in any order and with any coefficients (Laplacian_19, a Laplacian with
// reordered terms, a user stencil): the tile and z walk of k_zmarch, with the
// source planes staged through an LDS ring of 4 planes (one barrier per z
// step) and every tap read from LDS at its offset, accumulated in the
// stencil's order with one rounding per multiply and per add, as hostApply
// does.  Loads are bounds-checked against the source box (a stencil without
// z taps, e.g. 2-D, never reads the planes above and below).  PLANAR: every
// tap has dz = 0 (2-D data, one plane): a single LDS plane.
template<typename T, int MODE, int RY, int NT, int PLANAR, int RED>
__global__ __launch_bounds__(256) void k_zgen(ZArgs<T> a)
{
    using G = ZGeom<RY>;
    constexpr int NL = G::NL, LX = G::LX, NE = G::NE;
    constexpr int NS = PLANAR ? 1 : 4;
    __shared__ T sm[NS][G::NP];
    __shared__ T red[16];
    T rv = RED ? red_init<T>(a.op) : (T)0;  // RED: this thread's fold of the values it writes

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    unsigned b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = b % a.ntx;
    b /= a.ntx;
    const int ty = b % a.nty;
    const int tzi = b / a.nty;
    const int x0 = a.ibox.lo[0] + tx * G::TX, y0 = a.ibox.lo[1] + ty * G::TY;
    const int zb = a.ibox.lo[2] + tzi * a.tz;
    const int ze = min(zb + a.tz, a.ibox.hi[2] + 1);  // outputs z in [zb, ze)
    const long sx = box_nx(a.sbox), sxy = sx * box_ny(a.sbox);
    long eoff[NL];
    bool eok[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = tid + 256 * k;
        const int row = e / LX, col = e - (e / LX) * LX;
        const int gx = x0 - 1 + col, gy = y0 - 1 + row;
        eok[k] = (e < NE) && gx >= a.sbox.lo[0] && gx <= a.sbox.hi[0] && gy >= a.sbox.lo[1] && gy <= a.sbox.hi[1];
        eoff[k] = (long)(gx - a.sbox.lo[0]) + sx * (long)(gy - a.sbox.lo[1]);
    }
    T pre[NL];
    auto load = [&](int z) {
        const bool zin = z >= a.sbox.lo[2] && z <= a.sbox.hi[2];
        const T* base = a.src + sxy * (long)(z - a.sbox.lo[2]);
#pragma unroll
        for (int k = 0; k < NL; k++) { pre[k] = (zin && eok[k]) ? base[eoff[k]] : (T)0; }
    };
    auto slot = [&](int z) { return PLANAR ? 0 : (z & 3); };
    auto store = [&](int z) {
        T* pl = sm[slot(z)];
#pragma unroll
        for (int k = 0; k < NL; k++) { pl[tid + 256 * k] = pre[k]; }
    };
    const int gx = x0 + lane;
    const bool xok = gx <= a.ibox.hi[0];
    const int lc = (w * RY + 1) * LX + lane + 1;  // LDS element of this thread's first row, centre
    const Box3& ob = a.dbox;
    auto compute = [&](int z) {
        T acc[RY];
        bool ok[RY];
#pragma unroll
        for (int o = 0; o < RY; o++) {
            const int gy = y0 + w * RY + o;
            ok[o] = xok && gy <= a.ibox.hi[1];
            acc[o] = (MODE == ZM_ACCUM && ok[o]) ? a.dst[box_index(ob, gx, gy, z)] : (T)0;
        }
        for (int j = 0; j < a.ntap; j++) {
            const int t = a.tap[j];
            const int dz = (t >> 16) - 1;
            const int off = (((t >> 8) & 255) - 1) * LX + (t & 255) - 1;
            const T* p = sm[slot(z + dz)] + lc + off;
            const T c = a.c[j];
#pragma unroll
            for (int o = 0; o < RY; o++) { acc[o] = acc[o] + c * p[o * LX]; }
        }
#pragma unroll
        for (int o = 0; o < RY; o++) {
            if (ok[o]) {
                put<NT>(&a.dst[box_index(ob, gx, y0 + w * RY + o, z)], acc[o]);
                if (RED) { rv = red_update(a.op, rv, acc[o]); }
            }
        }
    };
    // RED: one partial per block, a fixed tree (deterministic for the launch shape)
    auto finish = [&]() {
        if (RED) {
            const T v = block_reduce(a.op, rv, red);
            if (tid == 0) { a.partials[blockIdx.x] = v; }
        }
    };
    if (PLANAR) {
        for (int z = zb; z < ze; z++) {
            load(z);
            if (z > zb) { __syncthreads(); }  // the previous plane's taps are read
            store(z);
            __syncthreads();
            compute(z);
        }
        finish();
        return;
    }
    // ring of 4 planes: step z commits plane z + 1 into the slot of plane z - 3
    // (last read at step z - 2, before the barrier of step z - 1)
    load(zb - 1);
    store(zb - 1);
    load(zb);
    store(zb);
    load(zb + 1);
    for (int z = zb; z < ze; z++) {
        store(z + 1);
        if (z + 1 < ze) { load(z + 2); }
        __syncthreads();
        compute(z);
    }
    finish();
}

// The one-block second stage of k_zgen's reductions (fixed strided split and tree).
template<typename T>
__global__ __launch_bounds__(1024) void k_finish_op(int op, const T* __restrict__ partials, int nb, T* out,
                                                    int combine)
{
    __shared__ T red[16];
    T v = red_init<T>(op);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) { v = red_combine(op, v, partials[i]); }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { *out = combine ? red_combine(op, *out, v) : v; }
}

// Deterministic second stage over per-block {amax, sum} groups of PER values
// (one pair per operator): one workgroup per value q (blockIdx.x), a fixed
// strided split and a fixed tree, so the value's bits do not depend on the
// launch.
template<typename T, int PER>
__global__ __launch_bounds__(1024) void k_finish_pairs(const T* __restrict__ partials, int nblocks,
                                                       T* __restrict__ out)
{
    __shared__ T red[16];
    const int q = blockIdx.x;
    const int op = (q % 2 == 0) ? PM_RED_ABS : PM_RED_SUM;
    T v = 0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) { v = red_combine(op, v, partials[PER * i + q]); }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { out[q] = v; }
}

// r = f - r followed by {absMax, sum} partials (generic residual path).
template<typename T>
__global__ __launch_bounds__(256) void k_f_minus(const T* __restrict__ f, T* __restrict__ r, long n,
                                                 T* __restrict__ partials)
{
    __shared__ T red[16];
    T amax = 0, sum = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const T v = f[i] - r[i];
        r[i] = v;
        const T av = v > 0 ? v : -v;
        amax = amax < av ? av : amax;
        sum = sum + v;
    }
    T v = block_reduce(PM_RED_ABS, amax, red);
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = v; }
    v = block_reduce(PM_RED_SUM, sum, red);
    if (threadIdx.x == 0) { partials[2 * blockIdx.x + 1] = v; }
}

// ------------------------------------------------------------ host side
static bool covers(const pm_stencil* st, const Box3& sbox, const Box3& dbox, const Box3& ibox)
{
    // destination: p*dR + dS over idx box; source: p*sR + span
    for (int d = 0; d < 3; d++) {
        const long dlo = (long)ibox.lo[d] * st->dr[d] + st->ds[d];
        const long dhi = (long)ibox.hi[d] * st->dr[d] + st->ds[d];
        if (dlo < dbox.lo[d] || dhi > dbox.hi[d]) { return false; }
        const long slo = (long)ibox.lo[d] * st->sr[d] + st->span_lo[d];
        const long shi = (long)ibox.hi[d] * st->sr[d] + st->span_hi[d];
        if (slo < sbox.lo[d] || shi > sbox.hi[d]) { return false; }
    }
    return true;
}

template<typename T>
static int launch_generic(const pm_stencil* st, const void* src, const Box3& sbox, void* dst, const Box3& dbox,
                          const Box3& ibox, int ncomp, int replace, T scale, int ncoef, hipStream_t s)
{
    GenArgs<T> a;
    a.src = (const T*)src;
    a.dst = (T*)dst;
    a.sbox = sbox;
    a.dbox = dbox;
    a.ibox = ibox;
    a.coefs = (const T*)st->d_coefs;
    a.offs = st->d_offs;
    a.ncoef = ncoef;
    a.ncomp = ncomp;
    a.replace = replace;
    a.apply_scale = (scale != (T)1);
    a.scale = scale;
    for (int d = 0; d < 3; d++) { a.sr[d] = st->sr[d]; a.dr[d] = st->dr[d]; a.ds[d] = st->ds[d]; }
    const long n = box_size(ibox);
    long blocks = (n + 255) / 256;
    blocks = std::min(blocks, 1L << 20);
    hipLaunchKernelGGL(k_stencil_generic<T>, dim3((unsigned)blocks), dim3(256), 0, s, a);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

// ------------------------------------------------------------ tile configurations
struct ZCfg {
    int ry, pf, nt, tz;  // rows per thread, planes prefetched, NT stores, planes per block (0: auto)
};

static ZCfg default_cfg(int dtype) { return ZCfg{dtype == PM_F64 ? 2 : 4, 1, 0, 0}; }

// Tuned configurations per (operator kind, mode, dtype, index-box shape).
struct TuneKey {
    int kind, mode, dtype, nx, ny, nz;
    bool operator<(const TuneKey& o) const
    {
        const int a[6] = {kind, mode, dtype, nx, ny, nz}, b[6] = {o.kind, o.mode, o.dtype, o.nx, o.ny, o.nz};
        for (int k = 0; k < 6; k++) {
            if (a[k] != b[k]) { return a[k] < b[k]; }
        }
        return false;
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, ZCfg>* g_tune = nullptr;

static TuneKey tune_key(int kind, int mode, int dtype, const Box3& ib)
{
    return TuneKey{kind, mode, dtype, (int)box_nx(ib), (int)box_ny(ib), (int)box_nz(ib)};
}

static ZCfg lookup_cfg(int kind, int mode, int dtype, const Box3& ib)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune) {
        auto it = g_tune->find(tune_key(kind, mode, dtype, ib));
        if (it != g_tune->end()) { return it->second; }
    }
    return default_cfg(dtype);
}

static void store_cfg(const TuneKey& k, const ZCfg& c)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune) { g_tune = new std::map<TuneKey, ZCfg>(); }
    (*g_tune)[k] = c;
}

struct ZPlan {
    int ntx, nty, ntz, tz;
    unsigned blocks;
};

static ZPlan zplan(const Box3& ibox, const ZCfg& c)
{
    ZPlan p;
    const int nx = (int)box_nx(ibox), ny = (int)box_ny(ibox), nz = (int)box_nz(ibox);
    p.ntx = (nx + 63) / 64;
    p.nty = (ny + 4 * c.ry - 1) / (4 * c.ry);
    int tz = nz;
    if (c.tz > 0) {
        tz = std::min(c.tz, nz);
    } else {
        // enough blocks to fill 256 CUs several times over, but long z runs
        while (tz > 16 && (long)p.ntx * p.nty * ((nz + tz - 1) / tz) < 4096) { tz = (tz + 1) / 2; }
    }
    p.tz = tz;
    p.ntz = (nz + tz - 1) / tz;
    p.blocks = (unsigned)((long)p.ntx * p.nty * p.ntz);
    return p;
}

template<typename T, int SHAPE, int MODE>
using ZKernel = void (*)(ZArgs<T>);

template<typename T, int SHAPE, int MODE>
static ZKernel<T, SHAPE, MODE> zkernel(const ZCfg& c)
{
#define PM_ZK(RY, PF, NT) \
    if (c.ry == RY && c.pf == PF && c.nt == NT) { return k_zmarch<T, SHAPE, MODE, RY, PF, NT>; }
    PM_ZK(1, 1, 0) PM_ZK(1, 1, 1) PM_ZK(1, 2, 0) PM_ZK(1, 2, 1)
    PM_ZK(2, 1, 0) PM_ZK(2, 1, 1) PM_ZK(2, 2, 0) PM_ZK(2, 2, 1)
    PM_ZK(4, 1, 0) PM_ZK(4, 1, 1) PM_ZK(4, 2, 0) PM_ZK(4, 2, 1)
#undef PM_ZK
    return nullptr;
}

static const int kRY[3] = {1, 2, 4};

static bool valid_cfg(const ZCfg& c) { return zkernel<double, 7, ZM_REPLACE>(c) != nullptr && c.tz >= 0; }

template<typename T, int SHAPE, int MODE>
static int launch_zmarch(ZArgs<T>& a, const Box3& ibox, const ZCfg& c, hipStream_t s, unsigned* nblocks)
{
    ZKernel<T, SHAPE, MODE> k = zkernel<T, SHAPE, MODE>(c);
    PM_REQUIRE(k != nullptr, "no z-march kernel for rows %d, prefetch %d, nt %d", c.ry, c.pf, c.nt);
    ZPlan p = zplan(ibox, c);
    a.ntx = p.ntx;
    a.nty = p.nty;
    a.tz = p.tz;
    a.ibox = ibox;
    *nblocks = p.blocks;
    hipLaunchKernelGGL(k, dim3(p.blocks), dim3(256), 0, s, a);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

// Generic z-march (kind 1): taps of the stencil in its order; PLANAR when no
// tap leaves the plane.
static bool zgen_planar(const pm_stencil* st)
{
    for (int j = 0; j < st->ncoef; j++)
        if (st->offs[3 * j + 2] != 0) { return false; }
    return true;
}

template<typename T, int MODE>
using ZGKernel = void (*)(ZArgs<T>);

template<typename T, int MODE>
static ZGKernel<T, MODE> zgkernel(const ZCfg& c, bool planar, bool red)
{
#define PM_ZG(RY, NT)                                                                                    \
    if (c.ry == RY && c.nt == NT) {                                                                      \
        if (red) { return planar ? k_zgen<T, MODE, RY, NT, 1, 1> : k_zgen<T, MODE, RY, NT, 0, 1>; }      \
        return planar ? k_zgen<T, MODE, RY, NT, 1, 0> : k_zgen<T, MODE, RY, NT, 0, 0>;                   \
    }
    PM_ZG(1, 0) PM_ZG(1, 1) PM_ZG(2, 0) PM_ZG(2, 1) PM_ZG(4, 0) PM_ZG(4, 1)
#undef PM_ZG
    return nullptr;
}

// Workgroups k_zgen launches for ibox with configuration c.
static unsigned zgen_blocks(const Box3& ibox, const ZCfg& c) { return zplan(ibox, c).blocks; }

// red: a.partials receives one partial of a.op per workgroup (zgen_blocks)
template<typename T, int MODE>
static int launch_zgen(const pm_stencil* st, ZArgs<T>& a, const Box3& ibox, const ZCfg& c, hipStream_t s,
                       bool red = false)
{
    const bool planar = zgen_planar(st);
    ZGKernel<T, MODE> k = zgkernel<T, MODE>(c, planar, red);
    PM_REQUIRE(k != nullptr, "no generic z-march kernel for rows %d, nt %d", c.ry, c.nt);
    ZPlan p = zplan(ibox, c);
    a.ntx = p.ntx;
    a.nty = p.nty;
    a.tz = p.tz;
    a.ibox = ibox;
    a.ntap = st->ncoef;
    for (int j = 0; j < st->ncoef; j++) {
        const int* o = &st->offs[3 * j];
        a.tap[j] = (o[0] + 1) | ((o[1] + 1) << 8) | ((o[2] + 1) << 16);
    }
    hipLaunchKernelGGL(k, dim3(p.blocks), dim3(256), 0, s, a);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

template<typename T>
static void scaled_coefs(const pm_stencil* st, double scale, T* out)
{
    const T sc = (T)scale;
    for (int j = 0; j < st->ncoef; j++) {
        T c = (T)st->coefs[j];
        if (sc != (T)1) { c = c * sc; }
        out[j] = c;
    }
}

// k_zmarch stages the index box grown by one cell on every side (planes z - 1
// and z + 1 included, without bounds checks in z): the 7/27-point stencils'
// coverage guarantees it; a one-tap stencil's does not, so kind 2 takes the
// z-march only when the source box holds that halo.
static bool zmarch_halo(const Box3& sbox, const Box3& ibox)
{
    for (int d = 0; d < 3; d++) {
        if (sbox.lo[d] > ibox.lo[d] - 1 || sbox.hi[d] < ibox.hi[d] + 1) { return false; }
    }
    return true;
}

template<typename T>
static int apply_t(const pm_stencil* st, const void* src, const Box3& sbox, void* dst, const Box3& dbox,
                   const Box3& ibox, int ncomp, int replace, double scale, hipStream_t s)
{
    const T sc = (T)scale;
    if (sc == (T)0) {
        if (!replace) { return PM_OK; }
        return launch_generic<T>(st, src, sbox, dst, dbox, ibox, ncomp, 1, sc, 0, s);
    }
    if ((st->kind == 7 || st->kind == 27 || (st->kind == 2 && zmarch_halo(sbox, ibox))) && ncomp >= 1) {
        const long ssz = box_size(sbox), dsz = box_size(dbox);
        for (int c = 0; c < ncomp; c++) {
            ZArgs<T> a;
            std::memset(&a, 0, sizeof(a));
            a.src = (const T*)src + c * ssz;
            a.dst = (T*)dst + c * dsz;
            a.sbox = sbox;
            a.dbox = dbox;
            scaled_coefs<T>(st, scale, a.c);
            unsigned nb;
            int rc;
            const int mode = replace ? ZM_REPLACE : ZM_ACCUM;
            const ZCfg cfg = lookup_cfg(st->kind, mode, st->dtype, ibox);
            if (st->kind == 7) {
                rc = replace ? launch_zmarch<T, 7, ZM_REPLACE>(a, ibox, cfg, s, &nb)
                             : launch_zmarch<T, 7, ZM_ACCUM>(a, ibox, cfg, s, &nb);
            } else if (st->kind == 2) {
                rc = replace ? launch_zmarch<T, 1, ZM_REPLACE>(a, ibox, cfg, s, &nb)
                             : launch_zmarch<T, 1, ZM_ACCUM>(a, ibox, cfg, s, &nb);
            } else {
                rc = replace ? launch_zmarch<T, 27, ZM_REPLACE>(a, ibox, cfg, s, &nb)
                             : launch_zmarch<T, 27, ZM_ACCUM>(a, ibox, cfg, s, &nb);
            }
            if (rc) { return rc; }
        }
        return PM_OK;
    }
    if (st->kind == 1 || st->kind == 2) {  // kind 2 here: no halo around the index box
        const long ssz = box_size(sbox), dsz = box_size(dbox);
        const ZCfg cfg = lookup_cfg(1, replace ? ZM_REPLACE : ZM_ACCUM, st->dtype, ibox);
        for (int c = 0; c < ncomp; c++) {
            ZArgs<T> a;
            std::memset(&a, 0, sizeof(a));
            a.src = (const T*)src + c * ssz;
            a.dst = (T*)dst + c * dsz;
            a.sbox = sbox;
            a.dbox = dbox;
            scaled_coefs<T>(st, scale, a.c);
            const int rc = replace ? launch_zgen<T, ZM_REPLACE>(st, a, ibox, cfg, s)
                                   : launch_zgen<T, ZM_ACCUM>(st, a, ibox, cfg, s);
            if (rc) { return rc; }
        }
        return PM_OK;
    }
    return launch_generic<T>(st, src, sbox, dst, dbox, ibox, ncomp, replace, sc, st->ncoef, s);
}

template<typename T>
static int residual_t(const pm_stencil* st, const void* u, const Box3& ubox, const void* f, void* r,
                      const Box3& rbox, double scale, void* red_out, hipStream_t s)
{
    const T sc = (T)scale;
    if ((st->kind == 7 || st->kind == 27) && sc != (T)0) {
        ZArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        a.src = (const T*)u;
        a.dst = (T*)r;
        a.f = (const T*)f;
        a.sbox = ubox;
        a.dbox = rbox;
        scaled_coefs<T>(st, scale, a.c);
        const ZCfg cfg = lookup_cfg(st->kind, ZM_RESIDUAL, st->dtype, rbox);
        ZPlan p = zplan(rbox, cfg);
        void* ws = nullptr;
        int rc = workspace(s, sizeof(T) * 2 * (size_t)p.blocks, &ws);
        if (rc) { return rc; }
        a.partials = (T*)ws;
        unsigned nb;
        rc = st->kind == 7 ? launch_zmarch<T, 7, ZM_RESIDUAL>(a, rbox, cfg, s, &nb)
                           : launch_zmarch<T, 27, ZM_RESIDUAL>(a, rbox, cfg, s, &nb);
        if (rc) { return rc; }
        if (red_out) {
            hipLaunchKernelGGL((k_finish_pairs<T, 2>), dim3(2), dim3(1024), 0, s, (const T*)ws, (int)nb, (T*)red_out);
            PM_CHECK_LAUNCH();
        }
        return PM_OK;
    }
    // generic path: Lu |= L(u) into r, then r = f - r with partials
    int rc = apply_t<T>(st, u, ubox, r, rbox, rbox, 1, 1, scale, s);
    if (rc) { return rc; }
    const long n = box_size(rbox);
    long blocks = std::min((n + 255) / 256, 2048L);
    if (blocks < 1) { blocks = 1; }
    void* ws = nullptr;
    rc = workspace(s, sizeof(T) * 2 * (size_t)blocks, &ws);
    if (rc) { return rc; }
    hipLaunchKernelGGL(k_f_minus<T>, dim3((unsigned)blocks), dim3(256), 0, s, (const T*)f, (T*)r, n, (T*)ws);
    PM_CHECK_LAUNCH();
    if (red_out) {
        hipLaunchKernelGGL((k_finish_pairs<T, 2>), dim3(2), dim3(1024), 0, s, (const T*)ws, (int)blocks, (T*)red_out);
        PM_CHECK_LAUNCH();
    }
    return PM_OK;
}

// Partials of the two-operator fused residual: 4 values per workgroup.
template<typename T>
static size_t residual2_partials_bytes(const pm_stencil* st1, const pm_stencil* st2, const Box3& rbox)
{
    const ZCfg cfg = lookup_cfg(st1->kind * 100 + st2->kind, ZM_RESIDUAL2, st1->dtype, rbox);
    return sizeof(T) * 4 * (size_t)zplan(rbox, cfg).blocks;
}

// ws: the partials buffer (residual2_partials_bytes), or null for the
// stream's scratch
template<typename T>
static int residual2_t(const pm_stencil* st1, const pm_stencil* st2, const void* u, const Box3& ubox,
                       const void* f, void* r1, void* r2, const Box3& rbox, double scale1, double scale2,
                       void* red_out, hipStream_t s, void* ws = nullptr)
{
    ZArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.src = (const T*)u;
    a.dst = (T*)r1;
    a.dst2 = (T*)r2;
    a.f = (const T*)f;
    a.sbox = ubox;
    a.dbox = rbox;
    scaled_coefs<T>(st1, scale1, a.c);
    scaled_coefs<T>(st2, scale2, a.c2);
    const ZCfg cfg = lookup_cfg(st1->kind * 100 + st2->kind, ZM_RESIDUAL2, st1->dtype, rbox);
    ZPlan p = zplan(rbox, cfg);
    int rc = PM_OK;
    if (!ws) {
        rc = workspace(s, sizeof(T) * 4 * (size_t)p.blocks, &ws);
        if (rc) { return rc; }
    }
    a.partials = (T*)ws;
    unsigned nb;
    rc = (st1->kind == 7) ? launch_zmarch<T, 7, ZM_RESIDUAL2>(a, rbox, cfg, s, &nb)
                          : launch_zmarch<T, 27, ZM_RESIDUAL2>(a, rbox, cfg, s, &nb);
    if (rc) { return rc; }
    if (red_out) {
        hipLaunchKernelGGL((k_finish_pairs<T, 4>), dim3(4), dim3(1024), 0, s, (const T*)ws, (int)nb, (T*)red_out);
        PM_CHECK_LAUNCH();
    }
    return PM_OK;
}

int stencil_info(const pm_stencil* st, int* kind, int* dtype, std::vector<double>* coefs)
{
    PM_REQUIRE(st, "null stencil");
    *kind = st->kind;
    *dtype = st->dtype;
    *coefs = st->coefs;
    return PM_OK;
}

// ------------------------------------------------------------ autotuner
template<typename T>
static int run_cfg(const pm_stencil* st, const pm_stencil* st2, int mode, ZArgs<T>& a, const Box3& ib, const ZCfg& c,
                   hipStream_t s)
{
    unsigned nb;
    const bool k7 = st->kind == 7;
    if (st->kind == 2) {  // modes 0 and 1 only (checked by the caller)
        return mode == ZM_REPLACE ? launch_zmarch<T, 1, ZM_REPLACE>(a, ib, c, s, &nb)
                                  : launch_zmarch<T, 1, ZM_ACCUM>(a, ib, c, s, &nb);
    }
    switch (mode) {
        case ZM_REPLACE: return k7 ? launch_zmarch<T, 7, ZM_REPLACE>(a, ib, c, s, &nb) : launch_zmarch<T, 27, ZM_REPLACE>(a, ib, c, s, &nb);
        case ZM_ACCUM: return k7 ? launch_zmarch<T, 7, ZM_ACCUM>(a, ib, c, s, &nb) : launch_zmarch<T, 27, ZM_ACCUM>(a, ib, c, s, &nb);
        case ZM_RESIDUAL: return k7 ? launch_zmarch<T, 7, ZM_RESIDUAL>(a, ib, c, s, &nb) : launch_zmarch<T, 27, ZM_RESIDUAL>(a, ib, c, s, &nb);
        default: return k7 ? launch_zmarch<T, 7, ZM_RESIDUAL2>(a, ib, c, s, &nb) : launch_zmarch<T, 27, ZM_RESIDUAL2>(a, ib, c, s, &nb);
    }
    (void)st2;
}

static int key_kind(const pm_stencil* st, const pm_stencil* st2, int mode)
{
    return mode == ZM_RESIDUAL2 ? st->kind * 100 + st2->kind : st->kind;
}

template<typename T>
static int autotune_t(const pm_stencil* st, const pm_stencil* st2, int mode, const Box3& sb, const Box3& ib, int reps,
                      ZCfg* best, float* best_ms, hipStream_t s)
{
    const size_t ns = (size_t)box_size(sb), ni = (size_t)box_size(ib);
    // the most blocks any candidate launches: one row per thread, 16 planes
    const size_t maxb = (size_t)((box_nx(ib) + 63) / 64) * (size_t)((box_ny(ib) + 3) / 4) * (size_t)((box_nz(ib) + 15) / 16);
    T *src = nullptr, *dst = nullptr, *dst2 = nullptr, *f = nullptr, *part = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PM_OK;
    auto cleanup = [&]() {
        hipFree(src); hipFree(dst); hipFree(dst2); hipFree(f); hipFree(part);
        if (e0) { hipEventDestroy(e0); }
        if (e1) { hipEventDestroy(e1); }
    };
    if (hipMalloc(&src, sizeof(T) * ns) != hipSuccess || hipMalloc(&dst, sizeof(T) * ni) != hipSuccess ||
        hipMalloc(&dst2, sizeof(T) * ni) != hipSuccess || hipMalloc(&f, sizeof(T) * ni) != hipSuccess ||
        hipMalloc(&part, sizeof(T) * 4 * maxb) != hipSuccess) {
        cleanup();
        set_error("pm_stencil_autotune: scratch allocation failed");
        return PM_ERR_NOMEM;
    }
    pm_fill_uniform(dtype_of<T>::value, src, ns, 17, 0, 1.0, 0.0, (pm_stream_t)s);
    pm_fill_uniform(dtype_of<T>::value, f, ni, 17, 1, 1.0, 0.0, (pm_stream_t)s);
    pm_fill_uniform(dtype_of<T>::value, dst, ni, 17, 2, 1.0, 0.0, (pm_stream_t)s);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        set_error("pm_stencil_autotune: event creation failed");
        return PM_ERR_HIP;
    }
    ZArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.src = src;
    a.dst = dst;
    a.dst2 = dst2;
    a.f = f;
    a.partials = part;
    a.sbox = sb;
    a.dbox = ib;
    scaled_coefs<T>(st, 1.0, a.c);
    if (st2) { scaled_coefs<T>(st2, 1.0, a.c2); }
    const int nz = (int)box_nz(ib);
    const int tzs[6] = {0, 16, 32, 64, 128, 256};
    float bms = 1e30f;
    ZCfg bc = default_cfg(dtype_of<T>::value);
    for (int ry : kRY)
        for (int pf = 1; pf <= 2; pf++)
            for (int nt = 0; nt <= 1; nt++)
                for (int tz : tzs) {
                    if (tz > nz) { continue; }
                    ZCfg c{ry, pf, nt, tz};
                    rc = run_cfg<T>(st, st2, mode, a, ib, c, s);
                    if (rc) { cleanup(); return rc; }
                    hipEventRecord(e0, s);
                    for (int r = 0; r < reps; r++) {
                        rc = run_cfg<T>(st, st2, mode, a, ib, c, s);
                        if (rc) { cleanup(); return rc; }
                    }
                    hipEventRecord(e1, s);
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        cleanup();
                        set_error("pm_stencil_autotune: timing failed");
                        return PM_ERR_HIP;
                    }
                    float ms = 0;
                    hipEventElapsedTime(&ms, e0, e1);
                    ms /= reps;
                    if (ms < bms) { bms = ms; bc = c; }
                }
    cleanup();
    store_cfg(tune_key(key_kind(st, st2, mode), mode, dtype_of<T>::value, ib), bc);
    *best = bc;
    *best_ms = bms;
    return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {

int pm_stencil_create(pm_stencil** out, int dtype, int dim, int ncoef, const void* coefs_host,
                      const int* offsets_host, const int src_ratio[3], const int dst_ratio[3],
                      const int dst_shift[3])
{
    PM_REQUIRE(out, "pm_stencil_create: null out");
    *out = nullptr;
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_stencil_create: bad dtype %d", dtype);
    PM_REQUIRE(dim == 2 || dim == 3, "pm_stencil_create: dim must be 2 or 3");
    PM_REQUIRE(ncoef >= 0 && ncoef <= 343, "pm_stencil_create: ncoef %d outside [0, 343]", ncoef);
    PM_REQUIRE(ncoef == 0 || (coefs_host && offsets_host), "pm_stencil_create: null coefficient data");
    pm_stencil* st = new pm_stencil();
    st->dtype = dtype;
    st->dim = dim;
    st->ncoef = ncoef;
    st->coefs.resize(ncoef);
    st->offs.assign(offsets_host, offsets_host + 3 * ncoef);
    for (int j = 0; j < ncoef; j++) {
        st->coefs[j] = dtype == PM_F64 ? ((const double*)coefs_host)[j] : (double)((const float*)coefs_host)[j];
    }
    for (int d = 0; d < 3; d++) {
        st->sr[d] = src_ratio ? src_ratio[d] : 1;
        st->dr[d] = dst_ratio ? dst_ratio[d] : 1;
        st->ds[d] = dst_shift ? dst_shift[d] : 0;
        if (d >= dim) {
            st->sr[d] = st->dr[d] = 1;
            st->ds[d] = 0;
        }
        if (st->sr[d] < 1 || st->dr[d] < 1) {
            delete st;
            set_error("pm_stencil_create: refinement ratios must be positive");
            return PM_ERR_INVALID;
        }
        st->span_lo[d] = 0;
        st->span_hi[d] = 0;
    }
    // the span is the bounding box of the offsets (m_span &= offset, Proto_StencilImplem.H:267)
    for (int j = 0; j < ncoef; j++) {
        for (int d = 0; d < 3; d++) {
            const int v = st->offs[3 * j + d];
            if (d >= dim && v != 0) {
                delete st;
                set_error("pm_stencil_create: offset %d has a component beyond dim", j);
                return PM_ERR_INVALID;
            }
            st->span_lo[d] = (j == 0) ? v : std::min(v, st->span_lo[d]);
            st->span_hi[d] = (j == 0) ? v : std::max(v, st->span_hi[d]);
        }
    }
    st->kind = detect_kind(st);
    st->d_coefs = nullptr;
    st->d_offs = nullptr;
    if (ncoef > 0) {
        const size_t es = dtype_size(dtype);
        std::vector<char> hc(es * ncoef);
        for (int j = 0; j < ncoef; j++) {
            if (dtype == PM_F64) { ((double*)hc.data())[j] = st->coefs[j]; }
            else { ((float*)hc.data())[j] = (float)st->coefs[j]; }
        }
        if (hipMalloc(&st->d_coefs, es * ncoef) != hipSuccess ||
            hipMalloc((void**)&st->d_offs, sizeof(int) * 3 * ncoef) != hipSuccess) {
            hipFree(st->d_coefs);
            delete st;
            set_error("pm_stencil_create: device allocation failed");
            return PM_ERR_NOMEM;
        }
        if (hipMemcpy(st->d_coefs, hc.data(), es * ncoef, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(st->d_offs, st->offs.data(), sizeof(int) * 3 * ncoef, hipMemcpyHostToDevice) != hipSuccess) {
            hipFree(st->d_coefs);
            hipFree(st->d_offs);
            delete st;
            set_error("pm_stencil_create: upload failed");
            return PM_ERR_HIP;
        }
    }
    *out = st;
    return PM_OK;
}

int pm_stencil_destroy(pm_stencil* st)
{
    if (!st) { return PM_OK; }
    if (st->d_coefs) { hipFree(st->d_coefs); }
    if (st->d_offs) { hipFree(st->d_offs); }
    delete st;
    return PM_OK;
}

int pm_stencil_kind(const pm_stencil* st, int* kind)
{
    PM_REQUIRE(st && kind, "pm_stencil_kind: null argument");
    *kind = st->kind;
    return PM_OK;
}

int pm_stencil_set_kind(pm_stencil* st, int kind)
{
    PM_REQUIRE(st, "pm_stencil_set_kind: null argument");
    const int natural = detect_kind(st);
    const bool ok = kind == 0 || kind == natural || (kind == 1 && (natural == 7 || natural == 27 || natural == 2));
    PM_REQUIRE(ok, "pm_stencil_set_kind: kind %d cannot serve this stencil (its own kind is %d)", kind, natural);
    st->kind = kind;
    return PM_OK;
}

int pm_stencil_apply(const pm_stencil* st, const void* src, const pm_box* src_box, void* dst,
                     const pm_box* dst_box, const pm_box* idx_box, int ncomp, int replace, double scale,
                     pm_stream_t stream)
{
    PM_REQUIRE(st && src_box && dst_box && idx_box, "pm_stencil_apply: null argument");
    const Box3 sb = box3(src_box), db = box3(dst_box), ib = box3(idx_box);
    if (st->ncoef == 0 || box_empty(ib)) { return PM_OK; }
    PM_REQUIRE(src && dst, "pm_stencil_apply: null data pointer");
    PM_REQUIRE(ncomp >= 1, "pm_stencil_apply: ncomp must be >= 1");
    PM_REQUIRE(covers(st, sb, db, ib),
               "pm_stencil_apply: index box [%d,%d,%d]-[%d,%d,%d] reads or writes outside the source/destination boxes",
               ib.lo[0], ib.lo[1], ib.lo[2], ib.hi[0], ib.hi[1], ib.hi[2]);
    if (st->dtype == PM_F64) { return apply_t<double>(st, src, sb, dst, db, ib, ncomp, replace, scale, S(stream)); }
    return apply_t<float>(st, src, sb, dst, db, ib, ncomp, replace, scale, S(stream));
}

int pm_stencil_apply_reduce(const pm_stencil* st, const void* src, const pm_box* src_box, void* dst,
                            const pm_box* dst_box, const pm_box* idx_box, int replace, double scale, int op,
                            void* out_dev, int combine, pm_stream_t stream)
{
    PM_REQUIRE(st && src_box && dst_box && idx_box && out_dev, "pm_stencil_apply_reduce: null argument");
    PM_REQUIRE(op >= PM_RED_MAX && op <= PM_RED_SUMSQ, "pm_stencil_apply_reduce: bad op %d", op);
    const Box3 sb = box3(src_box), db = box3(dst_box), ib = box3(idx_box);
    PM_REQUIRE(!box_empty(ib) && st->ncoef > 0, "pm_stencil_apply_reduce: empty index box or stencil");
    PM_REQUIRE(src && dst, "pm_stencil_apply_reduce: null data pointer");
    PM_REQUIRE(covers(st, sb, db, ib), "pm_stencil_apply_reduce: index box reads or writes outside the boxes");
    const hipStream_t s = S(stream);
    const bool f64 = st->dtype == PM_F64;
    const double sc_t = f64 ? scale : (double)(float)scale;
    if ((st->kind == 1 || st->kind == 2) && sc_t != 0.0) {
        // one pass: the generic z-march writes dst and a partial per workgroup
        const ZCfg cfg = lookup_cfg(1, replace ? ZM_REPLACE : ZM_ACCUM, st->dtype, ib);
        const unsigned nb = zgen_blocks(ib, cfg);
        void* ws = nullptr;
        int rc = workspace(s, (f64 ? 8 : 4) * (size_t)nb, &ws);
        if (rc) { return rc; }
        auto run = [&](auto tag) -> int {
            using T = decltype(tag);
            ZArgs<T> a;
            std::memset(&a, 0, sizeof(a));
            a.src = (const T*)src;
            a.dst = (T*)dst;
            a.sbox = sb;
            a.dbox = db;
            a.partials = (T*)ws;
            a.op = op;
            scaled_coefs<T>(st, scale, a.c);
            int r2 = replace ? launch_zgen<T, ZM_REPLACE>(st, a, ib, cfg, s, true)
                             : launch_zgen<T, ZM_ACCUM>(st, a, ib, cfg, s, true);
            if (r2) { return r2; }
            hipLaunchKernelGGL(k_finish_op<T>, dim3(1), dim3(1024), 0, s, op, (const T*)ws, (int)nb, (T*)out_dev,
                               combine);
            PM_CHECK_LAUNCH();
            return PM_OK;
        };
        return f64 ? run(0.0) : run(0.0f);
    }
    // other kernel families: the apply, then the reduction over the cells it wrote
    int rc = pm_stencil_apply(st, src, src_box, dst, dst_box, idx_box, 1, replace, scale, stream);
    if (rc) { return rc; }
    pm_box wr;
    for (int d = 0; d < 3; d++) {
        wr.lo[d] = ib.lo[d] * st->dr[d] + st->ds[d];
        wr.hi[d] = ib.hi[d] * st->dr[d] + st->ds[d];
    }
    bool dense = true;
    for (int d = 0; d < 3; d++) { dense = dense && (st->dr[d] == 1 || ib.lo[d] == ib.hi[d]); }
    PM_REQUIRE(dense, "pm_stencil_apply_reduce: destination-refined stencils write a strided set; use "
                      "pm_stencil_apply and pm_reduce");
    return pm_reduce_box(op, st->dtype, dst, dst_box, &wr, 0, out_dev, combine, stream);
}

int pm_stencil_residual(const pm_stencil* st, const void* u, const pm_box* u_box, const void* f, void* r,
                        const pm_box* r_box, double scale, void* red_out, pm_stream_t stream)
{
    PM_REQUIRE(st && u_box && r_box && u && f && r, "pm_stencil_residual: null argument");
    const Box3 ub = box3(u_box), rb = box3(r_box);
    PM_REQUIRE(!box_empty(rb), "pm_stencil_residual: empty box");
    PM_REQUIRE(st->ncoef > 0, "pm_stencil_residual: empty stencil");
    for (int d = 0; d < 3; d++) {
        PM_REQUIRE(st->sr[d] == 1 && st->dr[d] == 1 && st->ds[d] == 0,
                   "pm_stencil_residual: stencil must have unit ratios and no shift");
    }
    PM_REQUIRE(covers(st, ub, rb, rb), "pm_stencil_residual: u does not cover the stencil's domain of r_box");
    if (st->dtype == PM_F64) { return residual_t<double>(st, u, ub, f, r, rb, scale, red_out, S(stream)); }
    return residual_t<float>(st, u, ub, f, r, rb, scale, red_out, S(stream));
}

int pm_stencil_autotune(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* src_box,
                        const pm_box* idx_box, int reps, pm_tile_config* best, float* best_ms, pm_stream_t stream)
{
    PM_REQUIRE(st && src_box && idx_box && best && best_ms, "pm_stencil_autotune: null argument");
    PM_REQUIRE(mode >= 0 && mode <= 3, "pm_stencil_autotune: bad mode %d", mode);
    PM_REQUIRE(st->kind == 7 || st->kind == 27 || (st->kind == 2 && mode <= ZM_ACCUM),
               "pm_stencil_autotune: only the z-march kernels (7/27-point; the one-tap stencil's apply) are tunable");
    PM_REQUIRE(mode != ZM_RESIDUAL2 || (st2 && st2->kind == 27 && st2->dtype == st->dtype),
               "pm_stencil_autotune: mode 3 needs a 27-point second operator of the same dtype");
    PM_REQUIRE(reps >= 1, "pm_stencil_autotune: reps must be >= 1");
    const Box3 sb = box3(src_box), ib = box3(idx_box);
    PM_REQUIRE(!box_empty(ib) && covers(st, sb, ib, ib), "pm_stencil_autotune: source box does not cover the index box");
    PM_REQUIRE(st->kind != 2 || zmarch_halo(sb, ib),
               "pm_stencil_autotune: the one-tap z-march needs a source box one cell wider than the index box");
    ZCfg c;
    int rc = st->dtype == PM_F64 ? autotune_t<double>(st, st2, mode, sb, ib, reps, &c, best_ms, S(stream))
                                 : autotune_t<float>(st, st2, mode, sb, ib, reps, &c, best_ms, S(stream));
    if (rc) { return rc; }
    best->rows = c.ry;
    best->prefetch = c.pf;
    best->nt_store = c.nt;
    best->zplanes = c.tz;
    return PM_OK;
}

int pm_stencil_set_tile_config(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* idx_box,
                               const pm_tile_config* cfg)
{
    PM_REQUIRE(st && idx_box && cfg && mode >= 0 && mode <= 3, "pm_stencil_set_tile_config: bad argument");
    PM_REQUIRE(mode != ZM_RESIDUAL2 || st2, "pm_stencil_set_tile_config: mode 3 needs st2");
    ZCfg c{cfg->rows, cfg->prefetch, cfg->nt_store, cfg->zplanes};
    PM_REQUIRE(valid_cfg(c), "pm_stencil_set_tile_config: no such configuration");
    store_cfg(tune_key(key_kind(st, st2, mode), mode, st->dtype, box3(idx_box)), c);
    return PM_OK;
}

int pm_stencil_get_tile_config(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* idx_box,
                               pm_tile_config* cfg)
{
    PM_REQUIRE(st && idx_box && cfg && mode >= 0 && mode <= 3, "pm_stencil_get_tile_config: bad argument");
    PM_REQUIRE(mode != ZM_RESIDUAL2 || st2, "pm_stencil_get_tile_config: mode 3 needs st2");
    ZCfg c = lookup_cfg(key_kind(st, st2, mode), mode, st->dtype, box3(idx_box));
    cfg->rows = c.ry;
    cfg->prefetch = c.pf;
    cfg->nt_store = c.nt;
    cfg->zplanes = c.tz;
    return PM_OK;
}

int pm_tune_table_save(const char* path)
{
    PM_REQUIRE(path, "pm_tune_table_save: null path");
    FILE* fp = std::fopen(path, "w");
    PM_REQUIRE(fp, "pm_tune_table_save: cannot open %s", path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::fprintf(fp, "# kind mode dtype nx ny nz rows prefetch nt zplanes\n");
    if (g_tune) {
        for (auto& kv : *g_tune) {
            const TuneKey& k = kv.first;
            const ZCfg& c = kv.second;
            std::fprintf(fp, "%d %d %d %d %d %d %d %d %d %d\n", k.kind, k.mode, k.dtype, k.nx, k.ny, k.nz, c.ry, c.pf, c.nt,
                         c.tz);
        }
    }
    std::fclose(fp);
    return PM_OK;
}

int pm_tune_table_load(const char* path)
{
    PM_REQUIRE(path, "pm_tune_table_load: null path");
    FILE* fp = std::fopen(path, "r");
    PM_REQUIRE(fp, "pm_tune_table_load: cannot open %s", path);
    char line[256];
    int n = 0;
    while (std::fgets(line, sizeof(line), fp)) {
        if (line[0] == '#') { continue; }
        TuneKey k;
        ZCfg c;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &k.kind, &k.mode, &k.dtype, &k.nx, &k.ny, &k.nz, &c.ry,
                        &c.pf, &c.nt, &c.tz) == 10 &&
            valid_cfg(c)) {
            store_cfg(k, c);
            n++;
        }
    }
    std::fclose(fp);
    return PM_OK;
}

int pm_stencil_residual2(const pm_stencil* st1, const pm_stencil* st2, const void* u, const pm_box* u_box,
                         const void* f, void* r1, void* r2, const pm_box* r_box, double scale1, double scale2,
                         void* red_out, pm_stream_t stream)
{
    PM_REQUIRE(st1 && st2 && u && u_box && f && r1 && r2 && r_box, "pm_stencil_residual2: null argument");
    PM_REQUIRE(st1->dtype == st2->dtype, "pm_stencil_residual2: dtype mismatch");
    PM_REQUIRE((st1->kind == 7 || st1->kind == 27) && st2->kind == 27,
               "pm_stencil_residual2: needs a 7- or 27-point Laplacian shape and a 27-point one");
    const Box3 ub = box3(u_box), rb = box3(r_box);
    PM_REQUIRE(!box_empty(rb), "pm_stencil_residual2: empty box");
    PM_REQUIRE(covers(st1, ub, rb, rb) && covers(st2, ub, rb, rb),
               "pm_stencil_residual2: u does not cover the stencils' domain of r_box");
    PM_REQUIRE(scale1 != 0.0 && scale2 != 0.0, "pm_stencil_residual2: zero scale");
    if (st1->dtype == PM_F64) {
        return residual2_t<double>(st1, st2, u, ub, f, r1, r2, rb, scale1, scale2, red_out, S(stream));
    }
    return residual2_t<float>(st1, st2, u, ub, f, r1, r2, rb, scale1, scale2, red_out, S(stream));
}

}  // extern "C"

// ------------------------------------------------------------ allocation-free residual
// Fixed device scratch of pm_stencil_residual2_noalloc: the per-workgroup
// partials of the fused kernel (4 values each), 1 MiB, enough for 32768 FP64
// workgroups (512^3 with the default tiles launches 8192).  Each device holds
// its own copy for the life of the process.  Host code never passes the name
// `scratch` to a kernel (on the host it is only a stand-in for the symbol):
// the device address comes from hipGetSymbolAddress, once per device.
static constexpr size_t kScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace pm {

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices];

// Device address of this device's scratch, looked up on the first call made
// with the device current.
static int scratch_address(void** out)
{
    int dev = 0;
    PM_CHECK_HIP(hipGetDevice(&dev));
    PM_REQUIRE(dev >= 0 && dev < kMaxDevices, "pm_stencil_residual2_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("pm_stencil_residual2_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return PM_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return PM_OK;
}

}  // namespace pm

extern "C" {

int pm_stencil_residual2_noalloc(const pm_stencil* st1, const pm_stencil* st2, const void* u, const pm_box* u_box,
                                 const void* f, void* r1, void* r2, const pm_box* r_box, double scale1,
                                 double scale2, void* red_out, pm_stream_t stream)
{
    PM_REQUIRE(st1 && st2 && u && u_box && f && r1 && r2 && r_box, "pm_stencil_residual2_noalloc: null argument");
    PM_REQUIRE(st1->dtype == st2->dtype, "pm_stencil_residual2_noalloc: dtype mismatch");
    PM_REQUIRE((st1->kind == 7 || st1->kind == 27) && st2->kind == 27,
               "pm_stencil_residual2_noalloc: needs a 7- or 27-point Laplacian shape and a 27-point one");
    const Box3 ub = box3(u_box), rb = box3(r_box);
    PM_REQUIRE(!box_empty(rb), "pm_stencil_residual2_noalloc: empty box");
    PM_REQUIRE(covers(st1, ub, rb, rb) && covers(st2, ub, rb, rb),
               "pm_stencil_residual2_noalloc: u does not cover the stencils' domain of r_box");
    PM_REQUIRE(scale1 != 0.0 && scale2 != 0.0, "pm_stencil_residual2_noalloc: zero scale");
    const bool f64 = st1->dtype == PM_F64;
    const size_t need = f64 ? residual2_partials_bytes<double>(st1, st2, rb)
                            : residual2_partials_bytes<float>(st1, st2, rb);
    PM_REQUIRE(need <= kScratchBytes,
               "pm_stencil_residual2_noalloc: the box needs %zu bytes of partials, the fixed scratch holds %zu; "
               "use pm_stencil_residual2",
               need, kScratchBytes);
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != PM_OK) { return rc; }
    if (f64) {
        return residual2_t<double>(st1, st2, u, ub, f, r1, r2, rb, scale1, scale2, red_out, S(stream), ws);
    }
    return residual2_t<float>(st1, st2, u, ub, f, r1, r2, rb, scale1, scale2, red_out, S(stream), ws);
}

}  // extern "C"
