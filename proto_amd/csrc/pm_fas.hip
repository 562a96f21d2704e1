// Proto's library multigrid: LevelMultigridSolver<BoxOp_Laplace> (FAS,
// Jacobi relaxation, constant prolongation) on a periodic level, as launches
// over whole levels.
//
// Operator (Proto_BoxOpImplem.H:129-176 with BoxOp_Laplace.H:24-31 and
// setFluxScale(-1) at Proto_LevelMultigridSolverImplem.H:206):
//   F_d(face p) = sum_j a_j u(p + (j - 2) e_d), a_j = ({1,-15,15,-1}_j * (1/12)) * (1/dx)
//   divFlux     = sum over d = x, y, z of (1/dx) F_d(p + e_d) + (-1/dx) F_d(p)
//   L(u)        = (0 + divFlux) * (-1)
// accumulated in that order with one rounding per operation, so every level
// operator gives Proto's bits.  Relaxation (358-379): u += (-lambda) (L(u) + f),
// lambda = 0.5 / (2 DIM / dx^2), Jacobi (out of place, ping-pong with a
// scratch level).  V-cycle (408-433): 2 DIM pre / post sweeps, 10 at the
// bottom; crseState = AvgDown(u); crse0 = crseState; crseForce =
// AvgDown(L(u) + f) - L_c(crseState) (coarseForce 307-336); u += Constant
// interpolation of (crse0 - crseState) * -1 (fineCorrect 381-406).
#include <cmath>
#include <cstring>
#include <vector>

#include "pm_layout.hpp"

struct pm_fas {
    int nlev = 0;  // levels below the finest
    struct Level {
        pm_layout* layout = nullptr;
        double dx = 0;
        void* phi = nullptr;   // state, ghost 2 (level 0: the caller's)
        void* tmp = nullptr;   // Jacobi scratch, ghost 2
        void* f = nullptr;     // force, ghost 0 (level 0: the caller's)
        void* phi0 = nullptr;  // crseState_0, then the correction, ghost 0
    };
    std::vector<Level> lv;
    double* d_part = nullptr;  // per-block maxima of the residual norm
    double* d_norm = nullptr;
    size_t cap_part = 0;
    size_t bytes = 0;
};

namespace pm {

struct FasC {
    double a[4];   // face-flux coefficients
    double dp, dm; // flux-divergence coefficients
    double mlam;   // -lambda
};

static FasC fas_coefs(double dx)
{
    FasC c;
    const double k[4] = {1, -15, 15, -1};
    const double twelfth = 1.0 / 12.0;
    const double s = 1.0 / dx;
    for (int j = 0; j < 4; j++) { c.a[j] = (k[j] * twelfth) * s; }
    c.dp = 1.0 * s;
    c.dm = -1.0 * s;
    c.mlam = -(0.5 / ((2.0 * 3) / std::pow(dx, 2)));
    return c;
}

// L(u) at storage index i of a box with ghost-2 strides X, XY
__device__ inline double fas_L(const double* __restrict__ u, long i, long X, long XY, const FasC& c)
{
    const long st[3] = {1, X, XY};
    double acc = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const long s = st[d];
        double fr = 0;
        fr = fr + c.a[0] * u[i - s];
        fr = fr + c.a[1] * u[i];
        fr = fr + c.a[2] * u[i + s];
        fr = fr + c.a[3] * u[i + 2 * s];
        double fl = 0;
        fl = fl + c.a[0] * u[i - 2 * s];
        fl = fl + c.a[1] * u[i - s];
        fl = fl + c.a[2] * u[i];
        fl = fl + c.a[3] * u[i + s];
        acc = acc + c.dp * fr;
        acc = acc + c.dm * fl;
    }
    double o = 0;
    o = o + acc;
    return o * -1.0;
}

// one Jacobi sweep: out = u + (-lambda) (L(u) + f) on the valid cells
__global__ __launch_bounds__(256) void k_fas_relax(LayoutDev L, G3 gu, G3 gf, const double* __restrict__ u,
                                                   const double* __restrict__ f, double* __restrict__ out, FasC c)
{
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const double* ub = u + (long)b * gvol(L, gu);
    const double* fb = f + (long)b * gvol(L, gf);
    double* ob = out + (long)b * gvol(L, gu);
    const long X = L.bs[0] + 2 * gu.g[0], XY = X * (L.bs[1] + 2 * gu.g[1]);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gu, x, y, z);
        double res = fas_L(ub, i, X, XY, c);
        res = res + fb[rel_index(L, gf, x, y, z)];
        ob[i] = ub[i] + c.mlam * res;
    }
}

// The same operator on boxes 32..64 cells wide, marching z: a workgroup owns
// a 64 x 8 tile (4 wave64s, 2 rows each) of one box over a z range.  Each
// plane of u -- the tile with 2 halo cells on every side, all inside the box's
// own 2-ghost storage -- is staged once in LDS (a ring of 4 planes, one
// barrier per plane: plane p is last read in step p + 2 and its slot is
// rewritten in step p + 4, after every thread passed the barrier of step
// p + 3); each thread copies its own cells of every plane into a
// 5-plane register window for the z taps and reads the x and y taps of the
// output plane from LDS.  Loads of the next plane are issued before each plane
// is computed; results are stored at the start of the next step, after those
// loads.  The arithmetic is fas_L's, term by term.  MODE 0: Jacobi relaxation
// out = u + (-lambda)(L(u) + f); MODE 1: per-block max |L(u) + f| (and the
// residual itself if res_out).  NB = 1 (a periodic level covering its domain):
// the halo is staged from the neighbouring boxes' valid cells instead of the
// ghost layer -- the values an exchange would put there, with no exchange;
// planes beyond the box go through per-workgroup offset tables (s_dz: per side
// and xy neighbour position, the same for both planes of a side).
// Rows per wave: 2 measured fastest at 512^3 in 64^3 boxes (FAS solve 0.450 s,
// against 0.463 s with 1 row and 0.475 / 0.496 s with 4 / 3 rows): 4 rows held
// the relaxation at 3 waves per SIMD by both registers and LDS.
constexpr int FZ_RY = 2, FZ_TY = 4 * FZ_RY, FZ_NE = (FZ_TY + 4) * 68;

// IN = 2 (with NB): the staged input is u + the constant prolongation of the
// coarse correction crs (layout C, ghost gc), rounded as fineCorrect's
// InterpStencil::Constant rounds it (0 + delta), so the first relaxation after
// fineCorrect reads the cells the prolongation would have written into u.
// MODE 2 (with NB, even box sizes, box-aligned levels): crseForce, the coarse
// right-hand side sum over the 2x2x2 children (z, y, x nested, x fastest) of
// 0.125 (L(u) + f) minus L_c(crs) at the coarse cell (crs the coarse state on
// C with ghost gc, coefficients cc), into fco (ghost 0) -- k_fas_coarse_force's
// arithmetic and order; x neighbours come from the next lane, y pairs are a
// thread's two rows, z pairs consecutive planes.
template<int MODE, int NB, int IN = 0>
__global__ __launch_bounds__(256) void k_fas_op_z(LayoutDev L, const double* __restrict__ u,
                                                  const double* __restrict__ f, double* __restrict__ out, FasC c,
                                                  double* __restrict__ part, double* __restrict__ res_out, int zc,
                                                  LayoutDev C, G3 gc, const double* __restrict__ crs,
                                                  double* __restrict__ fco, FasC cc)
{
    constexpr int RY = FZ_RY, NL = (FZ_NE + 255) / 256;
    __shared__ double sm[4][FZ_NE];
    __shared__ double red[16];
    __shared__ int s_nb[NB ? 27 : 1], s_dz[NB ? 18 : 1], s_dc[IN == 2 ? 18 : 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b = blockIdx.y;
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2];
    const int nyt = (ny + FZ_TY - 1) / FZ_TY;
    const int y0 = (int)(blockIdx.x % nyt) * FZ_TY;
    const int z0 = (int)(blockIdx.x / nyt) * zc, z1 = min(z0 + zc, nz);
    const int LX = nx + 4, NE = (FZ_TY + 4) * LX;
    const long X = nx + 4, XY = X * (ny + 4), vol = XY * (nz + 4);
    const double* ub = u + (long)b * vol;
    const long nv = (long)nx * ny * nz;
    const double* fb = f + (long)b * nv;
    // NB: element of box-relative cell (x, y, z), possibly in a neighbour
    const int* lo = L.lows + 3 * b;
    auto at = [&](int x, int y, int z) {
        const int ox = x < 0 ? -1 : (x >= nx ? 1 : 0);
        const int oy = y < 0 ? -1 : (y >= ny ? 1 : 0);
        const int oz = z < 0 ? -1 : (z >= nz ? 1 : 0);
        const int o = s_nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        return (long)o * vol + (x - ox * nx + 2) + X * ((y - oy * ny + 2) + (long)(ny + 4) * (z - oz * nz + 2));
    };
    // IN == 2: coarse element under fine cell (x, y, z) of this box
    const long cvol = gvol(C, gc);
    const long XYc = (long)(C.bs[0] + 2 * gc.g[0]) * (C.bs[1] + 2 * gc.g[1]);
    auto atc = [&](int x, int y, int z) {
        int p[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
        for (int k = 0; k < 3; k++) {
            p[k] = ((p[k] % L.domain[k]) + L.domain[k]) % L.domain[k];
            p[k] >>= 1;
        }
        const int o = owner(C, p);
        const int* clo = C.lows + 3 * o;
        return (long)o * cvol + rel_index(C, gc, p[0] - clo[0], p[1] - clo[1], p[2] - clo[2]);
    };
    if constexpr (NB) {
        if (tid < 27) {
            int p[3] = {lo[0] + (tid % 3 - 1) * nx, lo[1] + ((tid / 3) % 3 - 1) * ny, lo[2] + (tid / 9 - 1) * nz};
            s_nb[tid] = owner(L, p);
        }
        __syncthreads();
        if (tid < 18) {
            const int side = tid / 9, k = tid % 9;
            const int x = k % 3 == 0 ? -1 : (k % 3 == 1 ? 0 : nx);
            const int y = k / 3 == 0 ? -1 : (k / 3 == 1 ? 0 : ny);
            const int z = side ? nz : -1;
            s_dz[tid] = (int)(at(x, y, z) - (XY * z + at(x, y, 0)));
            if constexpr (IN == 2) { s_dc[tid] = (int)(atc(x, y, z) - (XYc * (z >> 1) + atc(x, y, 0))); }
        }
        __syncthreads();
    }
    // staged element e: row y0 - 2 + e / LX, column e % LX - 2 of the box.
    // Ghost staging: storage row clamped to the last ghost row.  NB: level
    // offsets (fewer than 2^31 elements, checked by the host) and the xy
    // neighbour position, 4 bits per element.  Elements past the plane repeat
    // the last one.
    long goff[NL];
    unsigned long long kp = 0;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = min(tid + 256 * k, NE - 1);
        const int er = e / LX, ec = e - er * LX;
        if constexpr (NB) {
            const int xx = ec - 2, yy = y0 - 2 + er;
            goff[k] = (int)at(xx, yy, 0);
            kp |= (unsigned long long)((xx < 0 ? 0 : (xx >= nx ? 2 : 1)) + 3 * (yy < 0 ? 0 : (yy >= ny ? 2 : 1)))
                  << (4 * k);
        } else {
            goff[k] = X * (long)min(y0 + er, ny + 3) + ec;
        }
    }
    // IN == 2: coarse elements of the staged cells at plane 0; plane z adds
    // z >> 1 coarse planes inside the box (box lows even, levels box-aligned)
    constexpr int NCO = IN == 2 ? NL : 1;
    int offc[NCO];
    if constexpr (IN == 2) {
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int e = min(tid + 256 * k, NE - 1);
            const int er = e / LX, ec = e - er * LX;
            offc[k] = (int)atc(ec - 2, y0 - 2 + er, 0);
        }
    }
    double P[NL], PC[NCO];
    auto gload = [&](int z) {
        if constexpr (NB) {
            const bool inb = z >= 0 && z < nz;
            const int sd = z < 0 ? 0 : 9;
#pragma unroll
            for (int k = 0; k < NL; k++) {
                long o = XY * z + goff[k];
                long oc = 0;
                if constexpr (IN == 2) { oc = XYc * (z >> 1) + offc[k]; }
                if (!inb) {
                    const int ks = sd + (int)((kp >> (4 * k)) & 15);
                    o += s_dz[ks];
                    if constexpr (IN == 2) { oc += s_dc[ks]; }
                }
                P[k] = u[o];
                if constexpr (IN == 2) { PC[k] = crs[oc]; }
            }
        } else {
            const double* base = ub + XY * (long)(z + 2);
#pragma unroll
            for (int k = 0; k < NL; k++) { P[k] = base[goff[k]]; }
        }
    };
    auto lstore = [&](double* pl) {
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int e = tid + 256 * k;
            double v = P[k];
            if constexpr (IN == 2) {
                const double zero = 0;
                v = v + (zero + PC[k]);
            }
            if (e < NE) { pl[e] = v; }
        }
    };
    const int x = lane;
    const bool xok = x < nx;
    const int xc = xok ? x : 0;  // idle lanes read column 0 (results never stored)
    double zw[5][RY];             // own cells of planes zo-2 .. zo+2
    double fq[RY], res[RY];
    int zres = -1;
    auto fload = [&](int zo) {
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = min(y0 + w * RY + r, ny - 1);
            fq[r] = fb[xc + (long)nx * (y + (long)ny * zo)];
        }
    };
    double m = 0;
    auto put_res = [&]() {
        if (zres < 0 || !xok) { return; }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = y0 + w * RY + r;
            if (y >= ny) { continue; }
            if (MODE == 0) {
                // non-temporal: the 31-cycle 512^3 solve 0.406 vs 0.434 s
                __builtin_nontemporal_store(res[r], &out[(long)b * vol + XY * (zres + 2) + X * (y + 2) + (x + 2)]);
            } else if (res_out) {
                res_out[(long)b * nv + x + (long)nx * (y + (long)ny * zres)] = res[r];
            }
        }
    };
    // MODE 2: coarse box of this fine box (box-aligned levels) and the running
    // sums of the row pairs
    double cacc[RY / 2 > 0 ? RY / 2 : 1];
    int cbox = 0, cox = 0, coy = 0, coz = 0;
    const long CXc = C.bs[0] + 2 * gc.g[0], CXYc = CXc * (C.bs[1] + 2 * gc.g[1]);
    const long cv0 = (long)C.bs[0] * C.bs[1] * C.bs[2];
    if constexpr (MODE == 2) {
        int p[3] = {lo[0] >> 1, lo[1] >> 1, lo[2] >> 1};
        cbox = owner(C, p);
        const int* clo = C.lows + 3 * cbox;
        cox = (lo[0] >> 1) - clo[0];
        coy = (lo[1] >> 1) - clo[1];
        coz = (lo[2] >> 1) - clo[2];
    }
    const int kb = z0 - 2;  // first staged plane
    gload(kb);
    fload(z0);
    for (int zz = kb; zz <= z1 + 1; zz++) {
        double* sl = sm[(zz - kb) % 4];
        lstore(sl);
        if (zz + 1 <= z1 + 1) { gload(zz + 1); }
        put_res();
        zres = -1;
        __syncthreads();
        // shift the register window and take this thread's cells of plane zz
#pragma unroll
        for (int r = 0; r < RY; r++) {
#pragma unroll
            for (int q = 0; q < 4; q++) { zw[q][r] = zw[q + 1][r]; }
            zw[4][r] = sl[(w * RY + r + 2) * LX + (xc + 2)];
        }
        if (zz < z0 + 2) { continue; }
        const int zo = zz - 2;
        const double* s0 = sm[(zo - kb) % 4];
        double fcur[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) { fcur[r] = fq[r]; }
        if (zo + 1 < z1) { fload(zo + 1); }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int li = (w * RY + r + 2) * LX + (xc + 2);
            const double uc = zw[2][r];
            double acc = 0;
            double fr, fl;
            // x
            fr = 0; fr = fr + c.a[0] * s0[li - 1]; fr = fr + c.a[1] * uc; fr = fr + c.a[2] * s0[li + 1];
            fr = fr + c.a[3] * s0[li + 2];
            fl = 0; fl = fl + c.a[0] * s0[li - 2]; fl = fl + c.a[1] * s0[li - 1]; fl = fl + c.a[2] * uc;
            fl = fl + c.a[3] * s0[li + 1];
            acc = acc + c.dp * fr;
            acc = acc + c.dm * fl;
            // y
            fr = 0; fr = fr + c.a[0] * s0[li - LX]; fr = fr + c.a[1] * uc; fr = fr + c.a[2] * s0[li + LX];
            fr = fr + c.a[3] * s0[li + 2 * LX];
            fl = 0; fl = fl + c.a[0] * s0[li - 2 * LX]; fl = fl + c.a[1] * s0[li - LX]; fl = fl + c.a[2] * uc;
            fl = fl + c.a[3] * s0[li + LX];
            acc = acc + c.dp * fr;
            acc = acc + c.dm * fl;
            // z
            fr = 0; fr = fr + c.a[0] * zw[1][r]; fr = fr + c.a[1] * uc; fr = fr + c.a[2] * zw[3][r];
            fr = fr + c.a[3] * zw[4][r];
            fl = 0; fl = fl + c.a[0] * zw[0][r]; fl = fl + c.a[1] * zw[1][r]; fl = fl + c.a[2] * uc;
            fl = fl + c.a[3] * zw[3][r];
            acc = acc + c.dp * fr;
            acc = acc + c.dm * fl;
            double o = 0;
            o = o + acc;
            double rv = o * -1.0;
            rv = rv + fcur[r];
            if (MODE == 0) {
                res[r] = uc + c.mlam * rv;
            } else if (MODE == 1) {
                res[r] = rv;
                const bool ok = xok && (y0 + w * RY + r) < ny;
                const double a = rv > 0 ? rv : -rv;
                m = (ok && m < a) ? a : m;
            } else {
                res[r] = rv;
            }
        }
        if constexpr (MODE == 2) {
            static_assert(RY % 2 == 0, "row pairs");
            const bool first = (zo & 1) == 0;
#pragma unroll
            for (int q = 0; q < RY / 2; q++) {
                const double a0 = res[2 * q], a1 = __shfl_down(res[2 * q], 1, 64);
                const double b0 = res[2 * q + 1], b1 = __shfl_down(res[2 * q + 1], 1, 64);
                double t = first ? 0.0 : cacc[q];
                t = t + 0.125 * a0;
                t = t + 0.125 * a1;
                t = t + 0.125 * b0;
                t = t + 0.125 * b1;
                cacc[q] = t;
                const int y = y0 + w * RY + 2 * q;
                if (!first && xok && (x & 1) == 0 && y < ny) {
                    const long ci = rel_index(C, gc, cox + (x >> 1), coy + (y >> 1), coz + (zo >> 1));
                    const double lc = fas_L(crs + (long)cbox * cvol, ci, CXc, CXYc, cc);
                    fco[(long)cbox * cv0 + rel_index(C, G3{{0, 0, 0}}, cox + (x >> 1), coy + (y >> 1),
                                                     coz + (zo >> 1))] = t - lc;
                }
            }
        } else {
            zres = zo;
        }
    }
    put_res();
    if (MODE == 1) {
        m = block_reduce(PM_RED_MAX, m, red);
        if (tid == 0) { part[(long)b * gridDim.x + blockIdx.x] = m; }
    }
}

// All n Jacobi relaxations of a one-box periodic level of at most FS_MAX cells
// in one workgroup: u lives in LDS (periodic neighbours by index), each thread
// keeps f of its own cells in registers and computes its new values before a
// barrier and the write-back (out of place, as the ping-pong).  fas_L's terms
// in its order, then (L(u) + f) and u + (-lambda) res as k_fas_relax rounds
// them, so the same bits.  u has 2 ghost layers, f none.  Small levels are
// latency-bound: one launch instead of an exchange and a relaxation per sweep.
constexpr int FS_MAX = 4096, FS_PER = FS_MAX / 256;

__global__ __launch_bounds__(256) void k_fas_relax_small(LayoutDev L, double* __restrict__ u,
                                                         const double* __restrict__ f, FasC c, int nsw)
{
    __shared__ double us[FS_MAX];
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2], N = nx * ny * nz;
    const long X = nx + 4, XY = X * (ny + 4);
    double fv[FS_PER], nv[FS_PER];
#pragma unroll
    for (int j = 0; j < FS_PER; j++) {
        const int i = threadIdx.x + 256 * j;
        if (i < N) {
            const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            fv[j] = f[i];
            us[i] = u[(x + 2) + X * (y + 2) + XY * (z + 2)];
        }
    }
    // periodic fold of offsets up to +-2: a level 1 cell wide wraps more than once
    auto fold = [](int x, int n) {
        x %= n;
        return x < 0 ? x + n : x;
    };
    auto at = [&](int x, int y, int z) { return us[fold(x, nx) + nx * (fold(y, ny) + ny * fold(z, nz))]; };
    for (int sw = 0; sw < nsw; sw++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < FS_PER; j++) {
            const int i = threadIdx.x + 256 * j;
            if (i >= N) { continue; }
            const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            const double uc = us[i];
            double acc = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int ex = d == 0, ey = d == 1, ez = d == 2;
                double fr = 0;
                fr = fr + c.a[0] * at(x - ex, y - ey, z - ez);
                fr = fr + c.a[1] * uc;
                fr = fr + c.a[2] * at(x + ex, y + ey, z + ez);
                fr = fr + c.a[3] * at(x + 2 * ex, y + 2 * ey, z + 2 * ez);
                double fl = 0;
                fl = fl + c.a[0] * at(x - 2 * ex, y - 2 * ey, z - 2 * ez);
                fl = fl + c.a[1] * at(x - ex, y - ey, z - ez);
                fl = fl + c.a[2] * uc;
                fl = fl + c.a[3] * at(x + ex, y + ey, z + ez);
                acc = acc + c.dp * fr;
                acc = acc + c.dm * fl;
            }
            double o = 0;
            o = o + acc;
            double res = o * -1.0;
            res = res + fv[j];
            nv[j] = uc + c.mlam * res;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < FS_PER; j++) {
            const int i = threadIdx.x + 256 * j;
            if (i < N) { us[i] = nv[j]; }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += 256) {
        const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        u[(x + 2) + X * (y + 2) + XY * (z + 2)] = us[i];
    }
}

// per-block max |L(u) + f|
__global__ __launch_bounds__(256) void k_fas_resnorm(LayoutDev L, G3 gu, G3 gf, const double* __restrict__ u,
                                                     const double* __restrict__ f, FasC c, double* __restrict__ part,
                                                     double* __restrict__ res_out)
{
    __shared__ double red[16];
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const double* ub = u + (long)b * gvol(L, gu);
    const double* fb = f + (long)b * gvol(L, gf);
    const long X = L.bs[0] + 2 * gu.g[0], XY = X * (L.bs[1] + 2 * gu.g[1]);
    double m = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        double res = fas_L(ub, rel_index(L, gu, x, y, z), X, XY, c);
        res = res + fb[rel_index(L, gf, x, y, z)];
        if (res_out) { res_out[b * n + t] = res; }
        const double a = res > 0 ? res : -res;
        m = m < a ? a : m;
    }
    m = block_reduce(PM_RED_MAX, m, red);
    if (threadIdx.x == 0) { part[(long)b * gridDim.x + blockIdx.x] = m; }
}

__global__ __launch_bounds__(1024) void k_fas_max(const double* __restrict__ part, long n, double* __restrict__ out)
{
    __shared__ double red[16];
    double m = 0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) { m = m < part[i] ? part[i] : m; }
    m = block_reduce(PM_RED_MAX, m, red);
    if (threadIdx.x == 0) { *out = m; }
}

// crseState = AvgDown(u) (Box(2) order, x fastest) and crse0 = crseState
__global__ __launch_bounds__(256) void k_fas_avgdown(LayoutDev F, LayoutDev C, G3 gu, G3 gc, G3 g0,
                                                     const double* __restrict__ u, double* __restrict__ uc,
                                                     double* __restrict__ uc0)
{
    const int b = blockIdx.y;
    const int* clo = C.lows + 3 * b;
    const long n = (long)C.bs[0] * C.bs[1] * C.bs[2];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % C.bs[0]);
        const long r = t / C.bs[0];
        const int y = (int)(r % C.bs[1]);
        const int z = (int)(r / C.bs[1]);
        int p[3] = {2 * (clo[0] + x), 2 * (clo[1] + y), 2 * (clo[2] + z)};
        const int fb = owner(F, p);
        const int* flo = F.lows + 3 * fb;
        const double* ub = u + (long)fb * gvol(F, gu);
        const int fx = p[0] - flo[0], fy = p[1] - flo[1], fz = p[2] - flo[2];
        double acc = 0;
#pragma unroll
        for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
#pragma unroll
                for (int kx = 0; kx < 2; kx++) { acc = acc + 0.125 * ub[rel_index(F, gu, fx + kx, fy + ky, fz + kz)]; }
        uc[(long)b * gvol(C, gc) + rel_index(C, gc, x, y, z)] = acc;
        uc0[(long)b * gvol(C, g0) + rel_index(C, g0, x, y, z)] = acc;
    }
}

// crseForce = AvgDown(L_f(u) + f) - L_c(crseState); u and crseState exchanged
__global__ __launch_bounds__(256) void k_fas_coarse_force(LayoutDev F, LayoutDev C, G3 gu, G3 gf, G3 gc, G3 g0,
                                                          const double* __restrict__ u, const double* __restrict__ f,
                                                          const double* __restrict__ uc, double* __restrict__ fc,
                                                          FasC cf, FasC cc)
{
    const int b = blockIdx.y;
    const int* clo = C.lows + 3 * b;
    const long n = (long)C.bs[0] * C.bs[1] * C.bs[2];
    const long FX = F.bs[0] + 2 * gu.g[0], FXY = FX * (F.bs[1] + 2 * gu.g[1]);
    const long CX = C.bs[0] + 2 * gc.g[0], CXY = CX * (C.bs[1] + 2 * gc.g[1]);
    const double* ucb = uc + (long)b * gvol(C, gc);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % C.bs[0]);
        const long r = t / C.bs[0];
        const int y = (int)(r % C.bs[1]);
        const int z = (int)(r / C.bs[1]);
        int p[3] = {2 * (clo[0] + x), 2 * (clo[1] + y), 2 * (clo[2] + z)};
        const int fbx = owner(F, p);
        const int* flo = F.lows + 3 * fbx;
        const double* ub = u + (long)fbx * gvol(F, gu);
        const double* fb = f + (long)fbx * gvol(F, gf);
        const int fx = p[0] - flo[0], fy = p[1] - flo[1], fz = p[2] - flo[2];
        double acc = 0;
#pragma unroll
        for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
#pragma unroll
                for (int kx = 0; kx < 2; kx++) {
                    double res = fas_L(ub, rel_index(F, gu, fx + kx, fy + ky, fz + kz), FX, FXY, cf);
                    res = res + fb[rel_index(F, gf, fx + kx, fy + ky, fz + kz)];
                    acc = acc + 0.125 * res;
                }
        const double lc = fas_L(ucb, rel_index(C, gc, x, y, z), CX, CXY, cc);
        fc[(long)b * gvol(C, g0) + rel_index(C, g0, x, y, z)] = acc - lc;
    }
}

// correction = (crse0 - crseState) * -1, written over crse0
__global__ __launch_bounds__(256) void k_fas_delta(LayoutDev C, G3 gc, G3 g0, const double* __restrict__ uc,
                                                   double* __restrict__ uc0)
{
    const int b = blockIdx.y;
    const long n = (long)C.bs[0] * C.bs[1] * C.bs[2];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % C.bs[0]);
        const long r = t / C.bs[0];
        const int y = (int)(r % C.bs[1]);
        const int z = (int)(r / C.bs[1]);
        const long i0 = (long)b * gvol(C, g0) + rel_index(C, g0, x, y, z);
        const double d = uc0[i0] - uc[(long)b * gvol(C, gc) + rel_index(C, gc, x, y, z)];
        uc0[i0] = d * -1.0;
    }
}

static unsigned fas_blocks(long n) { return (unsigned)std::min<long>((n + 255) / 256, 1024); }

#define FAS_TRY(x)                    \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != PM_OK) return rc_; \
    } while (0)

static const int G2[3] = {2, 2, 2}, G0[3] = {0, 0, 0};

// boxes the z-marching operator kernel takes: 32..64 cells wide, at least a tile
// tall and deep
static bool fz_ok(const LayoutDev& D) { return D.bs[0] >= 32 && D.bs[0] <= 64 && D.bs[1] >= 16 && D.bs[2] >= 8; }
// levels whose z-marching operator can read the neighbouring boxes instead of
// the ghost layer: periodic, every box present, fewer than 2^31 elements
static bool fz_nb(const pm_layout* L)
{
    const LayoutDev D = ldev(L);
    if (!fz_ok(D) || L->nboxes != L->nb[0] * L->nb[1] * L->nb[2]) { return false; }
    for (int k = 0; k < 3; k++) {
        if (!L->periodic[k]) { return false; }
    }
    return gvol(D, g3(G2)) * L->nboxes < (1L << 31);
}

// its grid: 16-row tiles x z chunks (halved while fewer than 2048 workgroups) x boxes
static dim3 fz_grid(const LayoutDev& D, int* zc)
{
    const int nyt = (D.bs[1] + FZ_TY - 1) / FZ_TY;
    int z = D.bs[2];
    while (z > 8 && (long)nyt * ((D.bs[2] + z - 1) / z) * D.nboxes < 2048) { z = (z + 1) / 2; }
    *zc = z;
    return dim3((unsigned)(nyt * ((D.bs[2] + z - 1) / z)), (unsigned)D.nboxes);
}

// n Jacobi relaxations of level l.  delta (the coarse correction of level
// l + 1, ghost 0): fineCorrect's u += I(delta) is still to be applied; the first
// relaxation stages it where the levels allow, else it is applied first.
static int fas_relax(pm_fas* M, int l, void* u, const void* f, int n, hipStream_t s, const void* delta = nullptr)
{
    auto& Lv = M->lv[l];
    const LayoutDev D = ldev(Lv.layout);
    const FasC c = fas_coefs(Lv.dx);
    const long cells = (long)D.bs[0] * D.bs[1] * D.bs[2];
    void* buf[2] = {u, Lv.tmp};
    int zc = 0;
    const dim3 zgrid = fz_grid(D, &zc);
    const bool nb = fz_nb(Lv.layout);
    bool fold = false;
    if (delta) {
        const pm_layout* CL = M->lv[l + 1].layout;
        fold = nb && n > 0 && levels_box_aligned(Lv.layout, CL) &&
               gvol(ldev(CL), g3(G0)) * CL->nboxes < (1L << 31);
        if (!fold) { FAS_TRY(pm_level_prolong_add(Lv.layout, CL, PM_F64, delta, G0, u, G2, 0, s)); }
    }
    // a small one-box periodic level: every relaxation in one launch
    if (!nb && n > 0 && Lv.layout->nboxes == 1 && cells <= FS_MAX && Lv.layout->periodic[0] &&
        Lv.layout->periodic[1] && Lv.layout->periodic[2]) {
        hipLaunchKernelGGL(k_fas_relax_small, dim3(1), dim3(256), 0, s, D, (double*)u, (const double*)f, c, n);
        PM_CHECK_LAUNCH();
        return PM_OK;
    }
    for (int it = 0; it < n; it++) {
        if (nb && fold && it == 0) {
            hipLaunchKernelGGL((k_fas_op_z<0, 1, 2>), zgrid, dim3(256), 0, s, D, (const double*)u, (const double*)f,
                               (double*)buf[1], c, (double*)nullptr, (double*)nullptr, zc,
                               ldev(M->lv[l + 1].layout), g3(G0), (const double*)delta, (double*)nullptr, c);
            PM_CHECK_LAUNCH();
            continue;
        }
        if (nb) {
            // halo read from the neighbouring boxes of the input buffer (the
            // relaxation is out of place): no exchange
            hipLaunchKernelGGL((k_fas_op_z<0, 1>), zgrid, dim3(256), 0, s, D, (const double*)buf[it & 1],
                               (const double*)f, (double*)buf[(it + 1) & 1], c, (double*)nullptr, (double*)nullptr,
                               zc, D, g3(G0), (const double*)nullptr, (double*)nullptr, c);
            PM_CHECK_LAUNCH();
            continue;
        }
        FAS_TRY(pm_level_exchange(Lv.layout, PM_F64, buf[it & 1], G2, 1, s));
        if (fz_ok(D)) {
            hipLaunchKernelGGL((k_fas_op_z<0, 0>), zgrid, dim3(256), 0, s, D, (const double*)buf[it & 1],
                               (const double*)f, (double*)buf[(it + 1) & 1], c, (double*)nullptr, (double*)nullptr,
                               zc, D, g3(G0), (const double*)nullptr, (double*)nullptr, c);
        } else {
            hipLaunchKernelGGL(k_fas_relax, dim3(fas_blocks(cells), D.nboxes), dim3(256), 0, s, D, g3(G2), g3(G0),
                               (const double*)buf[it & 1], (const double*)f, (double*)buf[(it + 1) & 1], c);
        }
        PM_CHECK_LAUNCH();
    }
    if (n & 1) {
        FAS_TRY(pm_level_copy(Lv.layout, Lv.tmp, G2, Lv.layout, u, G2, PM_F64, 1, s));
    }
    return PM_OK;
}

static int fas_vcycle(pm_fas* M, int l, void* u, const void* f, hipStream_t s)
{
    if (l == M->nlev) { return fas_relax(M, l, u, f, 10, s); }
    auto& F = M->lv[l];
    auto& C = M->lv[l + 1];
    FAS_TRY(fas_relax(M, l, u, f, 6, s));
    const LayoutDev DF = ldev(F.layout), DC = ldev(C.layout);
    // crseForce from the z-marching operator kernel where the fine level reads
    // its neighbours (then u needs no exchange: averageDown reads valid cells)
    int zc = 0;
    const dim3 zgrid = fz_grid(DF, &zc);
    const bool zforce = fz_nb(F.layout) && levels_box_aligned(F.layout, C.layout) && zc % 2 == 0 &&
                        DF.bs[0] % 2 == 0 && DF.bs[1] % 2 == 0 && DF.bs[2] % 2 == 0 &&
                        gvol(DC, g3(G2)) * DC.nboxes < (1L << 31);
    if (!zforce) { FAS_TRY(pm_level_exchange(F.layout, PM_F64, u, G2, 1, s)); }
    const long cc = (long)DC.bs[0] * DC.bs[1] * DC.bs[2];
    const dim3 grid(fas_blocks(cc), DC.nboxes);
    hipLaunchKernelGGL(k_fas_avgdown, grid, dim3(256), 0, s, DF, DC, g3(G2), g3(G2), g3(G0), (const double*)u,
                       (double*)C.phi, (double*)C.phi0);
    PM_CHECK_LAUNCH();
    FAS_TRY(pm_level_exchange(C.layout, PM_F64, C.phi, G2, 1, s));
    if (zforce) {
        hipLaunchKernelGGL((k_fas_op_z<2, 1>), zgrid, dim3(256), 0, s, DF, (const double*)u, (const double*)f,
                           (double*)nullptr, fas_coefs(F.dx), (double*)nullptr, (double*)nullptr, zc, DC, g3(G2),
                           (const double*)C.phi, (double*)C.f, fas_coefs(C.dx));
    } else {
        hipLaunchKernelGGL(k_fas_coarse_force, grid, dim3(256), 0, s, DF, DC, g3(G2), g3(G0), g3(G2), g3(G0),
                           (const double*)u, (const double*)f, (const double*)C.phi, (double*)C.f, fas_coefs(F.dx),
                           fas_coefs(C.dx));
    }
    PM_CHECK_LAUNCH();
    FAS_TRY(fas_vcycle(M, l + 1, C.phi, C.f, s));
    hipLaunchKernelGGL(k_fas_delta, grid, dim3(256), 0, s, DC, g3(G2), g3(G0), (const double*)C.phi, (double*)C.phi0);
    PM_CHECK_LAUNCH();
    // fineCorrect (u += I(delta)) folded into the first post relaxation
    return fas_relax(M, l, u, f, 6, s, C.phi0);
}

// residual norm of level 0 (Proto's MG resnorm exchanges first).  in_solve:
// where the level qualifies, the norm kernel reads the neighbouring boxes and
// the exchange is skipped (pm_fas_solve exchanges once at its end).
static int fas_resnorm(pm_fas* M, void* u, const void* f, double* out, hipStream_t s, double* res_out = nullptr,
                       bool in_solve = false)
{
    auto& Lv = M->lv[0];
    const LayoutDev D = ldev(Lv.layout);
    const bool nb = in_solve && fz_nb(Lv.layout);
    if (!nb) { FAS_TRY(pm_level_exchange(Lv.layout, PM_F64, u, G2, 1, s)); }
    const long cells = (long)D.bs[0] * D.bs[1] * D.bs[2];
    unsigned bpb = fas_blocks(cells);
    if (fz_ok(D)) {
        int zc = 0;
        const dim3 zgrid = fz_grid(D, &zc);
        bpb = zgrid.x;  // no more partials than the generic kernel's (<= 64 per box)
        if (nb) {
            hipLaunchKernelGGL((k_fas_op_z<1, 1>), zgrid, dim3(256), 0, s, D, (const double*)u, (const double*)f,
                               (double*)nullptr, fas_coefs(Lv.dx), M->d_part, res_out, zc, D, g3(G0),
                               (const double*)nullptr, (double*)nullptr, fas_coefs(Lv.dx));
        } else {
            hipLaunchKernelGGL((k_fas_op_z<1, 0>), zgrid, dim3(256), 0, s, D, (const double*)u, (const double*)f,
                               (double*)nullptr, fas_coefs(Lv.dx), M->d_part, res_out, zc, D, g3(G0),
                               (const double*)nullptr, (double*)nullptr, fas_coefs(Lv.dx));
        }
    } else {
        hipLaunchKernelGGL(k_fas_resnorm, dim3(bpb, D.nboxes), dim3(256), 0, s, D, g3(G2), g3(G0),
                           (const double*)u, (const double*)f, fas_coefs(Lv.dx), M->d_part, res_out);
    }
    hipLaunchKernelGGL(k_fas_max, dim3(1), dim3(1024), 0, s, (const double*)M->d_part, (long)bpb * D.nboxes,
                       M->d_norm);
    PM_CHECK_LAUNCH();
    PM_CHECK_HIP(hipMemcpyAsync(out, M->d_norm, sizeof(double), hipMemcpyDeviceToHost, s));
    PM_CHECK_HIP(hipStreamSynchronize(s));
    return PM_OK;
}

static int fas_alloc(pm_fas* M, void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        set_error("pm_fas_create: allocation of %zu bytes failed", bytes);
        return PM_ERR_NOMEM;
    }
    M->bytes += bytes;
    return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {

int pm_fas_destroy(pm_fas* M)
{
    if (!M) { return PM_OK; }
    for (size_t l = 0; l < M->lv.size(); l++) {
        auto& L = M->lv[l];
        if (l > 0) {
            hipFree(L.phi);
            hipFree(L.f);
            hipFree(L.phi0);
        }
        hipFree(L.tmp);
        pm_layout_destroy(L.layout);
    }
    hipFree(M->d_part);
    hipFree(M->d_norm);
    delete M;
    return PM_OK;
}

int pm_fas_create(pm_fas** out, const pm_layout* fine, int num_levels, double dx)
{
    PM_REQUIRE(out && fine && dx > 0, "pm_fas_create: bad argument");
    *out = nullptr;
    PM_REQUIRE(num_levels >= 1 && num_levels <= 24, "pm_fas_create: num_levels must be 1..24");
    PM_REQUIRE(fine->dim == 3, "pm_fas_create: 3-D layouts only");
    for (int k = 0; k < 3; k++) {
        PM_REQUIRE(fine->periodic[k], "pm_fas_create: the solver needs a periodic domain");
        PM_REQUIRE(fine->domain[k] % (1 << (num_levels - 1)) == 0, "pm_fas_create: domain not coarsenable");
        PM_REQUIRE(fine->bs[k] >= 2, "pm_fas_create: boxes need 2+ cells per side");
    }
    PM_REQUIRE(fine->nboxes == fine->nb[0] * fine->nb[1] * fine->nb[2],
               "pm_fas_create: the layout must cover its domain");
    pm_fas* M = new pm_fas();
    M->nlev = num_levels - 1;
    M->lv.resize(num_levels);
    int rc = pm_layout_create(&M->lv[0].layout, 3, fine->domain, fine->bs, fine->periodic, fine->nboxes,
                              fine->lows.data());
    M->lv[0].dx = dx;
    // coarse layouts (MGLevel::define, 213-239): domain halves, box size kept
    // when it divides the coarse domain, else one box
    int dom[3] = {fine->domain[0], fine->domain[1], fine->domain[2]};
    int box[3] = {fine->bs[0], fine->bs[1], fine->bs[2]};
    for (int l = 1; l < num_levels && rc == PM_OK; l++) {
        bool divisible = true;
        for (int k = 0; k < 3; k++) {
            dom[k] /= 2;
            if (dom[k] % box[k] != 0) { divisible = false; }
        }
        if (!divisible) {
            for (int k = 0; k < 3; k++) { box[k] = dom[k]; }
        }
        std::vector<int> lows;
        for (int z = 0; z < dom[2] / box[2]; z++)
            for (int y = 0; y < dom[1] / box[1]; y++)
                for (int x = 0; x < dom[0] / box[0]; x++) {
                    lows.push_back(x * box[0]);
                    lows.push_back(y * box[1]);
                    lows.push_back(z * box[2]);
                }
        rc = pm_layout_create(&M->lv[l].layout, 3, dom, box, fine->periodic, (int)(lows.size() / 3), lows.data());
        M->lv[l].dx = M->lv[l - 1].dx * 2;
    }
    size_t maxpart = 0;
    for (int l = 0; l < num_levels && rc == PM_OK; l++) {
        auto& L = M->lv[l];
        const LayoutDev D = ldev(L.layout);
        const size_t e2 = (size_t)gvol(D, g3(G2)) * D.nboxes, e0 = (size_t)gvol(D, g3(G0)) * D.nboxes;
        rc = fas_alloc(M, &L.tmp, 8 * e2);
        if (l > 0 && rc == PM_OK) { rc = fas_alloc(M, &L.phi, 8 * e2); }
        if (l > 0 && rc == PM_OK) { rc = fas_alloc(M, &L.f, 8 * e0); }
        if (l > 0 && rc == PM_OK) { rc = fas_alloc(M, &L.phi0, 8 * e0); }
        if (l == 0) { maxpart = (size_t)fas_blocks((long)e0 / D.nboxes) * D.nboxes; }
    }
    if (rc == PM_OK) { rc = fas_alloc(M, (void**)&M->d_part, 8 * std::max<size_t>(maxpart, 1)); }
    if (rc == PM_OK) { rc = fas_alloc(M, (void**)&M->d_norm, 64); }
    if (rc != PM_OK) {
        pm_fas_destroy(M);
        return rc;
    }
    *out = M;
    return PM_OK;
}

int pm_fas_vcycle(pm_fas* M, void* state, const void* force, pm_stream_t stream)
{
    PM_REQUIRE(M && state && force, "pm_fas_vcycle: null argument");
    return fas_vcycle(M, 0, state, force, (hipStream_t)stream);
}

int pm_fas_resnorm(pm_fas* M, void* state, const void* force, double* out_host, pm_stream_t stream)
{
    PM_REQUIRE(M && state && force && out_host, "pm_fas_resnorm: null argument");
    return fas_resnorm(M, state, force, out_host, (hipStream_t)stream);
}

int pm_fas_residual(pm_fas* M, void* state, const void* force, void* residual, double* out_host, pm_stream_t stream)
{
    PM_REQUIRE(M && state && force && residual && out_host, "pm_fas_residual: null argument");
    return fas_resnorm(M, state, force, out_host, (hipStream_t)stream, (double*)residual);
}

int pm_fas_solve(pm_fas* M, void* state, const void* force, int max_iter, double tol, double* hist, int* iters,
                 pm_stream_t stream)
{
    PM_REQUIRE(M && state && force && max_iter >= 0, "pm_fas_solve: bad argument");
    const hipStream_t s = (hipStream_t)stream;
    double res0 = 0, r = 0;
    FAS_TRY(fas_resnorm(M, state, force, &res0, s, nullptr, true));
    if (hist) { hist[0] = res0; }
    r = res0;
    int it = 0;
    for (; it < max_iter; it++) {
        FAS_TRY(fas_vcycle(M, 0, state, force, s));
        FAS_TRY(fas_resnorm(M, state, force, &r, s, nullptr, true));
        if (hist) { hist[it + 1] = r; }
        if (r < tol * res0) {
            it++;
            break;
        }
    }
    if (iters) { *iters = it; }
    // the ghost cells of the state as Proto's last resnorm leaves them
    FAS_TRY(pm_level_exchange(M->lv[0].layout, PM_F64, state, G2, 1, s));
    return PM_OK;
}

int pm_fas_bytes(const pm_fas* M, size_t* bytes)
{
    PM_REQUIRE(M && bytes, "pm_fas_bytes: null argument");
    *bytes = M->bytes;
    return PM_OK;
}

}  // extern "C"
