// Level (LevelBoxData over a DisjointBoxLayout) kernels: periodic exchange,
// GSRB / Jacobi relaxation, residual norm, fused residual + restriction into
// the coarse layout, prolongation, copy between layouts.
//
// Every box of a layout has the same size and lies on the box grid of the
// domain, so the owner of any cell is one table lookup; each launch covers all
// boxes of the level (blockIdx.y = box), instead of one launch per box and
// per motion item as Proto's Copier/LevelCopierOp does.
#include <algorithm>
#include <vector>

#include "pm_internal.hpp"

#include "pm_layout.hpp"

namespace pm {

// The z-marching level kernels split z into chunks until a launch has enough
// workgroups; small levels, whose launches are bound by the latency of the z
// walk, go down to chunks of kZcFloor planes.
constexpr int kZcFloor = 2;
// waves the row sweep aims at per launch: 16 per CU (four per SIMD, what its
// FP32 register count allows); fewer z chunks mean fewer warm-up planes
constexpr long kRowsWaves = 4096;


// ------------------------------------------------------------ exchange
template<typename T>
__global__ __launch_bounds__(256) void k_exchange(LayoutDev L, G3 g, T* data, int ncomp)
{
    const int b = blockIdx.y;
    const int X = L.bs[0] + 2 * g.g[0], Y = L.bs[1] + 2 * g.g[1];
    const long A = (long)X * Y * g.g[2];
    const long B = (long)X * g.g[1] * L.bs[2];
    const long C = (long)g.g[0] * L.bs[1] * L.bs[2];
    const long n = 2 * (A + B + C);
    const int* lo = L.lows + 3 * b;
    const long vol = gvol(L, g);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        long u = t;
        if (u < 2 * A) {
            const bool hi = u >= A;
            if (hi) { u -= A; }
            x = (int)(u % X) - g.g[0];
            y = (int)((u / X) % Y) - g.g[1];
            const int zz = (int)(u / ((long)X * Y));
            z = hi ? L.bs[2] + zz : zz - g.g[2];
        } else if ((u -= 2 * A) < 2 * B) {
            const bool hi = u >= B;
            if (hi) { u -= B; }
            x = (int)(u % X) - g.g[0];
            const int yy = (int)((u / X) % g.g[1]);
            z = (int)(u / ((long)X * g.g[1]));
            y = hi ? L.bs[1] + yy : yy - g.g[1];
        } else {
            u -= 2 * B;
            const bool hi = u >= C;
            if (hi) { u -= C; }
            const int xx = (int)(u % g.g[0]);
            y = (int)((u / g.g[0]) % L.bs[1]);
            z = (int)(u / ((long)g.g[0] * L.bs[1]));
            x = hi ? L.bs[0] + xx : xx - g.g[0];
        }
        int p[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
        const int o = owner(L, p);
        if (o < 0) { continue; }
        const int* olo = L.lows + 3 * o;
        const long si = rel_index(L, g, p[0] - olo[0], p[1] - olo[1], p[2] - olo[2]);
        const long di = rel_index(L, g, x, y, z);
        for (int c = 0; c < ncomp; c++) {
            data[((long)b * ncomp + c) * vol + di] = data[((long)o * ncomp + c) * vol + si];
        }
    }
}

// ------------------------------------------------------------ GSRB
// Coefficients of Laplacian() (Proto_StencilDefs.H:77-87) in order.
__host__ __device__ constexpr int lapo(int j, int d)
{
    constexpr int t[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    return t[j][d];
}

template<typename T>
struct Lap {
    T c[7];
};

template<typename T, int DIMN>
__global__ __launch_bounds__(256) void k_gsrb(LayoutDev L, G3 gp, G3 gr, T* __restrict__ phi,
                                              const T* __restrict__ rhs, Lap<T> c, T lambda, int color)
{
    constexpr int NC = 2 * DIMN + 1;
    const int b = blockIdx.y;
    const int* lo = L.lows + 3 * b;
    const int hx = (L.bs[0] + 1) / 2;
    const long n = (long)hx * L.bs[1] * L.bs[2];
    T* ph = phi + (long)b * gvol(L, gp);
    const T* rh = rhs + (long)b * gvol(L, gr);
    const long X = L.bs[0] + 2 * gp.g[0], XY = X * (L.bs[1] + 2 * gp.g[1]);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int xh = (int)(t % hx);
        const long r = t / hx;
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const int par = (color + lo[0] + lo[1] + y + lo[2] + z) & 1;
        const int x = 2 * xh + par;
        if (x >= L.bs[0]) { continue; }
        const long i = rel_index(L, gp, x, y, z);
        T s = 0;
#pragma unroll
        for (int j = 0; j < NC; j++) {
            s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * X + lapo(j, 2) * XY];
        }
        s = s + rh[rel_index(L, gr, x, y, z)];
        ph[i] = ph[i] + lambda * s;
    }
}


// ------------------------------------------------------------ fused GSRB sweep
// One full red-black sweep in one pass, out of place (pin is only read, pout
// only written), with no exchange: a workgroup owns columns [x0, x1) x rows
// [y0, y0+R) x planes [z0, z1) of one box (column tiles of tw cells, at most
// 64 wide) and walks z through an LDS ring of 4 planes holding rows
// [y0-2, y1+2) x columns [x0-2, x1+2).  Cells outside the box come from
// their owner box's valid cells (periodic images folded in).  At step k the
// red cells of plane k+1 are updated over the tile plus a one-cell halo (the
// halo values equal what the neighbouring tiles compute for those cells:
// same inputs, same order), then the black cells of plane k over the tile,
// which is written out.  Same arithmetic and order as k_gsrb's two colours.
// IN selects what the staged input is: 0 pin; 1 zero (pin is not read: the
// sweep of a zero initial guess); 2 pin + the constant prolongation of the
// coarse level crs (layout C, ghost gc), rounded as k_prolong_const_box
// rounds it, so the sweep sees the cells fineInterp would have written.
// OUT = 1: instead of storing its result, the sweep adds it to the FP64 array
// pacc (layout and ghosts of pout) as k_add_correction adds a correction, and
// pout is not written (the last sweep of the FP32 preconditioner).
template<typename T, int MAXE, int IN, int OUT>
__global__ __launch_bounds__(256) void k_gsrb_fused(LayoutDev L, G3 gin, G3 gout, G3 gr, const T* __restrict__ pin,
                                                    T* __restrict__ pout, const T* __restrict__ rhs, Lap<T> c,
                                                    T lambda, int rows, int zc, int tw, LayoutDev C, G3 gc,
                                                    const T* __restrict__ crs, double* __restrict__ pacc)
{
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ int s_nb[27];  // box owning each of the 27 neighbouring box positions (13 = this box)
    // offset corrections of planes beyond the box (below: 0..8, above: 9..17),
    // per xy neighbour position k = (ox + 1) + 3 (oy + 1) of the staged cell:
    // phi, rhs and (IN == 2) coarse element
    __shared__ int s_dz[18], s_dr[18], s_dc[18];
    const int b = blockIdx.y;
    const int* lo = L.lows + 3 * b;
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2];
    const int nyt = (ny + rows - 1) / rows, nxt = (nx + tw - 1) / tw;
    const int bt = blockIdx.x / nxt;
    const int x0 = (blockIdx.x % nxt) * tw, x1 = min(x0 + tw, nx);
    const int y0 = (bt % nyt) * rows, y1 = min(y0 + rows, ny);
    const int z0 = (bt / nyt) * zc, z1 = min(z0 + zc, nz);
    const int TWc = x1 - x0;                         // columns of this tile
    const int W = TWc + 4, NP = (y1 - y0 + 4) * W;   // phi plane: rows [y0-2, y1+2) x columns [x0-2, x1+2)
    const int WR = TWc + 2, NR = (y1 - y0 + 2) * WR; // rhs plane: rows [y0-1, y1+1) x columns [x0-1, x1+1)
    T* sm = reinterpret_cast<T*>(smraw);
    T* rsm = sm + 4 * (long)(rows + 4) * (tw + 4);
    const int losum = lo[0] + lo[1] + lo[2];
    const long vin = gvol(L, gin), vr = gvol(L, gr);
    const long XYp = (long)(nx + 2 * gin.g[0]) * (ny + 2 * gin.g[1]);
    const long XYr = (long)(nx + 2 * gr.g[0]) * (ny + 2 * gr.g[1]);
    const long cvol = gvol(C, gc);
    const long XYc = (long)(C.bs[0] + 2 * gc.g[0]) * (C.bs[1] + 2 * gc.g[1]);
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        int p[3] = {lo[0] + (t % 3 - 1) * nx, lo[1] + ((t / 3) % 3 - 1) * ny, lo[2] + (t / 9 - 1) * nz};
        s_nb[t] = owner(L, p);
    }
    __syncthreads();
    // element of (box-local, possibly outside the box) cell (x, y, z) in data with ghost g
    auto at = [&](const G3& g, long vol, int x, int y, int z) {
        const int ox = x < 0 ? -1 : (x >= nx ? 1 : 0);
        const int oy = y < 0 ? -1 : (y >= ny ? 1 : 0);
        const int oz = z < 0 ? -1 : (z >= nz ? 1 : 0);
        const int o = s_nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        return (long)o * vol + rel_index(L, g, x - ox * nx, y - oy * ny, z - oz * nz);
    };
    // coarse element under fine cell (x, y, z) (IN == 2)
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
    if (threadIdx.x < 18) {
        // A plane beyond the box is the nearest planes of the box below /
        // above, so for every staged cell of neighbour position k the element
        // differs from the in-box formula (plane z * plane size + in-plane
        // offset) by the same amount for both planes on that side: the
        // difference at a representative cell of k (for the coarse element,
        // both fine planes of a side lie in one coarse plane, box lows even).
        const int t = threadIdx.x, side = t / 9, k = t % 9;
        const int x = k % 3 == 0 ? -1 : (k % 3 == 1 ? 0 : nx);
        const int y = k / 3 == 0 ? -1 : (k / 3 == 1 ? 0 : ny);
        const int z = side ? nz : -1;
        s_dz[t] = (int)(at(gin, vin, x, y, z) - (XYp * z + at(gin, vin, x, y, 0)));
        s_dr[t] = (int)(at(gr, vr, x, y, z) - (XYr * z + at(gr, vr, x, y, 0)));
        s_dc[t] = IN == 2 ? (int)(atc(x, y, z) - (XYc * (z >> 1) + atc(x, y, 0))) : 0;
    }
    __syncthreads();
    T* ob = pout + (long)b * gvol(L, gout);
    auto slot = [&](int z) { return sm + (long)(((z % 4) + 4) % 4) * NP; };
    auto rslot = [&](int z) { return rsm + (long)(z & 1) * NR; };
    // in-plane offsets of this thread's elements for planes inside the box
    // (z = 0; plane z adds z * plane size; planes beyond the box add s_dz /
    // s_dr / s_dc of the cell's neighbour position, packed 4 bits per element
    // in kp / kr).  Elements past the plane repeat the last one, so every
    // load is issued unconditionally, straight into the register that holds
    // it until its commit: the compiler's waits then only ever cover loads a
    // step old.
    int offp[MAXE], offr[MAXE];  // level arrays hold fewer than 2^31 elements (checked by the host)
    unsigned long long kp = 0, kr = 0;
    auto kpos = [&](int x, int y) {
        return (x < 0 ? 0 : (x >= nx ? 2 : 1)) + 3 * (y < 0 ? 0 : (y >= ny ? 2 : 1));
    };
#pragma unroll
    for (int i = 0; i < MAXE; i++) {
        const int e = min((int)threadIdx.x + 256 * i, NP - 1), er = e / W, ec = e - er * W;
        offp[i] = (int)at(gin, vin, x0 + ec - 2, y0 - 2 + er, 0);
        kp |= (unsigned long long)kpos(x0 + ec - 2, y0 - 2 + er) << (4 * i);
        const int f = min((int)threadIdx.x + 256 * i, NR - 1), fr = f / WR, fc = f - fr * WR;
        offr[i] = (int)at(gr, vr, x0 + fc - 1, y0 - 1 + fr, 0);
        kr |= (unsigned long long)kpos(x0 + fc - 1, y0 - 1 + fr) << (4 * i);
    }
    // IN == 2: coarse element under each staged fine cell.  Inside the box
    // every staged cell's owner lies in this box's z layer of boxes, whose
    // coarse boxes share one z range, so plane z adds (z >> 1) coarse planes.
    constexpr int NCO = IN == 2 ? MAXE : 1;
    int offc[NCO];  // the coarse level holds fewer than 2^31 elements (checked by the host)
    if constexpr (IN == 2) {
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = min((int)threadIdx.x + 256 * i, NP - 1), er = e / W, ec = e - er * W;
            offc[i] = (int)atc(x0 + ec - 2, y0 - 2 + er, 0);
        }
    }
    T pre[MAXE], prr[MAXE], prc[NCO];
    auto issue = [&](int z) {
        if constexpr (IN == 1) { return; }
        const bool inb = z >= 0 && z < nz;
        const int sd = z < 0 ? 0 : 9;
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            long o = XYp * z + offp[i];
            long oc = 0;
            if constexpr (IN == 2) { oc = XYc * (z >> 1) + offc[i]; }
            if (!inb) {
                const int k = sd + (int)((kp >> (4 * i)) & 15);
                o += s_dz[k];
                if constexpr (IN == 2) { oc += s_dc[k]; }
            }
            pre[i] = pin[o];
            if constexpr (IN == 2) { prc[i] = crs[oc]; }
        }
    };
    auto commit = [&](int z) {
        T* pl = slot(z);
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = threadIdx.x + 256 * i;
            T v;
            if constexpr (IN == 1) {
                v = 0;
            } else if constexpr (IN == 2) {
                const T zero = 0;
                v = pre[i] + (zero + prc[i]);
            } else {
                v = pre[i];
            }
            if (e < NP) { pl[e] = v; }
        }
    };
    auto issue_r = [&](int z) {
        const bool inb = z >= 0 && z < nz;
        const int sd = z < 0 ? 0 : 9;
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            long o = XYr * z + offr[i];
            if (!inb) { o += s_dr[sd + (int)((kr >> (4 * i)) & 15)]; }
            prr[i] = rhs[o];
        }
    };
    auto commit_r = [&](int z) {
        T* pl = rslot(z);
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = threadIdx.x + 256 * i;
            if (e < NR) { pl[e] = prr[i]; }
        }
    };
    // cell index q = tid + 256 i of a colour update <-> (row ya + q / hw, column
    // pair q % hw), walked without divisions from the split of tid and 256 by hw
    struct Walk {
        int hw, q0, r0, qs, rs;
    };
    auto mk_walk = [&](int hw) {
        const int t = (int)threadIdx.x;
        return Walk{hw, t / hw, t % hw, 256 / hw, 256 % hw};
    };
    const Walk wred = mk_walk((TWc + 2 + 1) / 2), wblk = mk_walk((TWc + 1) / 2);
    auto update = [&](int z, int ya, int yb, int xa, int xb, int colour, const Walk& wk) {
        T* zc_ = slot(z);
        const T* zm = slot(z - 1);
        const T* zp = slot(z + 1);
        const T* rz = rslot(z);
        const int hw = wk.hw;
        const int ncell = (yb - ya) * hw;
        int qq = wk.q0, rr = wk.r0;
        for (int q = threadIdx.x; q < ncell; q += 256) {
            const int yy = ya + qq;
            const int xx = xa + 2 * rr + ((xa + yy + z + colour + losum) & 1);
            qq += wk.qs;
            rr += wk.rs;
            if (rr >= hw) {
                rr -= hw;
                qq++;
            }
            if (xx >= xb) { continue; }
            const int li = (yy - (y0 - 2)) * W + (xx - x0 + 2);
            T s = 0;
            s = s + c.c[0] * zc_[li];
            s = s + c.c[1] * zc_[li + 1];
            s = s + c.c[2] * zc_[li - 1];
            s = s + c.c[3] * zc_[li + W];
            s = s + c.c[4] * zc_[li - W];
            s = s + c.c[5] * zp[li];
            s = s + c.c[6] * zm[li];
            s = s + rz[(yy - (y0 - 1)) * WR + (xx - x0 + 1)];
            zc_[li] = zc_[li] + lambda * s;
        }
    };
    // phi planes z0-2 .. z1+1 and rhs planes z0-1 .. z1 pass through registers
    // one step ahead of their use
    issue(z0 - 2);
    commit(z0 - 2);
    issue(z0 - 1);
    commit(z0 - 1);
    issue(z0);
    issue_r(z0 - 1);
    // OUT == 1: this thread's output cells (q = tid + 256 i < cnt) as box
    // offsets at plane 0 and LDS offsets; their pacc values of the plane
    // finished next are loaded one step ahead, clamped like the staging loads
    constexpr int NO = OUT == 1 ? MAXE : 1;
    int oo[NO], ol[NO];
    double acc_in[NO];
    const int cnt = (y1 - y0) * TWc;
    const long XYo = (long)(nx + 2 * gout.g[0]) * (ny + 2 * gout.g[1]);
    double* oacc = pacc + (long)b * gvol(L, gout);
    if constexpr (OUT == 1) {
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int q = min((int)threadIdx.x + 256 * i, cnt - 1), r = q / TWc, xx = q - r * TWc;
            oo[i] = (int)rel_index(L, gout, x0 + xx, y0 + r, 0);
            ol[i] = (r + 2) * W + xx + 2;
        }
    }
    auto issue_o = [&](int z) {
        if constexpr (OUT == 1) {
#pragma unroll
            for (int i = 0; i < MAXE; i++) { acc_in[i] = oacc[XYo * z + oo[i]]; }
        }
    };
    auto write_out = [&](int z) {
        const T* pl = slot(z);
        if constexpr (OUT == 1) {
#pragma unroll
            for (int i = 0; i < MAXE; i++) {
                if ((int)threadIdx.x + 256 * i < cnt) { oacc[XYo * z + oo[i]] = acc_in[i] + (double)pl[ol[i]]; }
            }
        } else {
            for (int e = threadIdx.x; e < cnt; e += 256) {
                const int r = e / TWc, xx = e - r * TWc;
                // non-temporal in FP64: 0.736 vs 0.756 ms for the 512^3 sweep
                // (FP32 measured 1 % slower with it)
                T* o = &ob[rel_index(L, gout, x0 + xx, y0 + r, z)];
                if constexpr (sizeof(T) == 8) {
                    __builtin_nontemporal_store(pl[(r + 2) * W + xx + 2], o);
                } else {
                    *o = pl[(r + 2) * W + xx + 2];
                }
            }
        }
    };
    // Memory operations of a step are all issued at its start (loads of the
    // planes ahead, then the write-out of finished plane k-1), so the wait for
    // the staged planes at the next step's start covers only operations a
    // whole step old: neither load nor store latency is exposed.
    for (int k = z0 - 2; k < z1; k++) {
        commit(k + 2);    // into the slot of plane k-2, last read before the previous barrier
        commit_r(k + 1);  // into the slot of rhs plane k-1, likewise
        if (k + 3 <= z1 + 1) { issue(k + 3); }
        if (k + 2 <= z1) { issue_r(k + 2); }
        if (k - 1 >= z0) { write_out(k - 1); }  // slot of plane k-1: only read during this step
        if (k >= z0) { issue_o(k); }           // OUT == 1: pacc of plane k, written next step
        __syncthreads();
        update(k + 1, y0 - 1, y1 + 1, x0 - 1, x1 + 1, 0, wred);
        __syncthreads();
        if (k >= z0) { update(k, y0, y1, x0, x1, 1, wblk); }
        __syncthreads();
    }
    write_out(z1 - 1);
}

// ------------------------------------------------------------ row-register sweep

// One full red-black sweep of a periodic level of boxes 64 cells wide that
// covers its domain, out of place, in one pass, the rows held in registers.
// A workgroup is one wave per box column of the domain (at most 8): lane l of
// wave w holds column 64 w + l, so the workgroup holds whole rows of the
// domain.  It owns rows [Y0, Y0 + RY) of one box row and planes [Z0, Z1) of
// one box layer; each thread keeps its column of rows [Y0 - 2, Y0 + RY + 2)
// of the planes in use in registers and walks z.  Step k: the red cells of
// plane k + 1 over the tile rows and one halo row on each side (the values
// the neighbouring tiles compute for those cells: same inputs, same order),
// then the black cells of plane k, which is written out.  x neighbours are
// the adjacent lanes (DPP wave shifts); at a box edge, the adjacent wave's
// edge lane, passed through LDS (the last wave wraps to the first): one
// barrier per step.  Addresses are uniform over a wave: per plane and array
// three row bases (the two rows above the tile, the tile, the two rows below,
// which may lie in the neighbouring boxes) are scalar, and rows add
// compile-time offsets (the ghost widths GP, GR, GO of phi in, rhs and phi
// out are template arguments).  Arithmetic, order and colouring are k_gsrb's;
// IN (0 phi, 1 zero guess, 2 phi + the constant prolongation of a coarse level
// with one ghost layer) and OUT as k_gsrb_fused.
template<typename T, int RY, int IN, int OUT, int NT, int GP, int GR, int GO, int PF>
__global__ __launch_bounds__(512) void k_gsrb_rows(LayoutDev L, const T* __restrict__ pin, T* __restrict__ pout,
                                                   const T* __restrict__ rhs, Lap<T> c, T lambda, int zc,
                                                   double* __restrict__ pacc, LayoutDev C,
                                                   const T* __restrict__ crs)
{
    constexpr int NR = RY + 4;  // phi rows held: Y0 - 2 .. Y0 + RY + 1
    constexpr int NQ = RY + 2;  // rhs rows: Y0 - 1 .. Y0 + RY
    constexpr long XP = 64 + 2 * GP, XR = 64 + 2 * GR, XO = 64 + 2 * GO;
    using TO = typename std::conditional<OUT == 1, double, T>::type;  // what the output array holds
    __shared__ T eo[2][8][2][NQ];   // old values of a plane's edge lanes (0, 63), rows Y0 - 1 .. Y0 + RY
    __shared__ T er[2][8][2][RY];   // red-updated edge values, rows Y0 .. Y0 + RY - 1
    __shared__ TO ef[2][8][2][GO == 1 ? RY : 1];  // final output values of the edge lanes (GO == 1, OUT == 0)
    // every argument the loops use, as scalars (a reference to a by-value
    // kernel argument would copy it to private memory)
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = L.nb[0], nb1 = L.nb[1], nb2 = L.nb[2];
    const int bs1 = L.bs[1], bs2 = L.bs[2], dom1 = L.domain[1];
    const int wl = w == 0 ? nw - 1 : w - 1, wr = w == nw - 1 ? 0 : w + 1;
    const int nyt = dom1 / RY;
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const int Y0 = (int)(t % nyt) * RY, Z0 = (int)(t / nyt) * zc, Z1 = Z0 + zc;
    const int by = Y0 / bs1, bz = Z0 / bs2, y0 = Y0 - by * bs1;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    // the two rows above the tile lie in the box above when the tile starts its
    // box, the two below in the box below when it ends it
    const int ytop = y0 == 0 ? -1 : 0, ybot = y0 + RY == bs1 ? 1 : 0;
    const int* table = L.table;
    auto box_at = [&](int ys, int zs) __attribute__((always_inline)) {
        const int yb = (by + ys + nb1) % nb1, zb = (bz + zs + nb2) % nb2;
        return table[w + nw * (yb + nb1 * zb)];
    };
    // boxes of the three row classes for the planes below / in / above the box layer
    const int tm = box_at(ytop, -1), t0 = box_at(ytop, 0), tp = box_at(ytop, 1);
    const int mm = box_at(0, -1), m0 = box_at(0, 0), mp = box_at(0, 1);
    const int bm = box_at(ybot, -1), b0 = box_at(ybot, 0), bp = box_at(ybot, 1);
    // row offsets of the classes within their box: row i sits at class base + X (i - 2)
    const int yt = y0 == 0 ? bs1 : y0, yb = ybot ? -RY : y0;
    // strides of the three arrays
    const long XYP = XP * (bs1 + 2 * GP), VP = XYP * (bs2 + 2 * GP), OP = GP + XP * GP + XYP * GP;
    const long XYR = XR * (bs1 + 2 * GR), VR = XYR * (bs2 + 2 * GR), OR = GR + XR * GR + XYR * GR;
    const long XYO = XO * (bs1 + 2 * GO), VO = XYO * (bs2 + 2 * GO), OO = GO + XO * GO + XYO * GO;
    // pointers to the first tile row (row i = 2) of each row class of plane Z:
    // row i of the class is at that pointer + X (i - 2); all uniform
#define PM_ROW_BASES(X, XY, V, O0, Z, PT, PM, PB)                                              \
    {                                                                                          \
        int it_ = t0, im_ = m0, ib_ = b0, zz_ = (Z) - bz * bs2;                                \
        if (zz_ < 0) {                                                                         \
            it_ = tm;                                                                          \
            im_ = mm;                                                                          \
            ib_ = bm;                                                                          \
            zz_ += bs2;                                                                        \
        } else if (zz_ >= bs2) {                                                               \
            it_ = tp;                                                                          \
            im_ = mp;                                                                          \
            ib_ = bp;                                                                          \
            zz_ -= bs2;                                                                        \
        }                                                                                      \
        const long zo_ = (O0) + (XY) * zz_;                                                    \
        PT = (long)it_ * (V) + zo_ + (X) * yt;                                                 \
        PM = (long)im_ * (V) + zo_ + (X) * y0;                                                 \
        PB = (long)ib_ * (V) + zo_ + (X) * yb;                                                 \
    }

    // IN == 2: the coarse level (ghost 1, boxes on its box grid).  Fine row i of
    // class (top / tile / bottom) over coarse row class base + CX ((i - 2) >> 1),
    // column lane / 2 of the coarse box under this wave's box; the class base of
    // plane Z adds CXY (Z wrapped / 2).  Every fine box's coarsened region lies
    // in one coarse box.
    long CX = 0, CXY = 0, ctm = 0, ct0 = 0, ctp = 0, cmm = 0, cm0 = 0, cmp = 0, cbm = 0, cb0 = 0, cbp = 0;
    const int dom2 = L.domain[2];
    if constexpr (IN == 2) {
        const int cbs0 = C.bs[0], cbs1 = C.bs[1], cbs2 = C.bs[2], cnb0 = C.nb[0], cnb1 = C.nb[1];
        const int* ctable = C.table;
        const long cX = cbs0 + 2, cXY = cX * (cbs1 + 2), cvol = cXY * (cbs2 + 2), co0 = 1 + cX + cXY;
        CX = cX;
        CXY = cXY;
        // coarse rows of the classes' first rows (row i = 2 position), as in the fine sweep
        const int yct = (((Y0 - 2 + dom1) % dom1) >> 1) + 1, ycm = Y0 >> 1, ycb = (((Y0 + RY) % dom1) >> 1) - RY / 2;
        auto cof = [&](int ys, int zs, int yc) {
            const int ybx = (by + ys + nb1) % nb1, zbx = (bz + zs + nb2) % nb2;
            const int cx = (w * 64 / 2) / cbs0, cy = (ybx * bs1 / 2) / cbs1, cz = (zbx * bs2 / 2) / cbs2;
            const int cid = ctable[cx + cnb0 * (cy + cnb1 * cz)];
            return (long)cid * cvol + co0 + (w * 32 - cx * cbs0) + cX * (long)(yc - cy * cbs1) - cXY * (long)(cz * cbs2);
        };
        ctm = cof(ytop, -1, yct), ct0 = cof(ytop, 0, yct), ctp = cof(ytop, 1, yct);
        cmm = cof(0, -1, ycm), cm0 = cof(0, 0, ycm), cmp = cof(0, 1, ycm);
        cbm = cof(ybot, -1, ycb), cb0 = cof(ybot, 0, ycb), cbp = cof(ybot, 1, ycb);
    }
    // planes are loaded PF steps ahead of their first use: a ring of 4 + PF slots
    constexpr int R = 4 + PF;
    T P[R][NR];  // phi planes by slot: plane Z0 - 2 + s + R m in slot s
    T Q[R][NQ];  // rhs planes, same slots
    auto loadP = [&](T (&d)[NR], int Z) {
        if constexpr (IN == 1) {
#pragma unroll
            for (int i = 0; i < NR; i++) { d[i] = (T)0; }
        } else {
            long ot, om, ob;
            PM_ROW_BASES(XP, XYP, VP, OP, Z, ot, om, ob)
#pragma unroll
            for (int i = 0; i < NR; i++) { d[i] = pin[(i < 2 ? ot : (i < RY + 2 ? om : ob)) + XP * (i - 2) + lane]; }
            if constexpr (IN == 2) {
                // plus the constant prolongation of the coarse correction, rounded as
                // k_prolong_const_box rounds it: the cells fineInterp would leave
                const int zw = Z < 0 ? Z + dom2 : (Z >= dom2 ? Z - dom2 : Z);
                const long zpart = CXY * (long)(zw >> 1);
                // uniform branches, not selects (a select chain of these became a
                // private-memory table)
                long ct = ct0, cm = cm0, cb = cb0;
                if (Z - bz * bs2 < 0) {
                    ct = ctm;
                    cm = cmm;
                    cb = cbm;
                } else if (Z - bz * bs2 >= bs2) {
                    ct = ctp;
                    cm = cmp;
                    cb = cbp;
                }
                ct += zpart;
                cm += zpart;
                cb += zpart;
                const T zero = 0;
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    const T v = crs[(i < 2 ? ct : (i < RY + 2 ? cm : cb)) + CX * ((i - 2) >> 1) + (lane >> 1)];
                    d[i] = d[i] + (zero + v);
                }
            }
        }
    };
    auto loadQ = [&](T (&d)[NQ], int Z) {
        long ot, om, ob;
        PM_ROW_BASES(XR, XYR, VR, OR, Z, ot, om, ob)
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            d[j] = rhs[(i < 2 ? ot : (i < RY + 2 ? om : ob)) + XR * (i - 2) + lane];
        }
    };
    const bool edge = lane == 0 || lane == 63;
    const int side = lane == 63 ? 1 : 0;
    auto pub_old = [&](const T (&d)[NR], int Z) {
        if (edge) {
#pragma unroll
            for (int j = 0; j < NQ; j++) { eo[Z & 1][w][side][j] = d[j + 1]; }
        }
    };
    auto pub_red = [&](const T (&d)[NR], int Z) {
        if (edge) {
#pragma unroll
            for (int r = 0; r < RY; r++) { er[Z & 1][w][side][r] = d[r + 2]; }
        }
    };
    // red (colour 0) where x + y + z is even; lane parity against the row's
    const bool lodd = lane & 1;
    auto is_red = [&](int i, int Z) __attribute__((always_inline)) { return lodd == (((Y0 - 2 + i + Z) & 1) != 0); };
    auto relax = [&](T cur, T left, T right, T up, T down, T zp, T zm, T f) __attribute__((always_inline)) {
        T s = 0;
        s = s + c0 * cur;
        s = s + c1 * right;
        s = s + c2 * left;
        s = s + c3 * up;
        s = s + c4 * down;
        s = s + c5 * zp;
        s = s + c6 * zm;
        s = s + f;
        return cur + lambda * s;
    };
    // GO == 1: after a plane is written, the x ghost cells of its rows receive
    // the periodic images (the neighbouring boxes' edge values, through LDS: what
    // an exchange would copy there), so that every 128-byte line of a row is
    // written whole and the partial-line writes at the row ends merge in L2.
    // Not for OUT == 1 (the FP64 accumulation): there it measured slower, 0.836
    // against 0.728 ms at 512^3
    auto ghost_cols = [&](int k) {
        if constexpr (GO == 1 && OUT == 0) {
            if (edge) {
                long ot, om, ob;
                PM_ROW_BASES(XO, XYO, VO, OO, k, ot, om, ob)
                (void)ot;
                (void)ob;
#pragma unroll
                for (int r = 0; r < RY; r++) {
                    const long o = om + XO * r + (side ? 64 : -1);
                    const TO v = side ? ef[k & 1][wr][0][r] : ef[k & 1][wl][1][r];
                    if constexpr (OUT == 1) {
                        pacc[o] = v;
                    } else {
                        pout[o] = v;
                    }
                }
            }
        }
    };
    // the step of plane k with slot S of plane k (k - (Z0 - 2) = S mod R)
    auto step = [&](auto sc, int k) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        constexpr int S0 = S, S1 = (S + 1) % R, S2 = (S + 2) % R, SM = (S + R - 1) % R;
        constexpr int LP = (S + 2 + PF) % R, LQ = (S + 1 + PF) % R;
        __syncthreads();
        if (k - 1 >= Z0) { ghost_cols(k - 1); }
        // the next planes (clamped to the chunk's last halo planes: a valid address)
        loadP(P[LP], min(k + 2 + PF, Z1 + 1));
        loadQ(Q[LQ], min(k + 1 + PF, Z1));
        // red cells of plane k + 1, rows Y0 - 1 .. Y0 + RY
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            const T cur = P[S1][i];
            const T le = eo[(k + 1) & 1][wl][1][j], re = eo[(k + 1) & 1][wr][0][j];
            const T nv = relax(cur, lane_shift<kShr1>(le, cur), lane_shift<kShl1>(re, cur), P[S1][i + 1],
                               P[S1][i - 1], P[S2][i], P[S0][i], Q[S1][j]);
            P[S1][i] = is_red(i, k + 1) ? nv : cur;
        }
        if (k >= Z0) {
            // black cells of plane k, rows Y0 .. Y0 + RY - 1, then plane k out
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + 2;
                const T cur = P[S0][i];
                const T le = er[k & 1][wl][1][r], re = er[k & 1][wr][0][r];
                const T nv = relax(cur, lane_shift<kShr1>(le, cur), lane_shift<kShl1>(re, cur), P[S0][i + 1],
                                   P[S0][i - 1], P[S1][i], P[SM][i], Q[S0][r + 1]);
                P[S0][i] = is_red(i, k) ? cur : nv;
            }
            // the tile rows of the output plane are the middle class
            long ot, om, ob;
            PM_ROW_BASES(XO, XYO, VO, OO, k, ot, om, ob)
            (void)ot;
            (void)ob;
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const long o = om + XO * r + lane;
                if constexpr (OUT == 1) {
                    pacc[o] = pacc[o] + (double)P[S0][r + 2];
                } else {
                    if constexpr (NT && GO == 0) {
                        __builtin_nontemporal_store(P[S0][r + 2], &pout[o]);
                    } else {
                        pout[o] = P[S0][r + 2];
                    }
                    if (GO == 1 && edge) { ef[k & 1][w][side][r] = P[S0][r + 2]; }
                }
            }
        }
        pub_red(P[S1], k + 1);
        pub_old(P[S2], k + 2);
    };
    // planes Z0 - 2 .. Z0 - 1 + PF (slots 0 .. PF + 1), rhs Z0 - 1 .. Z0 - 2 + PF
    // (slots 1 .. PF); the first step loads plane Z0 + PF and rhs Z0 - 1 + PF
    static_for<0, PF + 2>([&](auto sc) { loadP(P[decltype(sc)::value], min(Z0 - 2 + decltype(sc)::value, Z1 + 1)); });
    static_for<1, PF + 1>([&](auto sc) { loadQ(Q[decltype(sc)::value], min(Z0 - 2 + decltype(sc)::value, Z1)); });
    pub_old(P[1], Z0 - 1);
    int k = Z0 - 2;
    bool done = false;
    for (;;) {
        static_for<0, R>([&](auto sc) {
            if (!done) {
                step(sc, k);
                done = ++k >= Z1;
            }
        });
        if (done) { break; }
    }
    __syncthreads();
    ghost_cols(Z1 - 1);
#undef PM_ROW_BASES
}

// The one-pass sweep of k_gsrb_rows from phi with one ghost layer (IN 0), the
// zero initial guess (IN 1) or phi plus the constant prolongation of a coarse
// level with one ghost layer (IN 2) into a ghost-free scratch level, FP32, two
// cells per lane: a wave holds two neighbouring boxes along x, lane l columns
// 2 (l mod 32) (cell A) and 2 (l mod 32) + 1 (cell B) of box column 2 w + l / 32,
// so each load and store moves 8 bytes a lane, as the FP64 sweep's do, for the
// same instructions.  In a row of a plane one of A, B is red and the other
// black, which the plane's parity fixes: the steps are unrolled over two turns
// of the plane ring, so each row updates one named cell with no select.  A's
// left neighbour is the previous lane's B, B's right neighbour the next lane's
// A (DPP wave shifts; lanes 31 and 32 meet at the boundary between the wave's
// two boxes); lanes 0 and 63 take theirs from the neighbouring waves through
// LDS.  Arithmetic, order and colouring are k_gsrb's; z chunks start on even
// planes.
template<int RY, int IN, int NT, int PF>
__global__ __launch_bounds__(256) void k_gsrb_rows2(LayoutDev L, const float* __restrict__ pin,
                                                    float* __restrict__ pout, const float* __restrict__ rhs,
                                                    Lap<float> c, float lambda, int zc, LayoutDev C,
                                                    const float* __restrict__ crs)
{
    using T = float;
    // a pair of cells; rows of phi with one ghost layer put pairs on 4-byte boundaries
    typedef float V __attribute__((ext_vector_type(2), aligned(4)));
    constexpr int NR = RY + 4;          // phi rows held: Y0 - 2 .. Y0 + RY + 1
    constexpr int NQ = RY + 2;          // rhs rows: Y0 - 1 .. Y0 + RY
    constexpr long X = 64, XP = 66;     // row pitch of the scratch and rhs / of phi
    constexpr int R = 4 + PF;           // plane ring: k - 1 .. k + 2 + PF at step k (PF planes loaded ahead)
    __shared__ T eo[2][4][2][NQ];  // old values of lane 0's A (side 0) and lane 63's B (side 1)
    __shared__ T er[2][4][2][RY];  // their red-updated values, tile rows
    const int lane = threadIdx.x & 63, q = lane & 31, hb = lane >> 5;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nb0 = L.nb[0], nw = nb0 >> 1, nb1 = L.nb[1], nb2 = L.nb[2];
    const int bs1 = L.bs[1], bs2 = L.bs[2], dom1 = L.domain[1];
    const int wl = w == 0 ? nw - 1 : w - 1, wr = w == nw - 1 ? 0 : w + 1;
    const int bxl = 2 * w + hb;  // this lane's box column
    const int nyt = dom1 / RY;
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const int Y0 = (int)(t % nyt) * RY, Z0 = (int)(t / nyt) * zc, Z1 = Z0 + zc;
    const int by = Y0 / bs1, bz = Z0 / bs2, y0 = Y0 - by * bs1;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    const int ytop = y0 == 0 ? -1 : 0, ybot = y0 + RY == bs1 ? 1 : 0;
    const int* table = L.table;
    auto box_at = [&](int ys, int zs) __attribute__((always_inline)) {
        const int yb = (by + ys + nb1) % nb1, zb = (bz + zs + nb2) % nb2;
        return table[bxl + nb0 * (yb + nb1 * zb)];
    };
    // this lane's boxes of the three row classes for the planes below / in / above the box layer
    const int tm = box_at(ytop, -1), t0 = box_at(ytop, 0), tp = box_at(ytop, 1);
    const int mm = box_at(0, -1), m0 = box_at(0, 0), mp = box_at(0, 1);
    const int bm = box_at(ybot, -1), b0 = box_at(ybot, 0), bp = box_at(ybot, 1);
    const int yt = y0 == 0 ? bs1 : y0, yb = ybot ? -RY : y0;
    const long XYP = XP * (bs1 + 2), VP = XYP * (bs2 + 2), OP = 1 + XP + XYP;
    const long XY = X * bs1, VV = XY * bs2;
    // element of column 2q of the first tile row (row i = 2) of each row class of
    // plane Z in the lane's own boxes (the two boxes of a wave need not be stored
    // side by side; with uniform addresses for boxes that are, the sweep measured
    // 3 % slower); row i of the class adds X (i - 2)
#define PM_ROW_BASES2(XX, XXY, VB, O0, Z, PT, PM, PB)                                          \
    {                                                                                          \
        int it_ = t0, im_ = m0, ib_ = b0, zz_ = (Z) - bz * bs2;                                \
        if (zz_ < 0) {                                                                         \
            it_ = tm;                                                                          \
            im_ = mm;                                                                          \
            ib_ = bm;                                                                          \
            zz_ += bs2;                                                                        \
        } else if (zz_ >= bs2) {                                                               \
            it_ = tp;                                                                          \
            im_ = mp;                                                                          \
            ib_ = bp;                                                                          \
            zz_ -= bs2;                                                                        \
        }                                                                                      \
        const long zo_ = (O0) + (XXY) * zz_ + 2 * q;                                           \
        PT = (long)it_ * (VB) + zo_ + (XX) * yt;                                               \
        PM = (long)im_ * (VB) + zo_ + (XX) * y0;                                               \
        PB = (long)ib_ * (VB) + zo_ + (XX) * yb;                                               \
    }
    // IN == 2: coarse rows under the classes' first rows, as k_gsrb_rows, for
    // this lane's box column (coarse column q of its 32-column block); 32-bit
    // (coarse levels hold fewer than 2^31 elements)
    long CX = 0, CXY = 0;
    int ctm = 0, ct0 = 0, ctp = 0, cmm = 0, cm0 = 0, cmp = 0, cbm = 0, cb0 = 0, cbp = 0;
    const int dom2 = L.domain[2];
    if constexpr (IN == 2) {
        const int cbs0 = C.bs[0], cbs1 = C.bs[1], cbs2 = C.bs[2], cnb0 = C.nb[0], cnb1 = C.nb[1];
        const int* ctable = C.table;
        const long cX = cbs0 + 2, cXY = cX * (cbs1 + 2), cvol = cXY * (cbs2 + 2), co0 = 1 + cX + cXY;
        CX = cX;
        CXY = cXY;
        const int yct = (((Y0 - 2 + dom1) % dom1) >> 1) + 1, ycm = Y0 >> 1, ycb = (((Y0 + RY) % dom1) >> 1) - RY / 2;
        auto cof = [&](int ys, int zs, int yc) {
            const int ybx = (by + ys + nb1) % nb1, zbx = (bz + zs + nb2) % nb2;
            const int cx = (bxl * 64 / 2) / cbs0, cy = (ybx * bs1 / 2) / cbs1, cz = (zbx * bs2 / 2) / cbs2;
            const int cid = ctable[cx + cnb0 * (cy + cnb1 * cz)];
            return (int)((long)cid * cvol + co0 + (bxl * 32 - cx * cbs0 + q) + cX * (long)(yc - cy * cbs1) -
                         cXY * (long)(cz * cbs2));
        };
        ctm = cof(ytop, -1, yct), ct0 = cof(ytop, 0, yct), ctp = cof(ytop, 1, yct);
        cmm = cof(0, -1, ycm), cm0 = cof(0, 0, ycm), cmp = cof(0, 1, ycm);
        cbm = cof(ybot, -1, ycb), cb0 = cof(ybot, 0, ycb), cbp = cof(ybot, 1, ycb);
    }
    T A[R][NR], B[R][NR];    // phi planes by slot (plane Z0 - 2 + s + R m in slot s), cells A / B
    T QA[R][NQ], QB[R][NQ];  // rhs planes, same slots
    auto loadP = [&](T (&a)[NR], T (&b)[NR], int Z) {
        if constexpr (IN == 1) {
#pragma unroll
            for (int i = 0; i < NR; i++) {
                a[i] = (T)0;
                b[i] = (T)0;
            }
        } else {
            long ot, om, ob;
            PM_ROW_BASES2(XP, XYP, VP, OP, Z, ot, om, ob)
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const V v = *reinterpret_cast<const V*>(pin + (i < 2 ? ot : (i < RY + 2 ? om : ob)) + XP * (i - 2));
                a[i] = v.x;
                b[i] = v.y;
            }
            if constexpr (IN == 2) {
                // plus the constant prolongation, rounded as k_prolong_const_box
                // rounds it; cells A and B lie over the same coarse cell
                const int zw = Z < 0 ? Z + dom2 : (Z >= dom2 ? Z - dom2 : Z);
                const long zpart = CXY * (long)(zw >> 1);
                long ct = ct0, cm = cm0, cb = cb0;
                if (Z - bz * bs2 < 0) {
                    ct = ctm;
                    cm = cmm;
                    cb = cbm;
                } else if (Z - bz * bs2 >= bs2) {
                    ct = ctp;
                    cm = cmp;
                    cb = cbp;
                }
                ct += zpart;
                cm += zpart;
                cb += zpart;
                const T zero = 0;
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    const T v = crs[(i < 2 ? ct : (i < RY + 2 ? cm : cb)) + CX * ((i - 2) >> 1)];
                    a[i] = a[i] + (zero + v);
                    b[i] = b[i] + (zero + v);
                }
            }
        }
    };
    auto loadQ = [&](T (&a)[NQ], T (&b)[NQ], int Z) {
        long ot, om, ob;
        PM_ROW_BASES2(X, XY, VV, 0, Z, ot, om, ob)
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            const V v = *reinterpret_cast<const V*>(rhs + (i < 2 ? ot : (i < RY + 2 ? om : ob)) + X * (i - 2));
            a[j] = v.x;
            b[j] = v.y;
        }
    };
    auto pub_old = [&](const T (&a)[NR], const T (&b)[NR], int Z) {
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NQ; j++) { eo[Z & 1][w][0][j] = a[j + 1]; }
        } else if (lane == 63) {
#pragma unroll
            for (int j = 0; j < NQ; j++) { eo[Z & 1][w][1][j] = b[j + 1]; }
        }
    };
    auto pub_red = [&](const T (&a)[NR], const T (&b)[NR], int Z) {
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < RY; r++) { er[Z & 1][w][0][r] = a[r + 2]; }
        } else if (lane == 63) {
#pragma unroll
            for (int r = 0; r < RY; r++) { er[Z & 1][w][1][r] = b[r + 2]; }
        }
    };
    auto relax = [&](T cur, T left, T right, T up, T down, T zp, T zm, T f) __attribute__((always_inline)) {
        T s = 0;
        s = s + c0 * cur;
        s = s + c1 * right;
        s = s + c2 * left;
        s = s + c3 * up;
        s = s + c4 * down;
        s = s + c5 * zp;
        s = s + c6 * zm;
        s = s + f;
        return cur + lambda * s;
    };
    // row i of a plane: relax its A cell (left neighbour: the previous lane's B,
    // lane 0 the left wave's lane 63 B, e) or its B cell (right neighbour: the
    // next lane's A, lane 63 the right wave's lane 0 A, e)
    auto cellA = [&](T (&a)[NR], const T (&b)[NR], const T (&ap)[NR], const T (&am)[NR], T f, T e, int i)
        __attribute__((always_inline)) {
        a[i] = relax(a[i], lane_shift<kShr1>(e, b[i]), b[i], a[i + 1], a[i - 1], ap[i], am[i], f);
    };
    auto cellB = [&](const T (&a)[NR], T (&b)[NR], const T (&bp)[NR], const T (&bm)[NR], T f, T e, int i)
        __attribute__((always_inline)) {
        b[i] = relax(b[i], a[i], lane_shift<kShl1>(e, a[i]), b[i + 1], b[i - 1], bp[i], bm[i], f);
    };
    // step k with slot S of plane k and ZP the parity of plane k + 1: red cells of
    // plane k + 1 (rows Y0 - 1 .. Y0 + RY), black cells of plane k, plane k out.
    // The red cell of row i of plane k + 1 is A when i + ZP is even (x of A and Y0
    // even); so is the black cell of row i of plane k.
    auto step = [&](auto sc, int k) __attribute__((always_inline)) {
        constexpr int U = decltype(sc)::value;  // step within the two ring turns
        constexpr int S = U % R, ZP = (U + 1) & 1;
        constexpr int S0 = S, S1 = (S + 1) % R, S2 = (S + 2) % R, SM = (S + R - 1) % R;
        constexpr int LP = (S + 2 + PF) % R, LQ = (S + 1 + PF) % R;
        __syncthreads();
        loadP(A[LP], B[LP], min(k + 2 + PF, Z1 + 1));
        loadQ(QA[LQ], QB[LQ], min(k + 1 + PF, Z1));
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            if (((i + ZP) & 1) == 0) {
                cellA(A[S1], B[S1], A[S2], A[S0], QA[S1][j], eo[(k + 1) & 1][wl][1][j], i);
            } else {
                cellB(A[S1], B[S1], B[S2], B[S0], QB[S1][j], eo[(k + 1) & 1][wr][0][j], i);
            }
        }
        if (k >= Z0) {
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + 2;
                if (((i + ZP) & 1) == 0) {
                    cellA(A[S0], B[S0], A[S1], A[SM], QA[S0][r + 1], er[k & 1][wl][1][r], i);
                } else {
                    cellB(A[S0], B[S0], B[S1], B[SM], QB[S0][r + 1], er[k & 1][wr][0][r], i);
                }
            }
            long ot, om, ob;
            PM_ROW_BASES2(X, XY, VV, 0, k, ot, om, ob)
            (void)ot;
            (void)ob;
#pragma unroll
            for (int r = 0; r < RY; r++) {
                V v;
                v.x = A[S0][r + 2];
                v.y = B[S0][r + 2];
                V* o = reinterpret_cast<V*>(pout + om + X * r);
                if constexpr (NT) {
                    __builtin_nontemporal_store(v, o);
                } else {
                    *o = v;
                }
            }
        }
        pub_red(A[S1], B[S1], k + 1);
        pub_old(A[S2], B[S2], k + 2);
    };
    // planes Z0 - 2 .. Z0 - 1 + PF (slots 0 .. PF + 1), rhs Z0 - 1 .. Z0 - 2 + PF
    // (slots 1 .. PF); the first step loads plane Z0 + PF and rhs Z0 - 1 + PF
    static_for<0, PF + 2>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        loadP(A[s], B[s], min(Z0 - 2 + s, Z1 + 1));
    });
    static_for<1, PF + 1>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        loadQ(QA[s], QB[s], min(Z0 - 2 + s, Z1));
    });
    pub_old(A[1], B[1], Z0 - 1);
    int k = Z0 - 2;
    bool done = false;
    for (;;) {
        static_for<0, 2 * R>([&](auto sc) {
            if (!done) {
                step(sc, k);
                done = ++k >= Z1;
            }
        });
        if (done) { break; }
    }
#undef PM_ROW_BASES2
}

// The second pre-smoothing sweep of a V-cycle level with the residual
// restriction that follows it (coarseResidual: r = f - L phi, averaged onto the
// coarse level) in the same pass: k_gsrb_rows for input scratch (no ghosts) ->
// phi (one ghost layer), plus, one plane behind the sweep, the residual of the
// swept phi and its 2 x 2 x 2 averages written into the coarse right-hand side,
// with k_residual_nb's arithmetic and order.  The residual of the tile's rows
// needs the swept values one row above and below the tile and one plane beyond
// the z chunk at each end, which neighbouring tiles own: the tile recomputes
// them (red cells of rows Y0 - 2 .. Y0 + RY + 1, black cells of rows Y0 - 1 ..
// Y0 + RY, planes Z0 - 1 .. Z1), so each thread holds RY + 6 rows and the ring
// six planes.  Rows and planes outside the tile are never written.
template<typename T, int RY>
__global__ __launch_bounds__(512) void k_gsrb_rows_res(LayoutDev L, const T* __restrict__ pin, T* __restrict__ pout,
                                                       const T* __restrict__ rhs, Lap<T> c, T lambda, int zc,
                                                       Lap<T> cr, T avg, LayoutDev C, G3 gc, T* __restrict__ resc)
{
    constexpr int H = 3;                // halo rows held on each side
    constexpr int NR = RY + 2 * H;      // phi rows: Y0 - 3 .. Y0 + RY + 2
    constexpr int NQ = NR - 2;          // rhs rows and red rows: Y0 - 2 .. Y0 + RY + 1
    constexpr int NB = RY + 2;          // black rows: Y0 - 1 .. Y0 + RY
    constexpr long X = 64, XO = 66;     // rows of the ghost-free arrays / of phi
    constexpr int R = 6;                // phi plane ring: k - 2 .. k + 3 at step k
    constexpr int RQ = 3;               // rhs plane ring: k .. k + 2 at step k
    __shared__ T eo[2][8][2][NQ];  // old values of a plane's edge lanes (0, 63), red rows
    __shared__ T er[2][8][2][NB];  // red-updated edge values, black rows
    __shared__ T ef[2][8][2][RY];  // final edge values, tile rows
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = L.nb[0], nb1 = L.nb[1], nb2 = L.nb[2];
    const int bs1 = L.bs[1], bs2 = L.bs[2], dom1 = L.domain[1];
    const int wl = w == 0 ? nw - 1 : w - 1, wr = w == nw - 1 ? 0 : w + 1;
    const int nyt = dom1 / RY;
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const int Y0 = (int)(t % nyt) * RY, Z0 = (int)(t / nyt) * zc, Z1 = Z0 + zc;
    const int by = Y0 / bs1, bz = Z0 / bs2, y0 = Y0 - by * bs1;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    const T r0 = cr.c[0], r1 = cr.c[1], r2 = cr.c[2], r3 = cr.c[3], r4 = cr.c[4], r5 = cr.c[5], r6 = cr.c[6];
    // the three rows above the tile lie in the box above when the tile starts its
    // box, the three below in the box below when it ends it (box heights are
    // multiples of RY >= 4)
    const int ytop = y0 == 0 ? -1 : 0, ybot = y0 + RY == bs1 ? 1 : 0;
    const int* table = L.table;
    auto box_at = [&](int ys, int zs) __attribute__((always_inline)) {
        const int yb = (by + ys + nb1) % nb1, zb = (bz + zs + nb2) % nb2;
        return table[w + nw * (yb + nb1 * zb)];
    };
    const int tm = box_at(ytop, -1), t0 = box_at(ytop, 0), tp = box_at(ytop, 1);
    const int mm = box_at(0, -1), m0 = box_at(0, 0), mp = box_at(0, 1);
    const int bm = box_at(ybot, -1), b0 = box_at(ybot, 0), bp = box_at(ybot, 1);
    // row i (phi) sits at its class base + X (i - H)
    const int yt = y0 == 0 ? bs1 : y0, yb = ybot ? -RY : y0;
    const long XY = X * bs1, V = XY * bs2;
    const long XYO = XO * (bs1 + 2), VO = XYO * (bs2 + 2), OO = 1 + XO + XYO;
#define PM_ROW_BASES(XX, XXY, VV, O0, Z, PT, PM, PB)                                           \
    {                                                                                          \
        int it_ = t0, im_ = m0, ib_ = b0, zz_ = (Z) - bz * bs2;                                \
        if (zz_ < 0) {                                                                         \
            it_ = tm;                                                                          \
            im_ = mm;                                                                          \
            ib_ = bm;                                                                          \
            zz_ += bs2;                                                                        \
        } else if (zz_ >= bs2) {                                                               \
            it_ = tp;                                                                          \
            im_ = mp;                                                                          \
            ib_ = bp;                                                                          \
            zz_ -= bs2;                                                                        \
        }                                                                                      \
        const long zo_ = (O0) + (XXY) * zz_;                                                   \
        PT = (long)it_ * (VV) + zo_ + (XX) * yt;                                               \
        PM = (long)im_ * (VV) + zo_ + (XX) * y0;                                               \
        PB = (long)ib_ * (VV) + zo_ + (XX) * yb;                                               \
    }
    T P[R][NR];   // phi planes by slot: plane Z0 - 3 + s + R m in slot s
    T Q[RQ][NQ];  // rhs planes: plane Z0 - 3 + s + RQ m in slot s
    T F[2][RY];   // the tile rows of rhs plane k, kept for the residual at step k + 1 (slot k - Z0 + 3 mod 2)
    auto loadP = [&](T (&d)[NR], int Z) {
        long ot, om, ob;
        PM_ROW_BASES(X, XY, V, 0, Z, ot, om, ob)
#pragma unroll
        for (int i = 0; i < NR; i++) { d[i] = pin[(i < H ? ot : (i < RY + H ? om : ob)) + X * (i - H) + lane]; }
    };
    auto loadQ = [&](T (&d)[NQ], int Z) {
        long ot, om, ob;
        PM_ROW_BASES(X, XY, V, 0, Z, ot, om, ob)
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            d[j] = rhs[(i < H ? ot : (i < RY + H ? om : ob)) + X * (i - H) + lane];
        }
    };
    const bool edge = lane == 0 || lane == 63;
    const int side = lane == 63 ? 1 : 0;
    auto pub_old = [&](const T (&d)[NR], int Z) {
        if (edge) {
#pragma unroll
            for (int j = 0; j < NQ; j++) { eo[Z & 1][w][side][j] = d[j + 1]; }
        }
    };
    auto pub_red = [&](const T (&d)[NR], int Z) {
        if (edge) {
#pragma unroll
            for (int b = 0; b < NB; b++) { er[Z & 1][w][side][b] = d[b + H - 1]; }
        }
    };
    const bool lodd = lane & 1;
    auto is_red = [&](int i, int Z) __attribute__((always_inline)) { return lodd == (((Y0 - H + i + Z) & 1) != 0); };
    auto relax = [&](T cur, T left, T right, T up, T down, T zp, T zm, T f) __attribute__((always_inline)) {
        T s = 0;
        s = s + c0 * cur;
        s = s + c1 * right;
        s = s + c2 * left;
        s = s + c3 * up;
        s = s + c4 * down;
        s = s + c5 * zp;
        s = s + c6 * zm;
        s = s + f;
        return cur + lambda * s;
    };
    // the x ghost cells of plane k's output rows: the periodic images (as k_gsrb_rows)
    auto ghost_cols = [&](int k) {
        if (edge) {
            long ot, om, ob;
            PM_ROW_BASES(XO, XYO, VO, OO, k, ot, om, ob)
            (void)ot;
            (void)ob;
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const long o = om + XO * r + (side ? 64 : -1);
                pout[o] = side ? ef[k & 1][wr][0][r] : ef[k & 1][wl][1][r];
            }
        }
    };
    // the coarse level C (ghost gc): the tile's coarsened cells lie in one
    // coarse box (box-aligned levels); the element of coarse row Y0 / 2 and
    // plane 0 of the box's frame under this lane's even column, then rows add
    // CX, planes CXY
    long cbase, CX, CXY;
    {
        const int cbs0 = C.bs[0], cbs1 = C.bs[1], cbs2 = C.bs[2];
        const int cx = (w * 32) / cbs0, cy = (Y0 >> 1) / cbs1, cz = (Z0 >> 1) / cbs2;
        const int cid = C.table[cx + C.nb[0] * (cy + C.nb[1] * cz)];
        const int g0 = gc.g[0], g1 = gc.g[1], g2 = gc.g[2];
        CX = cbs0 + 2 * g0;
        CXY = CX * (cbs1 + 2 * g1);
        cbase = (long)cid * CXY * (cbs2 + 2 * g2) + g0 + (w * 32 - cx * cbs0 + (lane >> 1)) +
                CX * (long)(g1 + (Y0 >> 1) - cy * cbs1) + CXY * (long)(g2 - cz * cbs2);
    }
    T acc[RY / 2];
#pragma unroll
    for (int q = 0; q < RY / 2; q++) { acc[q] = (T)0; }
    // step k, slot S of plane k: red cells of plane k + 1, black cells of plane k
    // (written when it is a tile plane), the residual of plane k - 1
    auto step = [&](auto sc, int k) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        constexpr int S0 = S, S1 = (S + 1) % R, S2 = (S + 2) % R, S3 = (S + 3) % R;
        constexpr int SM = (S + R - 1) % R, SMM = (S + R - 2) % R;
        constexpr int Q0 = S % RQ, Q1 = (S + 1) % RQ, Q2 = (S + 2) % RQ, F0 = S & 1, FM = (S + 1) & 1;
        __syncthreads();
        if (k - 1 >= Z0 && k - 1 < Z1) { ghost_cols(k - 1); }
        // the next planes (clamped to the chunk's last halo planes: a valid address)
        loadP(P[S3], min(k + 3, Z1 + 2));
        loadQ(Q[Q2], min(k + 2, Z1 + 1));  // into the slot of plane k - 1 (its tile rows kept in F)
        // red cells of plane k + 1, rows Y0 - 2 .. Y0 + RY + 1
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int i = j + 1;
            const T cur = P[S1][i];
            const T le = eo[(k + 1) & 1][wl][1][j], re = eo[(k + 1) & 1][wr][0][j];
            const T nv = relax(cur, lane_shift<kShr1>(le, cur), lane_shift<kShl1>(re, cur), P[S1][i + 1],
                               P[S1][i - 1], P[S2][i], P[S0][i], Q[Q1][j]);
            P[S1][i] = is_red(i, k + 1) ? nv : cur;
        }
        if (k >= Z0 - 1) {
            // black cells of plane k, rows Y0 - 1 .. Y0 + RY
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int i = b + H - 1;
                const T cur = P[S0][i];
                const T le = er[k & 1][wl][1][b], re = er[k & 1][wr][0][b];
                const T nv = relax(cur, lane_shift<kShr1>(le, cur), lane_shift<kShl1>(re, cur), P[S0][i + 1],
                                   P[S0][i - 1], P[S1][i], P[SM][i], Q[Q0][i - 1]);
                P[S0][i] = is_red(i, k) ? cur : nv;
            }
#pragma unroll
            for (int r = 0; r < RY; r++) { F[F0][r] = Q[Q0][r + 2]; }
            if (k >= Z0 && k < Z1) {
                long ot, om, ob;
                PM_ROW_BASES(XO, XYO, VO, OO, k, ot, om, ob)
                (void)ot;
                (void)ob;
#pragma unroll
                for (int r = 0; r < RY; r++) {
                    pout[om + XO * r + lane] = P[S0][r + H];
                    if (edge) { ef[k & 1][w][side][r] = P[S0][r + H]; }
                }
            }
        }
        if (k - 1 >= Z0 && k - 1 < Z1) {
            // residual of plane kr = k - 1 (k_residual_nb's order), then its 2 x 2
            // averages onto the sums of plane pair kr & ~1
            const int kr = k - 1;
            T res[RY];
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + H;
                const T cur = P[SM][i];
                const T le = ef[kr & 1][wl][1][r], re = ef[kr & 1][wr][0][r];
                T sv = 0;
                sv = sv + F[FM][r];
                sv = sv + r0 * cur;
                sv = sv + r1 * lane_shift<kShl1>(re, cur);
                sv = sv + r2 * lane_shift<kShr1>(le, cur);
                sv = sv + r3 * P[SM][i + 1];
                sv = sv + r4 * P[SM][i - 1];
                sv = sv + r5 * P[S0][i];
                sv = sv + r6 * P[SMM][i];
                res[r] = sv;
            }
            const bool first = (kr & 1) == 0;
#pragma unroll
            for (int q = 0; q < RY / 2; q++) {
                // the right neighbour's value (lane 63's, which no lane writes, its own)
                const T a0 = res[2 * q], a1 = lane_shift<kShl1>(res[2 * q], res[2 * q]);
                const T b0v = res[2 * q + 1], b1v = lane_shift<kShl1>(res[2 * q + 1], res[2 * q + 1]);
                T tq = first ? (T)0 : acc[q];
                tq = tq + avg * a0;
                tq = tq + avg * a1;
                tq = tq + avg * b0v;
                tq = tq + avg * b1v;
                acc[q] = tq;
            }
            if (!first && !lodd) {
#pragma unroll
                for (int q = 0; q < RY / 2; q++) {
                    resc[cbase + CX * q + CXY * (kr >> 1)] = acc[q];
                }
            }
        }
        pub_red(P[S1], k + 1);
        pub_old(P[S2], k + 2);
    };
    // planes Z0 - 3 .. Z0 - 1 (slots 0..2), rhs Z0 - 2 (slot 1); the first step
    // loads plane Z0 and rhs Z0 - 1
    loadP(P[0], Z0 - 3);
    loadP(P[1], Z0 - 2);
    loadP(P[2], Z0 - 1);
    loadQ(Q[1], Z0 - 2);
    pub_old(P[1], Z0 - 2);
    int k = Z0 - 3;
    bool done = false;
    for (;;) {
        static_for<0, R>([&](auto sc) {
            if (!done) {
                step(sc, k);
                done = ++k > Z1;
            }
        });
        if (done) { break; }
    }
#undef PM_ROW_BASES
}

// ------------------------------------------------------------ small levels
// All n GSRB sweeps of a one-box periodic level of at most GS_MAX cells in one
// workgroup: phi lives in LDS for the whole relaxation (periodic neighbours by
// index), each thread keeps the right-hand side of its own cells in registers,
// and each colour is one pass with a barrier.  The arithmetic, the order and
// the colouring are k_gsrb's, so the result has the bits of n exchange / red /
// exchange / black sweeps.  IN 1: the input is the zero guess (phi not read);
// IN 2: phi + the constant prolongation of the one-box coarse level crs,
// rounded as k_prolong_const_box rounds it.  Tiny levels are latency-bound:
// one launch instead of n (or 2n) saves about 15 us per launch.
constexpr int GS_MAX = 4096, GS_PER = GS_MAX / 256;

template<typename T, int IN>
__global__ __launch_bounds__(256) void k_gsrb_small(LayoutDev L, G3 gp, G3 gr, T* __restrict__ phi,
                                                    const T* __restrict__ rhs, Lap<T> c, T lambda, int nsw,
                                                    LayoutDev C, G3 gc, const T* __restrict__ crs)
{
    __shared__ T u[GS_MAX];
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2], N = nx * ny * nz;
    const int* lo = L.lows;
    const int losum = lo[0] + lo[1] + lo[2];
    T f[GS_PER];
#pragma unroll
    for (int j = 0; j < GS_PER; j++) {
        const int i = threadIdx.x + 256 * j;
        if (i < N) {
            const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            f[j] = rhs[rel_index(L, gr, x, y, z)];
            T v = 0;
            if constexpr (IN != 1) { v = phi[rel_index(L, gp, x, y, z)]; }
            if constexpr (IN == 2) {
                const T zero = 0;
                v = v + (zero + crs[rel_index(C, gc, x >> 1, y >> 1, z >> 1)]);
            }
            u[i] = v;
        }
    }
    for (int sw = 0; sw < nsw; sw++) {
        for (int colour = 0; colour < 2; colour++) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < GS_PER; j++) {
                const int i = threadIdx.x + 256 * j;
                if (i >= N) { continue; }
                const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
                if (((x + y + z + losum + colour) & 1) != 0) { continue; }
                const int xp = x + 1 == nx ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
                const int yp = y + 1 == ny ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
                const int zp = z + 1 == nz ? 0 : z + 1, zm = z == 0 ? nz - 1 : z - 1;
                const int row = nx * (y + ny * z);
                T s = 0;
                s = s + c.c[0] * u[i];
                s = s + c.c[1] * u[xp + row];
                s = s + c.c[2] * u[xm + row];
                s = s + c.c[3] * u[x + nx * (yp + ny * z)];
                s = s + c.c[4] * u[x + nx * (ym + ny * z)];
                s = s + c.c[5] * u[x + nx * (y + ny * zp)];
                s = s + c.c[6] * u[x + nx * (y + ny * zm)];
                s = s + f[j];
                u[i] = u[i] + lambda * s;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += 256) {
        const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        phi[rel_index(L, gp, x, y, z)] = u[i];
    }
}

// ------------------------------------------------------------ Jacobi
template<typename T, int DIMN>
__global__ __launch_bounds__(256) void k_jacobi_tmp(LayoutDev L, G3 gp, G3 gr, const T* __restrict__ phi,
                                                    const T* __restrict__ rhs, T* __restrict__ tmp, Lap<T> c,
                                                    T dcoef)
{
    constexpr int NC = 2 * DIMN + 1;
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const T* ph = phi + (long)b * gvol(L, gp);
    const T* rh = rhs + (long)b * gvol(L, gr);
    T* tm = tmp + (long)b * n;
    const long X = L.bs[0] + 2 * gp.g[0], XY = X * (L.bs[1] + 2 * gp.g[1]);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gp, x, y, z);
        T s = 0;
#pragma unroll
        for (int j = 0; j < NC; j++) {
            s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * X + lapo(j, 2) * XY];
        }
        s = s + dcoef * rh[rel_index(L, gr, x, y, z)];
        tm[t] = s;
    }
}

template<typename T>
__global__ __launch_bounds__(256) void k_jacobi_add(LayoutDev L, G3 gp, T* __restrict__ phi, const T* __restrict__ tmp)
{
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    T* ph = phi + (long)b * gvol(L, gp);
    const T* tm = tmp + (long)b * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gp, x, y, z);
        ph[i] = ph[i] + tm[t];
    }
}

// ------------------------------------------------------------ residual norm
template<typename T, int DIMN>
__global__ __launch_bounds__(256) void k_resnorm(LayoutDev L, G3 gp, G3 gr, const T* __restrict__ phi,
                                                 const T* __restrict__ rhs, Lap<T> c, T* __restrict__ partials)
{
    constexpr int NC = 2 * DIMN + 1;
    __shared__ T red[16];
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const T* ph = phi + (long)b * gvol(L, gp);
    const T* rh = rhs + (long)b * gvol(L, gr);
    const long X = L.bs[0] + 2 * gp.g[0], XY = X * (L.bs[1] + 2 * gp.g[1]);
    T m = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gp, x, y, z);
        T s = 0;
        s = s - rh[rel_index(L, gr, x, y, z)];
#pragma unroll
        for (int j = 0; j < NC; j++) {
            s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * X + lapo(j, 2) * XY];
        }
        const T a = s > 0 ? s : -s;
        m = m < a ? a : m;
    }
    m = block_reduce(PM_RED_ABS, m, red);
    if (threadIdx.x == 0) { partials[(long)b * gridDim.x + blockIdx.x] = m; }
}

// ------------------------------------------------------------ restriction
template<typename T, int DIMN>
__global__ __launch_bounds__(256) void k_restrict_residual(LayoutDev F, LayoutDev C, G3 gp, G3 gr, G3 gc,
                                                           const T* __restrict__ phi, const T* __restrict__ rhs,
                                                           T* __restrict__ resc, Lap<T> c, T avg)
{
    constexpr int NC = 2 * DIMN + 1;
    const int b = blockIdx.y;
    const int* clo = C.lows + 3 * b;
    const long n = (long)C.bs[0] * C.bs[1] * C.bs[2];
    const long FX = F.bs[0] + 2 * gp.g[0], FXY = FX * (F.bs[1] + 2 * gp.g[1]);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % C.bs[0]);
        const long r = t / C.bs[0];
        const int y = (int)(r % C.bs[1]);
        const int z = (int)(r / C.bs[1]);
        int p[3] = {2 * (clo[0] + x), 2 * (clo[1] + y), DIMN == 3 ? 2 * (clo[2] + z) : 0};
        const int fb = owner(F, p);
        const int* flo = F.lows + 3 * fb;
        const T* ph = phi + (long)fb * gvol(F, gp);
        const T* rh = rhs + (long)fb * gvol(F, gr);
        const int fx = p[0] - flo[0], fy = p[1] - flo[1], fz = p[2] - flo[2];
        T acc = 0;
#pragma unroll
        for (int kz = 0; kz < (DIMN == 3 ? 2 : 1); kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
#pragma unroll
                for (int kx = 0; kx < 2; kx++) {
                    const long i = rel_index(F, gp, fx + kx, fy + ky, fz + kz);
                    T s = 0;
                    s = s + rh[rel_index(F, gr, fx + kx, fy + ky, fz + kz)];
#pragma unroll
                    for (int j = 0; j < NC; j++) {
                        s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * FX + lapo(j, 2) * FXY];
                    }
                    acc = acc + avg * s;
                }
        resc[(long)b * gvol(C, gc) + rel_index(C, gc, x, y, z)] = acc;
    }
}

// The same residual restriction when every fine box's coarsened region lies in
// one coarse box (box sizes of a multigrid hierarchy, see box_aligned): one
// workgroup grid per fine box, the coarse box found once per workgroup, 32-bit
// cell indices.  Same arithmetic and order as k_restrict_residual.
template<typename T>
__global__ __launch_bounds__(256) void k_restrict_residual_box(LayoutDev F, LayoutDev C, G3 gp, G3 gr, G3 gc,
                                                               const T* __restrict__ phi, const T* __restrict__ rhs,
                                                               T* __restrict__ resc, Lap<T> c, T avg)
{
    const int b = blockIdx.y;  // fine box
    const int* flo = F.lows + 3 * b;
    const int hx = F.bs[0] >> 1, hy = F.bs[1] >> 1, hz = F.bs[2] >> 1;
    int p[3] = {flo[0] >> 1, flo[1] >> 1, flo[2] >> 1};
    const int o = owner(C, p);
    const int* clo = C.lows + 3 * o;
    const int ox = (flo[0] >> 1) - clo[0], oy = (flo[1] >> 1) - clo[1], oz = (flo[2] >> 1) - clo[2];
    const T* ph = phi + (long)b * gvol(F, gp);
    const T* rh = rhs + (long)b * gvol(F, gr);
    T* rc = resc + (long)o * gvol(C, gc);
    const long FX = F.bs[0] + 2 * gp.g[0], FXY = FX * (F.bs[1] + 2 * gp.g[1]);
    const int n = hx * hy * hz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int x = t % hx, r = t / hx, y = r % hy, z = r / hy;
        const int fx = 2 * x, fy = 2 * y, fz = 2 * z;
        T acc = 0;
#pragma unroll
        for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
#pragma unroll
                for (int kx = 0; kx < 2; kx++) {
                    const long i = rel_index(F, gp, fx + kx, fy + ky, fz + kz);
                    T s = 0;
                    s = s + rh[rel_index(F, gr, fx + kx, fy + ky, fz + kz)];
#pragma unroll
                    for (int j = 0; j < 7; j++) {
                        s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * FX + lapo(j, 2) * FXY];
                    }
                    acc = acc + avg * s;
                }
        rc[rel_index(C, gc, ox + x, oy + y, oz + z)] = acc;
    }
}

// Constant prolongation on the same box-aligned hierarchies: the coarse box is
// found once per workgroup and each fine row is walked from one row base of
// the fine and the coarse storage.  Same arithmetic as k_prolong_add<.., 0>.
template<typename T>
__global__ __launch_bounds__(256) void k_prolong_const_box(LayoutDev F, LayoutDev C, G3 gd, G3 gp,
                                                           const T* __restrict__ delta, T* __restrict__ phi)
{
    const int b = blockIdx.y;
    const int* flo = F.lows + 3 * b;
    const int nx = F.bs[0], ny = F.bs[1], nz = F.bs[2];
    int p[3] = {flo[0] >> 1, flo[1] >> 1, flo[2] >> 1};
    const int o = owner(C, p);
    const int* clo = C.lows + 3 * o;
    const T* d = delta + (long)o * gvol(C, gd);
    T* ph = phi + (long)b * gvol(F, gp);
    const long FX = nx + 2 * gp.g[0], FY = ny + 2 * gp.g[1];
    const long CX = C.bs[0] + 2 * gd.g[0], CY = C.bs[1] + 2 * gd.g[1];
    const T one = 1;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int rows = ny * nz;
    for (int row = blockIdx.x * 4 + ty; row < rows; row += gridDim.x * 4) {
        const int y = row % ny, z = row / ny;
        const int J = ((flo[1] + y) >> 1) - clo[1], K = ((flo[2] + z) >> 1) - clo[2];
        T* prow = ph + FX * ((y + gp.g[1]) + FY * (z + gp.g[2])) + gp.g[0];
        const T* drow = d + CX * ((J + gd.g[1]) + CY * (K + gd.g[2])) + gd.g[0];
        for (int x = tx; x < nx; x += 64) {
            const T dv = drow[((flo[0] + x) >> 1) - clo[0]];
            T s0 = 0;
            s0 = s0 + one * dv;
            T s1 = 0;
            s1 = s1 + one * s0;
            T s2 = 0;
            s2 = s2 + one * s1;
            prow[x] = prow[x] + s2;
        }
    }
}

// The same residual on boxes 32..64 cells wide, marching z: a workgroup owns
// a 64 x 16 tile (4 wave64s, 4 rows each) of one box over a z range.  Each
// plane of phi (tile + 1-cell halo, inside the box's ghosted storage) is
// staged once in LDS (a ring of 3 planes: plane p is last read in step p + 1
// and rewritten in step p + 3, after every thread passed the barrier of step
// p + 2); each thread keeps its own cells of planes z-1, z, z+1 in registers
// for the z taps.  Next-plane loads are issued before each plane is computed;
// results are stored at the start of the next step.  NEG = 0: the arithmetic
// and order of k_residual (0 + rhs + sum c_j phi); NEG = 1: those of k_resnorm
// (0 - rhs + sum c_j phi).  out may be null (norm only).
constexpr int RZ_RY = 4, RZ_TY = 4 * RZ_RY, RZ_NE = (RZ_TY + 2) * 66;

template<typename T, typename TO, int NEG>
__global__ __launch_bounds__(256) void k_residual_z(LayoutDev L, G3 gp, G3 gr, G3 go, const T* __restrict__ phi,
                                                    const T* __restrict__ rhs, TO* __restrict__ out, Lap<T> c,
                                                    T* __restrict__ partials, int zc)
{
    constexpr int RY = RZ_RY, NL = (RZ_NE + 255) / 256;
    __shared__ T sm[3][RZ_NE];
    __shared__ T red[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b = blockIdx.y;
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2];
    const int nyt = (ny + RZ_TY - 1) / RZ_TY;
    const int y0 = (int)(blockIdx.x % nyt) * RZ_TY;
    const int z0 = (int)(blockIdx.x / nyt) * zc, z1 = min(z0 + zc, nz);
    const int LX = nx + 2, NE = (RZ_TY + 2) * LX;
    const long X = nx + 2 * gp.g[0], XY = X * (ny + 2 * gp.g[1]);
    const T* ph = phi + (long)b * gvol(L, gp);
    const T* rh = rhs + (long)b * gvol(L, gr);
    // staged element e: box row y0 - 1 + e / LX (clamped to the last storage
    // row), box column e % LX - 1; elements past the plane repeat the last one
    long goff[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = min(tid + 256 * k, NE - 1);
        const int er = e / LX, ec = e - er * LX;
        const int sr = min(y0 - 1 + er + gp.g[1], ny + 2 * gp.g[1] - 1);
        goff[k] = X * (long)sr + (ec - 1 + gp.g[0]);
    }
    T P[NL];
    auto gload = [&](int z) {
        const T* base = ph + XY * (long)(z + gp.g[2]);
#pragma unroll
        for (int k = 0; k < NL; k++) { P[k] = base[goff[k]]; }
    };
    auto lstore = [&](T* pl) {
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int e = tid + 256 * k;
            if (e < NE) { pl[e] = P[k]; }
        }
    };
    const int x = lane;
    const bool xok = x < nx;
    const int xc = xok ? x : 0;
    T zw[3][RY];
    T fq[RY], res[RY];
    int zres = -1;
    auto fload = [&](int zo) {
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = min(y0 + w * RY + r, ny - 1);
            fq[r] = rh[rel_index(L, gr, xc, y, zo)];
        }
    };
    T m = 0;
    auto put_res = [&]() {
        if (zres < 0 || !xok || !out) { return; }
        TO* ot = out + (long)b * gvol(L, go);
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = y0 + w * RY + r;
            if (y < ny) { ot[rel_index(L, go, x, y, zres)] = (TO)res[r]; }
        }
    };
    const int kb = z0 - 1;
    gload(kb);
    fload(z0);
    for (int zz = kb; zz <= z1; zz++) {
        T* sl = sm[(zz - kb) % 3];
        lstore(sl);
        if (zz + 1 <= z1) { gload(zz + 1); }
        put_res();
        zres = -1;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RY; r++) {
            zw[0][r] = zw[1][r];
            zw[1][r] = zw[2][r];
            zw[2][r] = sl[(w * RY + r + 1) * LX + (xc + 1)];
        }
        if (zz < z0 + 1) { continue; }
        const int zo = zz - 1;
        const T* s0 = sm[(zo - kb) % 3];
        T fcur[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) { fcur[r] = fq[r]; }
        if (zo + 1 < z1) { fload(zo + 1); }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int li = (w * RY + r + 1) * LX + (xc + 1);
            T sv = 0;
            sv = NEG ? sv - fcur[r] : sv + fcur[r];
            sv = sv + c.c[0] * zw[1][r];
            sv = sv + c.c[1] * s0[li + 1];
            sv = sv + c.c[2] * s0[li - 1];
            sv = sv + c.c[3] * s0[li + LX];
            sv = sv + c.c[4] * s0[li - LX];
            sv = sv + c.c[5] * zw[2][r];
            sv = sv + c.c[6] * zw[0][r];
            res[r] = sv;
            const bool ok = xok && (y0 + w * RY + r) < ny;
            const T a = sv > 0 ? sv : -sv;
            m = (ok && m < a) ? a : m;
        }
        zres = zo;
    }
    put_res();
    m = block_reduce(PM_RED_ABS, m, red);
    if (tid == 0) { partials[(long)b * gridDim.x + blockIdx.x] = m; }
}

// The z-marching residual of k_residual_z on a periodic level covering its
// domain, with the halo staged from the neighbouring boxes' valid cells
// (periodic images folded in) instead of the ghost layer: the values an
// exchange would have put there, with no exchange.  Planes beyond the box go
// through per-workgroup offset tables (s_dz, as in k_gsrb_fused).  MODE 0:
// r = 0 + rhs + sum c_j phi into out (TO) and max|r| partials; 1: the norm of
// 0 - rhs + sum c_j phi only; 2: residual-restriction, the coarse right-hand
// side sum over the 2x2x2 children (z, y, x nested, x fastest) of avg * r into
// resc on the box-aligned coarse level C (ghost gc), in the order of
// k_restrict_residual_box; x neighbours come from the next lane.
template<typename T, typename TO, int MODE>
__global__ __launch_bounds__(256) void k_residual_nb(LayoutDev L, G3 gp, G3 gr, G3 go, const T* __restrict__ phi,
                                                     const T* __restrict__ rhs, TO* __restrict__ out, Lap<T> c,
                                                     T* __restrict__ partials, int zc, LayoutDev C, G3 gc,
                                                     T* __restrict__ resc, T avg)
{
    constexpr int RY = RZ_RY, NL = (RZ_NE + 255) / 256;
    __shared__ T sm[3][RZ_NE];
    __shared__ T red[16];
    __shared__ int s_nb[27], s_dz[18];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b = blockIdx.y;
    const int* lo = L.lows + 3 * b;
    const int nx = L.bs[0], ny = L.bs[1], nz = L.bs[2];
    const int nyt = (ny + RZ_TY - 1) / RZ_TY;
    const int y0 = (int)(blockIdx.x % nyt) * RZ_TY;
    const int z0 = (int)(blockIdx.x / nyt) * zc, z1 = min(z0 + zc, nz);
    const int LX = nx + 2, NE = (RZ_TY + 2) * LX;
    const long vol = gvol(L, gp);
    const long XY = (long)(nx + 2 * gp.g[0]) * (ny + 2 * gp.g[1]);
    const T* rh = rhs + (long)b * gvol(L, gr);
    if (tid < 27) {
        int p[3] = {lo[0] + (tid % 3 - 1) * nx, lo[1] + ((tid / 3) % 3 - 1) * ny, lo[2] + (tid / 9 - 1) * nz};
        s_nb[tid] = owner(L, p);
    }
    __syncthreads();
    auto at = [&](int x, int y, int z) {
        const int ox = x < 0 ? -1 : (x >= nx ? 1 : 0);
        const int oy = y < 0 ? -1 : (y >= ny ? 1 : 0);
        const int oz = z < 0 ? -1 : (z >= nz ? 1 : 0);
        const int o = s_nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        return (long)o * vol + rel_index(L, gp, x - ox * nx, y - oy * ny, z - oz * nz);
    };
    if (tid < 18) {
        // element of a plane beyond the box minus the in-box formula, per
        // side and xy neighbour position (the same for every cell of it)
        const int side = tid / 9, k = tid % 9;
        const int x = k % 3 == 0 ? -1 : (k % 3 == 1 ? 0 : nx);
        const int y = k / 3 == 0 ? -1 : (k / 3 == 1 ? 0 : ny);
        const int z = side ? nz : -1;
        s_dz[tid] = (int)(at(x, y, z) - (XY * z + at(x, y, 0)));
    }
    __syncthreads();
    // staged element e: row y0 - 1 + e / LX, column e % LX - 1 (cells beyond
    // the box belong to its neighbours); elements past the plane repeat the
    // last one.  Level offsets fit 32 bits (checked by the host).
    int goff[NL];
    unsigned long long kp = 0;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = min(tid + 256 * k, NE - 1);
        const int er = e / LX, ec = e - er * LX;
        const int x = ec - 1, y = y0 - 1 + er;
        goff[k] = (int)at(x, y, 0);
        kp |= (unsigned long long)((x < 0 ? 0 : (x >= nx ? 2 : 1)) + 3 * (y < 0 ? 0 : (y >= ny ? 2 : 1))) << (4 * k);
    }
    T P[NL];
    auto gload = [&](int z) {
        const bool inb = z >= 0 && z < nz;
        const int sd = z < 0 ? 0 : 9;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            long o = XY * z + goff[k];
            if (!inb) { o += s_dz[sd + (int)((kp >> (4 * k)) & 15)]; }
            P[k] = phi[o];
        }
    };
    auto lstore = [&](T* pl) {
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int e = tid + 256 * k;
            if (e < NE) { pl[e] = P[k]; }
        }
    };
    const int x = lane;
    const bool xok = x < nx;
    const int xc = xok ? x : 0;
    T zw[3][RY];
    T fq[RY], res[RY];
    int zres = -1;
    auto fload = [&](int zo) {
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = min(y0 + w * RY + r, ny - 1);
            fq[r] = rh[rel_index(L, gr, xc, y, zo)];
        }
    };
    T m = 0;
    auto put_res = [&]() {
        if (MODE != 0 || zres < 0 || !xok) { return; }
        TO* ot = out + (long)b * gvol(L, go);
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int y = y0 + w * RY + r;
            if (y < ny) { ot[rel_index(L, go, x, y, zres)] = (TO)res[r]; }
        }
    };
    // MODE 2: coarse box of this fine box and the running sums of row pairs
    T acc[RY / 2];
    T* rc = resc;
    int cox = 0, coy = 0, coz = 0;
    if constexpr (MODE == 2) {
        int p[3] = {lo[0] >> 1, lo[1] >> 1, lo[2] >> 1};
        const int o = owner(C, p);
        const int* clo = C.lows + 3 * o;
        cox = (lo[0] >> 1) - clo[0];
        coy = (lo[1] >> 1) - clo[1];
        coz = (lo[2] >> 1) - clo[2];
        rc = resc + (long)o * gvol(C, gc);
    }
    const int kb = z0 - 1;
    gload(kb);
    fload(z0);
    for (int zz = kb; zz <= z1; zz++) {
        T* sl = sm[(zz - kb) % 3];
        lstore(sl);
        if (zz + 1 <= z1) { gload(zz + 1); }
        put_res();
        zres = -1;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RY; r++) {
            zw[0][r] = zw[1][r];
            zw[1][r] = zw[2][r];
            zw[2][r] = sl[(w * RY + r + 1) * LX + (xc + 1)];
        }
        if (zz < z0 + 1) { continue; }
        const int zo = zz - 1;
        const T* s0 = sm[(zo - kb) % 3];
        T fcur[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) { fcur[r] = fq[r]; }
        if (zo + 1 < z1) { fload(zo + 1); }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int li = (w * RY + r + 1) * LX + (xc + 1);
            T sv = 0;
            sv = MODE == 1 ? sv - fcur[r] : sv + fcur[r];
            sv = sv + c.c[0] * zw[1][r];
            sv = sv + c.c[1] * s0[li + 1];
            sv = sv + c.c[2] * s0[li - 1];
            sv = sv + c.c[3] * s0[li + LX];
            sv = sv + c.c[4] * s0[li - LX];
            sv = sv + c.c[5] * zw[2][r];
            sv = sv + c.c[6] * zw[0][r];
            res[r] = sv;
            if constexpr (MODE != 2) {
                const bool ok = xok && (y0 + w * RY + r) < ny;
                const T a = sv > 0 ? sv : -sv;
                m = (ok && m < a) ? a : m;
            }
        }
        if constexpr (MODE == 2) {
            // children (x, y), (x+1, y), (x, y+1), (x+1, y+1) of plane zo, in
            // that order, onto the sums of plane pair zo & ~1
            const bool first = (zo & 1) == 0;
#pragma unroll
            for (int q = 0; q < RY / 2; q++) {
                const T a0 = res[2 * q], a1 = __shfl_down(res[2 * q], 1, 64);
                const T b0 = res[2 * q + 1], b1 = __shfl_down(res[2 * q + 1], 1, 64);
                T t = first ? (T)0 : acc[q];
                t = t + avg * a0;
                t = t + avg * a1;
                t = t + avg * b0;
                t = t + avg * b1;
                acc[q] = t;
            }
            if (!first && xok && (x & 1) == 0) {
#pragma unroll
                for (int q = 0; q < RY / 2; q++) {
                    const int y = y0 + w * RY + 2 * q;
                    if (y < ny) { rc[rel_index(C, gc, cox + (x >> 1), coy + (y >> 1), coz + (zo >> 1))] = acc[q]; }
                }
            }
        } else {
            zres = zo;
        }
    }
    put_res();
    if constexpr (MODE != 2) {
        m = block_reduce(PM_RED_ABS, m, red);
        if (tid == 0) { partials[(long)b * gridDim.x + blockIdx.x] = m; }
    }
}

// boxes and ghosts the z-marching residual takes, and its grid (16-row tiles x
// z chunks, halved while fewer than 2048 workgroups)
static bool rz_ok(const pm_layout* L, const G3& gp)
{
    return L->dim == 3 && L->bs[0] >= 32 && L->bs[0] <= 64 && L->bs[1] >= 16 && L->bs[2] >= 4 && gp.g[0] >= 1 &&
           gp.g[1] >= 1 && gp.g[2] >= 1;
}
static dim3 rz_grid(const pm_layout* L, int* zc)
{
    const int nyt = (L->bs[1] + RZ_TY - 1) / RZ_TY;
    int z = L->bs[2];
    while (z > kZcFloor && (long)nyt * ((L->bs[2] + z - 1) / z) * L->nboxes < 2048) { z = (z + 1) / 2; }
    *zc = z;
    return dim3((unsigned)(nyt * ((L->bs[2] + z - 1) / z)), (unsigned)L->nboxes);
}

// ------------------------------------------------------------ prolongation
template<typename T, int DIMN, int LINEAR>
__global__ __launch_bounds__(256) void k_prolong_add(LayoutDev F, LayoutDev C, G3 gd, G3 gp,
                                                     const T* __restrict__ delta, T* __restrict__ phi)
{
    const int b = blockIdx.y;
    const int* flo = F.lows + 3 * b;
    const long n = (long)F.bs[0] * F.bs[1] * F.bs[2];
    T* ph = phi + (long)b * gvol(F, gp);
    const T one = 1, half = 0.5, zero = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % F.bs[0]);
        const long r = t / F.bs[0];
        const int y = (int)(r % F.bs[1]);
        const int z = (int)(r / F.bs[1]);
        const int fx = flo[0] + x, fy = flo[1] + y, fz = flo[2] + z;
        const int I = fx >> 1, J = fy >> 1, K = (DIMN == 3) ? (fz >> 1) : 0;
        T v;
        if (!LINEAR) {
            T s0 = 0;
            s0 = s0 + one * level_value(C, gd, delta, I, J, K);
            T s1 = 0;
            s1 = s1 + one * s0;
            v = s1;
            if (DIMN == 3) {
                T s2 = 0;
                s2 = s2 + one * s1;
                v = s2;
            }
        } else {
            // vertex-style weights of InterpStencil::Linear (ratio 2): shift 0 -> (1, 0), shift 1 -> (1/2, 1/2)
            const T qx = (fx & 1) ? half : one, px = (fx & 1) ? half : zero;
            const T qy = (fy & 1) ? half : one, py = (fy & 1) ? half : zero;
            const T qz = (fz & 1) ? half : one, pz = (fz & 1) ? half : zero;
            T s1v[2];
#pragma unroll
            for (int kk = 0; kk < (DIMN == 3 ? 2 : 1); kk++) {
                T s0v[2];
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    T a = 0;
                    a = a + qx * level_value(C, gd, delta, I, J + jj, K + kk);
                    a = a + px * level_value(C, gd, delta, I + 1, J + jj, K + kk);
                    s0v[jj] = a;
                }
                T a = 0;
                a = a + qy * s0v[0];
                a = a + py * s0v[1];
                s1v[kk] = a;
            }
            if (DIMN == 3) {
                T a = 0;
                a = a + qz * s1v[0];
                a = a + pz * s1v[1];
                v = a;
            } else {
                v = s1v[0];
            }
        }
        const long i = rel_index(F, gp, x, y, z);
        ph[i] = ph[i] + one * v;
    }
}

// ------------------------------------------------------------ copy / setval
template<typename T>
__global__ __launch_bounds__(256) void k_level_copy(LayoutDev S, G3 gs, const T* __restrict__ src, LayoutDev D, G3 gd,
                                                    T* __restrict__ dst, int ncomp)
{
    const int b = blockIdx.y;
    const int* lo = D.lows + 3 * b;
    const long n = (long)D.bs[0] * D.bs[1] * D.bs[2];
    const long dv = gvol(D, gd), sv = gvol(S, gs);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % D.bs[0]);
        const long r = t / D.bs[0];
        const int y = (int)(r % D.bs[1]);
        const int z = (int)(r / D.bs[1]);
        int p[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
        const int o = owner(S, p);
        if (o < 0) { continue; }
        const int* slo = S.lows + 3 * o;
        const long si = rel_index(S, gs, p[0] - slo[0], p[1] - slo[1], p[2] - slo[2]);
        const long di = rel_index(D, gd, x, y, z);
        for (int c = 0; c < ncomp; c++) {
            dst[((long)b * ncomp + c) * dv + di] = src[((long)o * ncomp + c) * sv + si];
        }
    }
}

// LevelBoxData::copyTo with Proto's regions.  mode 0 (LevelCopier between
// different layouts, Proto_LevelBoxDataImplem.H:795-870): every cell of each
// destination box grown by its ghosts whose point (or its periodic image) lies
// in a source patch takes that patch's valid value; other cells are left alone.
// mode 1 (copyToSimple between compatible layouts, 346-356: BoxData::copyTo box
// by box): every cell of the destination's ghosted box that also lies in the
// same box's ghosted source box takes the source value at that cell; the
// caller then exchanges the destination.
template<typename T>
__global__ __launch_bounds__(256) void k_level_copy_to(LayoutDev S, G3 gs, const T* __restrict__ src, LayoutDev D,
                                                       G3 gd, T* __restrict__ dst, int ncomp, int mode)
{
    const int b = blockIdx.y;
    const int* lo = D.lows + 3 * b;
    const int X = D.bs[0] + 2 * gd.g[0], Y = D.bs[1] + 2 * gd.g[1], Z = D.bs[2] + 2 * gd.g[2];
    const long n = (long)X * Y * Z;
    const long dv = gvol(D, gd), sv = gvol(S, gs);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % X) - gd.g[0];
        const long r = t / X;
        const int y = (int)(r % Y) - gd.g[1];
        const int z = (int)(r / Y) - gd.g[2];
        int o;
        long si;
        if (mode == 1) {
            if (x < -gs.g[0] || x >= S.bs[0] + gs.g[0] || y < -gs.g[1] || y >= S.bs[1] + gs.g[1] ||
                z < -gs.g[2] || z >= S.bs[2] + gs.g[2]) { continue; }
            o = b;
            si = rel_index(S, gs, x, y, z);
        } else {
            int p[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
            o = owner(S, p);
            if (o < 0) { continue; }
            const int* slo = S.lows + 3 * o;
            si = rel_index(S, gs, p[0] - slo[0], p[1] - slo[1], p[2] - slo[2]);
        }
        const long di = rel_index(D, gd, x, y, z);
        for (int c = 0; c < ncomp; c++) {
            dst[((long)b * ncomp + c) * dv + di] = src[((long)o * ncomp + c) * sv + si];
        }
    }
}

template<typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ data, long n, T v)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        data[t] = v;
    }
}

template<typename T>
__global__ __launch_bounds__(1024) void k_max_partials(const T* __restrict__ partials, long n, T* out, int combine)
{
    __shared__ T red[16];
    T v = 0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) { v = red_combine(PM_RED_ABS, v, partials[i]); }
    v = block_reduce(PM_RED_ABS, v, red);
    if (threadIdx.x == 0) { *out = combine ? red_combine(PM_RED_ABS, *out, v) : v; }
}

// ------------------------------------------------------------ host helpers
static unsigned bx_blocks(long n, long cap = 1024)
{
    long b = (n + 255) / 256;
    return (unsigned)std::max(1L, std::min(b, cap));
}

// Workgroups per box of a grid-stride launch over every box of a level: about
// 16384 workgroups for the level however many boxes it has (a 512-box level
// with one 256-cell pass per workgroup would dispatch 524288 workgroups, and
// dispatch, not memory, would set the time), at most one per 256 items.
static unsigned lvl_blocks(long n, int nboxes, long cap = 1024)
{
    return bx_blocks(n, std::min(cap, std::max(1L, 16384L / std::max(1, nboxes))));
}

// Workgroups per box of the generic residual / norm kernels (one max-norm
// partial each): 64 per box, and on levels of few large boxes enough for about
// 16384 over the level (a single 256^3 box with 64 workgroups took 289 us for
// its norm, 16 times its streaming time).
static unsigned res_blocks(long n, int nboxes)
{
    return bx_blocks(n, std::max(64L, 16384L / std::max(1, nboxes)));
}

template<typename T>
static Lap<T> lap_coefs(int dim, double m)
{
    // Laplacian() coefficients (-2 DIM, then 1s) times m, rounded once in T
    Lap<T> c;
    for (int j = 0; j < 7; j++) { c.c[j] = 0; }
    for (int j = 0; j < 2 * dim + 1; j++) {
        const T lj = (T)(j == 0 ? -2.0 * dim : 1.0);
        c.c[j] = lj * (T)m;
    }
    return c;
}

int layout_shape(const pm_layout* L, int* dim, int* domain, int* bs, int* periodic, int* nboxes)
{
    PM_REQUIRE(L, "null layout");
    *dim = L->dim;
    *nboxes = L->nboxes;
    for (int k = 0; k < 3; k++) {
        domain[k] = L->domain[k];
        bs[k] = L->bs[k];
        periodic[k] = L->periodic[k];
    }
    return PM_OK;
}

int layout_lows(const pm_layout* L, int* lows)
{
    PM_REQUIRE(L && lows, "null argument");
    for (size_t i = 0; i < L->lows.size(); i++) { lows[i] = L->lows[i]; }
    return PM_OK;
}

// 3D, both levels fully covered, fine boxes of even size and coarse boxes
// whose size is a multiple of half the fine box size: then every fine box's
// coarsened region lies in one coarse box, and the fine boxes' regions tile
// the coarse level
static bool box_aligned(const pm_layout* F, const pm_layout* C)
{
    if (F->dim != 3 || C->nboxes != C->nb[0] * C->nb[1] * C->nb[2] || F->nboxes != F->nb[0] * F->nb[1] * F->nb[2]) {
        return false;
    }
    for (int k = 0; k < 3; k++) {
        if (F->bs[k] % 2 != 0 || C->bs[k] % (F->bs[k] / 2) != 0) { return false; }
    }
    return true;
}

static bool same_layout_shape(const pm_layout* a, const pm_layout* b)
{
    for (int k = 0; k < 3; k++) {
        if (a->domain[k] != b->domain[k]) { return false; }
    }
    return a->dim == b->dim;
}

}  // namespace pm

using namespace pm;

extern "C" {

int pm_layout_create(pm_layout** out, int dim, const int domain[3], const int box_size[3], const int periodic[3],
                     int nboxes, const int* box_lows_host)
{
    PM_REQUIRE(out && domain && box_size, "pm_layout_create: null argument");
    *out = nullptr;
    PM_REQUIRE(dim == 2 || dim == 3, "pm_layout_create: dim must be 2 or 3");
    PM_REQUIRE(nboxes >= 1 && box_lows_host, "pm_layout_create: no boxes");
    pm_layout* L = new pm_layout();
    L->dim = dim;
    L->nboxes = nboxes;
    long ncell = 1;
    for (int k = 0; k < 3; k++) {
        L->domain[k] = (k < dim) ? domain[k] : 1;
        L->bs[k] = (k < dim) ? box_size[k] : 1;
        L->periodic[k] = (k < dim && periodic) ? (periodic[k] != 0) : 0;
        if (L->bs[k] < 1 || L->domain[k] < 1 || L->domain[k] % L->bs[k] != 0) {
            delete L;
            set_error("pm_layout_create: box size must divide the domain");
            return PM_ERR_INVALID;
        }
        L->nb[k] = L->domain[k] / L->bs[k];
        ncell *= L->nb[k];
    }
    std::vector<int> table(ncell, -1);
    L->lows.resize(3 * nboxes);
    for (int b = 0; b < nboxes; b++) {
        int gi[3];
        for (int k = 0; k < 3; k++) {
            const int lo = (k < dim) ? box_lows_host[3 * b + k] : 0;
            L->lows[3 * b + k] = lo;
            if (lo < 0 || lo >= L->domain[k] || lo % L->bs[k] != 0) {
                delete L;
                set_error("pm_layout_create: box %d low corner not on the box grid", b);
                return PM_ERR_INVALID;
            }
            gi[k] = lo / L->bs[k];
        }
        int& slot = table[gi[0] + (long)L->nb[0] * (gi[1] + (long)L->nb[1] * gi[2])];
        if (slot >= 0) {
            delete L;
            set_error("pm_layout_create: boxes %d and %d overlap", slot, b);
            return PM_ERR_INVALID;
        }
        slot = b;
    }
    if (hipMalloc((void**)&L->d_lows, sizeof(int) * 3 * nboxes) != hipSuccess ||
        hipMalloc((void**)&L->d_table, sizeof(int) * ncell) != hipSuccess) {
        delete L;
        set_error("pm_layout_create: device allocation failed");
        return PM_ERR_NOMEM;
    }
    if (hipMemcpy(L->d_lows, L->lows.data(), sizeof(int) * 3 * nboxes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L->d_table, table.data(), sizeof(int) * ncell, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(L->d_lows);
        hipFree(L->d_table);
        delete L;
        set_error("pm_layout_create: upload failed");
        return PM_ERR_HIP;
    }
    *out = L;
    return PM_OK;
}

int pm_layout_destroy(pm_layout* L)
{
    if (!L) { return PM_OK; }
    hipFree(L->d_lows);
    hipFree(L->d_table);
    delete L;
    return PM_OK;
}

int pm_layout_box_elems(const pm_layout* L, const int ghost[3], size_t* elems)
{
    PM_REQUIRE(L && elems, "pm_layout_box_elems: null argument");
    *elems = (size_t)gvol(ldev(L), g3(ghost));
    return PM_OK;
}

#define PM_DISPATCH_T(dtype, CALL)                                         \
    do {                                                                   \
        if ((dtype) == PM_F64) { using T = double; CALL; }                 \
        else if ((dtype) == PM_F32) { using T = float; CALL; }             \
        else { PM_REQUIRE(false, "bad dtype %d", (int)(dtype)); }          \
    } while (0)

int pm_level_linear_size(const pm_layout* L, int dtype, const int ghost[3], int ncomp, size_t* bytes)
{
    PM_REQUIRE(L && bytes && ncomp >= 1 && (dtype == PM_F32 || dtype == PM_F64), "pm_level_linear_size: bad argument");
    *bytes = (size_t)gvol(ldev(L), g3(ghost)) * ncomp * L->nboxes * (dtype == PM_F64 ? 8 : 4);
    return PM_OK;
}

int pm_level_linear_out(const pm_layout* L, int dtype, const void* data, const int ghost[3], int ncomp, void* buffer,
                        int to_host, pm_stream_t stream)
{
    PM_REQUIRE(data && buffer, "pm_level_linear_out: null argument");
    size_t bytes = 0;
    int rc = pm_level_linear_size(L, dtype, ghost, ncomp, &bytes);
    if (rc) { return rc; }
    PM_CHECK_HIP(hipMemcpyAsync(buffer, data, bytes, to_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                                S(stream)));
    if (to_host) { PM_CHECK_HIP(hipStreamSynchronize(S(stream))); }
    return PM_OK;
}

int pm_level_linear_in(const pm_layout* L, int dtype, void* data, const int ghost[3], int ncomp, const void* buffer,
                       int from_host, pm_stream_t stream)
{
    PM_REQUIRE(data && buffer, "pm_level_linear_in: null argument");
    size_t bytes = 0;
    int rc = pm_level_linear_size(L, dtype, ghost, ncomp, &bytes);
    if (rc) { return rc; }
    PM_CHECK_HIP(hipMemcpyAsync(data, buffer, bytes, from_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                                S(stream)));
    if (from_host) { PM_CHECK_HIP(hipStreamSynchronize(S(stream))); }
    return PM_OK;
}

int pm_level_exchange(const pm_layout* L, int dtype, void* data, const int ghost[3], int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(L && data && ncomp >= 1, "pm_level_exchange: bad argument");
    const G3 g = g3(ghost);
    if (g.g[0] == 0 && g.g[1] == 0 && g.g[2] == 0) { return PM_OK; }
    for (int k = 0; k < 3; k++) {
        // a ghost layer wider than a box takes images of boxes further out (the
        // owner lookup folds any periodic distance), as Proto's NeighborIterator
        // over bx0.grow(ghost) does (Proto_LevelBoxDataImplem.H:905-925)
        PM_REQUIRE(g.g[k] >= 0, "pm_level_exchange: negative ghost");
        PM_REQUIRE(k < L->dim || g.g[k] == 0, "pm_level_exchange: ghost beyond dim");
    }
    const LayoutDev D = ldev(L);
    const long X = L->bs[0] + 2 * g.g[0], Y = L->bs[1] + 2 * g.g[1];
    const long nghost = 2 * (X * Y * g.g[2] + X * g.g[1] * L->bs[2] + (long)g.g[0] * L->bs[1] * L->bs[2]);
    dim3 grid(lvl_blocks(nghost, L->nboxes, 256), L->nboxes);
    PM_DISPATCH_T(dtype, hipLaunchKernelGGL(k_exchange<T>, grid, dim3(256), 0, S(stream), D, g, (T*)data, ncomp));
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_gsrb(const pm_layout* L, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                  const int rhs_ghost[3], double dx, int color, pm_stream_t stream)
{
    PM_REQUIRE(L && phi && rhs && (color == 0 || color == 1), "pm_level_gsrb: bad argument");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost);
    for (int k = 0; k < L->dim; k++) { PM_REQUIRE(gp.g[k] >= 1, "pm_level_gsrb: phi needs a ghost layer"); }
    const LayoutDev D = ldev(L);
    const long n = (long)((L->bs[0] + 1) / 2) * L->bs[1] * L->bs[2];
    dim3 grid(bx_blocks(n), L->nboxes);  // one cell per thread measured fastest for this kernel
    // LaplaceOp::init (LaplaceOp.H:56-73): lambda = 2 * (-(dx*dx)/(4 DIM)); coefficients Lap_j * (-1/(dx*dx))
    double lam = -(dx * dx) / (4.0 * L->dim);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx);
    PM_DISPATCH_T(dtype, {
        Lap<T> c = lap_coefs<T>(L->dim, m);
        if (dtype == PM_F64) {
            for (int j = 0; j < 2 * L->dim + 1; j++) { c.c[j] = (T)((j == 0 ? -2.0 * L->dim : 1.0) * m); }
        }
        if (L->dim == 3) {
            hipLaunchKernelGGL((k_gsrb<T, 3>), grid, dim3(256), 0, S(stream), D, gp, gr, (T*)phi, (const T*)rhs, c,
                               (T)(1.0 * lam), color);
        } else {
            hipLaunchKernelGGL((k_gsrb<T, 2>), grid, dim3(256), 0, S(stream), D, gp, gr, (T*)phi, (const T*)rhs, c,
                               (T)(1.0 * lam), color);
        }
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // extern "C"

namespace pm {

#define PM_SWEEP_LAUNCH(E, M, O)                                                                                 \
    hipLaunchKernelGGL((k_gsrb_fused<T, E, M, O>), grid, dim3(256), lds, stream, D, gi, go, gr, pi, (T*)phi_out, \
                       (const T*)rhs, c, lt, rows, zc, tw, DC, gc, cr, acc)
#define PM_SWEEP_MODES(E)         \
    if (acc) {                    \
        PM_SWEEP_LAUNCH(E, 0, 1); \
    } else if (in_mode == 0) {    \
        PM_SWEEP_LAUNCH(E, 0, 0); \
    } else if (in_mode == 1) {    \
        PM_SWEEP_LAUNCH(E, 1, 0); \
    } else {                      \
        PM_SWEEP_LAUNCH(E, 2, 0); \
    }

// Sweep kernel of the one-pass sweep (tests and benchmarks compare them): 0
// automatic (the row-register sweep where the level qualifies), 1 the tile
// sweep, 2 / 3 / 4 the row-register sweep with 4 / 8 / 2 rows per thread.
static int g_sweep_form = 0;

// the row-register sweep takes boxes 64 cells wide, at most 8 across the
// domain, whose height the rows per thread divide
static bool rows_ok(const pm_layout* L, int ry)
{
    return L->bs[0] == 64 && L->nb[0] <= 8 && L->bs[1] % ry == 0 && L->bs[1] >= 2 && L->bs[2] >= 2;
}

// planes per workgroup of the row sweep: a divisor of the box depth, small
// enough for about 8192 waves over the level
// tuning overrides (pm_level_sweep_tune): largest planes per workgroup (0:
// automatic), non-temporal write-out (-1: automatic, FP64 only)
static int g_rows_zc_max = 0, g_rows_nt = -1;

static int rows_zc(const pm_layout* L, int ry, long waves = kRowsWaves)
{
    int zc = L->bs[2];
    while (zc % 2 == 0 && zc > kZcFloor && (long)(L->domain[1] / ry) * (L->domain[2] / zc) * L->nb[0] < waves) { zc /= 2; }
    while (g_rows_zc_max > 0 && zc > g_rows_zc_max && zc % 2 == 0) { zc /= 2; }
    return zc;
}

// ghost layers (phi in, phi out) the row sweep is built for: the multigrid
// driver's ping-pong phi (ghost 1) -> scratch (ghost 0) -> phi, rhs without ghosts
static int rows_ghosts(const int in_ghost[3], const int out_ghost[3], const int rhs_ghost[3], int in_mode)
{
    const G3 gi = g3(in_ghost), go = g3(out_ghost), gq = g3(rhs_ghost);
    bool in1 = true, in0 = true, out1 = true, out0 = true, rhs0 = true;
    for (int k = 0; k < 3; k++) {
        in1 = in1 && gi.g[k] == 1;
        in0 = in0 && gi.g[k] == 0;
        out1 = out1 && go.g[k] == 1;
        out0 = out0 && go.g[k] == 0;
        rhs0 = rhs0 && gq.g[k] == 0;
    }
    if (!rhs0) { return -1; }
    if ((in1 || in_mode == 1) && out0) { return 10; }
    if ((in0 || in_mode == 1) && out1) { return 1; }
    return -1;
}

template<typename T, int RY, int PF>
static void launch_rows(const pm_layout* L, int in_mode, int ghosts, const T* pi, T* po, const T* rhs,
                        const Lap<T>& c, T lt, double* acc, const pm_layout* C, const T* crs, hipStream_t stream)
{
    const int zc = rows_zc(L, RY);
    const dim3 grid((unsigned)((L->domain[1] / RY) * (L->domain[2] / zc))), block((unsigned)(64 * L->nb[0]));
    const LayoutDev D = ldev(L), DC = in_mode == 2 ? ldev(C) : D;
    // non-temporal write-out where the output has no ghost layer (512^3 FP32
    // sweep 0.385 vs 0.392 ms, FP64 0.672 vs 0.691 ms); into ghost-1 arrays
    // the stores stay cached, so that the row-end ghost cells written a step
    // later complete the same lines in L2
    const bool nt = g_rows_nt < 0 ? true : g_rows_nt == 1;
#define PM_ROWS(I, O, N, GP, GO)                                                                                  \
    hipLaunchKernelGGL((k_gsrb_rows<T, RY, I, O, N, GP, 0, GO, PF>), grid, block, 0, stream, D, pi, po, rhs, c, lt, zc, \
                       acc, DC, crs)
#define PM_ROWS_NT(I, O, GP, GO) \
    if (nt) {                    \
        PM_ROWS(I, O, 1, GP, GO); \
    } else {                     \
        PM_ROWS(I, O, 0, GP, GO); \
    }
    if (ghosts == 10) {  // phi (ghost 1) -> scratch (ghost 0)
        if (in_mode == 0) {
            PM_ROWS_NT(0, 0, 1, 0);
        } else if (in_mode == 1) {
            PM_ROWS_NT(1, 0, 1, 0);
        } else {
            PM_ROWS_NT(2, 0, 1, 0);
        }
    } else {  // scratch (ghost 0) -> phi (ghost 1), or added into the FP64 solution
        if (acc) {
            PM_ROWS(0, 1, 0, 0, 1);
        } else if (in_mode == 0) {
            PM_ROWS_NT(0, 0, 0, 1);
        } else {
            PM_ROWS_NT(1, 0, 0, 1);
        }
    }
#undef PM_ROWS_NT
#undef PM_ROWS
}

// the row-register sweep with the coefficients and weight of pm_level_gsrb
template<typename T>
static void sweep_rows(const pm_layout* L, int ry, int ghosts, const void* phi_in, void* phi_out,
                       const void* rhs, double dx, int in_mode, const pm_layout* C, const void* coarse,
                       hipStream_t stream, double* acc)
{
    double lam = -(dx * dx) / (4.0 * L->dim);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx);
    Lap<T> c = lap_coefs<T>(L->dim, m);
    if (sizeof(T) == 8) {
        for (int j = 0; j < 7; j++) { c.c[j] = (T)((j == 0 ? -2.0 * L->dim : 1.0) * m); }
    }
    const T* pi = (const T*)(in_mode == 1 ? phi_out : phi_in);
    const T lt = (T)(1.0 * lam);
    if constexpr (sizeof(T) == 4) {
        // FP32 into the scratch level: two cells per lane where the level is an
        // even number of boxes wide (z chunks start on even planes)
        const int zc = rows_zc(L, 4);
        if (g_sweep_form == 0 && ry == 4 && ghosts == 10 && !acc && L->nb[0] % 2 == 0 && zc % 2 == 0) {
            const dim3 grid((unsigned)((L->domain[1] / 4) * (L->domain[2] / zc))), block((unsigned)(32 * L->nb[0]));
            const LayoutDev D = ldev(L), DC = in_mode == 2 ? ldev(C) : D;
            const bool nt = g_rows_nt < 0 ? true : g_rows_nt == 1;
            // planes loaded one step ahead (two: 0.364-0.371 ms, the same)
#define PM_ROWS2(I, N)                                                                                          \
    hipLaunchKernelGGL((k_gsrb_rows2<4, I, N, 1>), grid, block, 0, stream, D, pi, (T*)phi_out, (const T*)rhs, c, lt, zc, \
                       DC, (const T*)coarse)
            if (in_mode == 0) {
                if (nt) { PM_ROWS2(0, 1); } else { PM_ROWS2(0, 0); }
            } else if (in_mode == 1) {
                if (nt) { PM_ROWS2(1, 1); } else { PM_ROWS2(1, 0); }
            } else {
                if (nt) { PM_ROWS2(2, 1); } else { PM_ROWS2(2, 0); }
            }
#undef PM_ROWS2
            return;
        }
    }
    // PF 1: planes loaded one step ahead (two steps ahead measured no faster:
    // 512^3 FP32 sweep 0.387 vs 0.390 ms, back into ghost 1 0.401 vs 0.393 ms,
    // FP64 0.692 vs 0.680 ms)
    if (ry == 4) {
        launch_rows<T, 4, 1>(L, in_mode, ghosts, pi, (T*)phi_out, (const T*)rhs, c, lt, acc, C, (const T*)coarse, stream);
    } else if (ry == 2) {
        launch_rows<T, 2, 1>(L, in_mode, ghosts, pi, (T*)phi_out, (const T*)rhs, c, lt, acc, C, (const T*)coarse, stream);
    } else {
        launch_rows<T, 8, 1>(L, in_mode, ghosts, pi, (T*)phi_out, (const T*)rhs, c, lt, acc, C, (const T*)coarse, stream);
    }
}

// The multigrid driver's second pre-smoothing sweep, scratch (ghost 0) -> phi
// (ghost 1), with the residual restriction onto the coarse right-hand side in
// the same pass (k_gsrb_rows_res): the bits of pm_level_gsrb_sweep followed by
// level_restrict_nb, one pass over the level fewer.  PM_ERR_UNSUPPORTED where
// the level does not qualify or a sweep form other than the automatic one is
// chosen.
int level_sweep_restrict(const pm_layout* L, const pm_layout* C, int dtype, const void* phi_in,
                         const int in_ghost[3], void* phi_out, const int out_ghost[3], const void* rhs,
                         const int rhs_ghost[3], double dx, void* resc, const int resc_ghost[3], hipStream_t stream)
{
    PM_REQUIRE(L && C && phi_in && phi_out && rhs && resc && phi_in != phi_out, "pm_level_gsrb_sweep: bad argument");
    // 8 rows per thread where the boxes allow (512^3 in 64^3 boxes: 0.503 ms
    // against 0.644 ms with 4 rows, whose halo rows are a larger share of the
    // work and loads)
    const int RY = L->bs[1] % 8 == 0 ? 8 : 4;
    auto unsupported = [](const char* why) {
        set_error("sweep with residual restriction: %s", why);
        return PM_ERR_UNSUPPORTED;
    };
    if (g_sweep_form != 0) { return unsupported("a sweep form is forced"); }
    if (L->dim != 3 || rows_ghosts(in_ghost, out_ghost, rhs_ghost, 0) != 1 || !rows_ok(L, 4) || L->bs[1] < 4 ||
        L->bs[2] < 4 || !box_aligned(L, C)) {
        return unsupported("level shape");
    }
    for (int k = 0; k < 3; k++) {
        if (!L->periodic[k] || C->domain[k] * 2 != L->domain[k]) { return unsupported("level shape"); }
    }
    const G3 gc = g3(resc_ghost);
    const LayoutDev D = ldev(L), DC = ldev(C);
    if (gvol(D, g3(out_ghost)) * L->nboxes >= (1L << 31) || gvol(DC, gc) * C->nboxes >= (1L << 31)) {
        return unsupported("level size");
    }
    // FP32 only: the FP64 rings (phi 6 x 10 rows, rhs 3 x 8) exceed the 256
    // registers a wave of a 512-thread workgroup has, and spill
    if (dtype != PM_F32) { return unsupported("FP32 only"); }
    // half the row sweep's waves: each z chunk recomputes a plane beyond each
    // end, so longer chunks pay (the 256^3 level in 16-plane chunks: FP32-IR
    // 0.02798-0.02806 against 0.02805-0.02819 s with 8)
    const int zc = rows_zc(L, RY, kRowsWaves / 2);
    if (zc % 2) { return unsupported("odd z chunk"); }
    // relaxation coefficients and weight of sweep_rows, residual ones of level_restrict_nb
    double lam = -(dx * dx) / (4.0 * L->dim);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx), hsqinv = 1. / (dx * dx);
    const dim3 grid((unsigned)((L->domain[1] / RY) * (L->domain[2] / zc))), block((unsigned)(64 * L->nb[0]));
    {
        using T = float;
        const Lap<T> c = lap_coefs<T>(L->dim, m);
        Lap<T> cr;
        for (int j = 0; j < 7; j++) { cr.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)(-hsqinv); }
        if (RY == 8) {
            hipLaunchKernelGGL((k_gsrb_rows_res<T, 8>), grid, block, 0, stream, D, (const T*)phi_in, (T*)phi_out,
                               (const T*)rhs, c, (T)(1.0 * lam), zc, cr, (T)(1.0 / 8), DC, gc, (T*)resc);
        } else {
            hipLaunchKernelGGL((k_gsrb_rows_res<T, 4>), grid, block, 0, stream, D, (const T*)phi_in, (T*)phi_out,
                               (const T*)rhs, c, (T)(1.0 * lam), zc, cr, (T)(1.0 / 8), DC, gc, (T*)resc);
        }
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int level_gsrb_sweep_in(const pm_layout* L, int dtype, const void* phi_in, const int in_ghost[3], void* phi_out,
                        const int out_ghost[3], const void* rhs, const int rhs_ghost[3], double dx, int in_mode,
                        const pm_layout* C, const void* coarse, const int coarse_ghost[3], hipStream_t stream,
                        double* acc)
{
    PM_REQUIRE(L && phi_out && rhs && (phi_in || in_mode == 1), "pm_level_gsrb_sweep: bad argument");
    PM_REQUIRE(!acc || (dtype == PM_F32 && in_mode == 0), "pm_level_gsrb_sweep: the FP64 accumulation is for FP32 sweeps");
    PM_REQUIRE(in_mode == 1 || phi_in != phi_out, "pm_level_gsrb_sweep: the sweep is out of place (phi_in != phi_out)");
    PM_REQUIRE(in_mode >= 0 && in_mode <= 2, "pm_level_gsrb_sweep: bad input mode");
    if (L->dim != 3) {
        set_error("pm_level_gsrb_sweep: 3-D layouts only");
        return PM_ERR_UNSUPPORTED;
    }
    const int full = L->nb[0] * L->nb[1] * L->nb[2];
    if (L->nboxes != full || !L->periodic[0] || !L->periodic[1] || !L->periodic[2]) {
        set_error("pm_level_gsrb_sweep: needs a periodic layout covering its domain");
        return PM_ERR_UNSUPPORTED;
    }
    {
        const LayoutDev D = ldev(L);
        const long big = std::max(gvol(D, g3(in_ghost)), gvol(D, g3(rhs_ghost))) * L->nboxes;
        if (big >= (1L << 31)) {
            set_error("pm_level_gsrb_sweep: levels of 2^31 or more elements are not supported");
            return PM_ERR_UNSUPPORTED;
        }
    }
    if (in_mode == 2) {
        PM_REQUIRE(C && coarse, "pm_level_gsrb_sweep: the prolongation input needs the coarse level");
        const G3 gcc = g3(coarse_ghost);
        if (!box_aligned(L, C) || (long)C->nboxes * gvol(ldev(C), gcc) >= (1L << 31)) {
            set_error("pm_level_gsrb_sweep: the prolongation input needs box-aligned levels");
            return PM_ERR_UNSUPPORTED;
        }
    }
    for (int k = 0; k < 3; k++) { PM_REQUIRE(L->bs[k] >= 2, "pm_level_gsrb_sweep: boxes need 2+ cells per side"); }
    {
        // 4 rows per thread: 512^3 in 64^3 boxes swept in 0.385 ms (FP32) and
        // 0.682 ms (FP64) against 0.392 / 1.24 ms with 8 rows (FP64 spills) and
        // 0.500 / 0.729 ms with the tile sweep.  FP32 back into a ghost-1 phi
        // (the plain store form) takes 8 rows: 0.367 against 0.390 ms alone, and
        // the FP32-IR solve 0.0280-0.0281 against 0.0282-0.0283 s on one box
        const int ghosts = rows_ghosts(in_ghost, out_ghost, rhs_ghost, in_mode);
        const bool back32 = g_sweep_form == 0 && dtype == PM_F32 && ghosts == 1 && !acc && rows_ok(L, 8);
        const int ry = g_sweep_form == 3 || back32 ? 8 : (g_sweep_form == 4 ? 2 : 4);
        bool coarse_ok = in_mode != 2 || ghosts == 10;
        if (in_mode == 2) {
            // the prolongation input reads a coarse level with one ghost layer (checked box-aligned above)
            const G3 gcc = g3(coarse_ghost);
            for (int k = 0; k < 3; k++) { coarse_ok = coarse_ok && gcc.g[k] == 1; }
        }
        if (g_sweep_form != 1 && coarse_ok && ghosts > 0 && (!acc || ghosts == 1) && rows_ok(L, ry)) {
            if (dtype == PM_F64) {
                sweep_rows<double>(L, ry, ghosts, phi_in, phi_out, rhs, dx, in_mode, C, coarse, stream, acc);
            } else {
                sweep_rows<float>(L, ry, ghosts, phi_in, phi_out, rhs, dx, in_mode, C, coarse, stream, acc);
            }
            PM_CHECK_LAUNCH();
            return PM_OK;
        }
    }
    const size_t esz = dtype == PM_F64 ? 8 : 4;
    // column tiles of at most 64 cells (boxes wider than 64 split evenly), so
    // that 8-row tiles fit the LDS however wide the box is
    const int nxt = (L->bs[0] + 63) / 64, tw = (L->bs[0] + nxt - 1) / nxt;
    auto lds_of = [&](int r) {
        return (4 * (size_t)(r + 4) * (tw + 4) + 2 * (size_t)(r + 2) * (tw + 2)) * esz;
    };
    // 8-row tiles: measured fastest for both precisions on 64^3 boxes (FP32
    // 0.52 ms vs 0.61 ms at 16 rows, FP64 0.73 ms vs 0.92 ms at 16 rows, for
    // the 512^3 level); taller tiles re-read fewer halo rows but hold fewer
    // workgroups per CU
    int rows = 8;
    while (rows > 1 && lds_of(rows) > 64 * 1024 - 256) { rows /= 2; }
    const size_t lds = lds_of(rows);
    if (lds > 64 * 1024 - 256) {
        set_error("pm_level_gsrb_sweep: tiles %d cells wide do not fit the LDS ring", tw);
        return PM_ERR_UNSUPPORTED;
    }
    const int nyt = (L->bs[1] + rows - 1) / rows;
    int zc = L->bs[2];
    while (zc > kZcFloor && (long)L->nboxes * nxt * nyt * ((L->bs[2] + zc - 1) / zc) < 2048) { zc = (zc + 1) / 2; }
    const G3 gi = g3(in_ghost), go = g3(out_ghost), gr = g3(rhs_ghost), gc = g3(coarse_ghost);
    const LayoutDev D = ldev(L);
    const LayoutDev DC = in_mode == 2 ? ldev(C) : D;
    dim3 grid((unsigned)(nxt * nyt * ((L->bs[2] + zc - 1) / zc)), L->nboxes);
    double lam = -(dx * dx) / (4.0 * L->dim);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx);
    PM_DISPATCH_T(dtype, {
        Lap<T> c = lap_coefs<T>(L->dim, m);
        if (dtype == PM_F64) {
            for (int j = 0; j < 7; j++) { c.c[j] = (T)((j == 0 ? -2.0 * L->dim : 1.0) * m); }
        }
        const long np = (long)(rows + 4) * (tw + 4);
        const T* pi = (const T*)(in_mode == 1 ? phi_out : phi_in);
        const T* cr = (const T*)coarse;
        const T lt = (T)(1.0 * lam);
        if (np <= 1024) {
            PM_SWEEP_MODES(4)
        } else if (np <= 2048) {
            PM_SWEEP_MODES(8)
        } else {
            PM_SWEEP_MODES(16)
        }
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

#undef PM_SWEEP_MODES
#undef PM_SWEEP_LAUNCH

bool levels_box_aligned(const pm_layout* F, const pm_layout* C) { return box_aligned(F, C); }

int level_gsrb_small(const pm_layout* L, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                     const int rhs_ghost[3], double dx, int nsweeps, int in_mode, const pm_layout* C,
                     const void* coarse, const int coarse_ghost[3], hipStream_t stream)
{
    PM_REQUIRE(L && phi && rhs && nsweeps >= 0 && in_mode >= 0 && in_mode <= 2, "pm_level_gsrb: bad argument");
    const long N = (long)L->bs[0] * L->bs[1] * L->bs[2];
    bool ok = L->dim == 3 && L->nboxes == 1 && L->nb[0] * L->nb[1] * L->nb[2] == 1 && N <= GS_MAX &&
              L->periodic[0] && L->periodic[1] && L->periodic[2] &&
              (size_t)GS_MAX * (dtype == PM_F64 ? 8 : 4) <= 64 * 1024 - 256;
    if (in_mode == 2) {
        ok = ok && C && coarse && C->nboxes == 1 && C->nb[0] * C->nb[1] * C->nb[2] == 1;
        for (int k = 0; k < 3 && ok; k++) { ok = C->bs[k] * 2 == L->bs[k]; }
    }
    if (!ok) {
        set_error("pm_level_gsrb: not a small one-box periodic level");
        return PM_ERR_UNSUPPORTED;
    }
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost), gc = g3(coarse_ghost);
    const LayoutDev D = ldev(L);
    const LayoutDev DC = in_mode == 2 ? ldev(C) : D;
    double lam = -(dx * dx) / (4.0 * L->dim);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx);
    PM_DISPATCH_T(dtype, {
        Lap<T> c = lap_coefs<T>(L->dim, m);
        if (dtype == PM_F64) {
            for (int j = 0; j < 7; j++) { c.c[j] = (T)((j == 0 ? -2.0 * L->dim : 1.0) * m); }
        }
        const T lt = (T)(1.0 * lam);
        if (in_mode == 0) {
            hipLaunchKernelGGL((k_gsrb_small<T, 0>), dim3(1), dim3(256), 0, stream, D, gp, gr, (T*)phi,
                               (const T*)rhs, c, lt, nsweeps, DC, gc, (const T*)coarse);
        } else if (in_mode == 1) {
            hipLaunchKernelGGL((k_gsrb_small<T, 1>), dim3(1), dim3(256), 0, stream, D, gp, gr, (T*)phi,
                               (const T*)rhs, c, lt, nsweeps, DC, gc, (const T*)coarse);
        } else {
            hipLaunchKernelGGL((k_gsrb_small<T, 2>), dim3(1), dim3(256), 0, stream, D, gp, gr, (T*)phi,
                               (const T*)rhs, c, lt, nsweeps, DC, gc, (const T*)coarse);
        }
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // namespace pm

extern "C" {

int pm_level_sweep_form(int form)
{
    PM_REQUIRE(form >= 0 && form <= 4, "pm_level_sweep_form: form must be 0..4");
    g_sweep_form = form;
    return PM_OK;
}

int pm_level_sweep_tune(int planes_max, int nt_store)
{
    PM_REQUIRE(planes_max >= 0 && nt_store >= -1 && nt_store <= 1, "pm_level_sweep_tune: bad argument");
    g_rows_zc_max = planes_max;
    g_rows_nt = nt_store;
    return PM_OK;
}

int pm_level_gsrb_sweep(const pm_layout* L, int dtype, const void* phi_in, const int in_ghost[3], void* phi_out,
                        const int out_ghost[3], const void* rhs, const int rhs_ghost[3], double dx, pm_stream_t stream)
{
    PM_REQUIRE(phi_in, "pm_level_gsrb_sweep: bad argument");
    return level_gsrb_sweep_in(L, dtype, phi_in, in_ghost, phi_out, out_ghost, rhs, rhs_ghost, dx, 0, nullptr, nullptr,
                               nullptr, S(stream), nullptr);
}

int pm_level_jacobi(const pm_layout* L, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                    const int rhs_ghost[3], void* tmp, double dx, pm_stream_t stream)
{
    PM_REQUIRE(L && phi && rhs && tmp, "pm_level_jacobi: bad argument");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    dim3 grid(lvl_blocks(n, L->nboxes), L->nboxes);
    // pointRelax (LevelMultigrid.cpp:99-123): weight 1/(4 DIM), lambda = dx^2/(4 DIM)
    const double weight = 1.0 / (4 * L->dim);
    const double lambda = dx * dx / (4 * L->dim);
    PM_DISPATCH_T(dtype, {
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = 0; }
        for (int j = 0; j < 2 * L->dim + 1; j++) { c.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)weight; }
        if (L->dim == 3) {
            hipLaunchKernelGGL((k_jacobi_tmp<T, 3>), grid, dim3(256), 0, S(stream), D, gp, gr, (const T*)phi,
                               (const T*)rhs, (T*)tmp, c, (T)(-lambda));
        } else {
            hipLaunchKernelGGL((k_jacobi_tmp<T, 2>), grid, dim3(256), 0, S(stream), D, gp, gr, (const T*)phi,
                               (const T*)rhs, (T*)tmp, c, (T)(-lambda));
        }
        hipLaunchKernelGGL(k_jacobi_add<T>, grid, dim3(256), 0, S(stream), D, gp, (T*)phi, (const T*)tmp);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_resnorm(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                     const int rhs_ghost[3], double dx, void* out_dev, int combine, pm_stream_t stream)
{
    PM_REQUIRE(L && phi && rhs && out_dev, "pm_level_resnorm: bad argument");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    unsigned bpb = res_blocks(n, L->nboxes);
    dim3 grid(bpb, L->nboxes);
    int zc = 0;
    const bool z = rz_ok(L, gp);
    if (z) {
        grid = rz_grid(L, &zc);
        bpb = grid.x;
    }
    const double hsqinv = 1. / (dx * dx);
    void* ws = nullptr;
    PM_DISPATCH_T(dtype, {
        int rc = workspace(S(stream), sizeof(T) * (size_t)bpb * L->nboxes, &ws);
        if (rc) { return rc; }
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = 0; }
        for (int j = 0; j < 2 * L->dim + 1; j++) { c.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)hsqinv; }
        if (z) {
            hipLaunchKernelGGL((k_residual_z<T, T, 1>), grid, dim3(256), 0, S(stream), D, gp, gr, gr, (const T*)phi,
                               (const T*)rhs, (T*)nullptr, c, (T*)ws, zc);
        } else if (L->dim == 3) {
            hipLaunchKernelGGL((k_resnorm<T, 3>), grid, dim3(256), 0, S(stream), D, gp, gr, (const T*)phi,
                               (const T*)rhs, c, (T*)ws);
        } else {
            hipLaunchKernelGGL((k_resnorm<T, 2>), grid, dim3(256), 0, S(stream), D, gp, gr, (const T*)phi,
                               (const T*)rhs, c, (T*)ws);
        }
        hipLaunchKernelGGL(k_max_partials<T>, dim3(1), dim3(1024), 0, S(stream), (const T*)ws,
                           (long)bpb * L->nboxes, (T*)out_dev, combine);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_restrict_residual(const pm_layout* F, const pm_layout* C, int dtype, const void* phi,
                               const int phi_ghost[3], const void* rhs, const int rhs_ghost[3], double dx,
                               void* resc, const int resc_ghost[3], pm_stream_t stream)
{
    PM_REQUIRE(F && C && phi && rhs && resc, "pm_level_restrict_residual: bad argument");
    PM_REQUIRE(F->dim == C->dim, "pm_level_restrict_residual: dim mismatch");
    for (int k = 0; k < F->dim; k++) {
        PM_REQUIRE(C->domain[k] * 2 == F->domain[k], "pm_level_restrict_residual: coarse domain is not fine/2");
        PM_REQUIRE(F->bs[k] % 2 == 0, "pm_level_restrict_residual: fine boxes must have even size");
    }
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost), gc = g3(resc_ghost);
    const LayoutDev DF = ldev(F), DC = ldev(C);
    const long n = (long)C->bs[0] * C->bs[1] * C->bs[2];
    dim3 grid(lvl_blocks(n, C->nboxes), C->nboxes);
    const double hsqinv = 1. / (dx * dx);
    PM_DISPATCH_T(dtype, {
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = 0; }
        for (int j = 0; j < 2 * F->dim + 1; j++) { c.c[j] = (T)(j == 0 ? -2.0 * F->dim : 1.0) * (T)(-hsqinv); }
        const T avg = (T)(1.0 / (F->dim == 3 ? 8 : 4));
        if (box_aligned(F, C)) {
            const long nh = (long)(F->bs[0] / 2) * (F->bs[1] / 2) * (F->bs[2] / 2);
            hipLaunchKernelGGL((k_restrict_residual_box<T>), dim3(lvl_blocks(nh, F->nboxes), F->nboxes), dim3(256), 0,
                               S(stream), DF, DC, gp, gr, gc, (const T*)phi, (const T*)rhs, (T*)resc, c, avg);
        } else if (F->dim == 3) {
            hipLaunchKernelGGL((k_restrict_residual<T, 3>), grid, dim3(256), 0, S(stream), DF, DC, gp, gr, gc,
                               (const T*)phi, (const T*)rhs, (T*)resc, c, avg);
        } else {
            hipLaunchKernelGGL((k_restrict_residual<T, 2>), grid, dim3(256), 0, S(stream), DF, DC, gp, gr, gc,
                               (const T*)phi, (const T*)rhs, (T*)resc, c, avg);
        }
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_prolong_add(const pm_layout* F, const pm_layout* C, int dtype, const void* delta, const int delta_ghost[3],
                         void* phi, const int phi_ghost[3], int linear, pm_stream_t stream)
{
    PM_REQUIRE(F && C && delta && phi, "pm_level_prolong_add: bad argument");
    PM_REQUIRE(F->dim == C->dim, "pm_level_prolong_add: dim mismatch");
    for (int k = 0; k < F->dim; k++) {
        PM_REQUIRE(C->domain[k] * 2 == F->domain[k], "pm_level_prolong_add: coarse domain is not fine/2");
        PM_REQUIRE(!linear || C->periodic[k], "pm_level_prolong_add: Linear needs a periodic coarse level");
    }
    const G3 gd = g3(delta_ghost), gp = g3(phi_ghost);
    const LayoutDev DF = ldev(F), DC = ldev(C);
    const long n = (long)F->bs[0] * F->bs[1] * F->bs[2];
    dim3 grid(lvl_blocks(n, F->nboxes), F->nboxes);
    PM_DISPATCH_T(dtype, {
        if (F->dim == 3) {
            if (linear) {
                hipLaunchKernelGGL((k_prolong_add<T, 3, 1>), grid, dim3(256), 0, S(stream), DF, DC, gd, gp,
                                   (const T*)delta, (T*)phi);
            } else if (box_aligned(F, C)) {
                // about 8192 workgroups over the level, each walking many rows: one
                // 4-row pass carries too little data to amortise a workgroup
                const long rows = (long)F->bs[1] * F->bs[2];
                const long per = std::max<long>(1, std::min<long>((rows + 3) / 4, 8192 / F->nboxes));
                hipLaunchKernelGGL((k_prolong_const_box<T>), dim3((unsigned)per, F->nboxes), dim3(256), 0,
                                   S(stream), DF, DC, gd, gp, (const T*)delta, (T*)phi);
            } else {
                hipLaunchKernelGGL((k_prolong_add<T, 3, 0>), grid, dim3(256), 0, S(stream), DF, DC, gd, gp,
                                   (const T*)delta, (T*)phi);
            }
        } else {
            if (linear) {
                hipLaunchKernelGGL((k_prolong_add<T, 2, 1>), grid, dim3(256), 0, S(stream), DF, DC, gd, gp,
                                   (const T*)delta, (T*)phi);
            } else {
                hipLaunchKernelGGL((k_prolong_add<T, 2, 0>), grid, dim3(256), 0, S(stream), DF, DC, gd, gp,
                                   (const T*)delta, (T*)phi);
            }
        }
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_setval(const pm_layout* L, int dtype, void* data, const int ghost[3], int ncomp, double value,
                    pm_stream_t stream)
{
    PM_REQUIRE(L && data && ncomp >= 1, "pm_level_setval: bad argument");
    const long n = gvol(ldev(L), g3(ghost)) * ncomp * L->nboxes;
    PM_DISPATCH_T(dtype, hipLaunchKernelGGL(k_fill<T>, dim3(bx_blocks(n, 1 << 14)), dim3(256), 0, S(stream), (T*)data,
                                            n, (T)value));
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_copy(const pm_layout* SL, const void* src, const int src_ghost[3], const pm_layout* DL, void* dst,
                  const int dst_ghost[3], int dtype, int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(SL && DL && src && dst && ncomp >= 1, "pm_level_copy: bad argument");
    PM_REQUIRE(same_layout_shape(SL, DL), "pm_level_copy: layouts of different domains");
    const LayoutDev DS = ldev(SL), DD = ldev(DL);
    const long n = (long)DL->bs[0] * DL->bs[1] * DL->bs[2];
    dim3 grid(lvl_blocks(n, DL->nboxes), DL->nboxes);
    PM_DISPATCH_T(dtype, hipLaunchKernelGGL(k_level_copy<T>, grid, dim3(256), 0, S(stream), DS, g3(src_ghost),
                                            (const T*)src, DD, g3(dst_ghost), (T*)dst, ncomp));
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_copy_to(const pm_layout* SL, const void* src, const int src_ghost[3], const pm_layout* DL, void* dst,
                     const int dst_ghost[3], int dtype, int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(SL && DL && src && dst && ncomp >= 1 && src_ghost && dst_ghost, "pm_level_copy_to: bad argument");
    PM_REQUIRE(same_layout_shape(SL, DL), "pm_level_copy_to: layouts of different domains");
    PM_REQUIRE(src != dst, "pm_level_copy_to: source and destination are the same array");
    // Proto's DisjointBoxLayout::compatible plus equal box sizes: same domain,
    // same patches in the same order
    bool compatible = SL->nboxes == DL->nboxes && SL->lows == DL->lows;
    for (int k = 0; k < 3; k++) {
        compatible = compatible && SL->bs[k] == DL->bs[k] && SL->periodic[k] == DL->periodic[k];
    }
    const LayoutDev DS = ldev(SL), DD = ldev(DL);
    const G3 gs = g3(src_ghost), gd = g3(dst_ghost);
    const long n = gvol(DD, gd);
    dim3 grid(lvl_blocks(n, DL->nboxes), DL->nboxes);
    const int mode = compatible ? 1 : 0;
    PM_DISPATCH_T(dtype, hipLaunchKernelGGL(k_level_copy_to<T>, grid, dim3(256), 0, S(stream), DS, gs,
                                            (const T*)src, DD, gd, (T*)dst, ncomp, mode));
    PM_CHECK_LAUNCH();
    if (compatible) { return pm_level_exchange(DL, dtype, dst, dst_ghost, ncomp, stream); }
    return PM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Residual into any dtype with max|r|, and the FP64 += FP32 correction.
namespace pm {

template<typename T, typename TO, int DIMN>
__global__ __launch_bounds__(256) void k_residual(LayoutDev L, G3 gp, G3 gr, G3 go, const T* __restrict__ phi,
                                                  const T* __restrict__ rhs, TO* __restrict__ out, Lap<T> c,
                                                  T* __restrict__ partials)
{
    constexpr int NC = 2 * DIMN + 1;
    __shared__ T red[16];
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const T* ph = phi + (long)b * gvol(L, gp);
    const T* rh = rhs + (long)b * gvol(L, gr);
    TO* ot = out + (long)b * gvol(L, go);
    const long X = L.bs[0] + 2 * gp.g[0], XY = X * (L.bs[1] + 2 * gp.g[1]);
    T m = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gp, x, y, z);
        T s = 0;
        s = s + rh[rel_index(L, gr, x, y, z)];
#pragma unroll
        for (int j = 0; j < NC; j++) {
            s = s + c.c[j] * ph[i + lapo(j, 0) + lapo(j, 1) * X + lapo(j, 2) * XY];
        }
        ot[rel_index(L, go, x, y, z)] = (TO)s;
        const T a = s > 0 ? s : -s;
        m = m < a ? a : m;
    }
    m = block_reduce(PM_RED_ABS, m, red);
    if (threadIdx.x == 0) { partials[(long)b * gridDim.x + blockIdx.x] = m; }
}

template<typename T, typename TE>
__global__ __launch_bounds__(256) void k_add_correction(LayoutDev L, G3 gp, G3 ge, T* __restrict__ phi,
                                                        const TE* __restrict__ e)
{
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    T* ph = phi + (long)b * gvol(L, gp);
    const TE* ee = e + (long)b * gvol(L, ge);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const long i = rel_index(L, gp, x, y, z);
        ph[i] = ph[i] + (T)ee[rel_index(L, ge, x, y, z)];
    }
}

}  // namespace pm

extern "C" {

int pm_level_residual(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                      const int rhs_ghost[3], double dx, int out_dtype, void* out, const int out_ghost[3],
                      void* norm_out, pm_stream_t stream)
{
    PM_REQUIRE(L && phi && rhs && out && norm_out, "pm_level_residual: bad argument");
    PM_REQUIRE(dtype == PM_F64 || (dtype == PM_F32 && out_dtype == PM_F32), "pm_level_residual: unsupported dtypes");
    PM_REQUIRE(out_dtype == PM_F32 || out_dtype == PM_F64, "pm_level_residual: bad out dtype");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost), go = g3(out_ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    unsigned bpb = res_blocks(n, L->nboxes);
    dim3 grid(bpb, L->nboxes);
    int zc = 0;
    const bool z = rz_ok(L, gp);
    if (z) {
        grid = rz_grid(L, &zc);
        bpb = grid.x;
    }
    const double hsqinv = 1. / (dx * dx);
    void* ws = nullptr;
    PM_DISPATCH_T(dtype, {
        int rc = workspace(S(stream), sizeof(T) * (size_t)bpb * L->nboxes, &ws);
        if (rc) { return rc; }
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = 0; }
        for (int j = 0; j < 2 * L->dim + 1; j++) { c.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)(-hsqinv); }
        if (z && out_dtype == PM_F32) {
            hipLaunchKernelGGL((k_residual_z<T, float, 0>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                               (const T*)phi, (const T*)rhs, (float*)out, c, (T*)ws, zc);
        } else if (z) {
            hipLaunchKernelGGL((k_residual_z<T, double, 0>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                               (const T*)phi, (const T*)rhs, (double*)out, c, (T*)ws, zc);
        } else if (out_dtype == PM_F32) {
            if (L->dim == 3) {
                hipLaunchKernelGGL((k_residual<T, float, 3>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                                   (const T*)phi, (const T*)rhs, (float*)out, c, (T*)ws);
            } else {
                hipLaunchKernelGGL((k_residual<T, float, 2>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                                   (const T*)phi, (const T*)rhs, (float*)out, c, (T*)ws);
            }
        } else {
            if (L->dim == 3) {
                hipLaunchKernelGGL((k_residual<T, double, 3>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                                   (const T*)phi, (const T*)rhs, (double*)out, c, (T*)ws);
            } else {
                hipLaunchKernelGGL((k_residual<T, double, 2>), grid, dim3(256), 0, S(stream), D, gp, gr, go,
                                   (const T*)phi, (const T*)rhs, (double*)out, c, (T*)ws);
            }
        }
        hipLaunchKernelGGL(k_max_partials<T>, dim3(1), dim3(1024), 0, S(stream), (const T*)ws,
                           (long)bpb * L->nboxes, (T*)norm_out, 0);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // extern "C"

namespace pm {

// Levels whose residual kernels can stage halos from the neighbouring boxes
// (k_residual_nb): 32..64-wide boxes as k_residual_z takes, periodic in every
// direction, every box present, fewer than 2^31 elements.
static bool nb_ok(const pm_layout* L, const G3& gp)
{
    if (!rz_ok(L, gp) || L->nboxes != L->nb[0] * L->nb[1] * L->nb[2]) { return false; }
    for (int k = 0; k < 3; k++) {
        if (!L->periodic[k]) { return false; }
    }
    return gvol(ldev(L), gp) * L->nboxes < (1L << 31);
}

int level_resnorm_nb(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                     const int rhs_ghost[3], double dx, void* out_dev, int combine, hipStream_t stream)
{
    PM_REQUIRE(L && phi && rhs && out_dev, "pm_level_resnorm: bad argument");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost);
    if (!nb_ok(L, gp)) {
        set_error("pm_level_resnorm: this level needs its ghost cells exchanged");
        return PM_ERR_UNSUPPORTED;
    }
    const LayoutDev D = ldev(L);
    int zc = 0;
    const dim3 grid = rz_grid(L, &zc);
    const double hsqinv = 1. / (dx * dx);
    void* ws = nullptr;
    PM_DISPATCH_T(dtype, {
        int rc = workspace(stream, sizeof(T) * (size_t)grid.x * L->nboxes, &ws);
        if (rc) { return rc; }
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)hsqinv; }
        hipLaunchKernelGGL((k_residual_nb<T, T, 1>), grid, dim3(256), 0, stream, D, gp, gr, gr, (const T*)phi,
                           (const T*)rhs, (T*)nullptr, c, (T*)ws, zc, D, gp, (T*)nullptr, (T)0);
        hipLaunchKernelGGL(k_max_partials<T>, dim3(1), dim3(1024), 0, stream, (const T*)ws,
                           (long)grid.x * L->nboxes, (T*)out_dev, combine);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int level_residual_nb(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                      const int rhs_ghost[3], double dx, int out_dtype, void* out, const int out_ghost[3],
                      void* norm_out, hipStream_t stream)
{
    PM_REQUIRE(L && phi && rhs && out && norm_out, "pm_level_residual: bad argument");
    PM_REQUIRE(dtype == PM_F64 || (dtype == PM_F32 && out_dtype == PM_F32), "pm_level_residual: unsupported dtypes");
    PM_REQUIRE(out_dtype == PM_F32 || out_dtype == PM_F64, "pm_level_residual: bad out dtype");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost), go = g3(out_ghost);
    if (!nb_ok(L, gp)) {
        set_error("pm_level_residual: this level needs its ghost cells exchanged");
        return PM_ERR_UNSUPPORTED;
    }
    const LayoutDev D = ldev(L);
    int zc = 0;
    const dim3 grid = rz_grid(L, &zc);
    const double hsqinv = 1. / (dx * dx);
    void* ws = nullptr;
    PM_DISPATCH_T(dtype, {
        int rc = workspace(stream, sizeof(T) * (size_t)grid.x * L->nboxes, &ws);
        if (rc) { return rc; }
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = (T)(j == 0 ? -2.0 * L->dim : 1.0) * (T)(-hsqinv); }
        if (out_dtype == PM_F32) {
            hipLaunchKernelGGL((k_residual_nb<T, float, 0>), grid, dim3(256), 0, stream, D, gp, gr, go,
                               (const T*)phi, (const T*)rhs, (float*)out, c, (T*)ws, zc, D, gp, (T*)nullptr, (T)0);
        } else {
            hipLaunchKernelGGL((k_residual_nb<T, double, 0>), grid, dim3(256), 0, stream, D, gp, gr, go,
                               (const T*)phi, (const T*)rhs, (double*)out, c, (T*)ws, zc, D, gp, (T*)nullptr, (T)0);
        }
        hipLaunchKernelGGL(k_max_partials<T>, dim3(1), dim3(1024), 0, stream, (const T*)ws,
                           (long)grid.x * L->nboxes, (T*)norm_out, 0);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int level_restrict_nb(const pm_layout* F, const pm_layout* C, int dtype, const void* phi, const int phi_ghost[3],
                      const void* rhs, const int rhs_ghost[3], double dx, void* resc, const int resc_ghost[3],
                      hipStream_t stream)
{
    PM_REQUIRE(F && C && phi && rhs && resc, "pm_level_restrict_residual: bad argument");
    const G3 gp = g3(phi_ghost), gr = g3(rhs_ghost), gc = g3(resc_ghost);
    if (!nb_ok(F, gp) || !box_aligned(F, C) || F->bs[0] % 2 || F->bs[1] % 2 || F->bs[2] % 2) {
        set_error("pm_level_restrict_residual: this level needs its ghost cells exchanged");
        return PM_ERR_UNSUPPORTED;
    }
    for (int k = 0; k < 3; k++) {
        PM_REQUIRE(C->domain[k] * 2 == F->domain[k], "pm_level_restrict_residual: coarse domain is not fine/2");
    }
    const LayoutDev DF = ldev(F), DC = ldev(C);
    int zc = 0;
    const dim3 grid = rz_grid(F, &zc);
    if (zc % 2) {
        set_error("pm_level_restrict_residual: odd z chunk");
        return PM_ERR_UNSUPPORTED;
    }
    const double hsqinv = 1. / (dx * dx);
    PM_DISPATCH_T(dtype, {
        Lap<T> c;
        for (int j = 0; j < 7; j++) { c.c[j] = (T)(j == 0 ? -2.0 * F->dim : 1.0) * (T)(-hsqinv); }
        const T avg = (T)(1.0 / 8);
        hipLaunchKernelGGL((k_residual_nb<T, T, 2>), grid, dim3(256), 0, stream, DF, gp, gr, gr, (const T*)phi,
                           (const T*)rhs, (T*)nullptr, c, (T*)nullptr, zc, DC, gc, (T*)resc, avg);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // namespace pm

extern "C" {

int pm_level_add_correction(const pm_layout* L, int dtype, void* phi, const int phi_ghost[3], int e_dtype,
                            const void* e, const int e_ghost[3], pm_stream_t stream)
{
    PM_REQUIRE(L && phi && e, "pm_level_add_correction: bad argument");
    const G3 gp = g3(phi_ghost), ge = g3(e_ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    dim3 grid(lvl_blocks(n, L->nboxes), L->nboxes);
    if (dtype == PM_F64 && e_dtype == PM_F32) {
        hipLaunchKernelGGL((k_add_correction<double, float>), grid, dim3(256), 0, S(stream), D, gp, ge, (double*)phi,
                           (const float*)e);
    } else if (dtype == PM_F64 && e_dtype == PM_F64) {
        hipLaunchKernelGGL((k_add_correction<double, double>), grid, dim3(256), 0, S(stream), D, gp, ge,
                           (double*)phi, (const double*)e);
    } else if (dtype == PM_F32 && e_dtype == PM_F32) {
        hipLaunchKernelGGL((k_add_correction<float, float>), grid, dim3(256), 0, S(stream), D, gp, ge, (float*)phi,
                           (const float*)e);
    } else {
        PM_REQUIRE(false, "pm_level_add_correction: unsupported dtypes");
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Level reductions and separable initialisation.
namespace pm {

template<typename T>
__global__ __launch_bounds__(256) void k_level_reduce1(LayoutDev L, G3 g, const T* __restrict__ data, int ncomp,
                                                       int comp, int op, T* __restrict__ partials)
{
    __shared__ T red[16];
    const int b = blockIdx.y;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    const T* d = data + ((long)b * ncomp + comp) * gvol(L, g);
    T v = red_init<T>(op);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        v = red_update(op, v, d[rel_index(L, g, x, y, z)]);
    }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { partials[(long)b * gridDim.x + blockIdx.x] = v; }
}

template<typename T>
__global__ __launch_bounds__(1024) void k_level_reduce2(const T* __restrict__ partials, long n, int op, T* out,
                                                        int combine)
{
    __shared__ T red[16];
    T v = red_init<T>(op);
    for (long i = threadIdx.x; i < n; i += blockDim.x) { v = red_combine(op, v, partials[i]); }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { *out = combine ? red_combine(op, *out, v) : v; }
}

template<typename T>
__global__ __launch_bounds__(256) void k_add_separable(LayoutDev L, G3 g, T* __restrict__ data, double coef,
                                                       const double* __restrict__ tx, const double* __restrict__ ty,
                                                       const double* __restrict__ tz)
{
    const int b = blockIdx.y;
    const int* lo = L.lows + 3 * b;
    const long n = (long)L.bs[0] * L.bs[1] * L.bs[2];
    T* d = data + (long)b * gvol(L, g);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % L.bs[0]);
        const long r = t / L.bs[0];
        const int y = (int)(r % L.bs[1]);
        const int z = (int)(r / L.bs[1]);
        const double v = (tx[lo[0] + x] * ty[lo[1] + y]) * tz[lo[2] + z];
        const long i = rel_index(L, g, x, y, z);
        d[i] = (T)((double)d[i] + coef * v);
    }
}

}  // namespace pm

extern "C" {

int pm_level_reduce(const pm_layout* L, int dtype, const void* data, const int ghost[3], int ncomp, int comp, int op,
                    void* out_dev, int combine, pm_stream_t stream)
{
    PM_REQUIRE(L && data && out_dev && ncomp >= 1 && comp >= 0 && comp < ncomp, "pm_level_reduce: bad argument");
    PM_REQUIRE(op >= PM_RED_MAX && op <= PM_RED_SUMSQ, "pm_level_reduce: bad op %d", op);
    const G3 g = g3(ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    const unsigned bpb = bx_blocks(n, 64);
    dim3 grid(bpb, L->nboxes);
    void* ws = nullptr;
    PM_DISPATCH_T(dtype, {
        int rc = workspace(S(stream), sizeof(T) * (size_t)bpb * L->nboxes, &ws);
        if (rc) { return rc; }
        hipLaunchKernelGGL(k_level_reduce1<T>, grid, dim3(256), 0, S(stream), D, g, (const T*)data, ncomp, comp, op,
                           (T*)ws);
        hipLaunchKernelGGL(k_level_reduce2<T>, dim3(1), dim3(1024), 0, S(stream), (const T*)ws,
                           (long)bpb * L->nboxes, op, (T*)out_dev, combine);
    });
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_level_add_separable(const pm_layout* L, int dtype, void* data, const int ghost[3], double coef,
                           const double* tx, const double* ty, const double* tz, pm_stream_t stream)
{
    PM_REQUIRE(L && data && tx && ty && tz, "pm_level_add_separable: null argument");
    const size_t nt = (size_t)L->domain[0] + L->domain[1] + L->domain[2];
    double* dt = nullptr;
    PM_CHECK_HIP(hipMalloc(&dt, sizeof(double) * nt));
    hipMemcpy(dt, tx, sizeof(double) * L->domain[0], hipMemcpyHostToDevice);
    hipMemcpy(dt + L->domain[0], ty, sizeof(double) * L->domain[1], hipMemcpyHostToDevice);
    hipMemcpy(dt + L->domain[0] + L->domain[1], tz, sizeof(double) * L->domain[2], hipMemcpyHostToDevice);
    const G3 g = g3(ghost);
    const LayoutDev D = ldev(L);
    const long n = (long)L->bs[0] * L->bs[1] * L->bs[2];
    dim3 grid(lvl_blocks(n, L->nboxes), L->nboxes);
    if (dtype == PM_F64) {
        hipLaunchKernelGGL(k_add_separable<double>, grid, dim3(256), 0, S(stream), D, g, (double*)data, coef, dt,
                           dt + L->domain[0], dt + L->domain[0] + L->domain[1]);
    } else {
        hipLaunchKernelGGL(k_add_separable<float>, grid, dim3(256), 0, S(stream), D, g, (float*)data, coef, dt,
                           dt + L->domain[0], dt + L->domain[0] + L->domain[1]);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(S(stream));
    hipFree(dt);
    if (le != hipSuccess || se != hipSuccess) {
        set_error("pm_level_add_separable: %s", hipGetErrorString(le != hipSuccess ? le : se));
        return PM_ERR_HIP;
    }
    return PM_OK;
}

}  // extern "C"
