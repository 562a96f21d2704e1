// Batches of independent patches of uneven size (AMR boxes): one persistent
// launch per operator for the whole batch.
//
// Proto applies an operator to a set of boxes as one launch per box
// (`for (auto iter : layout) ...`, e.g. Proto_LevelMultigridSolverImplem.H:
// 366-378).  Here a fixed grid of workgroups pulls tiles of every box from an
// atomic work queue (largest tiles first); the queue counters reset
// themselves at the end of each launch, so the launch can be captured in a
// hipGraph and replayed without a memset node.
//
// Storage: box b (valid extents n = sizes[b], g ghost layers) at
// ghosted_off[b] in the ghosted array (x fastest over [-g, n+g)) and at
// valid_off[b] in the valid-cells array (x fastest over [0, n)).
//
// Operators (bit-identical per box to the single-box paths):
//   apply: dst = sum_j c'_j src(p + o_j), Proto's coefficient order, for the
//          7-point Laplacian() and 27-point Laplacian_27() shapes;
//   gsrb:  one red-black Gauss-Seidel sweep with the ghosts held fixed (no
//          exchange between independent patches): every red cell ((x+y+z)
//          even, box-local indices) updated from old black neighbours, then
//          every black cell from the new red ones -- the arithmetic of
//          LaplaceOp.H:95-123.  Done in one pass, out of place (phi_in is
//          only read, phi_out only written): walking z, a workgroup updates
//          red cells of plane k+1 and then black cells of plane k in LDS,
//          recomputing the red cells of the rows just outside its tile
//          (identical values to the neighbouring tile's) so that no tile
//          waits for another.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pm_internal.hpp"

struct pm_batch {
    int nboxes, ghost;
    std::vector<int> sizes;
    std::vector<size_t> goff, voff;  // nboxes + 1
    void* d_boxes = nullptr;         // BatchBox[nboxes]
    void* d_items = nullptr;         // BatchItem[] (apply)
    int nitems_apply = 0, apply_rows = 0, apply_tz = 0;
    void* d_gitems = nullptr;        // BatchItem[] (gsrb)
    int nitems_gsrb = 0, gsrb_rows = 0;
    unsigned* d_counter = nullptr;   // 2 x (queue head, blocks done) per operator
    pm_batch_config cfg{0, 4, 8, 4};
    // the row-register sweep: every box 2..64 cells wide (rows_ok), its items
    // built for rows_ry rows per sub-tile; sweep_form as pm_batch_sweep_form
    bool rows_ok = false;
    void* d_ritems = nullptr;        // RowItem[] (row sweep)
    int nitems_rows = 0, rows_ry = 0, rows_wpb = 0, sweep_form = 0, sweep_waves = 0;
    // the two-cells-per-lane row sweep: every box even and 4..128 wide, even ghost
    bool pairs_ok = false;
    void* d_pitems = nullptr;        // RowItem[] (pair sweep)
    int nitems_pairs = 0, pairs_ry = 0, pairs_wpb = 0;
    // the four-cells-per-lane FP32 row sweep: ghost 2, every box 4..252 wide in
    // multiples of 4
    bool quads_ok = false;
    void* d_qitems = nullptr;        // RowItem[] (quad sweep, 2-row sub-tiles)
    int nitems_quads = 0;
};

namespace pm {

struct BatchBox {
    int nx, ny, nz, pad;
    long goff, voff;
};

struct BatchItem {
    int box, x0, y0, z0, z1, rows, xw, pad;
};

template<typename T>
struct BCoefs {
    T c[27];
};

// ------------------------------------------------------------ work queue
__device__ inline int next_item(unsigned* counter, int* s_item)
{
    __syncthreads();
    if (threadIdx.x == 0) { *s_item = (int)atomicAdd(&counter[0], 1u); }
    __syncthreads();
    return *s_item;
}

__device__ inline void queue_done(unsigned* counter)
{
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned done = atomicAdd(&counter[1], 1u);
        if (done == gridDim.x - 1) {
            // every workgroup has drawn its last item: reset for the next launch
            atomicExch(&counter[0], 0u);
            atomicExch(&counter[1], 0u);
            __threadfence();
        }
    }
}

// ------------------------------------------------------------ batched apply
// A tile is XW (<= 64) columns x TR thread-rows x BA_RY rows per thread of one
// box over a z range: thread t owns column t % XW and rows
// y0 + (t / XW) BA_RY + [0, BA_RY).  Each plane (tile + 1-cell halo, at most
// BA_NP elements) is staged once in LDS; each thread keeps the 3 x 3
// neighbourhoods of its rows for planes z-1, z, z+1 in registers, so one LDS
// pass per plane feeds all taps of BA_RY outputs.  Three LDS planes, three
// register planes and three staging sets rotate with a 3-way unrolled z loop.
// Every load is issued unconditionally (offsets clamped into the ghosted box;
// clamped elements only feed outputs outside the box, which are not stored),
// and a step's results are stored at the start of the next step, after its
// loads: the only waits are at the LDS commits and cover operations a whole
// step old.
constexpr int BA_RY = 4, BA_NL = 5, BA_NP = 256 * BA_NL;

template<typename T, int SHAPE, int NT>
__global__ __launch_bounds__(256) void k_batch_apply(const BatchBox* __restrict__ boxes,
                                                     const BatchItem* __restrict__ items, int nitems, int g,
                                                     const T* __restrict__ src, T* __restrict__ dst, BCoefs<T> c,
                                                     unsigned* counter)
{
    constexpr int NC = SHAPE == 7 ? 7 : 27, RY = BA_RY, NL = BA_NL;
    __shared__ T sm[3][BA_NP];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    for (;;) {
        const int it = next_item(counter, &s_item);
        if (it >= nitems) { break; }
        const BatchItem wi = items[it];
        const BatchBox bx = boxes[wi.box];
        const int XW = wi.xw, TR = wi.rows, LX = XW + 2, NE = (TR * RY + 2) * LX;
        const long Gx = bx.nx + 2 * g, Gxy = Gx * (bx.ny + 2 * g);
        const T* s = src + bx.goff;
        T* d = dst + bx.voff;
        const int col = tid % XW, trow = tid / XW;
        const bool tact = trow < TR;
        const int x = wi.x0 + col, yb = wi.y0 + trow * RY;
        int eoff[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int e = min(tid + 256 * k, NE - 1);
            const int er = e / LX, ec = e - er * LX;
            const int gx = wi.x0 - 1 + ec, gy = min(wi.y0 - 1 + er, bx.ny + g - 1);
            eoff[k] = (gx + g) + (int)Gx * (gy + g);
        }
        T P[3][NL];
        auto gload = [&](T (&p)[NL], int z) {
            const T* base = s + Gxy * (long)(z + g);
#pragma unroll
            for (int k = 0; k < NL; k++) { p[k] = base[eoff[k]]; }
        };
        auto lstore = [&](const T (&p)[NL], T* pl) {
#pragma unroll
            for (int k = 0; k < NL; k++) { pl[tid + 256 * k] = p[k]; }
        };
        T q[3][RY + 2][3];
        const int lr0 = tact ? trow * RY : 0;  // idle threads read row 0 (never stored)
        auto lread = [&](T (&qq)[RY + 2][3], const T* pl) {
#pragma unroll
            for (int r = 0; r < RY + 2; r++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) { qq[r][cc] = pl[(lr0 + r) * LX + col + cc]; }
        };
        T res[RY];
        int zres = -1;  // plane of the results held in res (-1: none)
        auto put_res = [&]() {
            if (zres < 0 || !tact) { return; }
            T* o = d + x + (long)bx.nx * ((long)yb + (long)bx.ny * zres);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                if (yb + r < bx.ny) {
                    if (NT) { __builtin_nontemporal_store(res[r], o + (long)r * bx.nx); } else { o[(long)r * bx.nx] = res[r]; }
                }
            }
        };
        // one z step: stage plane zz (slot S), store the previous step's
        // results, emit plane zz - 1 from slots S+1 (zz-2), S+2 (zz-1), S (zz)
        auto step = [&](auto sc, int zz) {
            constexpr int S = decltype(sc)::value, SB = (S + 1) % 3, SC = (S + 2) % 3;
            lstore(P[S], sm[S]);
            gload(P[(S + 1) % 3], min(zz + 1, wi.z1));
            put_res();
            __syncthreads();
            lread(q[S], sm[S]);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                auto tap = [&](int dx, int dy, int dz) -> T {
                    return dz < 0 ? q[SB][r + 1 + dy][1 + dx]
                                  : (dz == 0 ? q[SC][r + 1 + dy][1 + dx] : q[S][r + 1 + dy][1 + dx]);
                };
                T acc = 0;
#pragma unroll
                for (int j = 0; j < NC; j++) {
                    acc = acc + c.c[j] * tap(shape_off<SHAPE>(j, 0), shape_off<SHAPE>(j, 1), shape_off<SHAPE>(j, 2));
                }
                res[r] = acc;
            }
            zres = zz - 1;
        };
        // planes z0 - 1 and z0: staged, nothing to emit yet
        gload(P[0], wi.z0 - 1);
        lstore(P[0], sm[0]);
        gload(P[1], wi.z0);
        __syncthreads();
        lread(q[0], sm[0]);
        lstore(P[1], sm[1]);
        gload(P[2], min(wi.z0 + 1, wi.z1));
        __syncthreads();
        lread(q[1], sm[1]);
        int zz = wi.z0 + 1;
        for (;;) {
            step(std::integral_constant<int, 2>{}, zz);
            if (++zz > wi.z1) { break; }
            step(std::integral_constant<int, 0>{}, zz);
            if (++zz > wi.z1) { break; }
            step(std::integral_constant<int, 1>{}, zz);
            if (++zz > wi.z1) { break; }
        }
        put_res();
    }
    queue_done(counter);
}

// ------------------------------------------------------------ batched GSRB
// LDS: a ring of 4 phi planes (k-1 .. k+2) of (rows + 4) x (nx + 2) cells and
// a ring of 2 rhs planes of (rows + 2) x nx cells, both filled from registers
// loaded one step ahead of their use.
template<typename T, int MAXE>
__global__ __launch_bounds__(256) void k_batch_gsrb(const BatchBox* __restrict__ boxes,
                                                    const BatchItem* __restrict__ items, int nitems, int g,
                                                    const T* __restrict__ phi_in, T* __restrict__ phi_out,
                                                    const T* __restrict__ rhs, BCoefs<T> c, T lambda,
                                                    unsigned* counter, int rhs_lds_off)
{
    extern __shared__ __align__(16) unsigned char smraw[];
    T* sm = reinterpret_cast<T*>(smraw);
    T* rsm = sm + rhs_lds_off;
    __shared__ int s_item;
    const int tid = threadIdx.x;
    for (;;) {
        const int it = next_item(counter, &s_item);
        if (it >= nitems) { break; }
        const BatchItem wi = items[it];
        const BatchBox bx = boxes[wi.box];
        const int R = wi.rows, LX = bx.nx + 2, NP = (R + 4) * LX;
        const int NR = (R + 2) * bx.nx;
        const long Gx = bx.nx + 2 * g, Gxy = Gx * (bx.ny + 2 * g);
        const T* p = phi_in + bx.goff;
        T* po = phi_out + bx.goff;
        const T* rh = rhs + bx.voff;
        const int y0 = wi.y0;
        // in-plane offsets of this thread's elements: phi element e -> row
        // y0 - 2 + e / LX, column -1 + e % LX of the ghosted plane (bit i of
        // pout: outside the ghosts, reads as 0); rhs element e -> row
        // y0 - 1 + e / nx (bit i of rout: outside the box).  Elements past the
        // plane repeat the last one; every load is issued unconditionally
        // (outside elements load element 0 and are zeroed at their commit), so
        // the waits only cover loads a step old.
        int offp[MAXE], offr[MAXE];
        unsigned pout = 0, rout = 0;
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = min(tid + 256 * i, NP - 1), er = e / LX, ec = e - er * LX;
            const int gy = y0 - 2 + er, gx = ec - 1;
            const bool pin_ = gy >= -g && gy < bx.ny + g;
            offp[i] = pin_ ? (gx + g) + (int)Gx * (gy + g) : 0;
            pout |= pin_ ? 0u : (1u << i);
            const int f = min(tid + 256 * i, NR - 1), fr = f / bx.nx, fc = f - fr * bx.nx;
            const int fy = y0 - 1 + fr;
            const bool rin = fy >= 0 && fy < bx.ny;
            offr[i] = rin ? fc + bx.nx * fy : 0;
            rout |= rin ? 0u : (1u << i);
        }
        T pre[MAXE], prr[MAXE];
        auto issue = [&](int z) {
            const T* base = p + Gxy * (long)(z + g);
#pragma unroll
            for (int i = 0; i < MAXE; i++) { pre[i] = base[offp[i]]; }
        };
        auto slot = [&](int z) { return sm + (long)(((z % 4) + 4) % 4) * NP; };
        auto commit = [&](int z) {
            T* pl = slot(z);
#pragma unroll
            for (int i = 0; i < MAXE; i++) {
                const int e = tid + 256 * i;
                if (e < NP) { pl[e] = (pout >> i) & 1u ? (T)0 : pre[i]; }
            }
        };
        auto issue_r = [&](int z) {
            const T* base = rh + (long)bx.nx * bx.ny * z;
#pragma unroll
            for (int i = 0; i < MAXE; i++) { prr[i] = base[offr[i]]; }
        };
        auto rslot = [&](int z) { return rsm + (long)(z & 1) * NR; };
        auto commit_r = [&](int z) {
            T* pl = rslot(z);
#pragma unroll
            for (int i = 0; i < MAXE; i++) {
                const int e = tid + 256 * i;
                if (e < NR) { pl[e] = (rout >> i) & 1u ? (T)0 : prr[i]; }
            }
        };
        // half sweep on plane z, rows [ya, yb), colour col: p += lambda (sum c_j p(nbr) + rhs).
        // Cell index c = tid + 256 i <-> (row ya + c / hx, column pair c % hx),
        // walked without divisions from the per-tile split of tid and 256.
        const int hx = (bx.nx + 1) / 2;
        const int q0 = tid / hx, r0 = tid - q0 * hx, qs = 256 / hx, rs = 256 - qs * hx;
        auto update = [&](int z, int ya, int yb, int colour) {
            if (z < 0 || z >= bx.nz) { return; }
            ya = max(ya, 0);
            yb = min(yb, bx.ny);
            const int ncell = (yb - ya) * hx;
            T* zc = slot(z);
            const T* zm = slot(z - 1);
            const T* zp = slot(z + 1);
            const T* rz = rslot(z);
            int q = q0, rr = r0;
            for (int cidx = tid; cidx < ncell; cidx += 256) {
                const int yy = ya + q;
                const int xx = 2 * rr + ((yy + z + colour) & 1);
                q += qs;
                rr += rs;
                if (rr >= hx) {
                    rr -= hx;
                    q++;
                }
                if (xx >= bx.nx) { continue; }
                const int li = (yy - (y0 - 2)) * LX + (xx + 1);
                T s = 0;
                s = s + c.c[0] * zc[li];
                s = s + c.c[1] * zc[li + 1];
                s = s + c.c[2] * zc[li - 1];
                s = s + c.c[3] * zc[li + LX];
                s = s + c.c[4] * zc[li - LX];
                s = s + c.c[5] * zp[li];
                s = s + c.c[6] * zm[li];
                s = s + rz[(yy - (y0 - 1)) * bx.nx + xx];
                zc[li] = zc[li] + lambda * s;
            }
        };
        // planes -1 .. nz (ghost planes included) and rhs planes 0 .. nz-1
        issue(-1);
        commit(-1);
        issue(0);
        commit(0);
        issue(1);
        issue_r(0);
        auto write_out = [&](int z) {
            const T* pl = slot(z);
            T* base = po + Gxy * (long)(z + g);
            const int ny_t = min(R, bx.ny - y0);
            // element e = tid + 256 j <-> (row y0 + e / nx, column e % nx),
            // walked without divisions; non-temporal stores (FP64 sweep 0.765
            // vs 0.827 ms on 2000 boxes, FP32 0.509 vs 0.521 ms)
            int yy = y0 + tid / bx.nx, xx = tid % bx.nx;
            const int ys = 256 / bx.nx, xs = 256 % bx.nx;
            for (int e = tid; e < ny_t * bx.nx; e += 256) {
                __builtin_nontemporal_store(pl[(yy - (y0 - 2)) * LX + xx + 1], &base[(xx + g) + Gx * (long)(yy + g)]);
                yy += ys;
                xx += xs;
                if (xx >= bx.nx) {
                    xx -= bx.nx;
                    yy++;
                }
            }
        };
        // every memory operation of a step is issued at its start (loads of
        // the planes ahead, then the write-out of finished plane k-1), so the
        // wait at the next step's start covers only operations a whole step old
        for (int k = -1; k < bx.nz; k++) {
            if (k + 2 <= bx.nz) { commit(k + 2); }  // into the slot of plane k-2, no longer read
            if (k + 1 < bx.nz) { commit_r(k + 1); }  // into the slot of rhs plane k-1, likewise
            if (k + 3 <= bx.nz) { issue(k + 3); }
            if (k + 2 < bx.nz) { issue_r(k + 2); }
            if (k >= 1) { write_out(k - 1); }        // slot of plane k-1: only read during this step
            __syncthreads();
            update(k + 1, y0 - 1, y0 + R + 1, 0);  // red: tile rows and one row each side
            __syncthreads();
            update(k, y0, y0 + R, 1);              // black: tile rows
            __syncthreads();
        }
        write_out(bx.nz - 1);
    }
    queue_done(counter);
}

// ------------------------------------------------------------ batched GSRB, rows in registers
// The same sweep (arithmetic, order and colouring of k_batch_gsrb) for boxes at
// most 64 cells wide, with no LDS and no barrier.  A workgroup is one wave and
// does one item: P = 64 / nx sub-tiles of one box packed side by side in the
// wave (lanes [p nx, (p + 1) nx) hold sub-tile p, lane p nx + x its column x),
// sub-tile p owning rows y0 + p RY + [0, RY) over the whole z extent.  Each
// lane keeps its column of rows [Y0 - 2, Y0 + RY + 2) of the live planes in
// registers and walks z as the level row sweep does: step k updates the red
// cells of plane k + 1 (tile rows and one halo row each side, recomputed
// rather than shared), then the black cells of plane k, and writes plane k.
// x neighbours are the adjacent lanes (DPP wave shifts); a sub-tile's first
// and last column take the fixed ghost value from an edge register instead
// (each lane also loads its own edge neighbour, or itself, per row).  Ghost
// rows and planes are loaded (clamped into the ghosted box) but never updated
// or written.  Lanes past the last sub-tile shadow lane 0 and store nothing.
struct RowItem {
    int box, y0, nsub, pad;
};

// waves per workgroup of the row and pair sweeps unless set otherwise (config
// 4, 2000 boxes: FP32 row sweep 0.397 ms with 4 waves, 0.416 with 1, 0.513
// with 8; FP64 pair sweep 0.696 ms with 1, 0.722 with 2, 0.803 with 4)
constexpr int kRowsWaves = 4, kPairsWaves = 1;

template<typename T, int RY, int NT, int WPB, int PF, int GW>
__global__ __launch_bounds__(64 * WPB) void k_batch_rows(const BatchBox* __restrict__ boxes,
                                                   const RowItem* __restrict__ items, int g,
                                                   const T* __restrict__ phi_in, T* __restrict__ phi_out,
                                                   const T* __restrict__ rhs, BCoefs<T> c, T lambda)
{
    constexpr int NR = RY + 4;  // phi rows held: Y0 - 2 .. Y0 + RY + 1
    constexpr int NQ = RY + 2;  // rhs and edge rows: Y0 - 1 .. Y0 + RY
    const RowItem wi = items[blockIdx.x];
    if (wi.box < 0) { return; }  // padding of a shorter XCD list
    const BatchBox bx = boxes[wi.box];
    const int nx = bx.nx, ny = bx.ny, nz = bx.nz;
    // wave w of the workgroup holds sub-tiles w P .. w P + P - 1 of the item
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, npw = 64 / nx;
    const int pl = lane / nx;
    int p = w * npw + pl, x = lane - pl * nx;
    const bool live = pl < npw && p < wi.nsub;
    if (!live) {
        p = 0;
        x = 0;
    }
    const int Y0 = wi.y0 + p * RY;
    const int Gx = nx + 2 * g;
    const long Gxy = (long)Gx * (ny + 2 * g), Vxy = (long)nx * ny;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    // in-plane element offsets of this lane's rows, clamped to the rows the
    // sweep reads (-1 .. ny: rows further out only neighbour ghost rows, which
    // are not updated, so they are not fetched)
    int ro[NR], qo[NQ];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int y = min(max(Y0 - 2 + i, -1), ny);
        ro[i] = (x + g) + Gx * (y + g);
    }
#pragma unroll
    for (int j = 0; j < NQ; j++) { qo[j] = x + nx * min(max(Y0 - 1 + j, 0), ny - 1); }
    const bool xl = x == 0, xr = x == nx - 1;
    const int ed = xl ? -1 : (xr ? 1 : 0);
    // rows i this lane may update: inside the box and of a live sub-tile
    unsigned upd = 0;
#pragma unroll
    for (int i = 1; i <= RY + 2; i++) {
        const int y = Y0 - 2 + i;
        upd |= (live && y >= 0 && y < ny) ? (1u << i) : 0u;
    }
    const int par = (x + Y0) & 1;  // red where (x + y + z) is even
    auto is_red = [&](int i, int z) __attribute__((always_inline)) { return ((par + i + z) & 1) == 0; };
    const T* pb = phi_in + bx.goff;
    T* po = phi_out + bx.goff;
    const T* qb = rhs + bx.voff;
    // planes are loaded PF steps ahead of their first use: a ring of 4 + PF slots
    constexpr int R = 4 + PF;
    T P[R][NR];  // phi planes by slot: plane z in slot (z + 1) mod R
    T Q[R][NQ];  // rhs planes, same slots
    T E[R][NQ];  // edge values (x - 1 for the first column, x + 1 for the last)
    // GW: the x ghost cells of every written row are written too (with phi_in's
    // values, which phi_out's ghosts hold already), so that rows fill whole
    // cache lines; the second ghost column (x - 2, x + 2) of the output rows
    // is loaded for it when g == 2 (other ghost widths write no ghost cells)
    T G[R][GW ? RY : 1];
    const bool gw = GW && g <= 2;
    auto loadP = [&](T (&d)[NR], int z) {
        const T* b = pb + Gxy * (long)(min(z, nz) + g);
#pragma unroll
        for (int i = 0; i < NR; i++) { d[i] = b[ro[i]]; }
    };
    auto loadE = [&](T (&d)[NQ], int z) {
        const T* b = pb + Gxy * (long)(min(z, nz) + g) + ed;
#pragma unroll
        for (int j = 0; j < NQ; j++) { d[j] = b[ro[j + 1]]; }
    };
    auto loadG = [&](T (&d)[GW ? RY : 1], int z) {
        if (GW && g == 2) {
            const T* b = pb + Gxy * (long)(min(z, nz) + g) + 2 * ed;
#pragma unroll
            for (int r = 0; r < RY; r++) { d[r] = b[ro[r + 2]]; }
        }
    };
    auto loadQ = [&](T (&d)[NQ], int z) {
        const T* b = qb + Vxy * (long)min(z, nz - 1);
#pragma unroll
        for (int j = 0; j < NQ; j++) { d[j] = b[qo[j]]; }
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
    auto left = [&](T v, T e) __attribute__((always_inline)) {
        const T s = lane_shift<kShr1>(e, v);
        return xl ? e : s;
    };
    auto right = [&](T v, T e) __attribute__((always_inline)) {
        const T s = lane_shift<kShl1>(e, v);
        return xr ? e : s;
    };
    auto step = [&](auto sc, int k) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        constexpr int S0 = S, S1 = (S + 1) % R, S2 = (S + 2) % R, SM = (S + R - 1) % R;
        constexpr int LP = (S + 2 + PF) % R, LQ = (S + 1 + PF) % R;
        if (WPB > 1) { __syncthreads(); }  // the item's waves walk z together
        loadP(P[LP], k + 2 + PF);
        loadQ(Q[LQ], k + 1 + PF);
        loadE(E[LQ], k + 1 + PF);
        loadG(G[LQ], k + 1 + PF);
        if (k + 1 < nz) {  // red cells of plane k + 1, rows Y0 - 1 .. Y0 + RY
#pragma unroll
            for (int j = 0; j < NQ; j++) {
                const int i = j + 1;
                const T cur = P[S1][i], e = E[S1][j];
                const T nv = relax(cur, left(cur, e), right(cur, e), P[S1][i + 1], P[S1][i - 1], P[S2][i], P[S0][i],
                                   Q[S1][j]);
                P[S1][i] = (((upd >> i) & 1u) && is_red(i, k + 1)) ? nv : cur;
            }
        }
        if (k >= 0) {  // black cells of plane k, rows Y0 .. Y0 + RY - 1, then plane k out
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + 2;
                const T cur = P[S0][i], e = E[S0][r + 1];
                const T nv = relax(cur, left(cur, e), right(cur, e), P[S0][i + 1], P[S0][i - 1], P[S1][i], P[SM][i],
                                   Q[S0][r + 1]);
                P[S0][i] = (((upd >> i) & 1u) && !is_red(i, k)) ? nv : cur;
            }
            T* ob = po + Gxy * (long)(k + g);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                if ((upd >> (r + 2)) & 1u) {
                    if (NT) {
                        __builtin_nontemporal_store(P[S0][r + 2], &ob[ro[r + 2]]);
                    } else {
                        ob[ro[r + 2]] = P[S0][r + 2];
                    }
                    if (GW && gw && ed != 0) {
                        ob[ro[r + 2] + ed] = E[S0][r + 1];
                        if (g == 2) { ob[ro[r + 2] + 2 * ed] = G[S0][r]; }
                    }
                }
            }
        }
    };
    // planes -1 .. PF (slots 0 .. PF + 1), rhs and edges of planes 0 .. PF - 1 (slots 1 .. PF)
    static_for<0, PF + 2>([&](auto zc) { loadP(P[decltype(zc)::value], decltype(zc)::value - 1); });
    static_for<0, PF>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        loadQ(Q[z + 1], z);
        loadE(E[z + 1], z);
        loadG(G[z + 1], z);
    });
    int k = -1;
    bool done = false;
    for (;;) {
        static_for<0, R>([&](auto sc) {
            if (!done) {
                step(sc, k);
                done = ++k >= nz;
            }
        });
        if (done) { break; }
    }
}

// Two cells per lane: lane q of a sub-row holds columns 2q (A) and 2q + 1 (B)
// of its rows, loaded and stored as pairs, so a 64-lane wave covers 128
// columns and each colour pass relaxes one register of the pair per row
// instead of computing every lane and discarding half.  Sub-tile heights RY
// are even and sub-tiles start on even rows, so which of A and B holds the
// row's cell of a colour depends only on the row and the plane parity: the z
// walk is unrolled ten times (five register slots by two parities) and every
// colour choice is made at compile time.  Boxes 4..128 cells wide and even,
// an even number of ghost layers (pairs aligned); P = 64 / (nx / 2) sub-tiles
// per wave.  Arithmetic, order and colouring are k_batch_gsrb's.
template<typename T>
struct Vec2;
template<>
struct Vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template<>
struct Vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template<typename T, int RY, int NT, int WPB, int GW>
__global__ __launch_bounds__(64 * WPB) void k_batch_rows2(const BatchBox* __restrict__ boxes,
                                                    const RowItem* __restrict__ items, int g,
                                                    const T* __restrict__ phi_in, T* __restrict__ phi_out,
                                                    const T* __restrict__ rhs, BCoefs<T> c, T lambda)
{
    using V = typename Vec2<T>::type;
    constexpr int NR = RY + 4;  // phi rows held: Y0 - 2 .. Y0 + RY + 1
    constexpr int NQ = RY + 2;  // rhs and edge rows: Y0 - 1 .. Y0 + RY
    static_assert(RY % 2 == 0, "sub-tiles start on even rows");
    const RowItem wi = items[blockIdx.x];
    if (wi.box < 0) { return; }  // padding of a shorter XCD list
    const BatchBox bx = boxes[wi.box];
    const int nx = bx.nx, ny = bx.ny, nz = bx.nz, h = nx >> 1;
    // wave w of the workgroup holds sub-tiles w npw .. w npw + npw - 1 of the item
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, npw = 64 / h;
    const int pl = lane / h;
    int p = w * npw + pl, q = lane - pl * h;
    const bool live = pl < npw && p < wi.nsub;
    if (!live) {
        p = 0;
        q = 0;
    }
    const int Y0 = wi.y0 + p * RY;
    const int Gx = nx + 2 * g;
    const long Gxy = (long)Gx * (ny + 2 * g), Vxy = (long)nx * ny;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    // in-plane byte offsets of column 2q of this lane's rows, clamped to the
    // rows the sweep reads (-1 .. ny); 32-bit and unsigned: the loads take a
    // scalar plane base plus a 32-bit lane offset
    unsigned ro[NR], qo[NQ];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int y = min(max(Y0 - 2 + i, -1), ny);
        ro[i] = (unsigned)(((2 * q + g) + Gx * (y + g)) * (int)sizeof(T));
    }
#pragma unroll
    for (int j = 0; j < NQ; j++) {
        qo[j] = (unsigned)((2 * q + nx * min(max(Y0 - 1 + j, 0), ny - 1)) * (int)sizeof(T));
    }
    const bool ql = q == 0, qr = q == h - 1;
    // the edge value a lane needs: column -1 for the first pair, nx for the last
    // (ro >= g sizeof(T) >= 2 sizeof(T), so the unsigned sum does not wrap)
    const unsigned edu = ql ? (unsigned)(-(int)sizeof(T)) : (qr ? 2u * (unsigned)sizeof(T) : 0u);
    unsigned upd = 0;
#pragma unroll
    for (int i = 1; i <= RY + 2; i++) {
        const int y = Y0 - 2 + i;
        upd |= (live && y >= 0 && y < ny) ? (1u << i) : 0u;
    }
    const T* pb = phi_in + bx.goff;
    T* po = phi_out + bx.goff;
    const T* qb = rhs + bx.voff;
    T A[5][NR], B[5][NR];    // phi planes by slot (plane z in slot (z + 1) mod 5), columns 2q / 2q + 1
    T QA[5][NQ], QB[5][NQ];  // rhs, same slots
    T E[5][NQ];              // edge values
    // GW (g == 2): the ghost pairs (x -2, -1) and (nx, nx + 1) of every written
    // row are written too (phi_in's values, which phi_out's ghosts hold
    // already), so that rows fill whole cache lines; the second ghost column
    // of the output rows is loaded for it
    T G[5][GW ? RY : 1];
    const bool gw = GW && g == 2 && (ql || qr);
    const unsigned gdu = ql ? (unsigned)(-2 * (int)sizeof(T)) : 3u * (unsigned)sizeof(T);
    auto loadP = [&](T (&a)[NR], T (&b)[NR], int z) {
        const char* base = reinterpret_cast<const char*>(pb + Gxy * (long)(min(z, nz) + g));
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const V v = *reinterpret_cast<const V*>(base + ro[i]);
            a[i] = v.x;
            b[i] = v.y;
        }
    };
    auto loadE = [&](T (&d)[NQ], int z) {
        const char* base = reinterpret_cast<const char*>(pb + Gxy * (long)(min(z, nz) + g));
#pragma unroll
        for (int j = 0; j < NQ; j++) { d[j] = *reinterpret_cast<const T*>(base + (ro[j + 1] + edu)); }
    };
    auto loadG = [&](T (&d)[GW ? RY : 1], int z) {
        if (GW && g == 2) {
            const char* base = reinterpret_cast<const char*>(pb + Gxy * (long)(min(z, nz) + g));
#pragma unroll
            for (int r = 0; r < RY; r++) { d[r] = *reinterpret_cast<const T*>(base + (ro[r + 2] + gdu)); }
        }
    };
    auto loadQ = [&](T (&a)[NQ], T (&b)[NQ], int z) {
        const char* base = reinterpret_cast<const char*>(qb + Vxy * (long)min(z, nz - 1));
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const V v = *reinterpret_cast<const V*>(base + qo[j]);
            a[j] = v.x;
            b[j] = v.y;
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
    // or the edge) or its B cell (right neighbour: the next lane's A, or the edge)
    auto cellA = [&](T (&a)[NR], T (&b)[NR], const T (&ap)[NR], const T (&am)[NR], T f, T e, int i, bool ok)
        __attribute__((always_inline)) {
        const T cur = a[i];
        const T sl = lane_shift<kShr1>(e, b[i]);
        const T nv = relax(cur, ql ? e : sl, b[i], a[i + 1], a[i - 1], ap[i], am[i], f);
        a[i] = ok ? nv : cur;
    };
    auto cellB = [&](T (&a)[NR], T (&b)[NR], const T (&bp)[NR], const T (&bm)[NR], T f, T e, int i, bool ok)
        __attribute__((always_inline)) {
        const T cur = b[i];
        const T sr = lane_shift<kShl1>(e, a[i]);
        const T nv = relax(cur, a[i], qr ? e : sr, b[i + 1], b[i - 1], bp[i], bm[i], f);
        b[i] = ok ? nv : cur;
    };
    // step k, slot S of plane k, ZP = parity of plane k + 1: red cells of plane
    // k + 1 (rows Y0 - 1 .. Y0 + RY), black cells of plane k, plane k out.  The
    // red cell of row i of plane k + 1 is A when i + ZP is even; so is the black
    // cell of row i of plane k (whose parity is 1 - ZP).
    auto step = [&](auto sc, auto zc, int k) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value, ZP = decltype(zc)::value;
        constexpr int S0 = S, S1 = (S + 1) % 5, S2 = (S + 2) % 5, S3 = (S + 3) % 5, SM = (S + 4) % 5;
        if (WPB > 1) { __syncthreads(); }  // the item's waves walk z together
        loadP(A[S3], B[S3], k + 3);
        loadQ(QA[S2], QB[S2], k + 2);
        loadE(E[S2], k + 2);
        loadG(G[S2], k + 2);
        if (k + 1 < nz) {
#pragma unroll
            for (int j = 0; j < NQ; j++) {
                const int i = j + 1;
                const bool ok = (upd >> i) & 1u;
                if (((i + ZP) & 1) == 0) {
                    cellA(A[S1], B[S1], A[S2], A[S0], QA[S1][j], E[S1][j], i, ok);
                } else {
                    cellB(A[S1], B[S1], B[S2], B[S0], QB[S1][j], E[S1][j], i, ok);
                }
            }
        }
        if (k >= 0) {
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + 2;
                const bool ok = (upd >> i) & 1u;
                if (((i + ZP) & 1) == 0) {
                    cellA(A[S0], B[S0], A[S1], A[SM], QA[S0][r + 1], E[S0][r + 1], i, ok);
                } else {
                    cellB(A[S0], B[S0], B[S1], B[SM], QB[S0][r + 1], E[S0][r + 1], i, ok);
                }
            }
            char* ob = reinterpret_cast<char*>(po + Gxy * (long)(k + g));
#pragma unroll
            for (int r = 0; r < RY; r++) {
                if ((upd >> (r + 2)) & 1u) {
                    V v;
                    v.x = A[S0][r + 2];
                    v.y = B[S0][r + 2];
                    V* dst = reinterpret_cast<V*>(ob + ro[r + 2]);
                    if (NT) {
                        __builtin_nontemporal_store(v, dst);
                    } else {
                        *dst = v;
                    }
                    if (GW && gw) {
                        // the 32-bit offset is summed before it meets the pointer: the first
                        // pair's offset steps back 2 elements and must wrap within 32 bits
                        V gv;
                        gv.x = ql ? G[S0][r] : E[S0][r + 1];
                        gv.y = ql ? E[S0][r + 1] : G[S0][r];
                        const unsigned go = ro[r + 2] + (ql ? (unsigned)(-2 * (int)sizeof(T)) : 2u * (unsigned)sizeof(T));
                        *reinterpret_cast<V*>(ob + go) = gv;
                    }
                }
            }
        }
    };
    using std::integral_constant;
    loadP(A[0], B[0], -1);
    loadP(A[1], B[1], 0);
    loadP(A[2], B[2], 1);
    loadQ(QA[1], QB[1], 0);
    loadE(E[1], 0);
    loadG(G[1], 0);
    int k = -1;
    for (;;) {
        step(integral_constant<int, 0>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 1>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 2>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 3>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 4>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 0>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 1>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 2>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 3>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 4>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
    }
}

// The row sweep with four cells per lane, FP32, two ghost layers: lane q of a
// sub-tile holds the 16-byte quad of elements 4q .. 4q + 3 of a row as stored,
// columns 4q - 2 .. 4q + 1, so each load and store moves 16 bytes a lane, as
// the FP64 pair sweep's do.  The first and last quads of a row hold its ghost
// columns (-2, -1 and nx, nx + 1; nx a multiple of 4): they are not updated, serve
// as the neighbours of columns 0 and nx - 1, and are written back with their
// input values, so output rows are written whole.  h = nx / 4 + 1 lanes per row,
// 64 / h sub-tiles of RY rows per wave.  The red cells of a row are elements 0
// and 2 or 1 and 3, which the row's and plane's parity fix at compile time.
// Element 0's left neighbour is the previous lane's element 3, element 3's right
// neighbour the next lane's element 0 (DPP wave shifts; at a sub-tile's ends
// those elements are ghost columns, never updated).  Arithmetic and order are
// k_gsrb's.
template<int RY, int NT>
__global__ __launch_bounds__(64) void k_batch_rows4(const BatchBox* __restrict__ boxes,
                                                   const RowItem* __restrict__ items, const float* __restrict__ phi_in,
                                                   float* __restrict__ phi_out, const float* __restrict__ rhs,
                                                   BCoefs<float> c, float lambda)
{
    using T = float;
    typedef float V4 __attribute__((ext_vector_type(4)));
    typedef float V2 __attribute__((ext_vector_type(2)));
    constexpr int g = 2;
    constexpr int NR = RY + 4;  // phi rows held: Y0 - 2 .. Y0 + RY + 1
    constexpr int NQ = RY + 2;  // rhs rows: Y0 - 1 .. Y0 + RY
    static_assert(RY % 2 == 0, "sub-tiles start on even rows");
    const RowItem wi = items[blockIdx.x];
    if (wi.box < 0) { return; }  // padding of a shorter XCD list
    const BatchBox bx = boxes[wi.box];
    const int nx = bx.nx, ny = bx.ny, nz = bx.nz, h = (nx >> 2) + 1;
    const int lane = threadIdx.x & 63, npw = 64 / h;
    const int pl = lane / h;
    int p = pl, q = lane - pl * h;
    const bool live = pl < npw && p < wi.nsub;
    if (!live) {
        p = 0;
        q = 0;
    }
    const int Y0 = wi.y0 + p * RY;
    const int Gx = nx + 2 * g;
    const long Gxy = (long)Gx * (ny + 2 * g), Vxy = (long)nx * ny;
    const T c0 = c.c[0], c1 = c.c[1], c2 = c.c[2], c3 = c.c[3], c4 = c.c[4], c5 = c.c[5], c6 = c.c[6];
    // in-plane byte offsets of the lane's quad in its rows, clamped to the rows
    // the sweep reads (-1 .. ny); rhs pairs of columns 4q - 2, 4q - 1 and 4q, 4q + 1,
    // clamped into the row where they are ghost columns (whose cells are not updated)
    unsigned ro[NR], qa[NQ], qb[NQ];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int y = min(max(Y0 - 2 + i, -1), ny);
        ro[i] = (unsigned)((4 * q + Gx * (y + g)) * (int)sizeof(T));
    }
    const int ca = q == 0 ? 0 : 4 * q - 2, cb = q == h - 1 ? nx - 2 : 4 * q;
#pragma unroll
    for (int j = 0; j < NQ; j++) {
        const int y = min(max(Y0 - 1 + j, 0), ny - 1);
        qa[j] = (unsigned)((ca + nx * y) * (int)sizeof(T));
        qb[j] = (unsigned)((cb + nx * y) * (int)sizeof(T));
    }
    // elements 0, 1 are cells where q > 0, elements 2, 3 where q < h - 1
    const bool lok = q > 0, hok = q < h - 1;
    unsigned upd = 0;
#pragma unroll
    for (int i = 1; i <= RY + 2; i++) {
        const int y = Y0 - 2 + i;
        upd |= (live && y >= 0 && y < ny) ? (1u << i) : 0u;
    }
    const T* pb = phi_in + bx.goff;
    T* po = phi_out + bx.goff;
    const T* qbase = rhs + bx.voff;
    V4 P[5][NR];  // phi planes by slot (plane z in slot (z + 1) mod 5)
    V4 Q[5][NQ];  // rhs, same slots (elements 0..3: columns 4q - 2 .. 4q + 1)
    auto loadP = [&](V4 (&d)[NR], int z) {
        const char* base = reinterpret_cast<const char*>(pb + Gxy * (long)(min(z, nz) + g));
#pragma unroll
        for (int i = 0; i < NR; i++) { d[i] = *reinterpret_cast<const V4*>(base + ro[i]); }
    };
    auto loadQ = [&](V4 (&d)[NQ], int z) {
        const char* base = reinterpret_cast<const char*>(qbase + Vxy * (long)min(z, nz - 1));
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const V2 a = *reinterpret_cast<const V2*>(base + qa[j]);
            const V2 b = *reinterpret_cast<const V2*>(base + qb[j]);
            d[j] = V4{a.x, a.y, b.x, b.y};
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
    // the cells of row i of a plane with elements E0 and E0 + 2 (E0 0 or 1)
    auto cells = [&](auto ec, V4 (&d)[NR], const V4 (&dp)[NR], const V4 (&dm)[NR], const V4& f, int i, bool ok)
        __attribute__((always_inline)) {
        constexpr int E0 = decltype(ec)::value;
        const V4 cur = d[i];
        const bool okl = ok && lok, okh = ok && hok;
        if constexpr (E0 == 0) {
            const T l0 = lane_shift<kShr1>(cur.w, cur.w);
            const T n0 = relax(cur.x, l0, cur.y, d[i + 1].x, d[i - 1].x, dp[i].x, dm[i].x, f.x);
            const T n2 = relax(cur.z, cur.y, cur.w, d[i + 1].z, d[i - 1].z, dp[i].z, dm[i].z, f.z);
            d[i].x = okl ? n0 : cur.x;
            d[i].z = okh ? n2 : cur.z;
        } else {
            const T r3 = lane_shift<kShl1>(cur.x, cur.x);
            const T n1 = relax(cur.y, cur.x, cur.z, d[i + 1].y, d[i - 1].y, dp[i].y, dm[i].y, f.y);
            const T n3 = relax(cur.w, cur.z, r3, d[i + 1].w, d[i - 1].w, dp[i].w, dm[i].w, f.w);
            d[i].y = okl ? n1 : cur.y;
            d[i].w = okh ? n3 : cur.w;
        }
    };
    // step k, slot S of plane k, ZP = parity of plane k + 1: red cells of plane
    // k + 1 (rows Y0 - 1 .. Y0 + RY), black cells of plane k, plane k out.  The
    // red cells of row i of plane k + 1 are elements 0 and 2 when i + ZP is even;
    // so are the black cells of row i of plane k.
    auto step = [&](auto sc, auto zc, int k) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value, ZP = decltype(zc)::value;
        constexpr int S0 = S, S1 = (S + 1) % 5, S2 = (S + 2) % 5, S3 = (S + 3) % 5, SM = (S + 4) % 5;
        loadP(P[S3], k + 3);
        loadQ(Q[S2], k + 2);
        if (k + 1 < nz) {
#pragma unroll
            for (int j = 0; j < NQ; j++) {
                const int i = j + 1;
                const bool ok = (upd >> i) & 1u;
                if (((i + ZP) & 1) == 0) {
                    cells(std::integral_constant<int, 0>{}, P[S1], P[S2], P[S0], Q[S1][j], i, ok);
                } else {
                    cells(std::integral_constant<int, 1>{}, P[S1], P[S2], P[S0], Q[S1][j], i, ok);
                }
            }
        }
        if (k >= 0) {
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int i = r + 2;
                const bool ok = (upd >> i) & 1u;
                if (((i + ZP) & 1) == 0) {
                    cells(std::integral_constant<int, 0>{}, P[S0], P[S1], P[SM], Q[S0][r + 1], i, ok);
                } else {
                    cells(std::integral_constant<int, 1>{}, P[S0], P[S1], P[SM], Q[S0][r + 1], i, ok);
                }
            }
            char* ob = reinterpret_cast<char*>(po + Gxy * (long)(k + g));
#pragma unroll
            for (int r = 0; r < RY; r++) {
                if ((upd >> (r + 2)) & 1u) {
                    V4* dst = reinterpret_cast<V4*>(ob + ro[r + 2]);
                    if (NT) {
                        __builtin_nontemporal_store(P[S0][r + 2], dst);
                    } else {
                        *dst = P[S0][r + 2];
                    }
                }
            }
        }
    };
    using std::integral_constant;
    loadP(P[0], -1);
    loadP(P[1], 0);
    loadP(P[2], 1);
    loadQ(Q[1], 0);
    int k = -1;
    for (;;) {
        step(integral_constant<int, 0>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 1>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 2>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 3>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 4>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 0>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 1>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 2>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 3>{}, integral_constant<int, 0>{}, k);
        if (++k >= nz) { break; }
        step(integral_constant<int, 4>{}, integral_constant<int, 1>{}, k);
        if (++k >= nz) { break; }
    }
}

}  // namespace pm

using namespace pm;

namespace {

int upload(void** dptr, const void* host, size_t bytes)
{
    if (*dptr) { hipFree(*dptr); *dptr = nullptr; }
    if (bytes == 0) { return PM_OK; }
    if (hipMalloc(dptr, bytes) != hipSuccess) {
        set_error("batch: device allocation of %zu bytes failed", bytes);
        return PM_ERR_NOMEM;
    }
    if (hipMemcpy(*dptr, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("batch: upload failed");
        return PM_ERR_HIP;
    }
    return PM_OK;
}

// Apply tiles: x in steps of 64; TR thread-rows of BA_RY rows each, TR the
// largest with TR * XW <= 256 threads, TR * BA_RY no more than the rows left
// (rounded up) and the staged plane (TR BA_RY + 2) (XW + 2) <= BA_NP; z chunks
// of tz (0 = whole box); largest first.
int build_apply_items(pm_batch* B, int tz)
{
    std::vector<BatchItem> items;
    for (int b = 0; b < B->nboxes; b++) {
        const int nx = B->sizes[3 * b], ny = B->sizes[3 * b + 1], nz = B->sizes[3 * b + 2];
        for (int x0 = 0; x0 < nx; x0 += 64) {
            const int xw = std::min(64, nx - x0);
            const int trmax = std::max(1, std::min(256 / xw, (BA_NP / (xw + 2) - 2) / BA_RY));
            for (int y0 = 0; y0 < ny; y0 += trmax * BA_RY) {
                const int tr = std::min(trmax, (ny - y0 + BA_RY - 1) / BA_RY);
                const int zc = tz > 0 ? tz : nz;
                for (int z0 = 0; z0 < nz; z0 += zc) {
                    items.push_back(BatchItem{b, x0, y0, z0, std::min(nz, z0 + zc), tr, xw, 0});
                }
            }
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const BatchItem& a, const BatchItem& b) {
        return (long)a.xw * a.rows * (a.z1 - a.z0) > (long)b.xw * b.rows * (b.z1 - b.z0);
    });  // rows = thread-rows: proportional to the tile's cells
    B->nitems_apply = (int)items.size();
    B->apply_tz = tz;
    return upload(&B->d_items, items.data(), sizeof(BatchItem) * items.size());
}

// LDS of a sweep tile: 4 phi planes of (rows + 4) x (nx + 2), 2 rhs planes of (rows + 2) x nx
size_t gsrb_lds(int rows, int maxnx, size_t esz)
{
    return (4 * (size_t)(rows + 4) * (maxnx + 2) + 2 * (size_t)(rows + 2) * maxnx) * esz;
}

// Sweep tiles: full box width, whole z extent, largest first.  The widest box
// gets `rows` rows (fewer when its FP64 tile would not fit the LDS); narrower
// boxes get proportionally more rows (rows x widest / width), as far as the
// widest tile's LDS planes hold them, so every tile has about the same number
// of cells per colour update and re-reads about the same share of halo rows.
int build_gsrb_items(pm_batch* B, int rows)
{
    int maxnx = 0;
    for (int b = 0; b < B->nboxes; b++) { maxnx = std::max(maxnx, B->sizes[3 * b]); }
    while (rows > 1 && gsrb_lds(rows, maxnx, 8) + 64 > 64 * 1024) { rows--; }
    const long pcap = (long)(rows + 4) * (maxnx + 2), rcap = (long)(rows + 2) * maxnx;
    std::vector<BatchItem> items;
    for (int b = 0; b < B->nboxes; b++) {
        const int nx = B->sizes[3 * b], ny = B->sizes[3 * b + 1], nz = B->sizes[3 * b + 2];
        int r = std::max(rows, (int)(((long)rows * maxnx) / nx));
        while (r > rows && ((long)(r + 4) * (nx + 2) > pcap || (long)(r + 2) * nx > rcap)) { r--; }
        r = std::min(r, ny);
        for (int y0 = 0; y0 < ny; y0 += r) {
            items.push_back(BatchItem{b, 0, y0, 0, nz, r, nx, 0});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const BatchItem& a, const BatchItem& b) {
        return (long)a.xw * a.rows * a.z1 > (long)b.xw * b.rows * b.z1;
    });
    B->nitems_gsrb = (int)items.size();
    B->gsrb_rows = rows;
    return upload(&B->d_gitems, items.data(), sizeof(BatchItem) * items.size());
}

// Row-sweep items: per box, strips of ry rows, 64 / (lanes per row) strips per
// wave (nx, nx / 2 lanes with one or two cells per lane; nx / 4 + 1 with four,
// whose rows include the two ghost columns on each side), wpb waves per workgroup (one item each; the waves of an item
// walk z together, so the halo rows they share are re-read while still
// cached).  Workgroup w runs on XCD w mod 8, and
// the strips of a box share halo rows, so every box goes whole to one XCD
// (whose L2 then serves the re-read rows): boxes are spread over the 8 XCDs
// by cell count (largest first, to the least loaded), each XCD's list takes
// its boxes longest z extent first, and the lists are interleaved into the
// launch order, shorter ones padded with empty items (box -1).
int build_row_items(pm_batch* B, int ry, int cpl, int wpb)
{
    constexpr int kXcd = 8;
    std::vector<int> order(B->nboxes);
    for (int b = 0; b < B->nboxes; b++) { order[b] = b; }
    auto cells = [&](int b) { return (long)B->sizes[3 * b] * B->sizes[3 * b + 1] * B->sizes[3 * b + 2]; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    std::vector<std::vector<int>> xb(kXcd);
    std::vector<long> load(kXcd, 0);
    for (int b : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        xb[x].push_back(b);
        load[x] += cells(b);
    }
    std::vector<std::vector<RowItem>> lists(kXcd);
    size_t nmax = 0;
    for (int x = 0; x < kXcd; x++) {
        std::stable_sort(xb[x].begin(), xb[x].end(),
                         [&](int a, int b) { return B->sizes[3 * a + 2] > B->sizes[3 * b + 2]; });
        for (int b : xb[x]) {
            const int nx = B->sizes[3 * b], ny = B->sizes[3 * b + 1];
            const int lanes = cpl == 4 ? nx / 4 + 1 : nx / cpl;
            const int P = wpb * (64 / lanes), nstrips = (ny + ry - 1) / ry;
            for (int s = 0; s < nstrips; s += P) { lists[x].push_back(RowItem{b, s * ry, std::min(P, nstrips - s), 0}); }
        }
        nmax = std::max(nmax, lists[x].size());
    }
    std::vector<RowItem> items(nmax * kXcd, RowItem{-1, 0, 0, 0});
    for (int x = 0; x < kXcd; x++)
        for (size_t m = 0; m < lists[x].size(); m++) { items[m * kXcd + x] = lists[x][m]; }
    if (cpl == 4) {
        B->nitems_quads = (int)items.size();
        return upload(&B->d_qitems, items.data(), sizeof(RowItem) * items.size());
    }
    if (cpl == 2) {
        B->nitems_pairs = (int)items.size();
        B->pairs_ry = ry;
        B->pairs_wpb = wpb;
        return upload(&B->d_pitems, items.data(), sizeof(RowItem) * items.size());
    }
    B->nitems_rows = (int)items.size();
    B->rows_ry = ry;
    B->rows_wpb = wpb;
    return upload(&B->d_ritems, items.data(), sizeof(RowItem) * items.size());
}

unsigned persistent_blocks(int blocks_per_cu)
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) {
            ncu = prop.multiProcessorCount;
        }
    }
    return (unsigned)(ncu * blocks_per_cu);
}

template<typename T>
void coefs_of(const std::vector<double>& c, double scale, BCoefs<T>& out)
{
    const T sc = (T)scale;
    for (size_t j = 0; j < c.size() && j < 27; j++) {
        T v = (T)c[j];
        if (sc != (T)1) { v = v * sc; }
        out.c[j] = v;
    }
}

}  // namespace

extern "C" {

int pm_batch_create(pm_batch** out, int nboxes, const int* sizes_host, int ghost)
{
    PM_REQUIRE(out && sizes_host && nboxes >= 1, "pm_batch_create: bad argument");
    PM_REQUIRE(ghost >= 1 && ghost <= 4, "pm_batch_create: ghost must be 1..4");
    *out = nullptr;
    pm_batch* B = new pm_batch();
    B->nboxes = nboxes;
    B->ghost = ghost;
    B->sizes.assign(sizes_host, sizes_host + 3 * nboxes);
    B->goff.assign(nboxes + 1, 0);
    B->voff.assign(nboxes + 1, 0);
    std::vector<BatchBox> boxes(nboxes);
    for (int b = 0; b < nboxes; b++) {
        const int nx = B->sizes[3 * b], ny = B->sizes[3 * b + 1], nz = B->sizes[3 * b + 2];
        if (nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096) {
            delete B;
            set_error("pm_batch_create: box %d has bad extents", b);
            return PM_ERR_INVALID;
        }
        boxes[b] = BatchBox{nx, ny, nz, 0, (long)B->goff[b], (long)B->voff[b]};
        B->goff[b + 1] = B->goff[b] + (size_t)(nx + 2 * ghost) * (ny + 2 * ghost) * (nz + 2 * ghost);
        B->voff[b + 1] = B->voff[b] + (size_t)nx * ny * nz;
    }
    int rc = upload(&B->d_boxes, boxes.data(), sizeof(BatchBox) * nboxes);
    if (rc == PM_OK && hipMalloc((void**)&B->d_counter, 4 * sizeof(unsigned)) != hipSuccess) { rc = PM_ERR_NOMEM; }
    if (rc == PM_OK && hipMemset(B->d_counter, 0, 4 * sizeof(unsigned)) != hipSuccess) { rc = PM_ERR_HIP; }
    if (rc == PM_OK) { rc = build_apply_items(B, 0); }
    if (rc == PM_OK) { rc = build_gsrb_items(B, B->cfg.gsrb_rows); }
    B->rows_ok = true;
    B->pairs_ok = ghost % 2 == 0;
    B->quads_ok = ghost == 2;
    for (int b = 0; b < nboxes; b++) {
        const int nx = B->sizes[3 * b];
        B->rows_ok = B->rows_ok && nx >= 2 && nx <= 64;
        B->pairs_ok = B->pairs_ok && nx >= 4 && nx <= 128 && nx % 2 == 0;
        B->quads_ok = B->quads_ok && nx >= 4 && nx <= 252 && nx % 4 == 0;
    }
    if (rc == PM_OK && B->rows_ok) { rc = build_row_items(B, 4, 1, kRowsWaves); }
    if (rc == PM_OK && B->pairs_ok) { rc = build_row_items(B, 4, 2, kPairsWaves); }
    if (rc == PM_OK && B->quads_ok) { rc = build_row_items(B, 2, 4, 1); }
    if (rc != PM_OK) {
        pm_batch_destroy(B);
        return rc;
    }
    *out = B;
    return PM_OK;
}

int pm_batch_destroy(pm_batch* B)
{
    if (!B) { return PM_OK; }
    hipFree(B->d_boxes);
    hipFree(B->d_items);
    hipFree(B->d_gitems);
    hipFree(B->d_ritems);
    hipFree(B->d_pitems);
    hipFree(B->d_qitems);
    hipFree(B->d_counter);
    delete B;
    return PM_OK;
}

int pm_batch_elems(const pm_batch* B, size_t* ghosted, size_t* valid)
{
    PM_REQUIRE(B && ghosted && valid, "pm_batch_elems: null argument");
    *ghosted = B->goff[B->nboxes];
    *valid = B->voff[B->nboxes];
    return PM_OK;
}

int pm_batch_offsets(const pm_batch* B, size_t* ghosted_off, size_t* valid_off)
{
    PM_REQUIRE(B && ghosted_off && valid_off, "pm_batch_offsets: null argument");
    for (int b = 0; b <= B->nboxes; b++) {
        ghosted_off[b] = B->goff[b];
        valid_off[b] = B->voff[b];
    }
    return PM_OK;
}

int pm_batch_configure(pm_batch* B, const pm_batch_config* c)
{
    PM_REQUIRE(B && c, "pm_batch_configure: null argument");
    PM_REQUIRE(c->apply_zplanes >= 0 && c->gsrb_rows >= 1 && c->gsrb_rows <= 32 && c->apply_blocks_per_cu >= 1 &&
                   c->apply_blocks_per_cu <= 16 && c->gsrb_blocks_per_cu >= 1 && c->gsrb_blocks_per_cu <= 16,
               "pm_batch_configure: bad configuration");
    int rc = PM_OK;
    if (c->apply_zplanes != B->apply_tz) { rc = build_apply_items(B, c->apply_zplanes); }
    if (rc == PM_OK && c->gsrb_rows != B->cfg.gsrb_rows) { rc = build_gsrb_items(B, c->gsrb_rows); }
    if (rc == PM_OK) { B->cfg = *c; }
    return rc;
}

int pm_batch_get_config(const pm_batch* B, pm_batch_config* c)
{
    PM_REQUIRE(B && c, "pm_batch_get_config: null argument");
    *c = B->cfg;
    return PM_OK;
}

int pm_batch_apply(pm_batch* B, const pm_stencil* st, const void* src, void* dst, double scale, pm_stream_t stream)
{
    PM_REQUIRE(B && st && src && dst, "pm_batch_apply: null argument");
    int kind = 0, dtype = 0;
    std::vector<double> cv;
    int rc = stencil_info(st, &kind, &dtype, &cv);
    if (rc) { return rc; }
    PM_REQUIRE(kind == 7 || kind == 27, "pm_batch_apply: needs the 7-point or 27-point Laplacian shape");
    const unsigned nb = persistent_blocks(B->cfg.apply_blocks_per_cu);
    if (dtype == PM_F64) {
        BCoefs<double> cc;
        coefs_of<double>(cv, scale, cc);
        if (kind == 7) {
            hipLaunchKernelGGL((k_batch_apply<double, 7, 1>), dim3(nb), dim3(256), 0, S(stream),
                               (const BatchBox*)B->d_boxes, (const BatchItem*)B->d_items, B->nitems_apply, B->ghost,
                               (const double*)src, (double*)dst, cc, B->d_counter);
        } else {
            hipLaunchKernelGGL((k_batch_apply<double, 27, 1>), dim3(nb), dim3(256), 0, S(stream),
                               (const BatchBox*)B->d_boxes, (const BatchItem*)B->d_items, B->nitems_apply, B->ghost,
                               (const double*)src, (double*)dst, cc, B->d_counter);
        }
    } else {
        BCoefs<float> cc;
        coefs_of<float>(cv, scale, cc);
        if (kind == 7) {
            hipLaunchKernelGGL((k_batch_apply<float, 7, 1>), dim3(nb), dim3(256), 0, S(stream),
                               (const BatchBox*)B->d_boxes, (const BatchItem*)B->d_items, B->nitems_apply, B->ghost,
                               (const float*)src, (float*)dst, cc, B->d_counter);
        } else {
            hipLaunchKernelGGL((k_batch_apply<float, 27, 1>), dim3(nb), dim3(256), 0, S(stream),
                               (const BatchBox*)B->d_boxes, (const BatchItem*)B->d_items, B->nitems_apply, B->ghost,
                               (const float*)src, (float*)dst, cc, B->d_counter);
        }
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_batch_sweep_form(pm_batch* B, int form, int waves)
{
    PM_REQUIRE(B && form >= 0 && form <= 8, "pm_batch_sweep_form: bad form");
    PM_REQUIRE(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8,
               "pm_batch_sweep_form: waves per workgroup must be 0 (automatic), 1, 2, 4 or 8");
    PM_REQUIRE(form < 2 || form > 4 || B->rows_ok, "pm_batch_sweep_form: the row sweep needs every box 2..64 cells wide");
    PM_REQUIRE(form < 5 || form > 6 || B->pairs_ok,
               "pm_batch_sweep_form: the pair sweep needs every box even and 4..128 cells wide, and an even ghost");
    PM_REQUIRE(form < 7 || B->quads_ok,
               "pm_batch_sweep_form: the quad sweep needs two ghost layers and every box 4..252 cells wide in "
               "multiples of 4");
    PM_REQUIRE(form < 7 || waves <= 1, "pm_batch_sweep_form: forms 7 and 8 run one wave per workgroup");
    // 2-row sub-tiles are measured with one wave per workgroup only
    const bool one = form == 3 || form == 6;
    PM_REQUIRE(!one || waves <= 1, "pm_batch_sweep_form: forms 3 and 6 run one wave per workgroup");
    PM_REQUIRE(form != 4 || waves == 0 || waves == 1 || waves == 4, "pm_batch_sweep_form: form 4 runs 1 or 4 waves");
    const int rw = one ? 1 : (waves ? waves : kRowsWaves), pw = one ? 1 : (waves ? waves : kPairsWaves);
    int rc = PM_OK;
    const int ry = form == 3 ? 2 : 4;
    if (B->rows_ok && (B->rows_ry != ry || B->rows_wpb != rw) && form != 1 && form < 5) {
        rc = build_row_items(B, ry, 1, rw);
    }
    const int pry = 4;
    if (rc == PM_OK && B->pairs_ok && (B->pairs_ry != pry || B->pairs_wpb != pw) && (form == 0 || form == 5 || form == 6)) {
        rc = build_row_items(B, pry, 2, pw);
    }
    if (rc == PM_OK) {
        B->sweep_form = form;
        B->sweep_waves = waves;
    }
    return rc;
}

int pm_batch_gsrb(pm_batch* B, int dtype, const void* phi_in, void* phi_out, const void* rhs, double dx,
                  pm_stream_t stream)
{
    PM_REQUIRE(B && phi_in && phi_out && rhs, "pm_batch_gsrb: null argument");
    PM_REQUIRE(phi_in != phi_out, "pm_batch_gsrb: the sweep is out of place (phi_in != phi_out)");
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_batch_gsrb: bad dtype");
    // automatic: with 2 ghost layers the pair sweep writing its rows' ghost pairs
    // (config 4: FP64 0.615 ms, FP32 0.376 ms; without the ghost writes 0.698 /
    // 0.435 ms, the row sweep 0.720 / 0.397 ms, the LDS ring 0.758 / 0.501 ms;
    // with non-temporal stores 0.620 / 0.401 ms); otherwise FP64 the pair sweep
    // and FP32 the row sweep, each falling back on the other, then the LDS ring.
    // The pair sweep needs the arrays aligned to two elements.
    const size_t al = 2 * (dtype == PM_F64 ? 8 : 4);
    const bool aligned = (uintptr_t)phi_in % al == 0 && (uintptr_t)phi_out % al == 0 && (uintptr_t)rhs % al == 0;
    const int sf = B->sweep_form;
    // FP32 with two ghost layers: the quad sweep (rows of phi start on 16-byte
    // quads with their ghost columns, rhs rows on 8-byte pairs)
    const bool aligned16 = (uintptr_t)phi_in % 16 == 0 && (uintptr_t)phi_out % 16 == 0 && (uintptr_t)rhs % 8 == 0;
    if (sf >= 7 || (sf == 0 && dtype == PM_F32 && B->quads_ok && aligned16)) {
        PM_REQUIRE(dtype == PM_F32 && aligned16,
                   "pm_batch_gsrb: the quad sweep is FP32, phi aligned to 16 bytes and rhs to 8");
        double lam = -(dx * dx) / (4.0 * 3);
        lam *= 2.0;
        const double m = -1.0 / (dx * dx);
        BCoefs<float> cc;
        for (int j = 0; j < 7; j++) { cc.c[j] = (float)(j == 0 ? -6.0 : 1.0) * (float)m; }
        const float lamT = (float)(1.0 * lam);
        auto run = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3((unsigned)B->nitems_quads), dim3(64), 0, S(stream),
                               (const BatchBox*)B->d_boxes, (const RowItem*)B->d_qitems, (const float*)phi_in,
                               (float*)phi_out, (const float*)rhs, cc, lamT);
        };
        // 2-row sub-tiles (244 registers, two waves per SIMD): config 4 FP32
        // 0.308 ms against 0.320 ms with 4 rows (which holds 376 registers, one
        // wave per SIMD) and 0.375 ms for the pair sweep writing ghost pairs
        if (sf == 8) {
            run(k_batch_rows4<2, 1>);
        } else {
            run(k_batch_rows4<2, 0>);
        }
        PM_CHECK_LAUNCH();
        return PM_OK;
    }
    if (sf >= 5) { PM_REQUIRE(aligned, "pm_batch_gsrb: the pair sweep needs arrays aligned to two elements"); }
    const bool pairs_auto = B->pairs_ok && aligned && (B->ghost == 2 || dtype == PM_F64 || !B->rows_ok);
    const bool ghost_pairs = sf == 6 || (sf == 0 && B->ghost == 2);
    if (sf >= 5 || (sf == 0 && pairs_auto)) {
        double lam = -(dx * dx) / (4.0 * 3);
        lam *= 2.0;
        const double m = -1.0 / (dx * dx);
        auto go = [&](auto tag) {
            using T = decltype(tag);
            BCoefs<T> cc;
            for (int j = 0; j < 7; j++) { cc.c[j] = (T)(j == 0 ? -6.0 : 1.0) * (T)m; }
            const T lamT = (T)(1.0 * lam);
            auto run = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3((unsigned)B->nitems_pairs), dim3(64 * B->pairs_wpb), 0, S(stream),
                                   (const BatchBox*)B->d_boxes, (const RowItem*)B->d_pitems, B->ghost,
                                   (const T*)phi_in, (T*)phi_out, (const T*)rhs, cc, lamT);
            };
            if (ghost_pairs) {
                run(k_batch_rows2<T, 4, 0, 1, 1>);
            } else if (B->pairs_wpb == 2) {
                run(k_batch_rows2<T, 4, 1, 2, 0>);
            } else if (B->pairs_wpb == 4) {
                run(k_batch_rows2<T, 4, 1, 4, 0>);
            } else if (B->pairs_wpb == 8) {
                run(k_batch_rows2<T, 4, 1, 8, 0>);
            } else {
                run(k_batch_rows2<T, 4, 1, 1, 0>);
            }
        };
        if (dtype == PM_F64) {
            go(0.0);
        } else {
            go(0.0f);
        }
        PM_CHECK_LAUNCH();
        return PM_OK;
    }
    if (B->rows_ok && B->sweep_form != 1) {
        // LaplaceOp::init: coefficients Laplacian()_j * (-1/(dx*dx)), lambda = 2 * (-(dx*dx)/(4 DIM))
        double lam = -(dx * dx) / (4.0 * 3);
        lam *= 2.0;
        const double m = -1.0 / (dx * dx);
        const int form = B->sweep_form;
        auto go = [&](auto tag) {
            using T = decltype(tag);
            BCoefs<T> cc;
            for (int j = 0; j < 7; j++) { cc.c[j] = (T)(j == 0 ? -6.0 : 1.0) * (T)m; }
            const T lamT = (T)(1.0 * lam);
            auto run = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3((unsigned)B->nitems_rows), dim3(64 * B->rows_wpb), 0, S(stream),
                                   (const BatchBox*)B->d_boxes, (const RowItem*)B->d_ritems, B->ghost,
                                   (const T*)phi_in, (T*)phi_out, (const T*)rhs, cc, lamT);
            };
            if (form == 3) {
                run(k_batch_rows<T, 2, 1, 1, 1, 0>);
            } else if (form == 4) {
                if (B->rows_wpb == 4) {
                    run(k_batch_rows<T, 4, 0, 4, 1, 1>);
                } else {
                    run(k_batch_rows<T, 4, 0, 1, 1, 1>);
                }
            } else if (B->rows_wpb == 2) {
                run(k_batch_rows<T, 4, 1, 2, 1, 0>);
            } else if (B->rows_wpb == 4) {
                run(k_batch_rows<T, 4, 1, 4, 1, 0>);
            } else if (B->rows_wpb == 8) {
                run(k_batch_rows<T, 4, 1, 8, 1, 0>);
            } else {
                run(k_batch_rows<T, 4, 1, 1, 1, 0>);
            }
        };
        if (dtype == PM_F64) {
            go(0.0);
        } else {
            go(0.0f);
        }
        PM_CHECK_LAUNCH();
        return PM_OK;
    }
    int maxnx = 0;
    for (int b = 0; b < B->nboxes; b++) { maxnx = std::max(maxnx, B->sizes[3 * b]); }
    PM_REQUIRE(maxnx <= 256, "pm_batch_gsrb: boxes wider than 256 cells are not supported");
    const int phi_elems = 4 * (B->gsrb_rows + 4) * (maxnx + 2);
    const size_t esz = dtype == PM_F64 ? 8 : 4;
    const size_t lds = gsrb_lds(B->gsrb_rows, maxnx, esz);
    PM_REQUIRE(lds + 64 <= 64 * 1024, "pm_batch_gsrb: tile too large for LDS (rows %d, width %d)", B->gsrb_rows, maxnx);
    const unsigned nb = persistent_blocks(B->cfg.gsrb_blocks_per_cu);
    const long np = (long)(B->gsrb_rows + 4) * (maxnx + 2);
    // LaplaceOp::init: coefficients Laplacian()_j * (-1/(dx*dx)), lambda = 2 * (-(dx*dx)/(4 DIM))
    double lam = -(dx * dx) / (4.0 * 3);
    lam *= 2.0;
    const double m = -1.0 / (dx * dx);
    auto launch = [&](auto tag, auto coef, auto lamT) {
        using T = decltype(tag);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), lds, S(stream), (const BatchBox*)B->d_boxes,
                               (const BatchItem*)B->d_gitems, B->nitems_gsrb, B->ghost, (const T*)phi_in,
                               (T*)phi_out, (const T*)rhs, coef, lamT, B->d_counter + 2, phi_elems);
        };
        if (np <= 1024) { go(k_batch_gsrb<T, 4>); }
        else if (np <= 2048) { go(k_batch_gsrb<T, 8>); }
        else { go(k_batch_gsrb<T, 16>); }
    };
    if (dtype == PM_F64) {
        BCoefs<double> cc;
        for (int j = 0; j < 7; j++) { cc.c[j] = (j == 0 ? -6.0 : 1.0) * m; }
        launch(0.0, cc, 1.0 * lam);
    } else {
        BCoefs<float> cc;
        for (int j = 0; j < 7; j++) { cc.c[j] = (float)(j == 0 ? -6.0 : 1.0) * (float)m; }
        launch(0.0f, cc, (float)(1.0 * lam));
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_batch_autotune(pm_batch* B, const pm_stencil* st, const void* src, void* dst, void* phi_out, const void* rhs,
                      double dx, int reps, pm_batch_config* best, float* best_ms, pm_stream_t stream)
{
    PM_REQUIRE(B && st && src && dst && phi_out && rhs && best && best_ms && reps >= 1,
               "pm_batch_autotune: bad argument");
    PM_REQUIRE(src != phi_out, "pm_batch_autotune: phi_out must differ from src");
    int kind = 0, dtype = 0;
    std::vector<double> cv;
    int rc = stencil_info(st, &kind, &dtype, &cv);
    if (rc) { return rc; }
    int maxnx = 0;
    for (int b = 0; b < B->nboxes; b++) { maxnx = std::max(maxnx, B->sizes[3 * b]); }
    const pm_batch_config keep = B->cfg;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("pm_batch_autotune: event creation failed");
        return PM_ERR_HIP;
    }
    auto time = [&](auto&& launch, float* ms) -> int {
        int r = launch();  // warm-up (also builds the work list)
        if (r) { return r; }
        hipEventRecord(e0, S(stream));
        for (int i = 0; i < reps && r == PM_OK; i++) { r = launch(); }
        hipEventRecord(e1, S(stream));
        if (r) { return r; }
        if (hipEventSynchronize(e1) != hipSuccess) { set_error("pm_batch_autotune: sync failed"); return PM_ERR_HIP; }
        hipEventElapsedTime(ms, e0, e1);
        *ms /= (float)reps;
        return PM_OK;
    };
    const int zs[] = {0, 8, 16, 32};
    const int bpcs[] = {1, 2, 4, 8};
    const int rows[] = {2, 4, 8, 16};
    pm_batch_config win = keep;
    float best_a = 1e30f, best_g = 1e30f;
    for (int z : zs) {
        for (int bpc : bpcs) {
            pm_batch_config c = keep;
            c.apply_zplanes = z;
            c.apply_blocks_per_cu = bpc;
            rc = pm_batch_configure(B, &c);
            float ms = 0;
            if (rc == PM_OK) { rc = time([&]() { return pm_batch_apply(B, st, src, dst, 1.0, stream); }, &ms); }
            if (rc) { break; }
            if (ms < best_a) { best_a = ms; win.apply_zplanes = z; win.apply_blocks_per_cu = bpc; }
        }
        if (rc) { break; }
    }
    const size_t esz = dtype == PM_F64 ? 8 : 4;
    for (int r : rows) {
        if (rc) { break; }
        if (gsrb_lds(r, maxnx, esz) + 64 > 64 * 1024) { continue; }
        for (int bpc : bpcs) {
            pm_batch_config c = keep;
            c.apply_zplanes = win.apply_zplanes;
            c.gsrb_rows = r;
            c.gsrb_blocks_per_cu = bpc;
            rc = pm_batch_configure(B, &c);
            float ms = 0;
            if (rc == PM_OK) {
                rc = time([&]() { return pm_batch_gsrb(B, dtype, src, phi_out, rhs, dx, stream); }, &ms);
            }
            if (rc) { break; }
            if (ms < best_g) { best_g = ms; win.gsrb_rows = r; win.gsrb_blocks_per_cu = bpc; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) {
        pm_batch_configure(B, &keep);
        return rc;
    }
    rc = pm_batch_configure(B, &win);
    *best = win;
    best_ms[0] = best_a;
    best_ms[1] = best_g;
    return rc;
}

}  // extern "C"
