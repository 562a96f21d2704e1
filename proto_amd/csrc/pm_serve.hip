// Continuous-batching Poisson solve server.
//
// Mode 0 keeps every active request's hierarchy resident and advances all of
// them together: one loop iteration is one FP32 GSRB V-cycle (the
// preconditioner), the FP64 correction phi += e, the FP64 residual
// r = rhs - L(phi) (rounded to FP32 as the next V-cycle's right-hand side) and
// one read-back of every request's max-norm.  Each level operator of the
// V-cycle is ONE launch whose work list holds the tiles of every request that
// has that level, so small requests ride along with large ones instead of
// paying a launch sequence of their own.
//
// Arithmetic per request is that of the multigrid driver with precision 1 on a
// single periodic box (pm_mg.hip, pm_level.hip): the GSRB update of
// LaplaceOp.H:95-123, the residual-restriction of LevelMultigrid.cpp:51-73,
// constant prolongation (InterpStencil::Constant, staged), the FP32
// coefficients Lap_j * (float)(-1/h^2) and lambda = (float)(2 * (-h^2 / 12)).
// The sweep is done in one out-of-place pass (red cells of plane k+1, then
// black cells of plane k, in an LDS ring of planes, red cells recomputed in
// the one-cell halo from periodic images), which gives the bits of the
// exchange / red / exchange / black sequence.  So a served result equals the
// driver's answer for the same problem alone, and never depends on the mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

constexpr int SR = 8;    // rows per sweep tile
constexpr int SZC = 32;  // planes per sweep tile
constexpr int SCELLS = 2048;  // cells per item of the pointwise kernels

struct SLevel {
    float* phi;
    float* tmp;
    float* rhs;
    double* acc;        // level 0: the request's FP64 solution (the last sweep adds e into it)
    int n;
    int rows;           // rows per sweep tile
    int tw;             // columns per sweep tile
    float c0, c1, lam;  // centre / neighbour coefficients, GSRB weight
};

// rows and columns per sweep tile of a level n cells wide, and the LDS of its
// tile.  Levels wider than 64 cells are split into column tiles of at most 64:
// the tile's LDS (about 18 KB in FP32) then admits 8 workgroups per CU, where
// 128-wide tiles (36 KB) admitted 4; measured 665 vs 605 saturated solves/s.
__host__ __device__ inline int sweep_rows(int) { return SR; }
__host__ __device__ inline int sweep_cols(int n)
{
    const int k = (n + 63) / 64;
    return (n + k - 1) / k;
}
inline size_t sweep_lds(int n)
{
    const int r = sweep_rows(n), w = sweep_cols(n);
    return sizeof(float) * (4 * (size_t)(r + 4) * (w + 4) + 2 * (size_t)(r + 2) * (w + 2));
}

struct SReq {
    double* phi;   // FP64 solution, n^3
    double* rhs;   // FP64 right-hand side, n^3
    int lev0;      // index of its finest level in the SLevel table
    int slot;      // norm slot
    int n;
    int pad;
    double d0, d1; // FP64 residual coefficients
};

struct SItem {
    int lev;  // SLevel index (sweep, restrict: fine level, prolong: fine level) or SReq index
    int a, b, c;
    int x0, x1;  // sweep tiles: columns [x0, x1)
};

__device__ inline int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// the same for an index at most one period outside [0, n): a compare and
// select instead of an integer remainder
__device__ inline int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// One GSRB sweep (both colours) of the tile rows [y0, y0+rows) x planes [z0, z1)
// of level L, reading `in` and writing `out` (dir 0: phi -> tmp, 1: tmp -> phi).
// phi planes (rows [y0-2, y1+2) x columns [-2, n+2)) pass through a ring of 4
// LDS planes and rhs planes (rows [y0-1, y1+1) x columns [-1, n+1)) through a
// ring of 2, both loaded into registers one step ahead of their use.  IN
// selects the staged input: 0 `in`; 1 zero (`in` is not read: the first sweep
// of a zero initial guess); 2 `in` + the constant prolongation of the next
// coarser level's phi, rounded as k_srv_prolong rounds it (the first sweep
// after fineInterp, which is then not run).  OUT = 1 (the last sweep of level
// 0): the result is added to the request's FP64 solution L.acc as
// k_srv_correct adds the correction, and `out` is not written.
template<int MAXE, int IN, int OUT>
__global__ __launch_bounds__(256) void k_srv_gsrb(const SLevel* __restrict__ lv, const SItem* __restrict__ items, int dir)
{
    extern __shared__ __align__(16) unsigned char smraw[];
    const SItem it = items[blockIdx.x];
    const SLevel L = lv[it.lev];
    const int n = L.n, x0 = it.x0, x1 = it.x1, W = x1 - x0 + 4, WR = x1 - x0 + 2;
    const int y0 = it.a, y1 = min(y0 + L.rows, n), z0 = it.b, z1 = it.c;
    const int NP = (y1 - y0 + 4) * W, NR = (y1 - y0 + 2) * WR;
    float* sm = reinterpret_cast<float*>(smraw);
    float* rsm = sm + 4 * (L.rows + 4) * (L.tw + 4);
    const float* in = dir ? L.tmp : L.phi;
    float* out = dir ? L.phi : L.tmp;
    const long nn = (long)n * n;
    auto slot = [&](int z) { return sm + (long)(((z % 4) + 4) % 4) * NP; };
    auto rslot = [&](int z) { return rsm + (long)(z & 1) * NR; };
    // in-plane offsets (periodic images folded in) of this thread's elements;
    // elements past the plane repeat the last one, so every load is issued
    // unconditionally and the waits only cover loads a step old
    int offp[MAXE], offr[MAXE];
#pragma unroll
    for (int i = 0; i < MAXE; i++) {
        const int e = min((int)threadIdx.x + 256 * i, NP - 1), er = e / W, ec = e - er * W;
        offp[i] = wrapi(x0 + ec - 2, n) + n * wrapi(y0 - 2 + er, n);
        const int f = min((int)threadIdx.x + 256 * i, NR - 1), fr = f / WR, fc = f - fr * WR;
        offr[i] = wrapi(x0 + fc - 1, n) + n * wrapi(y0 - 1 + fr, n);
    }
    // IN == 2: coarse cell under each staged fine cell (in-plane part)
    constexpr int NCO = IN == 2 ? MAXE : 1;
    int offc[NCO];
    const float* cphi = lv[it.lev + (IN == 2 ? 1 : 0)].phi;
    const int nc = lv[it.lev + (IN == 2 ? 1 : 0)].n;  // n / 2 (IN == 2)
    if constexpr (IN == 2) {
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = min((int)threadIdx.x + 256 * i, NP - 1), er = e / W, ec = e - er * W;
            offc[i] = (wrapi(x0 + ec - 2, n) >> 1) + nc * (wrapi(y0 - 2 + er, n) >> 1);
        }
    }
    float pre[MAXE], prr[MAXE], prc[NCO];
    auto issue = [&](int z) {
        if constexpr (IN == 1) { return; }
        const int zw = wrapi(z, n);
        const float* base = in + nn * zw;
#pragma unroll
        for (int i = 0; i < MAXE; i++) { pre[i] = base[offp[i]]; }
        if constexpr (IN == 2) {
            const float* cb = cphi + (long)nc * nc * (zw >> 1);
#pragma unroll
            for (int i = 0; i < MAXE; i++) { prc[i] = cb[offc[i]]; }
        }
    };
    auto commit = [&](int z) {
        float* pl = slot(z);
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = threadIdx.x + 256 * i;
            float v;
            if constexpr (IN == 1) {
                v = 0.0f;
            } else if constexpr (IN == 2) {
                const float zero = 0.0f;
                v = pre[i] + (zero + prc[i]);
            } else {
                v = pre[i];
            }
            if (e < NP) { pl[e] = v; }
        }
    };
    auto issue_r = [&](int z) {
        const float* base = L.rhs + nn * wrapi(z, n);
#pragma unroll
        for (int i = 0; i < MAXE; i++) { prr[i] = base[offr[i]]; }
    };
    auto commit_r = [&](int z) {
        float* pl = rslot(z);
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int e = threadIdx.x + 256 * i;
            if (e < NR) { pl[e] = prr[i]; }
        }
    };
    // colour cells ((x+y+z+colour) even) of plane z, rows [ya, yb), columns [xa, xb)
    // cell index c = tid + 256 i of a colour update <-> (row ya + c / hw, column
    // pair c % hw), walked without divisions from the split of tid and 256 by hw
    struct Walk {
        int hw, q0, r0, qs, rs;
    };
    auto mk_walk = [&](int hw) {
        const int t = (int)threadIdx.x;
        return Walk{hw, t / hw, t % hw, 256 / hw, 256 % hw};
    };
    const Walk wred = mk_walk((x1 - x0 + 2 + 1) / 2), wblk = mk_walk((x1 - x0 + 1) / 2);
    auto update = [&](int z, int ya, int yb, int xa, int xb, int colour, const Walk& wk) {
        float* zc = slot(z);
        const float* zm = slot(z - 1);
        const float* zp = slot(z + 1);
        const float* rz = rslot(z);
        const int hw = wk.hw;
        const int ncell = (yb - ya) * hw;
        int qq = wk.q0, rr = wk.r0;
        for (int c = threadIdx.x; c < ncell; c += 256) {
            const int yy = ya + qq;
            const int xx = xa + 2 * rr + ((xa + yy + z + colour) & 1);
            qq += wk.qs;
            rr += wk.rs;
            if (rr >= hw) {
                rr -= hw;
                qq++;
            }
            if (xx >= xb) { continue; }
            const int li = (yy - (y0 - 2)) * W + (xx - x0 + 2);
            float s = 0;
            s = s + L.c0 * zc[li];
            s = s + L.c1 * zc[li + 1];
            s = s + L.c1 * zc[li - 1];
            s = s + L.c1 * zc[li + W];
            s = s + L.c1 * zc[li - W];
            s = s + L.c1 * zp[li];
            s = s + L.c1 * zm[li];
            s = s + rz[(yy - (y0 - 1)) * WR + (xx - x0 + 1)];
            zc[li] = zc[li] + L.lam * s;
        }
    };
    issue(z0 - 2);
    commit(z0 - 2);
    issue(z0 - 1);
    commit(z0 - 1);
    issue(z0);
    issue_r(z0 - 1);
    // OUT == 1: this thread's output cells (q = tid + 256 i < cnt, clamped like
    // the staging loads) as in-plane offsets and LDS offsets; their FP64
    // values of the plane finished next are loaded one step ahead
    const int tw = x1 - x0, cnt = (y1 - y0) * tw;
    constexpr int NO = OUT == 1 ? MAXE : 1;
    int oo[NO], ol[NO];
    double acc_in[NO];
    if constexpr (OUT == 1) {
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int q = min((int)threadIdx.x + 256 * i, cnt - 1), r = q / tw, xx = q - r * tw;
            oo[i] = x0 + xx + n * (y0 + r);
            ol[i] = (r + 2) * W + xx + 2;
        }
    }
    auto issue_o = [&](int z) {
        if constexpr (OUT == 1) {
#pragma unroll
            for (int i = 0; i < MAXE; i++) { acc_in[i] = L.acc[nn * z + oo[i]]; }
        }
    };
    auto write_out = [&](int z) {
        const float* pl = slot(z);
        if constexpr (OUT == 1) {
#pragma unroll
            for (int i = 0; i < MAXE; i++) {
                if ((int)threadIdx.x + 256 * i < cnt) { L.acc[nn * z + oo[i]] = acc_in[i] + (double)pl[ol[i]]; }
            }
        } else {
            float* ob = out + nn * z;
            for (int e = threadIdx.x; e < cnt; e += 256) {
                const int r = e / tw, xx = e - r * tw;
                ob[x0 + xx + (long)n * (y0 + r)] = pl[(r + 2) * W + xx + 2];
            }
        }
    };
    // all memory operations of a step are issued at its start (loads of the
    // planes ahead, then the write-out of finished plane k-1), so the wait at
    // the next step's start covers only operations a whole step old
    for (int k = z0 - 2; k < z1; k++) {
        commit(k + 2);    // into the slot of plane k-2, last read before the previous barrier
        commit_r(k + 1);  // into the slot of rhs plane k-1, likewise
        if (k + 3 <= z1 + 1) { issue(k + 3); }
        if (k + 2 <= z1) { issue_r(k + 2); }
        if (k - 1 >= z0) { write_out(k - 1); }
        if (k >= z0) { issue_o(k); }  // OUT == 1: FP64 values of plane k, written next step
        __syncthreads();
        update(k + 1, y0 - 1, y1 + 1, x0 - 1, x1 + 1, 0, wred);  // red, one-cell halo included
        __syncthreads();
        if (k >= z0) { update(k, y0, y1, x0, x1, 1, wblk); }  // black, tile cells
        __syncthreads();
    }
    write_out(z1 - 1);
}

// All it.a sweeps of one small level (n^3 <= SMALL_CELLS) of one request in one
// workgroup: phi lives in LDS for the whole relaxation (periodic neighbours by
// index), the right-hand side of each thread's cells in registers, one pass
// per colour.  Same update, order and colouring as k_srv_gsrb, so the same
// bits; IN as k_srv_gsrb (1 zero guess, 2 + prolongation of the next coarser
// level).  Small levels are latency-bound: one launch instead of one per sweep.
constexpr int SMALL_CELLS = 4096, SMALL_PER = SMALL_CELLS / 256;

template<int IN>
__global__ __launch_bounds__(256) void k_srv_small(const SLevel* __restrict__ lv, const SItem* __restrict__ items)
{
    __shared__ float u[SMALL_CELLS];
    const SItem it = items[blockIdx.x];
    const SLevel L = lv[it.lev];
    const int n = L.n, N = n * n * n, nsw = it.a;
    const float* cphi = lv[it.lev + (IN == 2 ? 1 : 0)].phi;
    const int nc = n >> 1;
    float f[SMALL_PER];
#pragma unroll
    for (int j = 0; j < SMALL_PER; j++) {
        const int i = threadIdx.x + 256 * j;
        if (i < N) {
            const int x = i % n, y = (i / n) % n, z = i / (n * n);
            f[j] = L.rhs[i];
            float v = 0.0f;
            if constexpr (IN != 1) { v = L.phi[i]; }
            if constexpr (IN == 2) {
                const float zero = 0.0f;
                v = v + (zero + cphi[(x >> 1) + nc * ((y >> 1) + nc * (z >> 1))]);
            }
            u[i] = v;
        }
    }
    for (int sw = 0; sw < nsw; sw++) {
        for (int colour = 0; colour < 2; colour++) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SMALL_PER; j++) {
                const int i = threadIdx.x + 256 * j;
                if (i >= N) { continue; }
                const int x = i % n, y = (i / n) % n, z = i / (n * n);
                if (((x + y + z + colour) & 1) != 0) { continue; }
                const int row = n * (y + n * z);
                float s = 0;
                s = s + L.c0 * u[i];
                s = s + L.c1 * u[wrap1(x + 1, n) + row];
                s = s + L.c1 * u[wrap1(x - 1, n) + row];
                s = s + L.c1 * u[x + n * (wrap1(y + 1, n) + n * z)];
                s = s + L.c1 * u[x + n * (wrap1(y - 1, n) + n * z)];
                s = s + L.c1 * u[x + n * (y + n * wrap1(z + 1, n))];
                s = s + L.c1 * u[x + n * (y + n * wrap1(z - 1, n))];
                s = s + f[j];
                u[i] = u[i] + L.lam * s;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += 256) { L.phi[i] = u[i]; }
}

// Coarse right-hand side = AvgDown(rhs + sum_j c_j phi(p + o_j)) of the fine
// level (LevelMultigrid.cpp:51-73 order), and, with zero_phi, coarse phi = 0
// (otherwise the coarse level's first sweep takes the zero guess as its input).
__global__ __launch_bounds__(256) void k_srv_restrict(const SLevel* __restrict__ lv, const SItem* __restrict__ items,
                                                      int zero_phi)
{
    const SItem it = items[blockIdx.x];
    const SLevel F = lv[it.lev], C = lv[it.lev + 1];
    const int n = F.n, nc = C.n;
    const long nn = (long)n * n;
    const float avg = 0.125f;
    for (int t = it.a + threadIdx.x; t < it.a + it.b; t += 256) {
        const int X = t % nc, Y = (t / nc) % nc, Z = t / (nc * nc);
        float acc = 0;
#pragma unroll
        for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
#pragma unroll
                for (int kx = 0; kx < 2; kx++) {
                    const int x = 2 * X + kx, y = 2 * Y + ky, z = 2 * Z + kz;
                    const float* pz = F.phi + nn * z;
                    const long row = (long)n * y;
                    float s = 0;
                    s = s + F.rhs[x + row + nn * z];
                    s = s + F.c0 * pz[x + row];
                    s = s + F.c1 * pz[wrap1(x + 1, n) + row];
                    s = s + F.c1 * pz[wrap1(x - 1, n) + row];
                    s = s + F.c1 * pz[x + (long)n * wrap1(y + 1, n)];
                    s = s + F.c1 * pz[x + (long)n * wrap1(y - 1, n)];
                    s = s + F.c1 * F.phi[x + row + nn * wrap1(z + 1, n)];
                    s = s + F.c1 * F.phi[x + row + nn * wrap1(z - 1, n)];
                    acc = acc + avg * s;
                }
        C.rhs[t] = acc;
        if (zero_phi) { C.phi[t] = 0.0f; }
    }
}

// Fine phi += constant prolongation of the coarse phi (three staged passes of
// InterpStencil::Constant, each 0 + 1 * v).
__global__ __launch_bounds__(256) void k_srv_prolong(const SLevel* __restrict__ lv, const SItem* __restrict__ items)
{
    const SItem it = items[blockIdx.x];
    const SLevel F = lv[it.lev], C = lv[it.lev + 1];
    const int n = F.n, nc = C.n;
    const float one = 1.0f;
    for (int t = it.a + threadIdx.x; t < it.a + it.b; t += 256) {
        const int x = t % n, y = (t / n) % n, z = t / (n * n);
        float v = 0;
        v = v + one * C.phi[(x >> 1) + (long)nc * ((y >> 1) + (long)nc * (z >> 1))];
        float v1 = 0;
        v1 = v1 + one * v;
        float v2 = 0;
        v2 = v2 + one * v1;
        F.phi[t] = F.phi[t] + v2;
    }
}

// FP64 residual r = rhs + sum_j d_j phi(p + o_j) of the finest level: FP32 copy
// into the level-0 rhs, with zero_e level-0 phi (the correction) = 0, max|r|
// into the request's norm slot (non-negative doubles order like their bit
// patterns).
__global__ __launch_bounds__(256) void k_srv_residual(const SReq* __restrict__ rq, const SLevel* __restrict__ lv,
                                                      const SItem* __restrict__ items,
                                                      unsigned long long* __restrict__ norms, int zero_e)
{
    __shared__ double red[16];
    const SItem it = items[blockIdx.x];
    const SReq R = rq[it.lev];
    const SLevel L = lv[R.lev0];
    const int n = R.n;
    const long nn = (long)n * n;
    double m = 0;
    for (int t = it.a + threadIdx.x; t < it.a + it.b; t += 256) {
        const int x = t % n, y = (t / n) % n, z = t / (n * n);
        const double* pz = R.phi + nn * z;
        const long row = (long)n * y;
        double s = 0;
        s = s + R.rhs[t];
        s = s + R.d0 * pz[x + row];
        s = s + R.d1 * pz[wrap1(x + 1, n) + row];
        s = s + R.d1 * pz[wrap1(x - 1, n) + row];
        s = s + R.d1 * pz[x + (long)n * wrap1(y + 1, n)];
        s = s + R.d1 * pz[x + (long)n * wrap1(y - 1, n)];
        s = s + R.d1 * R.phi[x + row + nn * wrap1(z + 1, n)];
        s = s + R.d1 * R.phi[x + row + nn * wrap1(z - 1, n)];
        L.rhs[t] = (float)s;
        if (zero_e) { L.phi[t] = 0.0f; }
        const double a = s > 0 ? s : -s;
        m = m < a ? a : m;
    }
    m = block_reduce(PM_RED_MAX, m, red);
    if (threadIdx.x == 0) { atomicMax(norms + R.slot, (unsigned long long)__double_as_longlong(m)); }
}

// phi (FP64) += e (the level-0 FP32 correction).
__global__ __launch_bounds__(256) void k_srv_correct(const SReq* __restrict__ rq, const SLevel* __restrict__ lv,
                                                     const SItem* __restrict__ items)
{
    const SItem it = items[blockIdx.x];
    const SReq R = rq[it.lev];
    const float* e = lv[R.lev0].phi;
    for (int t = it.a + threadIdx.x; t < it.a + it.b; t += 256) { R.phi[t] = R.phi[t] + (double)e[t]; }
}

__global__ __launch_bounds__(256) void k_srv_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { p[i] = 0ull; }
}

__global__ __launch_bounds__(256) void k_srv_zero_f64(double* p, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) { p[i] = 0.0; }
}

// rhs(x,y,z) = sum_k ((a_k * tx_k[x]) * ty_k[y]) * tz_k[z]; tables k-major, n each.
__global__ __launch_bounds__(256) void k_srv_gauss(double* __restrict__ rhs, int n, int K, const double* __restrict__ tab)
{
    const long total = (long)n * n * n;
    const double* amp = tab;
    const double* tx = tab + K;
    const double* ty = tx + (long)K * n;
    const double* tz = ty + (long)K * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % n), y = (int)((t / n) % n), z = (int)(t / ((long)n * n));
        double v = 0;
        for (int k = 0; k < K; k++) {
            v = v + ((amp[k] * tx[(long)k * n + x]) * ty[(long)k * n + y]) * tz[(long)k * n + z];
        }
        rhs[t] = v;
    }
}

// Fixed-order mean removal: partial sums over chunks of 4096 cells (block
// tree), their sum in chunk order by one block, then rhs -= sum / n^3.
constexpr int MEAN_CHUNK = 4096;
__global__ __launch_bounds__(256) void k_srv_chunk_sums(const double* __restrict__ rhs, long total,
                                                        double* __restrict__ part)
{
    __shared__ double red[16];
    const long c0 = (long)blockIdx.x * MEAN_CHUNK;
    double s = 0;
    for (long i = c0 + threadIdx.x; i < c0 + MEAN_CHUNK && i < total; i += 256) { s = s + rhs[i]; }
    s = block_reduce(PM_RED_SUM, s, red);
    if (threadIdx.x == 0) { part[blockIdx.x] = s; }
}

__global__ __launch_bounds__(256) void k_srv_sum_parts(double* __restrict__ part, int nparts)
{
    __shared__ double red[16];
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) { s = s + part[i]; }
    s = block_reduce(PM_RED_SUM, s, red);
    if (threadIdx.x == 0) { part[nparts] = s; }
}

__global__ __launch_bounds__(256) void k_srv_sub_mean(double* __restrict__ rhs, long total,
                                                      const double* __restrict__ sum)
{
    const double mean = *sum / (double)total;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        rhs[t] = rhs[t] - mean;
    }
}

}  // namespace pm

using namespace pm;

namespace {

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<int> level_sizes(int n)
{
    std::vector<int> ns{n};
    while (ns.back() % 2 == 0 && ns.back() > 8) { ns.push_back(ns.back() / 2); }
    return ns;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// SplitMix64 stream of pm_fill_uniform: value i of (seed, stream) in [-1, 1).
double splitmix_u(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t z = seed + stream * 0x632BE59BD9B4E019ull + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) * 0x1.0p-53) * 2.0 - 1.0;
}

// One resident hierarchy of a mode-0 request (one allocation).
struct Hier {
    int n = 0;
    void* base = nullptr;
    size_t bytes = 0;
    double* phi64 = nullptr;
    double* rhs64 = nullptr;
    double* tab = nullptr;  // Gaussian tables and mean partials
    std::vector<float*> phi, tmp, rhs;
};

size_t tab_elems(int n) { return 32 + 3 * 32 * (size_t)n + ((size_t)n * n * n + MEAN_CHUNK - 1) / MEAN_CHUNK + 1; }

size_t hier_bytes(int n)
{
    const size_t n3 = (size_t)n * n * n;
    size_t b = 2 * align256(8 * n3) + align256(8 * tab_elems(n));
    for (int m : level_sizes(n)) { b += 3 * align256(4 * (size_t)m * m * m); }
    return b;
}

struct Active {
    int idx;
    Hier* h;
    std::vector<int> ns;
    double res0 = 0, r = 0;
    int iters = 0;
    bool fresh = true;     // its first iteration only computes the residual
    bool restored = false; // resumed from a checkpoint: that residual is the one it had, not res0
};

// A mode-0 run between two calls of pm_server_run: stopped after the set
// number of loop iterations, or restored from a checkpoint.  The next
// pm_server_run with the same requests continues it.
struct HeldRun {
    bool held = false;
    bool restored = false;     // from a checkpoint: right-hand sides still to be made
    int nreq = 0;
    uint64_t fingerprint = 0;  // of the request list
    double elapsed = 0;        // seconds of the run before the stop
    size_t next = 0;           // position in arrival order of the next arrival
    int done = 0;
    std::vector<int> waiting;
    std::vector<Active> act;
    std::vector<pm_request_stats> stats;
};

struct Launch {
    int kind;  // 0 sweep, 1 restrict, 2 prolong, 3 correct, 4 residual, 5 small-level relaxation
    int off, count, dir;
    int in;    // sweeps: staged input (k_srv_gsrb IN), 3 = IN 0 with OUT 1; restrict / residual: 1 = write the zero guess
    size_t lds;  // sweeps: dynamic LDS of the largest tile
    long np;     // sweeps: largest phi plane of a tile (elements)
};

}  // namespace

struct pm_server {
    pm_server_options opt;
    std::map<int, std::vector<Hier*>> pool;  // free hierarchies by n
    size_t resident = 0;
    // device tables
    SLevel* d_lv = nullptr;
    size_t cap_lv = 0;
    SReq* d_rq = nullptr;
    size_t cap_rq = 0;
    SItem* d_items = nullptr;
    size_t cap_items = 0;
    unsigned long long* d_norms = nullptr;
    unsigned long long* h_norms = nullptr;
    std::vector<Launch> plan;
    // kept results of the last run
    std::map<int, std::vector<double>> results;
    // mode 0: stop pm_server_run after this many loop iterations (0: never)
    int stop_after = 0;
    HeldRun run;
    // mode 1: one pm_mg (FP64) per size, with its phi / rhs
    struct Plain {
        pm_layout* layout = nullptr;
        pm_mg* mg = nullptr;
        double* phi = nullptr;  // ghost 1
        double* rhs = nullptr;
        double* tab = nullptr;
        size_t bytes = 0;
    };
    std::map<int, Plain> plain;
};

namespace {

int grow(void** p, size_t* cap, size_t need)
{
    if (need <= *cap) { return PM_OK; }
    if (*p) { hipFree(*p); *p = nullptr; }
    size_t c = std::max(need, *cap * 2);
    if (hipMalloc(p, c) != hipSuccess) {
        *cap = 0;
        set_error("server: device allocation of %zu bytes failed", c);
        return PM_ERR_NOMEM;
    }
    *cap = c;
    return PM_OK;
}

void free_hier(Hier* h)
{
    hipFree(h->base);
    delete h;
}

int alloc_hier(int n, Hier** out)
{
    Hier* h = new Hier();
    h->n = n;
    h->bytes = hier_bytes(n);
    if (hipMalloc(&h->base, h->bytes) != hipSuccess) {
        delete h;
        set_error("server: hierarchy allocation of %zu bytes failed", hier_bytes(n));
        return PM_ERR_NOMEM;
    }
    char* p = (char*)h->base;
    const size_t n3 = (size_t)n * n * n;
    h->phi64 = (double*)p;
    p += align256(8 * n3);
    h->rhs64 = (double*)p;
    p += align256(8 * n3);
    h->tab = (double*)p;
    p += align256(8 * tab_elems(n));
    for (int m : level_sizes(n)) {
        const size_t b = align256(4 * (size_t)m * m * m);
        h->phi.push_back((float*)p);
        p += b;
        h->tmp.push_back((float*)p);
        p += b;
        h->rhs.push_back((float*)p);
        p += b;
    }
    *out = h;
    return PM_OK;
}

// Gaussian tables on the host: amplitudes, then x, y, z factors exp(-d^2 / (2 sigma^2)),
// d the periodic minimum-image distance from the cell centre (i + 1/2) / n.
void gauss_tables(int n, int K, uint64_t seed, double sigma, std::vector<double>& tab)
{
    tab.assign((size_t)K + 3 * (size_t)K * n, 0.0);
    for (int k = 0; k < K; k++) {
        const double u = splitmix_u(seed, 7, 3 * (uint64_t)K + k);
        tab[k] = u >= 0 ? 1.0 : -1.0;
        for (int d = 0; d < 3; d++) {
            const double c = (splitmix_u(seed, 7, 3 * (uint64_t)k + d) + 1.0) * 0.5;
            double* t = tab.data() + K + ((size_t)d * K + k) * n;
            for (int i = 0; i < n; i++) {
                double dx = (i + 0.5) / n - c;
                dx -= std::nearbyint(dx);
                t[i] = std::exp(-(dx * dx) / (2.0 * sigma * sigma));
            }
        }
    }
}

// rhs (n^3 FP64, device) = Gaussian sum minus its mean; tab = device scratch.
int make_rhs(double* rhs, double* tab_dev, int n, const pm_request& q, double sigma, hipStream_t s)
{
    std::vector<double> tab;
    gauss_tables(n, q.ngauss, q.seed, sigma, tab);
    PM_CHECK_HIP(hipMemcpyAsync(tab_dev, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, s));
    PM_CHECK_HIP(hipStreamSynchronize(s));  // the host table goes out of scope
    const long total = (long)n * n * n;
    const int nparts = (int)((total + MEAN_CHUNK - 1) / MEAN_CHUNK);
    double* part = tab_dev + tab.size();
    const unsigned g = (unsigned)std::min<long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_srv_gauss, dim3(g), dim3(256), 0, s, rhs, n, q.ngauss, (const double*)tab_dev);
    hipLaunchKernelGGL(k_srv_chunk_sums, dim3(nparts), dim3(256), 0, s, (const double*)rhs, total, part);
    hipLaunchKernelGGL(k_srv_sum_parts, dim3(1), dim3(256), 0, s, part, nparts);
    hipLaunchKernelGGL(k_srv_sub_mean, dim3(g), dim3(256), 0, s, rhs, total, (const double*)(part + nparts));
    PM_CHECK_LAUNCH();
    return PM_OK;
}

void level_coefs(int m, float* c0, float* c1, float* lam)
{
    const double dx = 1.0 / m;
    const double mm = -1.0 / (dx * dx);
    *c0 = (float)(-6.0) * (float)mm;
    *c1 = (float)(1.0) * (float)mm;
    double l = -(dx * dx) / (4.0 * 3);
    l *= 2.0;
    *lam = (float)(1.0 * l);
}

void push_cells(std::vector<SItem>& items, int lev, long cells)
{
    for (long c = 0; c < cells; c += SCELLS) { items.push_back(SItem{lev, (int)c, (int)std::min<long>(SCELLS, cells - c), 0}); }
}

// Device tables and the launch plan of one loop iteration for the active set.
int build_plan(pm_server* S, std::vector<Active>& act, hipStream_t s)
{
    std::vector<SLevel> lv;
    std::vector<SReq> rq;
    std::vector<int> lev0(act.size());
    int maxd = 0;
    for (size_t r = 0; r < act.size(); r++) {
        Active& a = act[r];
        lev0[r] = (int)lv.size();
        for (size_t l = 0; l < a.ns.size(); l++) {
            SLevel L;
            L.phi = a.h->phi[l];
            L.tmp = a.h->tmp[l];
            L.rhs = a.h->rhs[l];
            L.n = a.ns[l];
            L.rows = sweep_rows(a.ns[l]);
            L.tw = sweep_cols(a.ns[l]);
            L.acc = l == 0 ? a.h->phi64 : nullptr;
            level_coefs(a.ns[l], &L.c0, &L.c1, &L.lam);
            lv.push_back(L);
        }
        const double hsqinv = 1. / ((1.0 / a.ns[0]) * (1.0 / a.ns[0]));
        SReq q;
        q.phi = a.h->phi64;
        q.rhs = a.h->rhs64;
        q.lev0 = lev0[r];
        q.slot = (int)r;
        q.n = a.ns[0];
        q.pad = 0;
        q.d0 = (-6.0) * (-hsqinv);
        q.d1 = (1.0) * (-hsqinv);
        rq.push_back(q);
        maxd = std::max(maxd, (int)a.ns.size() - 1);
    }
    std::vector<SItem> items;
    S->plan.clear();
    auto begin = [&](int kind, int dir, int in = 0) {
        S->plan.push_back(Launch{kind, (int)items.size(), 0, dir, in, 0, 0});
    };
    auto end = [&]() {
        Launch& L = S->plan.back();
        L.count = (int)items.size() - L.off;
        if (L.count == 0) { S->plan.pop_back(); }
    };
    auto sweep_items = [&](int levidx, int m) {
        const int r = sweep_rows(m), w = sweep_cols(m);
        for (int z0 = 0; z0 < m; z0 += SZC)
            for (int y0 = 0; y0 < m; y0 += r)
                for (int x0 = 0; x0 < m; x0 += w) {
                    items.push_back(SItem{levidx, y0, z0, std::min(m, z0 + SZC), x0, std::min(m, x0 + w)});
                }
        Launch& L = S->plan.back();
        L.lds = std::max(L.lds, sweep_lds(m));
        L.np = std::max(L.np, (long)(r + 4) * (w + 4));
    };
    const int pre = S->opt.pre, post = S->opt.post, bot = S->opt.bottom;
    // The zero initial guess of every level's correction is folded into that
    // level's first sweep when every level sweeps first (pre and bottom
    // counts > 0); otherwise restriction and residual write it.  The constant
    // prolongation is folded into the first post sweep when there is one.
    const bool zero_in = pre > 0 && bot > 0;
    const bool fold_prolong = post > 0;
    // levels below the finest of at most SMALL_CELLS cells relax in one launch
    // each (k_srv_small), all requests' such levels of one depth together
    auto small = [&](const Active& a, int d) { return d > 0 && (long)a.ns[d] * a.ns[d] * a.ns[d] <= SMALL_CELLS; };
    // `inner` V-cycles of every non-fresh request (fresh ones only get their first
    // residual): the first takes the zero initial guess on every level, the later
    // ones start level 0 from the previous cycle's correction; only the last
    // cycle's last sweep of level 0 adds e into the FP64 solution
    const int inner = S->opt.inner > 1 ? S->opt.inner : 1;
    for (int cyc = 0; cyc < inner; cyc++) {
    const bool last_cycle = cyc == inner - 1;
    for (int d = 0; d <= maxd; d++) {
        // level 0 of a later cycle reads its correction; coarser levels start from zero
        const bool zin = zero_in && (d > 0 || cyc == 0);
        begin(5, 0, zin ? 1 : 0);
        for (size_t r = 0; r < act.size(); r++) {
            const Active& a = act[r];
            const int bd = (int)a.ns.size() - 1;
            if (a.fresh || d > bd || !small(a, d)) { continue; }
            const int cnt = d < bd ? pre : bot;
            if (cnt > 0) { items.push_back(SItem{lev0[r] + d, cnt, 0, 0, 0, 0}); }
        }
        end();
        const int nsw = std::max(pre, bot);
        for (int sw = 0; sw < nsw; sw++) {
            begin(0, sw & 1, sw == 0 && zin ? 1 : 0);
            for (size_t r = 0; r < act.size(); r++) {
                const Active& a = act[r];
                const int bd = (int)a.ns.size() - 1;
                if (a.fresh || d > bd || small(a, d)) { continue; }
                if ((d < bd && sw < pre) || (d == bd && sw < bot)) { sweep_items(lev0[r] + d, a.ns[d]); }
            }
            end();
        }
        begin(1, 0, zero_in ? 0 : 1);
        for (size_t r = 0; r < act.size(); r++) {
            const Active& a = act[r];
            if (a.fresh || d >= (int)a.ns.size() - 1) { continue; }
            const long c = (long)a.ns[d + 1] * a.ns[d + 1] * a.ns[d + 1];
            push_cells(items, lev0[r] + d, c);
        }
        end();
    }
    for (int d = maxd - 1; d >= 0; d--) {
        if (!fold_prolong) {
            begin(2, 0);
            for (size_t r = 0; r < act.size(); r++) {
                const Active& a = act[r];
                if (a.fresh || d >= (int)a.ns.size() - 1) { continue; }
                push_cells(items, lev0[r] + d, (long)a.ns[d] * a.ns[d] * a.ns[d]);
            }
            end();
        }
        if (post > 0) {
            begin(5, 0, fold_prolong ? 2 : 0);
            for (size_t r = 0; r < act.size(); r++) {
                const Active& a = act[r];
                if (a.fresh || d >= (int)a.ns.size() - 1 || !small(a, d)) { continue; }
                items.push_back(SItem{lev0[r] + d, post, 0, 0, 0, 0});
            }
            end();
        }
        for (int sw = 0; sw < post; sw++) {
            // the last sweep of level 0 adds e into the FP64 solution
            const int in = sw == 0 && fold_prolong ? 2 : 0;
            begin(0, sw & 1, last_cycle && d == 0 && sw == post - 1 && in == 0 ? 3 : in);
            for (size_t r = 0; r < act.size(); r++) {
                const Active& a = act[r];
                if (a.fresh || d >= (int)a.ns.size() - 1 || small(a, d)) { continue; }
                sweep_items(lev0[r] + d, a.ns[d]);
            }
            end();
        }
    }
    }
    // phi += e where no post sweep of level 0 added it (single-level requests,
    // or no post sweeps, or a post count of 1 whose only sweep takes the
    // prolongation input)
    const bool folded_e = post >= 2;
    begin(3, 0);
    for (size_t r = 0; r < act.size(); r++) {
        const bool single = act[r].ns.size() == 1;
        if (!act[r].fresh && (single || !folded_e)) {
            push_cells(items, (int)r, (long)act[r].ns[0] * act[r].ns[0] * act[r].ns[0]);
        }
    }
    end();
    begin(4, 0, zero_in ? 0 : 1);
    for (size_t r = 0; r < act.size(); r++) { push_cells(items, (int)r, (long)act[r].ns[0] * act[r].ns[0] * act[r].ns[0]); }
    end();
    int rc = grow((void**)&S->d_lv, &S->cap_lv, sizeof(SLevel) * lv.size());
    if (rc == PM_OK) { rc = grow((void**)&S->d_rq, &S->cap_rq, sizeof(SReq) * rq.size()); }
    if (rc == PM_OK) { rc = grow((void**)&S->d_items, &S->cap_items, sizeof(SItem) * std::max<size_t>(1, items.size())); }
    if (rc) { return rc; }
    // the stream is idle here (the last norm read-back synchronised it)
    PM_CHECK_HIP(hipMemcpy(S->d_lv, lv.data(), sizeof(SLevel) * lv.size(), hipMemcpyHostToDevice));
    PM_CHECK_HIP(hipMemcpy(S->d_rq, rq.data(), sizeof(SReq) * rq.size(), hipMemcpyHostToDevice));
    if (!items.empty()) {
        PM_CHECK_HIP(hipMemcpy(S->d_items, items.data(), sizeof(SItem) * items.size(), hipMemcpyHostToDevice));
    }
    (void)s;
    return PM_OK;
}

int run_plan(pm_server* S, int nact, hipStream_t s)
{
    hipLaunchKernelGGL(k_srv_zero_u64, dim3((nact + 255) / 256), dim3(256), 0, s, S->d_norms, nact);
    for (const Launch& L : S->plan) {
        const SItem* it = S->d_items + L.off;
        switch (L.kind) {
        case 0: {
            const size_t lds = L.lds;
            const long np = L.np;
#define PM_SRV_LAUNCH(E, I, O) \
    hipLaunchKernelGGL((k_srv_gsrb<E, I, O>), dim3(L.count), dim3(256), lds, s, (const SLevel*)S->d_lv, it, L.dir)
#define PM_SRV_SWEEP(E)            \
    if (L.in == 0) {               \
        PM_SRV_LAUNCH(E, 0, 0);    \
    } else if (L.in == 1) {        \
        PM_SRV_LAUNCH(E, 1, 0);    \
    } else if (L.in == 2) {        \
        PM_SRV_LAUNCH(E, 2, 0);    \
    } else {                       \
        PM_SRV_LAUNCH(E, 0, 1);    \
    }
            if (np <= 1024) {
                PM_SRV_SWEEP(4)
            } else if (np <= 2048) {
                PM_SRV_SWEEP(8)
            } else {
                PM_SRV_SWEEP(16)
            }
#undef PM_SRV_SWEEP
#undef PM_SRV_LAUNCH
            break;
        }
        case 1:
            hipLaunchKernelGGL(k_srv_restrict, dim3(L.count), dim3(256), 0, s, (const SLevel*)S->d_lv, it, L.in);
            break;
        case 5:
            if (L.in == 0) {
                hipLaunchKernelGGL(k_srv_small<0>, dim3(L.count), dim3(256), 0, s, (const SLevel*)S->d_lv, it);
            } else if (L.in == 1) {
                hipLaunchKernelGGL(k_srv_small<1>, dim3(L.count), dim3(256), 0, s, (const SLevel*)S->d_lv, it);
            } else {
                hipLaunchKernelGGL(k_srv_small<2>, dim3(L.count), dim3(256), 0, s, (const SLevel*)S->d_lv, it);
            }
            break;
        case 2:
            hipLaunchKernelGGL(k_srv_prolong, dim3(L.count), dim3(256), 0, s, (const SLevel*)S->d_lv, it);
            break;
        case 3:
            hipLaunchKernelGGL(k_srv_correct, dim3(L.count), dim3(256), 0, s, (const SReq*)S->d_rq,
                               (const SLevel*)S->d_lv, it);
            break;
        default:
            hipLaunchKernelGGL(k_srv_residual, dim3(L.count), dim3(256), 0, s, (const SReq*)S->d_rq,
                               (const SLevel*)S->d_lv, it, S->d_norms, L.in);
        }
    }
    PM_CHECK_LAUNCH();
    PM_CHECK_HIP(hipMemcpyAsync(S->h_norms, S->d_norms, sizeof(unsigned long long) * nact, hipMemcpyDeviceToHost, s));
    PM_CHECK_HIP(hipStreamSynchronize(s));
    return PM_OK;
}

// Take a hierarchy for n from the pool or allocate one within the budget
// (evicting free hierarchies of other sizes when needed).  *h = null: wait.
int take_hier(pm_server* S, int n, Hier** h)
{
    *h = nullptr;
    auto& fl = S->pool[n];
    if (!fl.empty()) {
        *h = fl.back();
        fl.pop_back();
        return PM_OK;
    }
    const size_t need = hier_bytes(n);
    for (auto& kv : S->pool) {
        while (S->resident + need > S->opt.hbm_budget && !kv.second.empty()) {
            Hier* v = kv.second.back();
            kv.second.pop_back();
            S->resident -= v->bytes;
            free_hier(v);
        }
    }
    if (S->resident + need > S->opt.hbm_budget) { return PM_OK; }
    int rc = alloc_hier(n, h);
    if (rc == PM_OK) { S->resident += need; }
    return rc;
}

// FNV-1a over the request fields that decide a run; a held run continues only
// with the list it was started with.
uint64_t fingerprint(int nreq, const pm_request* reqs)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t b) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < b; i++) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(&nreq, sizeof(nreq));
    for (int i = 0; i < nreq; i++) {
        const pm_request& q = reqs[i];
        mix(&q.arrival_s, sizeof(q.arrival_s));
        mix(&q.n, sizeof(q.n));
        mix(&q.ngauss, sizeof(q.ngauss));
        mix(&q.seed, sizeof(q.seed));
        mix(&q.tol, sizeof(q.tol));
        mix(&q.keep_result, sizeof(q.keep_result));
    }
    return h;
}

// Hierarchies of in-flight requests back to the pool (a failed run, a
// discarded hold): pm_server_destroy then frees them.
void release_active(pm_server* S, std::vector<Active>& act)
{
    for (Active& a : act) {
        if (a.h) { S->pool[a.h->n].push_back(a.h); }
        a.h = nullptr;
    }
    act.clear();
}

int run_batched(pm_server* S, int nreq, const pm_request* reqs, pm_request_stats* st, hipStream_t s)
{
    std::vector<int> order(nreq);
    for (int i = 0; i < nreq; i++) { order[i] = i; }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    const uint64_t fp = fingerprint(nreq, reqs);
    HeldRun R;
    if (S->run.held) {
        PM_REQUIRE(S->run.nreq == nreq && S->run.fingerprint == fp,
                   "pm_server_run: the held run (stopped or restored) belongs to another request list; pass the same "
                   "list or drop it with pm_server_discard");
        R = std::move(S->run);
        S->run = HeldRun();
        std::copy(R.stats.begin(), R.stats.end(), st);
    }
    // on error the in-flight hierarchies go back to the pool
    auto fail = [&](int rc) {
        release_active(S, R.act);
        return rc;
    };
    // a restored run: the right-hand sides are regenerated from the requests
    for (Active& a : R.act) {
        if (!R.restored) { break; }
        if (reqs[a.idx].n != a.ns[0]) {
            set_error("pm_server_run: request %d does not match its checkpoint", a.idx);
            return fail(PM_ERR_INVALID);
        }
        const int rc = make_rhs(a.h->rhs64, a.h->tab, reqs[a.idx].n, reqs[a.idx], S->opt.sigma, s);
        if (rc) { return fail(rc); }
    }
    R.restored = false;
    bool dirty = true;
    int iterations = 0;
    const auto t0 = std::chrono::steady_clock::now() -
                    std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(R.elapsed));
    while (R.done < nreq) {
        const double t = now_s(t0);
        while (R.next < order.size() && reqs[order[R.next]].arrival_s <= t) {
            const int i = order[R.next++];
            if (hier_bytes(reqs[i].n) > S->opt.hbm_budget) {
                st[i].status = PM_SERVE_REJECTED;
                st[i].admit_s = st[i].finish_s = t;
                R.done++;
            } else {
                R.waiting.push_back(i);
            }
        }
        // admission, first come first served
        while (!R.waiting.empty() && (int)R.act.size() < S->opt.max_active) {
            const int i = R.waiting.front();
            Hier* h = nullptr;
            int rc = take_hier(S, reqs[i].n, &h);
            if (rc) { return fail(rc); }
            if (!h) { break; }
            R.waiting.erase(R.waiting.begin());
            Active a;
            a.idx = i;
            a.h = h;
            a.ns = level_sizes(reqs[i].n);
            R.act.push_back(a);
            const long total = (long)reqs[i].n * reqs[i].n * reqs[i].n;
            rc = make_rhs(h->rhs64, h->tab, reqs[i].n, reqs[i], S->opt.sigma, s);
            if (rc) { return fail(rc); }
            const unsigned g = (unsigned)std::min<long>((total + 255) / 256, 4096);
            hipLaunchKernelGGL(k_srv_zero_f64, dim3(g), dim3(256), 0, s, h->phi64, total);
            if (hipGetLastError() != hipSuccess) {
                set_error("server: zero fill launch failed");
                return fail(PM_ERR_HIP);
            }
            st[i].admit_s = now_s(t0);
            st[i].bytes = h->bytes;
            dirty = true;
        }
        if (R.act.empty()) {
            if (R.next < order.size()) {
                const double wait = reqs[order[R.next]].arrival_s - now_s(t0);
                if (wait > 0) { std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001))); }
            }
            continue;
        }
        if (dirty) {
            int rc = build_plan(S, R.act, s);
            if (rc) { return fail(rc); }
            dirty = false;
        }
        int rc = run_plan(S, (int)R.act.size(), s);
        if (rc) { return fail(rc); }
        const double tf = now_s(t0);
        std::vector<Active> keep;
        for (size_t r = 0; r < R.act.size(); r++) {
            Active& a = R.act[r];
            double v;
            std::memcpy(&v, &S->h_norms[r], sizeof(double));
            const pm_request& q = reqs[a.idx];
            bool fin = false;
            if (a.fresh) {
                a.fresh = false;
                if (a.restored) {
                    // the residual of the checkpointed solution: its norm is the
                    // one recorded at the checkpoint, and no cycle is counted
                    a.restored = false;
                    a.r = v;
                    fin = v < q.tol * a.res0 || a.iters >= S->opt.max_iter;
                } else {
                    a.res0 = a.r = v;
                    fin = v == 0.0;
                }
                dirty = true;  // its V-cycles start next iteration
            } else {
                a.iters++;
                a.r = v;
                fin = v < q.tol * a.res0 || a.iters >= S->opt.max_iter;
            }
            if (!fin) {
                keep.push_back(a);
                continue;
            }
            pm_request_stats& o = st[a.idx];
            o.iters = a.iters;
            o.rel_residual = a.res0 > 0 ? a.r / a.res0 : 0.0;
            o.status = (a.r < q.tol * a.res0 || a.res0 == 0.0) ? PM_SERVE_SOLVED : PM_SERVE_NOT_CONVERGED;
            o.finish_s = tf;
            if (q.keep_result && o.status == PM_SERVE_SOLVED) {
                std::vector<double>& out = S->results[a.idx];
                out.resize((size_t)q.n * q.n * q.n);
                if (hipMemcpy(out.data(), a.h->phi64, sizeof(double) * out.size(), hipMemcpyDeviceToHost) != hipSuccess) {
                    set_error("server: result read-back failed");
                    for (size_t k = r; k < R.act.size(); k++) { keep.push_back(R.act[k]); }
                    R.act.swap(keep);
                    return fail(PM_ERR_HIP);
                }
            }
            S->pool[a.h->n].push_back(a.h);
            R.done++;
            dirty = true;
        }
        R.act.swap(keep);
        if (S->stop_after > 0 && ++iterations >= S->stop_after && R.done < nreq) {
            R.held = true;
            R.nreq = nreq;
            R.fingerprint = fp;
            R.elapsed = now_s(t0);
            R.stats.assign(st, st + nreq);
            S->run = std::move(R);
            set_error("pm_server_run: stopped after %d loop iterations with %d of %d requests done; the run is held "
                      "for pm_server_checkpoint and the next pm_server_run",
                      iterations, S->run.done, nreq);
            return PM_ERR_CANCELLED;
        }
    }
    return PM_OK;
}

void free_plain(pm_server::Plain& P)
{
    pm_mg_destroy(P.mg);
    pm_layout_destroy(P.layout);
    hipFree(P.phi);
    hipFree(P.rhs);
    hipFree(P.tab);
    P = pm_server::Plain();
}

// Mode 1: one request at a time, FP64 GSRB V-cycles of the multigrid driver,
// within the HBM budget (solvers of other sizes are evicted; a size that does
// not fit alone is rejected).
int run_plain(pm_server* S, int nreq, const pm_request* reqs, pm_request_stats* st, hipStream_t s)
{
    std::vector<int> order(nreq);
    for (int i = 0; i < nreq; i++) { order[i] = i; }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    const auto t0 = std::chrono::steady_clock::now();
    for (int i : order) {
        const pm_request& q = reqs[i];
        while (now_s(t0) < q.arrival_s) {
            std::this_thread::sleep_for(std::chrono::duration<double>(std::min(q.arrival_s - now_s(t0), 0.001)));
        }
        const int n = q.n;
        auto it = S->plain.find(n);
        if (it == S->plain.end()) {
            pm_server::Plain P;
            const int nl = (int)level_sizes(n).size() - 1;
            int dom[3] = {n, n, n}, per[3] = {1, 1, 1}, lows[3] = {0, 0, 0};
            int rc = pm_layout_create(&P.layout, 3, dom, dom, per, 1, lows);
            if (rc) { return rc; }
            pm_mg_options o;
            pm_mg_default_options(&o);
            o.relax = 0;
            o.pre = S->opt.pre;
            o.post = S->opt.post;
            o.bottom = S->opt.bottom;
            o.precision = 0;
            o.max_iter = S->opt.max_iter;
            o.tol = q.tol;
            rc = pm_mg_create(&P.mg, P.layout, nl, 1.0 / n, &o);
            if (rc) {
                free_plain(P);
                return rc;
            }
            const size_t n3 = (size_t)n * n * n, ng = (size_t)(n + 2) * (n + 2) * (n + 2);
            if (hipMalloc(&P.phi, 8 * ng) != hipSuccess || hipMalloc(&P.rhs, 8 * n3) != hipSuccess ||
                hipMalloc(&P.tab, 8 * tab_elems(n)) != hipSuccess) {
                free_plain(P);
                set_error("server: plain-mode allocation for n = %d failed", n);
                return PM_ERR_NOMEM;
            }
            size_t mb = 0;
            pm_mg_bytes(P.mg, &mb);
            P.bytes = mb + 8 * (ng + n3 + tab_elems(n));
            // the HBM budget: evict the solvers of other sizes, else reject
            for (auto e = S->plain.begin(); e != S->plain.end() && S->resident + P.bytes > S->opt.hbm_budget;) {
                S->resident -= e->second.bytes;
                free_plain(e->second);
                e = S->plain.erase(e);
            }
            if (S->resident + P.bytes > S->opt.hbm_budget) {
                free_plain(P);
                st[i].status = PM_SERVE_REJECTED;
                st[i].admit_s = st[i].finish_s = now_s(t0);
                continue;
            }
            S->resident += P.bytes;
            it = S->plain.emplace(n, P).first;
        }
        pm_server::Plain& P = it->second;
        st[i].admit_s = now_s(t0);
        st[i].bytes = P.bytes;
        int rc = make_rhs(P.rhs, P.tab, n, q, S->opt.sigma, s);
        if (rc) { return rc; }
        const size_t ng = (size_t)(n + 2) * (n + 2) * (n + 2);
        hipLaunchKernelGGL(k_srv_zero_f64, dim3(4096), dim3(256), 0, s, P.phi, (long)ng);
        PM_CHECK_LAUNCH();
        std::vector<double> hist(S->opt.max_iter + 1, 0.0);
        int iters = 0;
        rc = pm_mg_set_tolerance(P.mg, q.tol, S->opt.max_iter);  // the request's own tolerance
        if (rc) { return rc; }
        rc = pm_mg_solve(P.mg, P.phi, P.rhs, hist.data(), &iters, s);
        if (rc != PM_OK && rc != PM_ERR_NOT_CONVERGED) { return rc; }
        st[i].finish_s = now_s(t0);
        st[i].iters = iters;
        st[i].rel_residual = hist[0] > 0 ? hist[iters] / hist[0] : 0.0;
        st[i].status = rc == PM_OK ? PM_SERVE_SOLVED : PM_SERVE_NOT_CONVERGED;
        if (q.keep_result && rc == PM_OK) {
            std::vector<double> g(ng);
            PM_CHECK_HIP(hipMemcpy(g.data(), P.phi, 8 * ng, hipMemcpyDeviceToHost));
            std::vector<double>& out = S->results[i];
            out.resize((size_t)n * n * n);
            const size_t X = n + 2;
            for (int z = 0; z < n; z++)
                for (int y = 0; y < n; y++)
                    for (int x = 0; x < n; x++) {
                        out[x + (size_t)n * (y + (size_t)n * z)] = g[(x + 1) + X * ((y + 1) + X * (z + 1))];
                    }
        }
    }
    return PM_OK;
}

// ---------------------------------------------------------------- checkpoint
// Buffer of a held mode-0 run (host memory, native byte order):
//   CkHead | pm_request_stats[nreq] | int32 waiting[] | CkActive[] | CkResult[]
//   | CkPool[] | (8-byte aligned) solution blocks of the active requests, then
//   the kept results.
// Each solution block is LevelBoxData<double, 1>::linearOut of the request's
// one-box periodic level without ghosts (Proto_LevelBoxDataImplem.H:465-477:
// every patch in layout order, C * |box| values x fastest): n^3 values, which
// a Proto program reads back with linearIn into a LevelBoxData on the same
// layout.  The FP32 hierarchy is not saved: each V-cycle starts it from zero,
// and the right-hand side is regenerated from the request.
constexpr char kCkMagic[8] = {'P', 'M', 'S', 'R', 'V', 'C', 'K', '1'};
constexpr uint32_t kCkVersion = 2;

struct CkHead {
    char magic[8];
    uint32_t version, stats_bytes;
    int32_t nreq, done, nwaiting, nactive, nresults, npool;
    uint64_t fingerprint, next;
    double elapsed;
    int32_t pre, post, bottom, max_iter;
    double sigma;
    int32_t inner, pad;
};
struct CkActive {
    int32_t idx, n, iters, fresh;
    double res0, r;
};
struct CkResult {
    int32_t idx, n;
};
struct CkPool {
    int32_t n, count;
};

struct CkLayout {
    size_t stats, waiting, active, results, pool, data, total;
    std::vector<CkPool> pools;
};

size_t align8(size_t b) { return (b + 7) & ~(size_t)7; }

CkLayout ck_layout(const pm_server* S)
{
    const HeldRun& R = S->run;
    CkLayout L;
    for (const auto& kv : S->pool) {
        if (!kv.second.empty()) { L.pools.push_back(CkPool{kv.first, (int32_t)kv.second.size()}); }
    }
    L.stats = sizeof(CkHead);
    L.waiting = L.stats + sizeof(pm_request_stats) * (size_t)R.nreq;
    L.active = L.waiting + sizeof(int32_t) * R.waiting.size();
    L.results = L.active + sizeof(CkActive) * R.act.size();
    L.pool = L.results + sizeof(CkResult) * S->results.size();
    L.data = align8(L.pool + sizeof(CkPool) * L.pools.size());
    size_t b = L.data;
    for (const Active& a : R.act) { b += 8 * (size_t)a.ns[0] * a.ns[0] * a.ns[0]; }
    for (const auto& kv : S->results) { b += 8 * kv.second.size(); }
    L.total = b;
    return L;
}

// Proto's linearOut / linearIn of a one-box level of n^3 cells without ghosts
int level_linear(int n, double* dev, char* host, bool out, hipStream_t s)
{
    pm_layout* lay = nullptr;
    int dom[3] = {n, n, n}, per[3] = {1, 1, 1}, lows[3] = {0, 0, 0}, g0[3] = {0, 0, 0};
    int rc = pm_layout_create(&lay, 3, dom, dom, per, 1, lows);
    if (rc) { return rc; }
    rc = out ? pm_level_linear_out(lay, PM_F64, dev, g0, 1, host, 1, (pm_stream_t)s)
             : pm_level_linear_in(lay, PM_F64, dev, g0, 1, host, 1, (pm_stream_t)s);
    pm_layout_destroy(lay);
    return rc;
}

int ck_write(pm_server* S, char* buf, const CkLayout& L, hipStream_t s)
{
    const HeldRun& R = S->run;
    CkHead h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kCkMagic, sizeof(kCkMagic));
    h.version = kCkVersion;
    h.stats_bytes = (uint32_t)sizeof(pm_request_stats);
    h.nreq = R.nreq;
    h.done = R.done;
    h.nwaiting = (int32_t)R.waiting.size();
    h.nactive = (int32_t)R.act.size();
    h.nresults = (int32_t)S->results.size();
    h.npool = (int32_t)L.pools.size();
    h.fingerprint = R.fingerprint;
    h.next = R.next;
    h.elapsed = R.elapsed;
    h.pre = S->opt.pre;
    h.post = S->opt.post;
    h.bottom = S->opt.bottom;
    h.max_iter = S->opt.max_iter;
    h.sigma = S->opt.sigma;
    h.inner = S->opt.inner > 1 ? S->opt.inner : 1;
    h.pad = 0;
    std::memcpy(buf, &h, sizeof(h));
    std::memcpy(buf + L.stats, R.stats.data(), sizeof(pm_request_stats) * R.stats.size());
    int32_t* w = (int32_t*)(buf + L.waiting);
    for (size_t k = 0; k < R.waiting.size(); k++) { w[k] = R.waiting[k]; }
    CkActive* ca = (CkActive*)(buf + L.active);
    size_t off = L.data;
    for (size_t k = 0; k < R.act.size(); k++) {
        const Active& a = R.act[k];
        ca[k] = CkActive{a.idx, a.ns[0], a.iters, a.fresh ? 1 : 0, a.res0, a.r};
        const int rc = level_linear(a.ns[0], a.h->phi64, buf + off, true, s);
        if (rc) { return rc; }
        off += 8 * (size_t)a.ns[0] * a.ns[0] * a.ns[0];
    }
    CkResult* cr = (CkResult*)(buf + L.results);
    size_t k = 0;
    for (const auto& kv : S->results) {
        int n = 1;
        while ((size_t)n * n * n < kv.second.size()) { n++; }
        cr[k++] = CkResult{kv.first, n};
        std::memcpy(buf + off, kv.second.data(), 8 * kv.second.size());
        off += 8 * kv.second.size();
    }
    std::memcpy(buf + L.pool, L.pools.data(), sizeof(CkPool) * L.pools.size());
    return PM_OK;
}

int ck_read(pm_server* S, const char* buf, size_t bytes, hipStream_t s)
{
    CkHead h;
    PM_REQUIRE(bytes >= sizeof(h), "pm_server_restore: buffer too short for a checkpoint header");
    std::memcpy(&h, buf, sizeof(h));
    PM_REQUIRE(std::memcmp(h.magic, kCkMagic, sizeof(kCkMagic)) == 0 && h.version == kCkVersion &&
                   h.stats_bytes == sizeof(pm_request_stats),
               "pm_server_restore: not a server checkpoint of this version");
    PM_REQUIRE(h.nreq >= 1 && h.done >= 0 && h.done <= h.nreq && h.nwaiting >= 0 && h.nactive >= 0 &&
                   h.nresults >= 0 && h.npool >= 0 && h.nwaiting + h.nactive + h.done <= h.nreq &&
                   h.next <= (uint64_t)h.nreq,
               "pm_server_restore: inconsistent checkpoint header");
    PM_REQUIRE(h.pre == S->opt.pre && h.post == S->opt.post && h.bottom == S->opt.bottom &&
                   h.max_iter == S->opt.max_iter && h.sigma == S->opt.sigma &&
                   h.inner == (S->opt.inner > 1 ? S->opt.inner : 1),
               "pm_server_restore: the checkpoint was written by a server with other sweep counts / max_iter / sigma "
               "/ inner cycles");
    PM_REQUIRE(h.nactive <= S->opt.max_active, "pm_server_restore: %d requests in flight, max_active is %d", h.nactive,
               S->opt.max_active);
    HeldRun R;
    R.nreq = h.nreq;
    R.done = h.done;
    R.fingerprint = h.fingerprint;
    R.next = (size_t)h.next;
    R.elapsed = h.elapsed;
    size_t off = sizeof(h);
    auto take = [&](size_t b) -> const char* {
        if (off + b > bytes) { return nullptr; }
        const char* p = buf + off;
        off += b;
        return p;
    };
    const char* ps = take(sizeof(pm_request_stats) * (size_t)h.nreq);
    const char* pw = ps ? take(sizeof(int32_t) * (size_t)h.nwaiting) : nullptr;
    const char* pa = pw ? take(sizeof(CkActive) * (size_t)h.nactive) : nullptr;
    const char* pr = pa ? take(sizeof(CkResult) * (size_t)h.nresults) : nullptr;
    const char* pp = pr ? take(sizeof(CkPool) * (size_t)h.npool) : nullptr;
    PM_REQUIRE(pp, "pm_server_restore: buffer too short for the checkpoint tables");
    off = align8(off);
    R.stats.resize(h.nreq);
    std::memcpy(R.stats.data(), ps, sizeof(pm_request_stats) * (size_t)h.nreq);
    for (int k = 0; k < h.nwaiting; k++) {
        int32_t i;
        std::memcpy(&i, pw + 4 * k, 4);
        PM_REQUIRE(i >= 0 && i < h.nreq, "pm_server_restore: bad waiting entry");
        R.waiting.push_back(i);
    }
    std::vector<CkActive> ca(h.nactive);
    std::vector<CkResult> cr(h.nresults);
    std::vector<CkPool> cp(h.npool);
    std::memcpy(ca.data(), pa, sizeof(CkActive) * ca.size());
    std::memcpy(cr.data(), pr, sizeof(CkResult) * cr.size());
    std::memcpy(cp.data(), pp, sizeof(CkPool) * cp.size());
    std::map<int, std::vector<double>> results;
    int rc = PM_OK;
    for (const CkActive& c : ca) {
        if (rc) { break; }
        if (!(c.idx >= 0 && c.idx < h.nreq && c.n >= 8 && c.n <= 1024 && c.n % 2 == 0 && c.iters >= 0)) {
            set_error("pm_server_restore: bad in-flight entry");
            rc = PM_ERR_INVALID;
            break;
        }
        const size_t nb = 8 * (size_t)c.n * c.n * c.n;
        const char* pd = take(nb);
        if (!pd) {
            set_error("pm_server_restore: buffer too short for the solution of request %d", c.idx);
            rc = PM_ERR_INVALID;
            break;
        }
        Hier* hh = nullptr;
        rc = take_hier(S, c.n, &hh);
        if (rc == PM_OK && !hh) {
            set_error("pm_server_restore: the HBM budget has no room for request %d (n = %d)", c.idx, c.n);
            rc = PM_ERR_NOMEM;
        }
        if (rc) { break; }
        Active a;
        a.idx = c.idx;
        a.h = hh;
        a.ns = level_sizes(c.n);
        a.iters = c.iters;
        a.res0 = c.res0;
        a.r = c.r;
        a.fresh = true;            // its first iteration recomputes the residual of the saved solution
        a.restored = c.fresh == 0; // a request still fresh at the checkpoint starts over
        R.act.push_back(a);
        rc = level_linear(c.n, hh->phi64, (char*)pd, false, s);
    }
    for (const CkResult& c : cr) {
        if (rc) { break; }
        const size_t n3 = (size_t)c.n * c.n * c.n;
        const char* pd = (c.idx >= 0 && c.idx < h.nreq && c.n >= 1 && c.n <= 1024) ? take(8 * n3) : nullptr;
        if (!pd) {
            set_error("pm_server_restore: bad or truncated kept result");
            rc = PM_ERR_INVALID;
            break;
        }
        std::vector<double>& v = results[c.idx];
        v.resize(n3);
        std::memcpy(v.data(), pd, 8 * n3);
    }
    if (rc) {
        release_active(S, R.act);
        return rc;
    }
    // the resident pool the run had: warm hierarchies, as far as the budget allows
    for (const CkPool& c : cp) {
        if (c.n < 8 || c.n > 1024 || c.n % 2 || c.count < 0) { continue; }
        for (int k = (int)S->pool[c.n].size(); k < c.count && S->resident + hier_bytes(c.n) <= S->opt.hbm_budget; k++) {
            Hier* hh = nullptr;
            if (alloc_hier(c.n, &hh) != PM_OK) { break; }
            S->resident += hh->bytes;
            S->pool[c.n].push_back(hh);
        }
    }
    R.held = true;
    R.restored = true;
    S->run = std::move(R);
    S->results = std::move(results);
    return PM_OK;
}

}  // namespace

extern "C" {

void pm_server_default_options(pm_server_options* o)
{
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->hbm_budget = (size_t)200 << 30;
    o->max_active = 64;
    o->mode = 0;
    o->pre = 2;
    o->post = 2;
    o->bottom = 10;
    o->max_iter = 30;
    o->sigma = 0.05;
    o->inner = 1;
}

int pm_server_create(pm_server** out, const pm_server_options* o)
{
    PM_REQUIRE(out && o, "pm_server_create: null argument");
    *out = nullptr;
    PM_REQUIRE(o->mode == 0 || o->mode == 1, "pm_server_create: mode must be 0 or 1");
    PM_REQUIRE(o->max_active >= 1 && o->max_active <= 4096, "pm_server_create: max_active must be 1..4096");
    PM_REQUIRE(o->pre >= 0 && o->post >= 0 && o->bottom >= 0 && o->pre % 2 == 0 && o->post % 2 == 0 &&
                   o->bottom % 2 == 0,
               "pm_server_create: sweep counts must be even");
    PM_REQUIRE(o->max_iter >= 1 && o->sigma > 0, "pm_server_create: bad max_iter / sigma");
    PM_REQUIRE(o->inner >= 0 && o->inner <= 4, "pm_server_create: inner must be 0..4");
    pm_server* S = new pm_server();
    S->opt = *o;
    if (hipMalloc((void**)&S->d_norms, sizeof(unsigned long long) * o->max_active) != hipSuccess ||
        hipHostMalloc((void**)&S->h_norms, sizeof(unsigned long long) * o->max_active, hipHostMallocDefault) !=
            hipSuccess) {
        pm_server_destroy(S);
        set_error("pm_server_create: allocation failed");
        return PM_ERR_NOMEM;
    }
    *out = S;
    return PM_OK;
}

int pm_server_destroy(pm_server* S)
{
    if (!S) { return PM_OK; }
    release_active(S, S->run.act);  // a held run's in-flight hierarchies
    for (auto& kv : S->pool)
        for (Hier* h : kv.second) { free_hier(h); }
    for (auto& kv : S->plain) { free_plain(kv.second); }
    hipFree(S->d_lv);
    hipFree(S->d_rq);
    hipFree(S->d_items);
    hipFree(S->d_norms);
    if (S->h_norms) { hipHostFree(S->h_norms); }
    delete S;
    return PM_OK;
}

int pm_server_request_bytes(const pm_server* S, int n, size_t* bytes)
{
    PM_REQUIRE(S && bytes && n >= 4 && n % 2 == 0, "pm_server_request_bytes: bad argument");
    *bytes = hier_bytes(n);
    return PM_OK;
}

int pm_server_resident_bytes(const pm_server* S, size_t* bytes)
{
    PM_REQUIRE(S && bytes, "pm_server_resident_bytes: null argument");
    *bytes = S->resident;
    return PM_OK;
}

int pm_server_run(pm_server* S, int nreq, const pm_request* reqs, pm_request_stats* st, pm_stream_t stream)
{
    PM_REQUIRE(S && reqs && st && nreq >= 1, "pm_server_run: bad argument");
    for (int i = 0; i < nreq; i++) {
        const pm_request& q = reqs[i];
        PM_REQUIRE(q.n >= 8 && q.n <= 1024 && q.n % 2 == 0, "pm_server_run: request %d: n must be even, 8..1024", i);
        const std::vector<int> ns = level_sizes(q.n);
        PM_REQUIRE(ns.back() >= 4, "pm_server_run: request %d: coarsest level below 4 cells", i);
        PM_REQUIRE(q.ngauss >= 1 && q.ngauss <= 32, "pm_server_run: request %d: ngauss must be 1..32", i);
        PM_REQUIRE(q.tol > 0 && q.arrival_s >= 0, "pm_server_run: request %d: bad tol / arrival", i);
        PM_REQUIRE(S->opt.mode == 1 || (size_t)q.n * q.n * q.n < ((size_t)1 << 31),
                   "pm_server_run: request %d too large", i);
    }
    if (!S->run.held) {
        // a held run keeps the results of the requests it finished before the stop
        S->results.clear();
    }
    for (int i = 0; i < nreq; i++) {
        pm_request_stats& o = st[i];
        std::memset(&o, 0, sizeof(o));
        o.arrival_s = reqs[i].arrival_s;
        o.status = -1;
    }
    const hipStream_t s = (hipStream_t)stream;
    return S->opt.mode == 0 ? run_batched(S, nreq, reqs, st, s) : run_plain(S, nreq, reqs, st, s);
}

int pm_server_set_stop(pm_server* S, int iterations)
{
    PM_REQUIRE(S && iterations >= 0, "pm_server_set_stop: bad argument");
    if (S->opt.mode != 0) {
        set_error("pm_server_set_stop: only the batched mode (0) stops between loop iterations");
        return PM_ERR_UNSUPPORTED;
    }
    S->stop_after = iterations;
    return PM_OK;
}

int pm_server_discard(pm_server* S)
{
    PM_REQUIRE(S, "pm_server_discard: null argument");
    release_active(S, S->run.act);
    S->run = HeldRun();
    S->results.clear();
    return PM_OK;
}

int pm_server_checkpoint_size(const pm_server* S, size_t* bytes)
{
    PM_REQUIRE(S && bytes, "pm_server_checkpoint_size: null argument");
    PM_REQUIRE(S->run.held, "pm_server_checkpoint_size: no held run (stop one with pm_server_set_stop)");
    *bytes = ck_layout(S).total;
    return PM_OK;
}

int pm_server_checkpoint(pm_server* S, void* buffer, size_t bytes, pm_stream_t stream)
{
    PM_REQUIRE(S && buffer, "pm_server_checkpoint: null argument");
    PM_REQUIRE(S->run.held, "pm_server_checkpoint: no held run (stop one with pm_server_set_stop)");
    const CkLayout L = ck_layout(S);
    PM_REQUIRE(bytes >= L.total, "pm_server_checkpoint: buffer of %zu bytes, %zu needed", bytes, L.total);
    return ck_write(S, (char*)buffer, L, (hipStream_t)stream);
}

int pm_server_restore(pm_server* S, const void* buffer, size_t bytes, pm_stream_t stream)
{
    PM_REQUIRE(S && buffer, "pm_server_restore: null argument");
    PM_REQUIRE(S->opt.mode == 0, "pm_server_restore: only a batched-mode (0) server resumes a checkpoint");
    PM_REQUIRE(!S->run.held, "pm_server_restore: the server holds a run (pm_server_discard it first)");
    return ck_read(S, (const char*)buffer, bytes, (hipStream_t)stream);
}

int pm_server_rhs(int n, int ngauss, uint64_t seed, double sigma, double* rhs_dev, pm_stream_t stream)
{
    PM_REQUIRE(rhs_dev && n >= 4 && n <= 1024 && ngauss >= 1 && ngauss <= 32 && sigma > 0,
               "pm_server_rhs: bad argument");
    double* tab = nullptr;
    PM_CHECK_HIP(hipMalloc((void**)&tab, sizeof(double) * tab_elems(n)));
    pm_request q;
    std::memset(&q, 0, sizeof(q));
    q.n = n;
    q.ngauss = ngauss;
    q.seed = seed;
    const hipStream_t s = (hipStream_t)stream;
    int rc = make_rhs(rhs_dev, tab, n, q, sigma, s);
    if (rc == PM_OK && hipStreamSynchronize(s) != hipSuccess) {
        set_error("pm_server_rhs: synchronisation failed");
        rc = PM_ERR_HIP;
    }
    hipFree(tab);
    return rc;
}

int pm_server_result(pm_server* S, int index, double* phi_host)
{
    PM_REQUIRE(S && phi_host, "pm_server_result: null argument");
    auto it = S->results.find(index);
    PM_REQUIRE(it != S->results.end(), "pm_server_result: request %d was not kept or not solved", index);
    std::memcpy(phi_host, it->second.data(), sizeof(double) * it->second.size());
    return PM_OK;
}

}  // extern "C"
