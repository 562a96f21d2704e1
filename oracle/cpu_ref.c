/* TEST INFRASTRUCTURE ONLY -- the oracle (see cpu_ref.h).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, no -march): every
 * multiply and every add is rounded once, as in Proto's x86-64 Release build.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_stencil_apply_*  - golden vectors written by Proto's own hostApply
 *                          (tests/golden/), the compiled reference where it is
 *                          built, Laplacian of a paraboloid = 2*DIM exactly,
 *                          constant input -> 0 / 4.9e-17, refinement-ratio and
 *                          destShift cases (AvgDown, GSRB sub-lattices).
 *   orc_infer_box        - the box-inference expectations of
 *                          tests/StencilTests.cpp:80-190 restated as cases.
 *   orc_reduce_*         - closed forms of tests/ReductionTests.cpp:17-31,
 *                          Max's numeric_limits<T>::min() quirk, golden sums.
 *   orc_interp_apply     - golden vectors, prolongation of 1 is 1, Constant
 *                          injection (tests/InterpTests.cpp:42-113), Linear
 *                          reproduces linear functions.
 *   orc_level_exchange   - golden dumps of LevelBoxData::exchange and a brute
 *                          force periodic image.
 *   orc_mg_solve & parts - golden residual histories of the archived Jacobi
 *                          solver (compiled from its own source) and of the
 *                          GSRB V-cycle composed from Proto Stencils; one GSRB
 *                          half-sweep zeroes the residual at the updated colour.
 *   orc_uniform_f64      - a numpy restatement and fixed known values.
 */
#include "cpu_ref.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Box::sindex (Proto_BoxImplem.H:76-86): x fastest. */
static long sindex(const int* p, const int* lo, const int* hi)
{
    long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    return (long)(p[0] - lo[0]) + nx * ((long)(p[1] - lo[1]) + ny * (long)(p[2] - lo[2]));
}

static int contains(const int* p, const int* lo, const int* hi)
{
    for (int d = 0; d < 3; d++) { if (p[d] < lo[d] || p[d] > hi[d]) { return 0; } }
    return 1;
}

/* Stencil::hostApply, Proto_StencilImplem.H:725-828.  Loop order as there:
 * zero pass over destination pencils (750-765), coefficients pre-multiplied by
 * scale when scale != 1 (771-779), offsets linearised against the source box
 * (781-801), then for each cross point (y,z), for each coefficient j, for each
 * x of the pencil: dst += c_j * src (805-827).  The only difference is that a
 * pencil whose start lies outside dst is skipped in the zero pass too, where
 * the reference would write out of bounds. */
#define STENCIL_APPLY_BODY(T)                                                        \
    if (ncoef == 0) { return; }                                                      \
    for (int d = 0; d < 3; d++) { if (bhi[d] < blo[d]) { return; } }                 \
    int npencil = bhi[0] - blo[0] + 1;                                               \
    if (replace)                                                                     \
    {                                                                                \
        for (int z = blo[2]; z <= bhi[2]; z++)                                       \
        for (int y = blo[1]; y <= bhi[1]; y++)                                       \
        {                                                                            \
            int pt[3] = {blo[0] * dr[0] + ds[0], y * dr[1] + ds[1], z * dr[2] + ds[2]}; \
            if (!contains(pt, dlo, dhi)) { continue; }                               \
            T* val = dst + sindex(pt, dlo, dhi);                                     \
            for (int ii = 0; ii < npencil; ii++) { val[ii * dr[0]] = 0; }            \
        }                                                                            \
    }                                                                                \
    if (scale == 0) { return; }                                                      \
    T* c = (T*)malloc(sizeof(T) * ncoef);                                            \
    long* off = (long*)malloc(sizeof(long) * ncoef);                                 \
    long f1 = shi[0] - slo[0] + 1, f2 = f1 * (shi[1] - slo[1] + 1);                  \
    for (int j = 0; j < ncoef; j++)                                                  \
    {                                                                                \
        c[j] = coefs[j];                                                             \
        if (scale != 1) { c[j] *= scale; }                                           \
        off[j] = offs[3 * j] + offs[3 * j + 1] * f1 + offs[3 * j + 2] * f2;          \
    }                                                                                \
    for (int z = blo[2]; z <= bhi[2]; z++)                                           \
    for (int y = blo[1]; y <= bhi[1]; y++)                                           \
    {                                                                                \
        int dpt[3] = {blo[0] * dr[0] + ds[0], y * dr[1] + ds[1], z * dr[2] + ds[2]}; \
        if (!contains(dpt, dlo, dhi)) { continue; }                                  \
        int spt[3] = {blo[0] * sr[0], y * sr[1], z * sr[2]};                         \
        for (int j = 0; j < ncoef; j++)                                              \
        {                                                                            \
            T* destVal = dst + sindex(dpt, dlo, dhi);                                \
            const T* srcVal = src + sindex(spt, slo, shi) + off[j];                  \
            for (int ii = 0; ii < npencil; ii++)                                     \
            {                                                                        \
                destVal[ii * dr[0]] += c[j] * srcVal[ii * sr[0]];                    \
            }                                                                        \
        }                                                                            \
    }                                                                                \
    free(c);                                                                         \
    free(off);

void orc_stencil_apply_f64(int ncoef, const double* coefs, const int* offs,
                           const int* sr, const int* dr, const int* ds,
                           const double* src, const int* slo, const int* shi,
                           double* dst, const int* dlo, const int* dhi,
                           const int* blo, const int* bhi, int replace, double scale)
{
    STENCIL_APPLY_BODY(double)
}

void orc_stencil_apply_f32(int ncoef, const float* coefs, const int* offs,
                           const int* sr, const int* dr, const int* ds,
                           const float* src, const int* slo, const int* shi,
                           float* dst, const int* dlo, const int* dhi,
                           const int* blo, const int* bhi, int replace, float scale)
{
    STENCIL_APPLY_BODY(float)
}

/* Stencil::addCoef (Proto_StencilImplem.H:237-276): a repeated offset adds
 * its coefficient into the first occurrence; new offsets append, so the order
 * is first-insertion order.  Returns the merged count. */
int orc_stencil_merge(int n, const double* coefs, const int* offs, double* mc, int* mo)
{
    int m = 0;
    for (int j = 0; j < n; j++)
    {
        int k = 0;
        for (; k < m; k++)
        {
            if (mo[3 * k] == offs[3 * j] && mo[3 * k + 1] == offs[3 * j + 1] && mo[3 * k + 2] == offs[3 * j + 2]) { break; }
        }
        if (k < m) { mc[k] += coefs[j]; }
        else
        {
            mc[m] = coefs[j];
            mo[3 * m] = offs[3 * j]; mo[3 * m + 1] = offs[3 * j + 1]; mo[3 * m + 2] = offs[3 * j + 2];
            m++;
        }
    }
    return m;
}

/* Box::taperCoarsen (Proto_BoxImplem.H:555-564): high by C++ integer division
 * (truncation toward zero), low by ceil of the real quotient. */
static void taper(int* lo, int* hi, const int* r, int dim)
{
    for (int d = 0; d < dim; d++)
    {
        hi[d] = hi[d] / r[d];
        lo[d] = (int)ceil(lo[d] / (double)r[d]);
    }
}

int orc_infer_box(int dim, const int* span_lo, const int* span_hi,
                  const int* sr, const int* dr, const int* ds,
                  const int* slo, const int* shi, const int* dlo, const int* dhi,
                  int* blo, int* bhi)
{
    /* indexDomain(dst): dst.shift(-destShift).taperCoarsen(destRatio) (330-338) */
    int alo[3], ahi[3], rlo[3], rhi[3];
    for (int d = 0; d < 3; d++)
    {
        alo[d] = dlo[d] - ds[d]; ahi[d] = dhi[d] - ds[d];
        rlo[d] = slo[d] - span_lo[d]; rhi[d] = shi[d] - span_hi[d];
    }
    taper(alo, ahi, dr, dim);
    /* indexRange(src): [lo - span.lo, hi - span.hi].taperCoarsen(srcRatio) (320-328) */
    taper(rlo, rhi, sr, dim);
    int empty = 0;
    for (int d = 0; d < 3; d++)
    {
        if (d >= dim) { blo[d] = 0; bhi[d] = 0; continue; }
        blo[d] = alo[d] > rlo[d] ? alo[d] : rlo[d];
        bhi[d] = ahi[d] < rhi[d] ? ahi[d] : rhi[d];
        if (bhi[d] < blo[d]) { empty = 1; }
    }
    return !empty;
}

/* Reduction init/update (Proto_ReductionImplem.H:64-102) folded sequentially
 * over the buffer (183-191).  Max starts at numeric_limits<T>::min(), the
 * smallest positive normal number, as the reference does. */
#define REDUCE_BODY(T, TMIN, TMAX)                                               \
    T v;                                                                         \
    switch (op)                                                                  \
    {                                                                            \
        case 0: v = TMIN; break;                                                 \
        case 1: v = TMAX; break;                                                 \
        default: v = 0; break;                                                   \
    }                                                                            \
    for (long i = 0; i < n; i++)                                                 \
    {                                                                            \
        T x = data[i];                                                           \
        switch (op)                                                              \
        {                                                                        \
            case 0: v = (v < x) ? x : v; break;                                  \
            case 1: v = (x < v) ? x : v; break;                                  \
            case 2: { T a = (x > 0) ? x : -x; v = (v < a) ? a : v; } break;      \
            case 3: v += x; break;                                               \
            case 4: v += (x > 0 ? x : -x); break;                                \
            case 5: v += x * x; break;                                           \
        }                                                                        \
    }                                                                            \
    return v;

double orc_reduce_f64(int op, const double* data, long n) { REDUCE_BODY(double, DBL_MIN, DBL_MAX) }
float orc_reduce_f32(int op, const float* data, long n) { REDUCE_BODY(float, FLT_MIN, FLT_MAX) }

/* One InterpStencil1D stage (Proto_InterpStencil1DImplem.H:90-113 and the
 * BoxData(LazyStencil&&) constructor, Proto_BoxDataImplem.H:113-127): the
 * output box is the union of the shift stencils' ranges; each shift s writes
 * out(r*i + s) = (0 + q_s*in(i)) [+ p_s*in(i+e)] over the index box
 * [in.lo - span.lo, in.hi - span.hi]. */
static double* interp_stage(int dim, int dir, int linear, int r, const double* in,
                            const int* ilo, const int* ihi, int* olo, int* ohi)
{
    for (int d = 0; d < 3; d++) { olo[d] = ilo[d]; ohi[d] = ihi[d]; }
    int span_hi = linear ? 1 : 0;
    /* range of shift s: [r*lo + s, r*(hi - span_hi) + s]; union over s */
    olo[dir] = r * ilo[dir];
    ohi[dir] = r * (ihi[dir] - span_hi) + (r - 1);
    long n = 1;
    for (int d = 0; d < 3; d++) { n *= (ohi[d] - olo[d] + 1); }
    double* out = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (long i = 0; i < n; i++) { out[i] = 7.77e77; } /* Proto's poison fill */
    (void)dim;
    for (int s = 0; s < r; s++)
    {
        double rr = r, ss = s;
        double p = ss / rr, q = (rr - ss) / rr;
        double c0 = linear ? q : 1.0, c1 = p;
        int bhi[3] = {ihi[0], ihi[1], ihi[2]};
        bhi[dir] -= span_hi;
        for (int z = ilo[2]; z <= bhi[2]; z++)
        for (int y = ilo[1]; y <= bhi[1]; y++)
        for (int x = ilo[0]; x <= bhi[0]; x++)
        {
            int ip[3] = {x, y, z};
            int op[3] = {x, y, z};
            op[dir] = r * ip[dir] + s;
            double acc = 0;
            acc += c0 * in[sindex(ip, ilo, ihi)];
            if (linear)
            {
                int jp[3] = {x, y, z};
                jp[dir] += 1;
                acc += c1 * in[sindex(jp, ilo, ihi)];
            }
            out[sindex(op, olo, ohi)] = acc;
        }
    }
    return out;
}

void orc_interp_apply(int dim, int linear, int ratio,
                      const double* crse, const int* clo, const int* chi,
                      double* fine, const int* flo, const int* fhi, int replace)
{
    /* InterpStencil::apply (Proto_InterpStencilImplem.H:83-112): stage per
     * direction, then output (|= or +=) 1.0*stage over range(in) & output. */
    int lo[3], hi[3], nlo[3], nhi[3];
    double* cur = NULL;
    const double* in = crse;
    for (int d = 0; d < 3; d++) { lo[d] = clo[d]; hi[d] = chi[d]; }
    for (int dir = 0; dir < dim; dir++)
    {
        double* nxt = interp_stage(dim, dir, linear, ratio, in, lo, hi, nlo, nhi);
        if (cur) { free(cur); }
        cur = nxt;
        in = cur;
        for (int d = 0; d < 3; d++) { lo[d] = nlo[d]; hi[d] = nhi[d]; }
    }
    int blo[3], bhi[3];
    for (int d = 0; d < 3; d++)
    {
        blo[d] = lo[d] > flo[d] ? lo[d] : flo[d];
        bhi[d] = hi[d] < fhi[d] ? hi[d] : fhi[d];
    }
    double one = 1.0;
    int z0[3] = {0, 0, 0}, o1[3] = {1, 1, 1};
    orc_stencil_apply_f64(1, &one, z0, o1, o1, z0, cur, lo, hi, fine, flo, fhi, blo, bhi, replace, 1.0);
    free(cur);
}

#define RESIDUAL_BODY(T)                                                          \
    long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;  \
    long n = nx * ny * nz;                                                        \
    T* Lu = (T*)malloc(sizeof(T) * n);                                            \
    int o1[3] = {1, 1, 1}, z0[3] = {0, 0, 0};                                     \
    orc_stencil_apply_##T(ncoef, coefs, offs, o1, o1, z0, u, ulo, uhi, Lu, lo, hi, lo, hi, 1, scale); \
    for (long i = 0; i < n; i++) { r[i] = f[i] - Lu[i]; }                         \
    free(Lu);

typedef double f64;
typedef float f32;
#define orc_stencil_apply_double orc_stencil_apply_f64
#define orc_stencil_apply_float orc_stencil_apply_f32

void orc_residual_f64(int ncoef, const double* coefs, const int* offs,
                      const double* u, const int* ulo, const int* uhi,
                      const double* f, double* r, const int* lo, const int* hi, double scale)
{
    RESIDUAL_BODY(double)
}

void orc_residual_f32(int ncoef, const float* coefs, const int* offs,
                      const float* u, const int* ulo, const int* uhi,
                      const float* f, float* r, const int* lo, const int* hi, float scale)
{
    RESIDUAL_BODY(float)
}

/* ------------------------------------------------------------------------
 * Periodic level.  On a fully periodic domain tiled by a DisjointBoxLayout,
 * exchange() fills every ghost cell with the periodic image of a valid cell,
 * so each per-box stencil evaluation of the reference reads exactly the values
 * that periodic (mod n) indexing of one global array reads, in the same order.
 * The functions below therefore work on one global n^3 array per level; the
 * per-point arithmetic is the reference's, operation for operation.
 * ------------------------------------------------------------------------ */

static inline long gidx(int n, int x, int y, int z)
{
    x = (x % n + n) % n; y = (y % n + n) % n; z = (z % n + n) % n;
    return (long)x + (long)n * ((long)y + (long)n * z);
}

void orc_level_exchange(int n, int bs, int g, const double* global, double* boxes)
{
    int nb = n / bs;
    int e = bs + 2 * g;
    long per = (long)e * e * e;
    for (int bz = 0; bz < nb; bz++)
    for (int by = 0; by < nb; by++)
    for (int bx = 0; bx < nb; bx++)
    {
        double* b = boxes + per * ((long)bx + (long)nb * ((long)by + (long)nb * bz));
        for (int z = 0; z < e; z++)
        for (int y = 0; y < e; y++)
        for (int x = 0; x < e; x++)
        {
            b[x + (long)e * (y + (long)e * z)] =
                global[gidx(n, bx * bs + x - g, by * bs + y - g, bz * bs + z - g)];
        }
    }
}

/* Laplacian() coefficient order (Proto_StencilDefs.H:77-87): centre, then
 * +e_d, -e_d for d = 0..DIM-1. */
static const int LAP_OFF[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const double LAP_C[7] = {-6.0, 1, 1, 1, 1, 1, 1};

/* GSRB, LaplaceOp.H:56-73 (stencils) and 95-123 (sweep).  At a point p of the
 * active colour (x+y+z parity, global indices):
 *   temp |= m_GSRB(phi):   t = 0 + c'_0 phi(p) + ... + c'_6 phi(p-e_z),
 *                          c'_j = c_j * (-1/(dx*dx))
 *   temp += m_GSRB_I(rhs): t = t + 1.0*rhs(p)
 *   phi  += m_GSRB_I(temp, lambda): phi = phi + (1.0*lambda)*t,
 *                          lambda = 2 * (-(dx*dx)/(4*DIM)).
 * The opposite colour is read only, so the in-place update equals the
 * reference's exchange-then-update per colour. */
void orc_gsrb_relax(int n, double dx, int niter, const double* rhs, double* phi)
{
    double lambda = -(dx * dx) / (4.0 * 3);
    lambda *= 2.0;
    double m = -1.0 / (dx * dx);
    double c[7];
    for (int j = 0; j < 7; j++) { c[j] = LAP_C[j] * m; }
    double one_lambda = 1.0 * lambda;
    for (int it = 0; it < niter; it++)
    for (int color = 0; color < 2; color++)
    for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
    for (int x = (y + z + color) & 1; x < n; x += 2)
    {
        double t = 0;
        for (int j = 0; j < 7; j++)
        {
            t += c[j] * phi[gidx(n, x + LAP_OFF[j][0], y + LAP_OFF[j][1], z + LAP_OFF[j][2])];
        }
        long p = gidx(n, x, y, z);
        t += 1.0 * rhs[p];
        phi[p] += one_lambda * t;
    }
}

/* pointRelax, LevelMultigrid.cpp:99-123: temp = L(phi, 1/(4 DIM));
 * temp += (-lambda)*rhs with lambda = dx^2/(4 DIM); phi += temp. */
void orc_jacobi_relax(int n, double dx, int niter, const double* rhs, double* phi)
{
    double weight = 1.0 / (4 * 3);
    double lambda = dx * dx / (4 * 3);
    double dcoef = -lambda;
    double c[7];
    for (int j = 0; j < 7; j++) { c[j] = LAP_C[j] * weight; }
    long N = (long)n * n * n;
    double* temp = (double*)malloc(sizeof(double) * N);
    for (int it = 0; it < niter; it++)
    {
        for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++)
        {
            double t = 0;
            for (int j = 0; j < 7; j++)
            {
                t += c[j] * phi[gidx(n, x + LAP_OFF[j][0], y + LAP_OFF[j][1], z + LAP_OFF[j][2])];
            }
            t += dcoef * rhs[gidx(n, x, y, z)];
            temp[gidx(n, x, y, z)] = t;
        }
        for (long i = 0; i < N; i++) { phi[i] += temp[i]; }
    }
    free(temp);
}

/* resnorm, LevelMultigrid.cpp:75-97: res = 0; res -= rhs; res += L(phi, 1/h^2);
 * absMax over the valid cells. */
double orc_resnorm(int n, double dx, const double* rhs, const double* phi)
{
    double hsqinv = 1. / (dx * dx);
    double c[7];
    for (int j = 0; j < 7; j++) { c[j] = LAP_C[j] * hsqinv; }
    double v = 0;
    for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++)
    {
        double r = 0;
        r -= rhs[gidx(n, x, y, z)];
        for (int j = 0; j < 7; j++)
        {
            r += c[j] * phi[gidx(n, x + LAP_OFF[j][0], y + LAP_OFF[j][1], z + LAP_OFF[j][2])];
        }
        double a = r > 0 ? r : -r;
        v = (v < a) ? a : v;
    }
    return v;
}

/* coarseResidual, LevelMultigrid.cpp:51-73: res = 0; res += rhs;
 * res += L(phi, -1/h^2); crse |= AvgDown(2)(res), whose coefficients are
 * 1/8 over Box(2) in x-fastest order (Proto_StencilDefs.H:302-327). */
void orc_coarse_residual(int n, double dx, const double* rhs, const double* phi, double* resc)
{
    double hsqinv = 1. / (dx * dx);
    double c[7];
    for (int j = 0; j < 7; j++) { c[j] = LAP_C[j] * (-hsqinv); }
    long N = (long)n * n * n;
    double* res = (double*)malloc(sizeof(double) * N);
    for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++)
    {
        double r = 0;
        r += rhs[gidx(n, x, y, z)];
        for (int j = 0; j < 7; j++)
        {
            r += c[j] * phi[gidx(n, x + LAP_OFF[j][0], y + LAP_OFF[j][1], z + LAP_OFF[j][2])];
        }
        res[gidx(n, x, y, z)] = r;
    }
    int nc = n / 2;
    double a = 1.0 / 8;
    for (int z = 0; z < nc; z++)
    for (int y = 0; y < nc; y++)
    for (int x = 0; x < nc; x++)
    {
        double acc = 0;
        for (int kz = 0; kz < 2; kz++)
        for (int ky = 0; ky < 2; ky++)
        for (int kx = 0; kx < 2; kx++)
        {
            acc += a * res[gidx(n, 2 * x + kx, 2 * y + ky, 2 * z + kz)];
        }
        resc[gidx(nc, x, y, z)] = acc;
    }
    free(res);
}

/* fineInterp, LevelMultigrid.cpp:125-140: phi += I(delta), I staged x, y, z
 * (Proto_InterpStencilImplem.H:83-112).  Constant: each stage is 0 + 1.0*in.
 * Linear (Proto_InterpStencilDefs.H:17-36): fine 2i+s <- (0 + q_s in(i)) +
 * p_s in(i+1), q = (2-s)/2, p = s/2, with in(i+1) the periodic image (the
 * reference reads it from the exchanged ghost layer). */
void orc_fine_interp(int n, int linear, const double* delta, double* phi)
{
    int nc = n / 2;
    long nf = (long)n * n * n;
    double* s0 = (double*)malloc(sizeof(double) * (long)n * nc * nc);
    double* s1 = (double*)malloc(sizeof(double) * (long)n * n * nc);
    double* s2 = (double*)malloc(sizeof(double) * nf);
    double q[2] = {1.0, 0.5}, p[2] = {0.0, 0.5};
    if (!linear) { q[1] = 1.0; }
    for (int z = 0; z < nc; z++)
    for (int y = 0; y < nc; y++)
    for (int x = 0; x < n; x++)
    {
        int i = x >> 1, s = x & 1;
        double acc = 0;
        acc += q[s] * delta[gidx(nc, i, y, z)];
        if (linear) { acc += p[s] * delta[gidx(nc, i + 1, y, z)]; }
        s0[x + (long)n * (y + (long)nc * z)] = acc;
    }
    for (int z = 0; z < nc; z++)
    for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++)
    {
        int i = y >> 1, s = y & 1;
        double acc = 0;
        acc += q[s] * s0[x + (long)n * (i + (long)nc * z)];
        if (linear) { acc += p[s] * s0[x + (long)n * (((i + 1) % nc) + (long)nc * z)]; }
        s1[x + (long)n * (y + (long)n * z)] = acc;
    }
    for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++)
    {
        int i = z >> 1, s = z & 1;
        double acc = 0;
        acc += q[s] * s1[x + (long)n * (y + (long)n * i)];
        if (linear) { acc += p[s] * s1[x + (long)n * (y + (long)n * ((i + 1) % nc))]; }
        s2[x + (long)n * (y + (long)n * z)] = acc;
    }
    for (long i = 0; i < nf; i++) { phi[i] += 1.0 * s2[i]; }
    free(s0); free(s1); free(s2);
}

static void vcycle(int n, int level, double dx, int relax, int pre, int post, int bottom, int linear,
                   const double* rhs, double* phi)
{
    if (level > 0)
    {
        if (relax == 0) { orc_gsrb_relax(n, dx, pre, rhs, phi); }
        else { orc_jacobi_relax(n, dx, pre, rhs, phi); }
        int nc = n / 2;
        long ncc = (long)nc * nc * nc;
        double* resc = (double*)malloc(sizeof(double) * ncc);
        double* delta = (double*)calloc(ncc, sizeof(double));
        orc_coarse_residual(n, dx, rhs, phi, resc);
        vcycle(nc, level - 1, 2 * dx, relax, pre, post, bottom, linear, resc, delta);
        orc_fine_interp(n, linear, delta, phi);
        free(resc);
        free(delta);
        if (relax == 0) { orc_gsrb_relax(n, dx, post, rhs, phi); }
        else { orc_jacobi_relax(n, dx, post, rhs, phi); }
    }
    else
    {
        if (relax == 0) { orc_gsrb_relax(n, dx, bottom, rhs, phi); }
        else { orc_jacobi_relax(n, dx, bottom, rhs, phi); }
    }
}

/* Solve loop of examples/_archived/LevelMultigrid/exec/main.cpp:84-103
 * (and LevelMultigridSolverImplem.H:82-132): stop when resnorm < tol*res0. */
int orc_mg_solve(int n, int nlev, int relax, int pre, int post, int bottom, int linear,
                 const double* rhs, double* phi, int maxiter, double tol, double* hist)
{
    double dx = 1.0 / n;
    if (relax == 1) { pre = 2 * 3; post = 2 * 3; bottom = 10; }
    double res0 = orc_resnorm(n, dx, rhs, phi);
    hist[0] = res0;
    int it = 0;
    for (; it < maxiter; it++)
    {
        vcycle(n, nlev, dx, relax, pre, post, bottom, linear, rhs, phi);
        double r = orc_resnorm(n, dx, rhs, phi);
        hist[it + 1] = r;
        if (r < tol * res0) { it++; break; }
    }
    return it;
}

#define PATCH_GSRB_BODY(T, C0, CN, LAM)                                                    \
    long Gx = nx + 2 * g, Gxy = Gx * (ny + 2 * g);                                          \
    T c[7] = {C0, CN, CN, CN, CN, CN, CN};                                                  \
    const long off[7] = {0, 1, -1, Gx, -Gx, Gxy, -Gxy};                                     \
    for (int color = 0; color < 2; color++)                                                 \
    for (int z = 0; z < nz; z++)                                                            \
    for (int y = 0; y < ny; y++)                                                            \
    for (int x = (y + z + color) & 1; x < nx; x += 2)                                       \
    {                                                                                       \
        long p = (x + g) + Gx * (y + g) + Gxy * (z + g);                                    \
        T t = 0;                                                                            \
        for (int j = 0; j < 7; j++) { t += c[j] * phi[p + off[j]]; }                        \
        t += (T)1.0 * rhs[x + (long)nx * (y + (long)ny * z)];                               \
        phi[p] += (LAM) * t;                                                                \
    }

/* ---- library FAS multigrid with BoxOp_Laplace ------------------------------
 * LevelMultigridSolver<BoxOp_Laplace> on a periodic cube, as one global array
 * per level (the per-box work with exchanged ghosts gives the same values):
 *   L(phi) = fluxScale * sum_d DIV_d(F_d) with fluxScale = -1
 *     (Proto_BoxOpImplem.H:129-176; setFluxScale(-1) at
 *     Proto_LevelMultigridSolverImplem.H:206),
 *   F_d = DiffCellToFace(d, Lo) applied with scale 1/dx: coefficients
 *     {1, -15, 15, -1} * (1/12) at offsets -2, -1, 0, +1, each then times 1/dx
 *     (BoxOp_Laplace.H:24-31, Proto_StencilDefs.H:232-268),
 *   DIV_d = 1.0 Shift(+e_d) - 1.0 Shift(0) with scale 1/dx, accumulated over
 *     d = x, y, z into a zeroed divFlux; output = (0 + divFlux) * (-1);
 *   residual = L(phi) + force; Jacobi relax phi += (-lambda) * residual with
 *     lambda = 0.5 / (2 DIM / dx^2) (LevelMultigridSolverImplem.H:207, 358-379);
 *   V-cycle (408-433): 2 DIM pre / post sweeps, 10 at the bottom;
 *     crseState = AvgDown(phi); crse0 = crseState;
 *     crseForce = AvgDown(L(phi) + force) - L_c(crseState) (coarseForce, 307-336);
 *     fine += Constant interp of (crse0 - crseState) * -1 (fineCorrect, 381-406);
 *   solve (85-132): res0 = max|L(phi) + force|, V-cycles until res < tol res0. */
static void fas_apply(int n, double dx, const double* u, double* out)
{
    const double k[4] = {1, -15, 15, -1};
    const double twelfth = 1.0 / 12.0;
    const double s = 1.0 / dx;
    double a[4];
    for (int j = 0; j < 4; j++) { a[j] = (k[j] * twelfth) * s; }
    const double dp = 1.0 * s, dm = -1.0 * s;
    for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++)
    {
        const int p[3] = {x, y, z};
        double acc = 0;
        for (int d = 0; d < 3; d++)
        {
            double v[6];  /* cells p-2e .. p+2e */
            for (int o = -2; o <= 2; o++)
            {
                int q[3] = {p[0], p[1], p[2]};
                q[d] = (q[d] + o + n) % n;
                v[o + 2] = u[gidx(n, q[0], q[1], q[2])];
            }
            double fr = 0;
            fr = fr + a[0] * v[1];
            fr = fr + a[1] * v[2];
            fr = fr + a[2] * v[3];
            fr = fr + a[3] * v[4];
            double fl = 0;
            fl = fl + a[0] * v[0];
            fl = fl + a[1] * v[1];
            fl = fl + a[2] * v[2];
            fl = fl + a[3] * v[3];
            acc = acc + dp * fr;
            acc = acc + dm * fl;
        }
        double o = 0;
        o = o + acc;
        o = o * -1.0;
        out[gidx(n, x, y, z)] = o;
    }
}

static double fas_lambda(double dx) { return 0.5 / ((2.0 * 3) / pow(dx, 2)); }

void orc_fas_relax(int n, double dx, int niter, const double* force, double* phi)
{
    long nc = (long)n * n * n;
    double* res = (double*)malloc(sizeof(double) * nc);
    const double ml = -fas_lambda(dx);
    for (int it = 0; it < niter; it++)
    {
        fas_apply(n, dx, phi, res);
        for (long i = 0; i < nc; i++)
        {
            double r = res[i] + force[i];
            phi[i] = phi[i] + ml * r;
        }
    }
    free(res);
}

double orc_fas_resnorm(int n, double dx, const double* force, const double* phi)
{
    long nc = (long)n * n * n;
    double* res = (double*)malloc(sizeof(double) * nc);
    fas_apply(n, dx, phi, res);
    double m = 0;
    for (long i = 0; i < nc; i++)
    {
        double r = res[i] + force[i];
        double a = r < 0 ? -r : r;
        if (a > m) { m = a; }
    }
    free(res);
    return m;
}

static void fas_avgdown(int n, const double* f, double* c)
{
    int nc = n / 2;
    for (int Z = 0; Z < nc; Z++)
    for (int Y = 0; Y < nc; Y++)
    for (int X = 0; X < nc; X++)
    {
        double acc = 0;
        for (int kz = 0; kz < 2; kz++)
        for (int ky = 0; ky < 2; ky++)
        for (int kx = 0; kx < 2; kx++) { acc = acc + 0.125 * f[gidx(n, 2 * X + kx, 2 * Y + ky, 2 * Z + kz)]; }
        c[gidx(nc, X, Y, Z)] = acc;
    }
}

static void fas_vcycle(int n, int level, double dx, const double* force, double* phi)
{
    if (level > 0)
    {
        long nf = (long)n * n * n;
        int nc = n / 2;
        long ncc = (long)nc * nc * nc;
        orc_fas_relax(n, dx, 6, force, phi);
        double* cs = (double*)malloc(sizeof(double) * ncc);
        double* c0 = (double*)malloc(sizeof(double) * ncc);
        double* cf = (double*)malloc(sizeof(double) * ncc);
        double* lc = (double*)malloc(sizeof(double) * ncc);
        double* res = (double*)malloc(sizeof(double) * nf);
        fas_avgdown(n, phi, cs);
        memcpy(c0, cs, sizeof(double) * ncc);
        fas_apply(n, dx, phi, res);
        for (long i = 0; i < nf; i++) { res[i] = res[i] + force[i]; }
        fas_avgdown(n, res, cf);
        fas_apply(nc, 2 * dx, cs, lc);
        for (long i = 0; i < ncc; i++) { cf[i] = cf[i] - lc[i]; }
        fas_vcycle(nc, level - 1, 2 * dx, cf, cs);
        for (long i = 0; i < ncc; i++)
        {
            double d = c0[i] - cs[i];
            c0[i] = d * -1.0;
        }
        orc_fine_interp(n, 0, c0, phi);
        orc_fas_relax(n, dx, 6, force, phi);
        free(cs); free(c0); free(cf); free(lc); free(res);
    }
    else
    {
        orc_fas_relax(n, dx, 10, force, phi);
    }
}

int orc_fas_solve(int n, int num_levels, const double* force, double* phi, int max_iter, double tol,
                  double* hist)
{
    double dx = 1.0 / n;
    double res0 = orc_fas_resnorm(n, dx, force, phi);
    hist[0] = res0;
    int it = 0;
    for (; it < max_iter; it++)
    {
        fas_vcycle(n, num_levels - 1, dx, force, phi);
        double r = orc_fas_resnorm(n, dx, force, phi);
        hist[it + 1] = r;
        if (r < tol * res0) { it++; break; }
    }
    return it;
}

void orc_patch_gsrb_f64(int nx, int ny, int nz, int g, double dx, const double* rhs, double* phi)
{
    double lam = -(dx * dx) / (4.0 * 3);
    lam *= 2.0;
    double m = -1.0 / (dx * dx);
    PATCH_GSRB_BODY(double, -6.0 * m, 1.0 * m, 1.0 * lam)
}

void orc_patch_gsrb_f32(int nx, int ny, int nz, int g, double dx, const float* rhs, float* phi)
{
    double lam = -(dx * dx) / (4.0 * 3);
    lam *= 2.0;
    double m = -1.0 / (dx * dx);
    float fm = (float)m, flam = (float)(1.0 * lam);
    PATCH_GSRB_BODY(float, -6.0f * fm, 1.0f * fm, flam)
}

/* SplitMix64 (Steele, Lea, Flood 2014), counter form: z_i = mix(s0 + (i+1)*G)
 * with s0 = seed + stream * 0x632BE59BD9B4E019. */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_uniform_f64(uint64_t seed, uint64_t stream, long n, double* out)
{
    uint64_t s0 = seed + stream * 0x632BE59BD9B4E019ULL;
    for (long i = 0; i < n; i++)
    {
        uint64_t z = mix64(s0 + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        out[i] = ((double)(z >> 11) * 0x1.0p-53) * 2.0 - 1.0;
    }
}
