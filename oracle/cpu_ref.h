/* TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A plain-C restatement of Proto's CPU hot path (Stencil::hostApply, host
 * Reduction, InterpStencil staging, periodic exchange, the correction-scheme
 * V-cycle with GSRB or damped-Jacobi relaxation).  Only tests/, the smoke()
 * check and bench.py's cpu_baseline leg may load it; the product path in
 * proto_amd/ never does.
 *
 * Boxes are inclusive [lo, hi] triples with x fastest; 2-D data uses
 * lo[2] = hi[2] = 0.  Each function cites the reference lines it restates.
 * Pinning (what checks each function beyond itself) is listed in the header
 * of cpu_ref.c.
 */
#ifndef PROTOMI_ORACLE_CPU_REF_H
#define PROTOMI_ORACLE_CPU_REF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stencil::hostApply (Proto_StencilImplem.H:725-828) for one component. */
void orc_stencil_apply_f64(int ncoef, const double* coefs, const int* offs,
                           const int* sr, const int* dr, const int* ds,
                           const double* src, const int* slo, const int* shi,
                           double* dst, const int* dlo, const int* dhi,
                           const int* blo, const int* bhi, int replace, double scale);
void orc_stencil_apply_f32(int ncoef, const float* coefs, const int* offs,
                           const int* sr, const int* dr, const int* ds,
                           const float* src, const int* slo, const int* shi,
                           float* dst, const int* dlo, const int* dhi,
                           const int* blo, const int* bhi, int replace, float scale);

/* Stencil::addCoef merge of repeated offsets (Proto_StencilImplem.H:237-276). */
int orc_stencil_merge(int n, const double* coefs, const int* offs, double* mc, int* mo);

/* Stencil::indexDomain(dst) & Stencil::indexRange(src)
 * (Proto_StencilImplem.H:320-343 with Box::taperCoarsen, Proto_BoxImplem.H:555-564).
 * span_lo/span_hi: the stencil's offset bounding box.  Returns 0 if empty. */
int orc_infer_box(int dim, const int* span_lo, const int* span_hi,
                  const int* sr, const int* dr, const int* ds,
                  const int* slo, const int* shi, const int* dlo, const int* dhi,
                  int* blo, int* bhi);

/* Reduction<T,OP,HOST>::reduce + fetchLocal (Proto_ReductionImplem.H:64-102,157-192).
 * op: 0 Max, 1 Min, 2 Abs, 3 Sum, 4 SumAbs, 5 SumSquare. */
double orc_reduce_f64(int op, const double* data, long n);
float orc_reduce_f32(int op, const float* data, long n);

/* fine (+)= I(crse) for InterpStencil::Constant / ::Linear with isotropic ratio
 * (Proto_InterpStencilImplem.H:83-112, Proto_InterpStencilDefs.H:1-36,
 *  Proto_InterpStencil1DImplem.H:66-113).  dim = 2 or 3. */
void orc_interp_apply(int dim, int linear, int ratio,
                      const double* crse, const int* clo, const int* chi,
                      double* fine, const int* flo, const int* fhi, int replace);

/* Config-2 fused residual r = f - L(u) as a Proto program writes it:
 * Lu |= L(u, scale) then forall r = f - Lu (Proto_ForallImplem.H:195-213).
 * u is on the ghosted box [ulo,uhi]; f, r on the interior box [lo,hi]. */
void orc_residual_f64(int ncoef, const double* coefs, const int* offs,
                      const double* u, const int* ulo, const int* uhi,
                      const double* f, double* r, const int* lo, const int* hi, double scale);
void orc_residual_f32(int ncoef, const float* coefs, const int* offs,
                      const float* u, const int* ulo, const int* uhi,
                      const float* f, float* r, const int* lo, const int* hi, float scale);

/* ---- periodic level (one global n^dim array; see cpu_ref.c) ---- */

/* LevelBoxData::exchange on a periodic n^3 domain tiled by boxes of edge bs
 * with g ghost layers, box order x-fastest over the box grid.  Input: global
 * valid data; output: nb ghosted boxes stacked. */
void orc_level_exchange(int n, int bs, int g, const double* global, double* boxes);

/* GSRB relaxation, niter iterations of (exchange, red update, exchange, black
 * update) on the periodic n^3 level with spacing h = 1/n * 2^k: LaplaceOp.H:95-123. */
void orc_gsrb_relax(int n, double dx, int niter, const double* rhs, double* phi);
/* Damped point Jacobi of LevelMultigrid.cpp:99-123. */
void orc_jacobi_relax(int n, double dx, int niter, const double* rhs, double* phi);
/* resnorm of LevelMultigrid.cpp:75-97: max |L(phi)/h^2 - rhs| */
double orc_resnorm(int n, double dx, const double* rhs, const double* phi);
/* coarseResidual of LevelMultigrid.cpp:51-73 (n fine -> n/2 coarse). */
void orc_coarse_residual(int n, double dx, const double* rhs, const double* phi, double* resc);
/* fineInterp of LevelMultigrid.cpp:125-140: phi += I(delta), Constant or Linear. */
void orc_fine_interp(int n, int linear, const double* delta, double* phi);

/* Full correction-scheme solve.  relax: 0 = GSRB (pre/post/bottom sweeps),
 * 1 = archived damped Jacobi (pre = post = 2*DIM, bottom = 10 forced).
 * nlev = number of coarsenings (Proto's numLevels argument).  hist gets the
 * max-norm residual before the first V-cycle and after each.  Returns #V-cycles. */
int orc_mg_solve(int n, int nlev, int relax, int pre, int post, int bottom, int linear,
                 const double* rhs, double* phi, int maxiter, double tol, double* hist);

/* One GSRB sweep (red then black) on an isolated patch of nx*ny*nz valid
 * cells with g ghost layers held fixed (no exchange), colours by box-local
 * (x+y+z) parity, the arithmetic of LaplaceOp.H:56-73, 95-123.  phi: ghosted
 * x-fastest storage; rhs: valid storage.  The f32 form uses coefficients
 * (float)Lap_j * (float)(-1/dx^2) and lambda (float)(2*(-(dx*dx)/12)). */
/* Library FAS multigrid, LevelMultigridSolver<BoxOp_Laplace> (Jacobi, 2 DIM / 2 DIM / 10
 * sweeps, constant prolongation) on a periodic n^3 cube with num_levels levels;
 * hist[0..iters]; returns the V-cycle count.  Pinned by Proto's golden history
 * and solution hash (multigrid "fas"). */
void orc_fas_relax(int n, double dx, int niter, const double* force, double* phi);
double orc_fas_resnorm(int n, double dx, const double* force, const double* phi);
int orc_fas_solve(int n, int num_levels, const double* force, double* phi, int max_iter, double tol,
                  double* hist);
void orc_patch_gsrb_f64(int nx, int ny, int nz, int g, double dx, const double* rhs, double* phi);
void orc_patch_gsrb_f32(int nx, int ny, int nz, int g, double dx, const float* rhs, float* phi);

/* ---- seeded inputs (SplitMix64; shared with the GPU input kernel) ---- */
/* value i of stream (seed, stream): u = ((z >> 11) * 2^-53) * 2 - 1 */
void orc_uniform_f64(uint64_t seed, uint64_t stream, long n, double* out);

#ifdef __cplusplus
}
#endif
#endif
