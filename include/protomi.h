/* protomi -- C ABI of the MI355X-native Proto hot path.
 *
 * Proto's device boundary is the template launch
 *   protoLaunchKernelMemAsyncT<MEM, Ker>(blocks, threads, smem, stream, args...)
 * (include/base/Proto_Macros.H:155-172) plus proto_malloc / proto_free /
 * proto_memcpy<MEM> (include/base/implem/Proto_MemoryImplem.H:45-123).  Every
 * Ker::gpu that the hot path launches is replaced by one entry point below,
 * written for gfx950 and called by the C++ API in proto_amd/cpp (Proto's
 * Stencil / BoxData / forall / Reduction / LevelBoxData / multigrid names).
 *
 * Conventions
 *   - Every function returns 0 (PM_OK) or a negative PM_ERR_* code; the text of
 *     the last failure on the calling thread is pm_last_error().  No C++
 *     exception crosses this boundary.
 *   - Boxes are inclusive [lo, hi] index triples with x fastest (Box::sindex,
 *     Proto_BoxImplem.H:76-86); 2-D data uses lo[2] = hi[2] = 0.
 *   - Arrays are BoxData storage: component c of a box B starts at
 *     c * |B| elements (SoA, Proto_BoxDataImplem.H:697-717).  Pointers are
 *     device pointers unless a name says host.
 *   - dtype is PM_F32 or PM_F64; "T" below is that type.  Scalars (scale,
 *     values) travel as double and are converted to T once.
 *   - stream is a hipStream_t (NULL = the null stream).  Work is asynchronous
 *     unless a function says it synchronises.  Handles are caller-owned; a
 *     handle may not be used from two threads at once.
 */
#ifndef PROTOMI_H
#define PROTOMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_OK 0
#define PM_ERR_INVALID (-1)     /* bad argument: shape, box, null pointer, dtype */
#define PM_ERR_HIP (-2)         /* HIP runtime error (message says which call)   */
#define PM_ERR_NOMEM (-3)       /* device allocation failed / over budget        */
#define PM_ERR_UNSUPPORTED (-4) /* valid request this build does not implement   */
#define PM_ERR_NOT_CONVERGED (-5)
#define PM_ERR_CANCELLED (-6)

enum pm_dtype { PM_F32 = 0, PM_F64 = 1 };

/* Reduction operators, in the order of Proto's enum Operation
 * (include/base/Proto_Reduction.H:15). */
enum pm_red_op { PM_RED_MAX = 0, PM_RED_MIN = 1, PM_RED_ABS = 2, PM_RED_SUM = 3,
                 PM_RED_SUMABS = 4, PM_RED_SUMSQ = 5 };

typedef struct pm_box { int lo[3]; int hi[3]; } pm_box;
typedef void* pm_stream_t;

/* ---------------------------------------------------------------- runtime */
const char* pm_last_error(void);
const char* pm_version(void);
/* Number of visible devices; 0 when HIP finds none (no error). */
int pm_device_count(int* count);

/* proto_malloc<DEVICE> / proto_free<DEVICE> (Proto_MemoryImplem.H:77-123). */
int pm_malloc(void** ptr, size_t bytes);
int pm_free(void* ptr);
/* proto_memcpy<SRC,DST> (Proto_MemoryImplem.H:45-75); kind: 0 H2D, 1 D2H, 2 D2D.
 * Asynchronous on `stream`; pageable host memory makes it synchronous. */
int pm_memcpy(void* dst, const void* src, size_t bytes, int kind, pm_stream_t stream);
int pm_memset_zero(void* ptr, size_t bytes, pm_stream_t stream);
/* Device-to-device copy by this library's own streaming kernel (16-byte
 * accesses, grid-stride; nontemporal != 0 streams the stores past the
 * caches).  Also the achievable-bandwidth probe of the benchmark.  bytes and
 * both pointers must be multiples of 16. */
int pm_copy_device(void* dst, const void* src, size_t bytes, int nontemporal, pm_stream_t stream);
/* Streaming-bandwidth probe (no counterpart in Proto; measures the ceilings the
 * kernels' roofline fractions are read against).  kind 0 reads `src` (dst may be
 * null or any 16-byte-aligned device word), 1 writes `dst` with non-temporal
 * stores, 2 copies src -> dst (grid-stride, non-temporal stores), 3 / 4 copy
 * src -> dst in 64 KB chunks per workgroup (non-temporal loads and stores /
 * plain); 16-byte accesses over 256 * blocks_per_cu 256-thread workgroups
 * (blocks_per_cu in [1, 64]).  bytes and the pointers must be multiples of 16. */
int pm_stream_probe(int kind, void* dst, const void* src, size_t bytes, int blocks_per_cu, pm_stream_t stream);
int pm_stream_create(pm_stream_t* stream);
int pm_stream_destroy(pm_stream_t stream);
/* Per-stream reduction scratch (no counterpart in Proto, which allocates per
 * call, Proto_ReductionImplem.H:161-179): its current size and the number of
 * superseded buffers kept alive because a graph capture may reference them
 * (see the graphs section). */
int pm_stream_scratch_info(pm_stream_t stream, size_t* bytes, int* retired);
int pm_stream_sync(pm_stream_t stream);
int pm_device_sync(void);
/* HIP events for timing work on a stream (hipEventRecord / hipEventElapsedTime). */
typedef void* pm_event_t;
int pm_event_create(pm_event_t* ev);
int pm_event_destroy(pm_event_t ev);
int pm_event_record(pm_event_t ev, pm_stream_t stream);
int pm_event_sync(pm_event_t ev);
int pm_event_elapsed_ms(pm_event_t start, pm_event_t stop, float* ms);

/* --------------------------------------------------------------- Stencil */
typedef struct pm_stencil pm_stencil;

/* Stencil<T> after construction (coefficients merged by addCoef,
 * Proto_StencilImplem.H:237-276) plus closeForDevice (539-556): uploads the
 * coefficient list (T, in Proto's order) and the offsets (ncoef x 3 ints).
 * src_ratio / dst_ratio / dst_shift: Stencil::srcRatio/destRatio/destShift. */
int pm_stencil_create(pm_stencil** out, int dtype, int dim, int ncoef, const void* coefs_host,
                      const int* offsets_host, const int src_ratio[3], const int dst_ratio[3],
                      const int dst_shift[3]);
int pm_stencil_destroy(pm_stencil* st);
/* Which kernel family serves this stencil (for tests and tools): 7 / 27 the
 * 7- / 27-point z-march (Proto's Laplacian / Laplacian_27 in Proto's order,
 * 3-D), 2 the same z-march with its single centre tap (a 3-D stencil of one
 * coefficient at the origin: Stencil<T>::Identity(), Proto_Stencil.H:620, and
 * its multiples; it stages the same halo'd planes as the 7-point kernel), 1
 * the generic z-march (any other unit-ratio stencil whose offsets lie in
 * [-1,1]^DIM, 2-D included: taps read from an LDS plane ring in the stencil's
 * order), 0 the per-cell kernel (refinement ratios, shifts, wider offsets). */
int pm_stencil_kind(const pm_stencil* st, int* kind);
/* Serve this stencil with another kernel family that can (tests and
 * benchmarks compare the families): 0 (any stencil), 1 (a 7/27-point or
 * one-tap one too) or its own kind.  Every family gives the same bits. */
int pm_stencil_set_kind(pm_stencil* st, int kind);

/* Stencil::apply -> protoApply<DEVICE> -> sIndexer3::gpu
 * (Proto_StencilImplem.H:412-431, 844-895, 565-614 and hostApply 725-828):
 *   for p in idx_box:  dst[p*dR + dS] = (replace ? 0 : dst) + sum_j c'_j src[p*sR + o_j]
 * accumulated in coefficient order with one rounding per multiply and per add,
 * c'_j = c_j * scale when scale != 1 (bit-identical to hostApply).  scale == 0
 * only zeroes (replace) or does nothing (accumulate).  Box inference
 * (indexDomain(dst) & indexRange(src)) is the caller's (the C++ API's). */
int pm_stencil_apply(const pm_stencil* st, const void* src, const pm_box* src_box, void* dst,
                     const pm_box* dst_box, const pm_box* idx_box, int ncomp, int replace,
                     double scale, pm_stream_t stream);

/* Stencil::apply followed by a Reduction of the values it wrote, in one
 * pass where the generic z-march serves the stencil (kinds 1 and 2, 2-D
 * included):
 * config 1's  dst |= L(u, scale); dst.absMax()  as one kernel plus a
 * one-block finish (Proto_StencilImplem.H:412-431, Proto_BoxDataImplem.H:
 * 1375-1414).  out_dev receives op over the written cells (combined with its
 * old value when combine != 0); fixed tree, so the bits repeat run to run.
 * Other families run pm_stencil_apply then pm_reduce_box over the same cells
 * (destination-refined stencils are refused: use the two calls).  ncomp 1. */
int pm_stencil_apply_reduce(const pm_stencil* st, const void* src, const pm_box* src_box, void* dst,
                            const pm_box* dst_box, const pm_box* idx_box, int replace, double scale, int op,
                            void* out_dev, int combine, pm_stream_t stream);

/* Fused form of   Lu |= L(u, scale);  r = forall(f - Lu);  r.absMax(); r.sum()
 * (config-2 program: Proto_StencilImplem.H:456-470, Proto_ForallImplem.H:195-213,
 * Proto_BoxDataImplem.H:1375-1414).  r, f share r_box; u covers the stencil's
 * domain of idx_box = r_box.  red_out (device, may be NULL) receives
 * {absMax(r), sum(r)} as two T values; the sum is a fixed-shape tree, so it is
 * deterministic but not in Proto's sequential order (bound: n*eps*sum|r|). */
int pm_stencil_residual(const pm_stencil* st, const void* u, const pm_box* u_box, const void* f,
                        void* r, const pm_box* r_box, double scale, void* red_out,
                        pm_stream_t stream);

/* The same for two operators in one pass over u and f (config 2 computes the
 * 7-point and the 27-point residual of the same u): r1 = f - L1(u, scale1),
 * r2 = f - L2(u, scale2); red_out = {absMax(r1), sum(r1), absMax(r2), sum(r2)}. */
int pm_stencil_residual2(const pm_stencil* st1, const pm_stencil* st2, const void* u,
                         const pm_box* u_box, const void* f, void* r1, void* r2,
                         const pm_box* r_box, double scale1, double scale2, void* red_out,
                         pm_stream_t stream);

/* pm_stencil_residual2 with no device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partials live in a fixed 1 MiB device array of the library
 * (one per GPU, held until the process ends) instead of the stream's scratch,
 * so r_box may launch at most 32768 FP64 workgroups (65536 FP32): 512^3 with
 * the default tiles launches 8192; larger boxes return PM_ERR_INVALID.  Same
 * kernels, same bits as pm_stencil_residual2.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (two streams, two threads) race. */
int pm_stencil_residual2_noalloc(const pm_stencil* st1, const pm_stencil* st2, const void* u,
                                 const pm_box* u_box, const void* f, void* r1, void* r2,
                                 const pm_box* r_box, double scale1, double scale2,
                                 void* red_out, pm_stream_t stream);

/* ------------------------------------------------------------- autotuner */
/* Tile configuration of the z-marching 7/27-point kernels: rows of the
 * 64-wide tile per thread (1, 2, 4), planes prefetched (1, 2), non-temporal
 * result stores (0, 1), z planes per workgroup (0 = automatic). */
typedef struct pm_tile_config { int rows; int prefetch; int nt_store; int zplanes; } pm_tile_config;
/* Times every tile configuration for this operator (st2: the second operator
 * of mode 3, else NULL), mode (0 apply |=, 1 apply +=, 2 fused residual,
 * 3 fused two-operator residual), dtype and index box on scratch data (the
 * source box src_box), keeps the fastest in the library's table -- later calls
 * with the same key use it -- and returns it with its time.  The z-march
 * kinds only (7 and 27 in every mode, 2 in modes 0 and 1).  No counterpart in
 * the reference, which has fixed launch shapes (Proto_StencilImplem.H:860-891). */
int pm_stencil_autotune(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* src_box,
                        const pm_box* idx_box, int reps, pm_tile_config* best, float* best_ms,
                        pm_stream_t stream);
int pm_stencil_set_tile_config(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* idx_box,
                               const pm_tile_config* cfg);
int pm_stencil_get_tile_config(const pm_stencil* st, const pm_stencil* st2, int mode, const pm_box* idx_box,
                               pm_tile_config* cfg);
/* The tuned table as text (one entry per line), to keep tuning across runs. */
int pm_tune_table_save(const char* path);
int pm_tune_table_load(const char* path);

/* ------------------------------------------------------------- Reduction */
/* Reduction<T,OP,DEVICE>::reduce (Proto_ReductionImplem.H:157-192) over n
 * contiguous values.  The result (one T) goes to out_dev; with combine != 0
 * it is folded into the value already there (Proto's m_deviTotal), else it
 * replaces it.  Deterministic for a given n. */
int pm_reduce(int op, int dtype, const void* data, size_t n, void* out_dev, int combine,
              pm_stream_t stream);
/* The same over the cells of `region` of component `comp` of an array on `box`
 * (LevelBoxData::reduce's per-box valid region, Proto_LevelBoxDataImplem.H:595-617). */
int pm_reduce_box(int op, int dtype, const void* data, const pm_box* box, const pm_box* region,
                  int comp, void* out_dev, int combine, pm_stream_t stream);
/* Reduction::init (Proto_ReductionImplem.H:64-76) written to out_dev. */
int pm_reduce_init(int op, int dtype, void* out_dev, pm_stream_t stream);

/* --------------------------------------------------- BoxData pointwise ops */
/* BoxData::copyTo(dst, region, shift) (Proto_BoxDataImplem.H:490-640):
 * dst[p + shift] = src[p] for p in region (source coordinates), all ncomp. */
int pm_boxdata_copy(int dtype, const void* src, const pm_box* src_box, void* dst,
                    const pm_box* dst_box, const pm_box* region, const int shift[3], int ncomp,
                    pm_stream_t stream);
/* dst op= src on region (BoxData operatorT, Proto_BoxDataImplem.H:937-983).
 * op: 0 add, 1 subtract, 2 multiply, 3 divide. */
int pm_boxdata_binop(int op, int dtype, void* dst, const pm_box* dst_box, const void* src,
                     const pm_box* src_box, const pm_box* region, int ncomp, pm_stream_t stream);
/* dst op= value on region: op 0 add, 1 subtract, 2 multiply, 3 divide, 4 set
 * (scalarIndexer / setVal, Proto_BoxDataImplem.H:430-486, 1260-1278). */
int pm_boxdata_scalar(int op, int dtype, void* dst, const pm_box* dst_box, const pm_box* region,
                      double value, int ncomp, pm_stream_t stream);
/* BoxData::increment: dst += src * scale (Proto_BoxDataImplem.H:1219-1234). */
int pm_boxdata_increment(int dtype, void* dst, const pm_box* dst_box, const void* src,
                         const pm_box* src_box, const pm_box* region, double scale, int ncomp,
                         pm_stream_t stream);

/* ------------------------------------------------------------------ level */
/* A DisjointBoxLayout of equal boxes on a ProblemDomain [0, domain-1]
 * (Proto_DisjointBoxLayout.H:52-110, Proto_ProblemDomainImplem.H:1-75): box
 * size bs divides the domain; nboxes boxes given by their low corners in
 * layout order (holes allowed).  Level data (LevelBoxData, one component set
 * per box) is one device allocation: box b, component c starts at
 * (b * ncomp + c) * |box grown by ghost| elements. */
typedef struct pm_layout pm_layout;
int pm_layout_create(pm_layout** out, int dim, const int domain[3], const int box_size[3],
                     const int periodic[3], int nboxes, const int* box_lows_host);
int pm_layout_destroy(pm_layout* layout);
/* Elements of one box's ghosted storage (one component). */
int pm_layout_box_elems(const pm_layout* layout, const int ghost[3], size_t* elems);

/* LevelBoxData::exchange (Proto_LevelBoxDataImplem.H:313-341, motion plans
 * of LevelExchangeCopier 898-951, periodic images ProblemDomainImplem.H:61-75):
 * every ghost cell that is the (periodic image of a) valid cell of some box
 * receives that cell's value.  One launch for the whole level. */
int pm_level_exchange(const pm_layout* layout, int dtype, void* data, const int ghost[3], int ncomp,
                      pm_stream_t stream);

/* Red-black Gauss-Seidel half sweep on the 2*DIM+1-point Laplacian at the
 * cells of one colour ((x+y+z) mod 2 == color, global indices), as
 * examples/_archived/AMRFAS_old/src/LaplaceOp.H:56-73, 95-123 composes it:
 *   t = 0 + sum_j c_j phi(p+o_j), c_j = Laplacian()_j * (-1/(dx*dx));
 *   t = t + rhs(p);  phi(p) = phi(p) + lambda*t,  lambda = 2*(-(dx*dx)/(4*DIM)).
 * phi's ghost cells must hold the neighbours (exchange first). */
int pm_level_gsrb(const pm_layout* layout, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                  const int rhs_ghost[3], double dx, int color, pm_stream_t stream);

/* One full GSRB sweep (red then black, LaplaceOp.H:95-123) in one pass and
 * out of place, with no exchange: phi_out (valid cells) = swept phi_in; cells
 * of other boxes are read from their owners.  Bitwise the exchange / red /
 * exchange / black sequence.  3-D, periodic layouts covering their domain;
 * PM_ERR_UNSUPPORTED otherwise (use pm_level_gsrb).  Levels of boxes 64 cells
 * wide (at most 8 across the domain) run the row-register kernel, others the
 * LDS tile kernel; both give the same bits. */
int pm_level_gsrb_sweep(const pm_layout* L, int dtype, const void* phi_in, const int in_ghost[3], void* phi_out,
                        const int out_ghost[3], const void* rhs, const int rhs_ghost[3], double dx,
                        pm_stream_t stream);
/* Which kernel runs the one-pass sweep here and in the multigrid driver
 * (process-wide; tests and benchmarks compare them): 0 automatic (levels of
 * 64-wide boxes: the row-register kernel; FP32 into a ghost-free output two
 * cells per lane where the level is an even number of boxes across, FP32 back
 * into a ghost-1 phi with 8 rows per thread; in the multigrid driver the FP32
 * second pre-smoothing sweep also writes the residual restriction), 1 the LDS
 * tile kernel, 2 / 3 / 4 the row-register kernel with 4 / 8 / 2 rows per
 * thread (one cell per lane, restriction in its own pass).
 * pm_level_sweep_tune: the row kernel's largest planes per workgroup (0
 * automatic) and non-temporal write-out into ghost-free outputs (-1
 * automatic: on, 0 off, 1 on); every setting gives the same bits. */
int pm_level_sweep_form(int form);
int pm_level_sweep_tune(int planes_max, int nt_store);
/* Damped point Jacobi of examples/_archived/LevelMultigrid/src/LevelMultigrid.cpp:99-123:
 * tmp = L(phi, 1/(4 DIM)) + (-(dx*dx)/(4 DIM)) rhs on all valid cells, then
 * phi += tmp.  tmp: level data without ghosts. */
int pm_level_jacobi(const pm_layout* layout, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                    const int rhs_ghost[3], void* tmp, double dx, pm_stream_t stream);

/* resnorm of LevelMultigrid.cpp:75-97: max over valid cells of
 * |(0 - rhs) + sum_j (c_j/dx^2) phi(p+o_j)| written to out_dev (one T);
 * combine != 0 folds into the value there. */
int pm_level_resnorm(const pm_layout* layout, int dtype, const void* phi, const int phi_ghost[3],
                     const void* rhs, const int rhs_ghost[3], double dx, void* out_dev, int combine,
                     pm_stream_t stream);

/* coarseResidual of LevelMultigrid.cpp:51-73 fused with the copy into the
 * coarse layout: res = (0 + rhs) + sum_j (c_j * (-1/dx^2)) phi(p+o_j) on the
 * fine level, resc(I) = 0 + sum_{k in Box(2), x fastest} (1/8) res(2I+k)
 * (AvgDown(2), Proto_StencilDefs.H:302-327), written straight into the boxes
 * of `crse` (LevelBoxData::copyTo, 343-386).  crse covers fine.coarsen(2). */
int pm_level_restrict_residual(const pm_layout* fine, const pm_layout* crse, int dtype, const void* phi,
                               const int phi_ghost[3], const void* rhs, const int rhs_ghost[3], double dx,
                               void* resc, const int resc_ghost[3], pm_stream_t stream);

/* fineInterp of LevelMultigrid.cpp:125-140: phi += I(delta) with
 * InterpStencil::Constant (linear = 0) or ::Linear (linear = 1) of ratio 2,
 * staged x, y, z as Proto_InterpStencilImplem.H:83-112 does; delta lives on
 * the coarse layout (periodic images read across boxes, so no exchange is
 * needed). */
int pm_level_prolong_add(const pm_layout* fine, const pm_layout* crse, int dtype, const void* delta,
                         const int delta_ghost[3], void* phi, const int phi_ghost[3], int linear,
                         pm_stream_t stream);

/* setVal on every cell (ghosts included) of level data. */
int pm_level_setval(const pm_layout* layout, int dtype, void* data, const int ghost[3], int ncomp,
                    double value, pm_stream_t stream);

/* LevelBoxData::copyTo between two layouts of the same domain
 * (Proto_LevelBoxDataImplem.H:343-386): every valid cell of dst receives
 * the valid cell of src at the same index. */
int pm_level_copy(const pm_layout* src_layout, const void* src, const int src_ghost[3],
                  const pm_layout* dst_layout, void* dst, const int dst_ghost[3], int dtype, int ncomp,
                  pm_stream_t stream);
/* LevelBoxData::copyTo with Proto's regions (Proto_LevelBoxDataImplem.H:346-386):
 * for compatible layouts (same domain, box size and patches) the ghosted box
 * data is copied box by box over the common ghosted region and the destination
 * then exchanged (copyToSimple); otherwise (LevelCopier, 795-870) every cell of
 * every ghosted destination box that lies in a source patch, or whose periodic
 * image does, takes the source's valid value, and all other cells are kept. */
int pm_level_copy_to(const pm_layout* src_layout, const void* src, const int src_ghost[3],
                     const pm_layout* dst_layout, void* dst, const int dst_ghost[3], int dtype, int ncomp,
                     pm_stream_t stream);

/* Residual of the correction scheme, r = (0 + rhs) + sum_j (c_j * (-1/dx^2)) phi(p+o_j)
 * (the order of coarseResidual, LevelMultigrid.cpp:64-66) on every valid cell,
 * written as `out_dtype` into out (level data, out_ghost), with max|r| to
 * norm_out (one value of phi's dtype; |r| equals resnorm's value bit for bit).
 * One pass serves both the norm test and the next right-hand side of the
 * FP32-preconditioned refinement.  phi must be exchanged. */
int pm_level_residual(const pm_layout* layout, int dtype, const void* phi, const int phi_ghost[3],
                      const void* rhs, const int rhs_ghost[3], double dx, int out_dtype, void* out,
                      const int out_ghost[3], void* norm_out, pm_stream_t stream);
/* phi += (T)e on valid cells (e of dtype e_dtype): the FP64 update of the
 * refinement with an FP32 correction. */
int pm_level_add_correction(const pm_layout* layout, int dtype, void* phi, const int phi_ghost[3], int e_dtype,
                            const void* e, const int e_ghost[3], pm_stream_t stream);

/* LevelBoxData::reduce over the valid cells of every box, component comp
 * (Proto_LevelBoxDataImplem.H:595-617); one T to out_dev (combine folds). */
int pm_level_reduce(const pm_layout* layout, int dtype, const void* data, const int ghost[3], int ncomp, int comp,
                    int op, void* out_dev, int combine, pm_stream_t stream);
/* data(p) += coef * (tx[p.x] * ty[p.y]) * tz[p.z] on the valid cells of every
 * box, p in global indices, tables of domain-length (host arrays; tz one 1.0
 * in 2-D).  Builds separable fields and sums of them on a level (the
 * forallInPlace_p initialisations of Proto's examples); all-ones tables
 * with coef = -mean subtract a mean. */
int pm_level_add_separable(const pm_layout* layout, int dtype, void* data, const int ghost[3], double coef,
                           const double* tx_host, const double* ty_host, const double* tz_host, pm_stream_t stream);

/* LevelBoxData::linearSize / linearOut / linearIn (Proto_LevelBoxDataImplem.H:
 * 412-507 over BoxData::linearOut, Proto_BoxDataImplem.H:1619-1648): the
 * buffer holds box after box in layout order, component after component,
 * each over the ghosted box x fastest -- this library's level storage -- so
 * both are one copy.  to_host / from_host: the buffer is host memory
 * (otherwise device memory, as Proto's DEVICE buffers are). */
int pm_level_linear_size(const pm_layout* layout, int dtype, const int ghost[3], int ncomp, size_t* bytes);
int pm_level_linear_out(const pm_layout* layout, int dtype, const void* data, const int ghost[3], int ncomp,
                        void* buffer, int to_host, pm_stream_t stream);
int pm_level_linear_in(const pm_layout* layout, int dtype, void* data, const int ghost[3], int ncomp,
                       const void* buffer, int from_host, pm_stream_t stream);

/* ------------------------------------------------------------- multigrid */
/* Correction-scheme V-cycle solver on a periodic level
 * (examples/_archived/LevelMultigrid/src/LevelMultigrid.cpp:13-160 with the
 * GSRB relaxation of examples/_archived/AMRFAS_old/src/LaplaceOp.H:95-123, or
 * its own damped Jacobi), solving Laplacian(phi)/dx^2 = rhs.  Coarse levels
 * follow LevelMultigrid::define (21-49): domain halves; box size kept when it
 * divides the coarse domain, else one box.  All level buffers are allocated
 * at create and stay resident. */
typedef struct pm_mg pm_mg;
typedef struct pm_mg_options {
    int relax;      /* 0 GSRB (pre/post/bottom sweeps), 1 damped Jacobi (2*DIM, 2*DIM, 10) */
    int pre, post, bottom;
    int interp;     /* 0 InterpStencil::Constant, 1 InterpStencil::Linear */
    int precision;  /* 0 plain FP64 V-cycles; 1 FP32 V-cycle preconditioner inside FP64
                       iterative refinement (residual, update and norm in FP64) */
    int max_iter;
    double tol;     /* stop when resnorm < tol * resnorm_0 (LevelMultigridSolverImplem.H:129) */
    int use_graph;  /* capture the V-cycle into a hipGraph and replay it */
    int fused;      /* GSRB sweeps as one out-of-place pass each (pm_level_gsrb_sweep, same
                       bits; 3-D periodic levels) instead of exchange / red / exchange / black;
                       one-box levels of up to 4096 cells run all sweeps of a relaxation in
                       one launch (same bits) */
    int inner;      /* precision 1: FP32 V-cycles per refinement step (the preconditioner; 0 or
                       1 = one, at most 4); each starts from the previous one's correction */
} pm_mg_options;
void pm_mg_default_options(pm_mg_options* opt);
/* nlevels = number of coarsenings (the numLevels argument of LevelMultigrid). */
int pm_mg_create(pm_mg** out, const pm_layout* fine, int nlevels, double dx, const pm_mg_options* opt);
int pm_mg_destroy(pm_mg* mg);
/* Device bytes the hierarchy holds (for admission by memory budget). */
int pm_mg_bytes(const pm_mg* mg, size_t* bytes);
/* Layout of level l (0 = finest) owned by the solver. */
int pm_mg_level_layout(const pm_mg* mg, int level, const pm_layout** layout);
/* Solve in place: phi (FP64 level data, ghost 1), rhs (FP64, ghost 0).
 * hist_host (max_iter + 1 doubles, may be NULL) gets the max-norm residual
 * before the first cycle and after each; *iters the cycles done.  Returns
 * PM_ERR_NOT_CONVERGED if max_iter cycles did not reach tol (phi holds the
 * last iterate).  Synchronises `stream` once per cycle (the norm test).  On
 * return the ghost cells of phi are exchanged, as Proto's last resnorm leaves
 * them; the cycles themselves read halos from the neighbouring boxes. */
int pm_mg_solve(pm_mg* mg, void* phi, const void* rhs, double* hist_host, int* iters, pm_stream_t stream);
/* One plain V-cycle on (phi, rhs) of the finest level, asynchronous. */
/* Stopping rule of later solves: max-norm residual < tol * initial, at most
 * max_iter cycles (LevelMultigridSolver::solve's a_tolerance / a_maxIter,
 * LevelMultigridSolverImplem.H:408-433). */
int pm_mg_set_tolerance(pm_mg* mg, double tol, int max_iter);
int pm_mg_vcycle(pm_mg* mg, void* phi, const void* rhs, pm_stream_t stream);
/* resnorm (LevelMultigrid.cpp:75-97) of (phi, rhs); synchronises. */
int pm_mg_resnorm(pm_mg* mg, void* phi, const void* rhs, double* out_host, pm_stream_t stream);

/* ---------------------------------------------------------------- batches */
/* A batch of independent patches of uneven size (AMR boxes).  Box b has
 * valid extents sizes[3b..3b+2] (box-local indices from 0) and `ghost` ghost
 * layers; ghosted storage holds the boxes one after another (x fastest over
 * [-ghost, n+ghost)), valid storage likewise over [0, n).  Each operator is
 * ONE persistent launch over all boxes (a work queue of tiles, largest first)
 * instead of Proto's launch per box (e.g. LevelMultigridSolverImplem.H:
 * 366-378); the queue resets itself, so the launch can be graph-captured. */
typedef struct pm_batch pm_batch;
int pm_batch_create(pm_batch** out, int nboxes, const int* sizes_host, int ghost);
int pm_batch_destroy(pm_batch* batch);
int pm_batch_elems(const pm_batch* batch, size_t* ghosted, size_t* valid);
/* nboxes + 1 element offsets of every box in ghosted / valid storage. */
int pm_batch_offsets(const pm_batch* batch, size_t* ghosted_off, size_t* valid_off);
/* Tiling knobs of the batched operators (what pm_batch_autotune searches):
 * z planes per apply tile (0 = whole box), persistent workgroups per CU for
 * the apply, rows per GSRB tile (fewer when an FP64 tile as wide as the widest
 * box would not fit the 64 KiB LDS of a workgroup), workgroups per CU for
 * the sweep. */
typedef struct pm_batch_config {
    int apply_zplanes, apply_blocks_per_cu, gsrb_rows, gsrb_blocks_per_cu;
} pm_batch_config;
int pm_batch_configure(pm_batch* batch, const pm_batch_config* cfg);
int pm_batch_get_config(const pm_batch* batch, pm_batch_config* cfg);
/* Times every candidate configuration (reps launches each, on `stream`) of
 * pm_batch_apply with `st` (src -> dst) and of pm_batch_gsrb (phi_in = src,
 * phi_out, rhs, dx; dtype from `st`), keeps the fastest of each operator and
 * returns it in *best; best_ms[0..1] = its mean ms per launch (apply, gsrb).
 * dst and phi_out are overwritten.  Results are identical under every
 * configuration (only the schedule changes). */
int pm_batch_autotune(pm_batch* batch, const pm_stencil* st, const void* src, void* dst, void* phi_out,
                      const void* rhs, double dx, int reps, pm_batch_config* best, float* best_ms,
                      pm_stream_t stream);
/* dst = L(src, scale) on every box (src ghosted storage, dst valid storage)
 * for the 7-point Laplacian() or 27-point Laplacian_27() shape: per box the
 * bits of pm_stencil_apply. */
int pm_batch_apply(pm_batch* batch, const pm_stencil* st, const void* src, void* dst, double scale,
                   pm_stream_t stream);
/* One red-black Gauss-Seidel sweep of the 7-point Laplacian on every box,
 * ghosts held fixed (LaplaceOp.H:95-123 without the exchange: the patches are
 * independent): phi_out (ghosted storage) receives the swept phi_in; the
 * ghosts of phi_out must equal those of phi_in (some kernels write the x ghost
 * cells of the rows they sweep, with phi_in's values; none writes any other
 * ghost cell). */
int pm_batch_gsrb(pm_batch* batch, int dtype, const void* phi_in, void* phi_out, const void* rhs, double dx,
                  pm_stream_t stream);
/* Which kernel runs pm_batch_gsrb (tests and benchmarks compare them; every
 * form gives the same bits): 0 automatic (FP32 with 2 ghost layers form 7
 * where every box is 4..252 cells wide in multiples of 4, phi 16-byte and rhs
 * 8-byte aligned; else with 2 ghost layers form 6 where every box is even and
 * 4..128 cells wide and the arrays pair-aligned; else
 * FP64 the pair sweep (even ghost), then the row sweep where every box is 2..64
 * cells wide; FP32 the row sweep first, then the pair sweep; else the LDS
 * ring), 1 the LDS-ring kernel, 2
 * the row sweep (a wave per 64 / nx packed sub-tiles of 4 rows, rows in
 * registers, one cell per lane), 3 the row sweep with 2-row sub-tiles, 4 the
 * row sweep that also writes the x ghost cells of its rows (ghost 1 or 2;
 * phi_in's values) with cached stores, 5 the pair sweep (two cells per lane,
 * 64 / (nx / 2) sub-tiles of 4 rows per wave; either dtype), 6 the pair sweep
 * also writing the ghost pairs of its rows (ghost 2; phi_in's values) with
 * cached stores, so that output rows fill whole cache lines, 7 the quad sweep
 * (FP32, ghost 2, boxes 4..252 wide in multiples of 4: four cells per lane, a
 * lane holding a 16-byte quad of a row as stored, the rows' ghost columns
 * included and written back with phi_in's values; 64 / (nx / 4 + 1) sub-tiles
 * of 2 rows per wave), 8 the quad sweep with non-temporal stores.  waves:
 * workgroup size of forms 0, 2, 4 and 5 in waves (0 automatic, 1, 2, 4 or 8;
 * form 4: 1 or 4; the waves of a workgroup take neighbouring sub-tiles of one
 * box and walk z together); forms 3, 6, 7 and 8 run one wave per workgroup.
 * No counterpart in the reference (one fixed kernel per box). */
int pm_batch_sweep_form(pm_batch* batch, int form, int waves);

/* ------------------------------------------------------ library multigrid */
/* Proto's LevelMultigridSolver<BoxOp_Laplace> (include/utils/
 * Proto_LevelMultigridSolver.H, implem 1-433; examples/_common/BoxOp_Laplace.H):
 * FAS V-cycles with the 4th-order flux-form Laplacian (DiffCellToFace fluxes,
 * flux scale -1), damped Jacobi relaxation (2 DIM / 2 DIM / 10 sweeps, lambda =
 * 0.5 / spectral radius), AvgDown restriction of state and residual and
 * constant prolongation of the correction, on a periodic 3-D layout covering
 * its domain.  num_levels counts the finest level (Proto's a_numLevels).
 * state: level storage with 2 ghost layers; force: no ghosts.  Each
 * operator is one launch per level; results are Proto's bits. */
typedef struct pm_fas pm_fas;
int pm_fas_create(pm_fas** out, const pm_layout* fine, int num_levels, double dx);
int pm_fas_destroy(pm_fas* fas);
int pm_fas_bytes(const pm_fas* fas, size_t* bytes);
/* MGLevel::vCycle (408-433) */
int pm_fas_vcycle(pm_fas* fas, void* state, const void* force, pm_stream_t stream);
/* LevelMultigridSolver::resnorm (134-160): max |L(state) + force| */
int pm_fas_resnorm(pm_fas* fas, void* state, const void* force, double* out_host, pm_stream_t stream);
/* the same, also writing the residual L(state) + force (valid cells, no ghosts) */
int pm_fas_residual(pm_fas* fas, void* state, const void* force, void* residual, double* out_host,
                    pm_stream_t stream);
/* LevelMultigridSolver::solve (85-132): V-cycles until resnorm < tol *
 * initial or max_iter; hist[0..*iters] the norms (hist may be NULL).  On
 * return the ghost cells of state are exchanged, as Proto's last resnorm
 * leaves them. */
int pm_fas_solve(pm_fas* fas, void* state, const void* force, int max_iter, double tol, double* hist, int* iters,
                 pm_stream_t stream);

/* ---------------------------------------------------------------- serving */
/* A continuous-batching Poisson solve server.  No counterpart in the reference
 * (Proto solves one problem per LevelMultigridSolver::solve call,
 * LevelMultigridSolverImplem.H:408-433).  Each request is a periodic cube of
 * n^3 cells (unit domain, h = 1/n) with rhs = sum of K Gaussians
 * (sigma, centres in [0,1)^3 and amplitudes +-1 from the request seed, periodic
 * minimum-image distance) minus its mean; it is solved for Laplacian(phi) =
 * rhs to max-norm residual < tol * initial.
 *
 * Mode 0 (batched FP32-IR): requests are admitted FIFO while their hierarchy
 * fits the HBM budget (hierarchies stay resident in a pool keyed by n) and a
 * slot is free; every loop iteration runs ONE FP32 GSRB V-cycle over all
 * active requests (each level operator is one launch covering every request
 * that has that level), the FP64 correction, the FP64 residual and one norm
 * read-back; converged requests retire and waiting ones join at once.  Per
 * request the arithmetic is that of pm_mg precision 1 on a one-box layout
 * (same FP32 coefficients, same order), so a result does not depend on which
 * other requests were in flight.
 * Mode 1 (plain FP64): one request at a time through pm_mg precision 0
 * (Proto's GSRB V-cycle), the reference point for the batched mode; one
 * solver per size stays resident within the HBM budget (solvers of other
 * sizes are evicted first; a size that cannot fit alone is rejected).
 * Levels: n is halved while it is even and larger than 8 (32 -> 8, 48 -> 6). */
typedef struct pm_server pm_server;
typedef struct pm_server_options {
    size_t hbm_budget;  /* bytes the resident hierarchies may hold */
    int max_active;     /* requests in flight together (mode 0) */
    int mode;           /* 0 batched FP32-IR, 1 plain FP64 one at a time */
    int pre, post, bottom; /* GSRB sweeps (even counts) */
    int max_iter;       /* outer iterations before PM_SERVE_NOT_CONVERGED */
    double sigma;       /* Gaussian width */
    int inner;          /* mode 0: FP32 V-cycles per refinement step (0 or 1 = one, at most 4),
                           as pm_mg_options.inner */
} pm_server_options;
typedef struct pm_request {
    double arrival_s;   /* seconds after the start of pm_server_run */
    int n;              /* cells per side */
    int ngauss;         /* K */
    uint64_t seed;      /* Gaussian centres / amplitudes */
    double tol;         /* relative max-norm residual */
    int keep_result;    /* keep phi for pm_server_result */
} pm_request;
enum { PM_SERVE_SOLVED = 0, PM_SERVE_NOT_CONVERGED = 1, PM_SERVE_REJECTED = 2 };
typedef struct pm_request_stats {
    double arrival_s, admit_s, finish_s; /* seconds after the start of the run */
    int iters;          /* outer iterations (V-cycles) */
    int status;         /* PM_SERVE_* */
    double rel_residual;
    size_t bytes;       /* resident hierarchy */
} pm_request_stats;
void pm_server_default_options(pm_server_options* opt);
int pm_server_create(pm_server** out, const pm_server_options* opt);
int pm_server_destroy(pm_server* srv);
/* Device bytes of one request's hierarchy for n (mode of the server). */
int pm_server_request_bytes(const pm_server* srv, int n, size_t* bytes);
/* Serves the requests in real time (arrivals at their arrival_s) on `stream`
 * and returns when every request is done; stats[i] for reqs[i]. */
int pm_server_run(pm_server* srv, int nreq, const pm_request* reqs, pm_request_stats* stats, pm_stream_t stream);
/* phi (n^3 values, x fastest) of request `index` of the last run, when it was
 * kept (keep_result) and solved. */
int pm_server_result(pm_server* srv, int index, double* phi_host);
/* The right-hand side a request (n, ngauss, seed) is served with, written to
 * rhs_dev (n^3 FP64, x fastest) -- to solve the same problem elsewhere. */
int pm_server_rhs(int n, int ngauss, uint64_t seed, double sigma, double* rhs_dev, pm_stream_t stream);
/* Device bytes resident in the server's pools. */
int pm_server_resident_bytes(const pm_server* srv, size_t* bytes);

/* Checkpoint / resume of a served run (mode 0).  No counterpart in the
 * reference; the solution blocks use LevelBoxData::linearOut's layout
 * (Proto_LevelBoxDataImplem.H:412-507): each in-flight request's FP64 solution
 * as the linearOut of its one-box periodic level (n^3 values, x fastest),
 * readable by LevelBoxData<double,1>::linearIn on that layout.
 *
 * pm_server_set_stop: pm_server_run stops after `iterations` loop iterations
 * (0 = never) and returns PM_ERR_CANCELLED, holding the run: in-flight
 * hierarchies, the waiting queue, the statistics and kept results so far.
 * The next pm_server_run with the same request list continues it (results
 * and iteration counts as if it had not stopped); pm_server_discard drops it.
 * pm_server_checkpoint writes a held run into a host buffer of
 * pm_server_checkpoint_size bytes (it synchronises `stream`);
 * pm_server_restore loads one into a server with the same sweep counts,
 * max_iter and sigma (another process, another GPU), allocating the in-flight
 * hierarchies within its HBM budget and the resident pool the run had, after
 * which pm_server_run with the same request list resumes it.  A resumed
 * request recomputes the FP64 residual of its saved solution, then continues
 * its V-cycles: same bits and the same cycle count as an uninterrupted run. */
int pm_server_set_stop(pm_server* srv, int iterations);
int pm_server_discard(pm_server* srv);
int pm_server_checkpoint_size(const pm_server* srv, size_t* bytes);
int pm_server_checkpoint(pm_server* srv, void* buffer_host, size_t bytes, pm_stream_t stream);
int pm_server_restore(pm_server* srv, const void* buffer_host, size_t bytes, pm_stream_t stream);

/* ----------------------------------------------------------------- graphs */
/* hipGraph capture of work issued on a (non-null) stream by any entry point
 * above, and replay: the V-cycle and the batched operators are launch-bound
 * sequences that replay as one submission.  No counterpart in the reference
 * (its prototype, experimental/cuda-graph/cuda-graph.H, is not in the
 * library).
 * Entry points that reduce keep per-block partials in a scratch buffer owned
 * by the stream.  A capture records that buffer's address, so a buffer seen
 * by a capture is never freed when a later, larger call grows the scratch:
 * it is kept until pm_stream_destroy, and graphs captured on a stream must
 * not be launched after that stream is destroyed.  A call that would need
 * more scratch than the stream holds while it is capturing fails with
 * PM_ERR_UNSUPPORTED (run it once on the stream before capturing). */
typedef void* pm_graph_t;
int pm_graph_begin(pm_stream_t stream);
int pm_graph_end(pm_stream_t stream, pm_graph_t* graph);
int pm_graph_launch(pm_graph_t graph, pm_stream_t stream);
int pm_graph_destroy(pm_graph_t graph);

/* ----------------------------------------------------------------- inputs */
/* Seeded inputs shared with the tests: value i of the SplitMix64 stream
 * (seed, stream) mapped to [-1, 1); written to data[i] for i < n, as
 * a*u + b when affine (used for config 2's 1e-3 noise). */
int pm_fill_uniform(int dtype, void* data, size_t n, uint64_t seed, uint64_t stream_id,
                    double a, double b, pm_stream_t stream);

/* Initialise a field the way Proto examples do with forallInPlace_p
 * (examples/_archived/LevelMultigrid/exec/main.cpp:18-25): for every cell p
 * of `region` of an array on `box`, with w = p folded into [0, period)
 * (period[d] > 0; use a large period for no folding),
 *   data[p] = (tx[w.x] * ty[w.y]) * tz[w.z]  +  amp * U(seed, stream_id, w.x + P.x (w.y + P.y w.z))
 * where tx/ty/tz are host tables of period[d] values (tz ignored in 2-D: use
 * a table of one 1.0) and U is the pm_fill_uniform stream.  amp = 0 skips
 * the noise.  Ghost cells so filled hold the periodic images. */
int pm_fill_separable(int dtype, void* data, const pm_box* box, const pm_box* region, const int period[3],
                      const double* tx_host, const double* ty_host, const double* tz_host, double amp,
                      uint64_t seed, uint64_t stream_id, pm_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PROTOMI_H */
