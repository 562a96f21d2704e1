// Correction-scheme multigrid on a periodic level, driven from the host over
// the level kernels.  Structure of examples/_archived/LevelMultigrid/src/
// LevelMultigrid.cpp:13-160 (define / coarseResidual / resnorm / fineInterp /
// vCycle) with the GSRB relaxation of examples/_archived/AMRFAS_old/src/
// LaplaceOp.H:95-123 or the archived damped Jacobi.  The whole hierarchy is
// allocated once at create and stays resident; a solve only launches kernels
// and reads back one norm per cycle.
//
// precision = 1: FP32 V-cycles as the preconditioner of an FP64 iterative
// refinement.  Each outer step computes r = rhs - L(phi) in FP64 (one pass that
// also yields the FP64 max-norm and writes r rounded to FP32), solves
// L(e) = r with one FP32 V-cycle from e = 0, and updates phi += e in FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include "pm_internal.hpp"

struct pm_mg {
    pm_mg_options opt;
    int nlev = 0;
    int dim = 3;
    struct Level {
        pm_layout* layout = nullptr;
        double dx = 0;
        void* phi = nullptr;    // FP64, ghost 1 (levels >= 1: the correction delta)
        void* rhs = nullptr;    // FP64, ghost 0 (levels >= 1: the restricted residual)
        void* tmp = nullptr;    // FP64 Jacobi scratch, ghost 0
        void* phi32 = nullptr;  // FP32 hierarchy (precision 1)
        void* rhs32 = nullptr;
        void* tmp32 = nullptr;
    };
    std::vector<Level> lv;
    void* d_norm = nullptr;  // device scalar for norms
    size_t bytes = 0;
    int g1[3] = {1, 1, 1};
    int g0[3] = {0, 0, 0};
    // captured V-cycle (use_graph)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const void* g_phi = nullptr;
    const void* g_rhs = nullptr;
    hipStream_t g_stream = nullptr;
    int g_prec = -1;
};

namespace pm {

static size_t level_elems(const pm_layout* L, const int* g)
{
    size_t e = 0;
    pm_layout_box_elems(L, g, &e);
    return e;
}

}  // namespace pm

namespace pm {

#define MG_TRY(x)                     \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != PM_OK) return rc_; \
    } while (0)

static int dalloc(pm_mg* mg, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) { return PM_OK; }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("multigrid hierarchy allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return PM_ERR_NOMEM;
    }
    mg->bytes += bytes;
    return PM_OK;
}

static void mg_free(pm_mg* mg)
{
    for (auto& l : mg->lv) {
        hipFree(l.phi);
        hipFree(l.rhs);
        hipFree(l.tmp);
        hipFree(l.phi32);
        hipFree(l.rhs32);
        hipFree(l.tmp32);
        pm_layout_destroy(l.layout);
    }
    hipFree(mg->d_norm);
    if (mg->gexec) { hipGraphExecDestroy(mg->gexec); }
    if (mg->graph) { hipGraphDestroy(mg->graph); }
}

// n relaxation sweeps of level l.  in_mode 1: phi is a zero initial guess not
// yet written (m_delta.setToZero()); 2: phi still lacks fineInterp's
// constant prolongation of cphi (level l + 1).  The first one-pass sweep reads
// such an input directly; otherwise it is written first, as Proto does.
// acc64 (FP32 level 0 of the refinement): the last sweep adds its result to
// the FP64 solution instead of storing it (*folded = true), which is then the
// correction phi += e.  crhs / restricted (pre-smoothing): where the level
// qualifies, the last sweep also writes the residual restriction into the
// coarse right-hand side crhs (*restricted = true).
static int relax(pm_mg* mg, int l, int dtype, void* phi, const void* rhs, void* tmp, int n, hipStream_t s,
                 int in_mode = 0, const void* cphi = nullptr, double* acc64 = nullptr, bool* folded = nullptr,
                 void* crhs = nullptr, bool* restricted = nullptr)
{
    const auto& L = mg->lv[l];
    auto prepare = [&]() -> int {
        if (in_mode == 1) {
            MG_TRY(pm_level_setval(L.layout, dtype, phi, mg->g1, 1, 0.0, s));
        } else if (in_mode == 2) {
            MG_TRY(pm_level_prolong_add(L.layout, mg->lv[l + 1].layout, dtype, cphi, mg->g1, phi, mg->g1, 0, s));
        }
        in_mode = 0;
        return PM_OK;
    };
    // tiny one-box levels: every sweep in one launch, phi held in LDS
    if (mg->opt.relax == 0 && mg->opt.fused && mg->dim == 3 && !acc64 && n > 0) {
        const int rc = level_gsrb_small(L.layout, dtype, phi, mg->g1, rhs, mg->g0, L.dx, n, in_mode,
                                        in_mode == 2 ? mg->lv[l + 1].layout : nullptr, cphi, mg->g1, s);
        if (rc == PM_OK) { return PM_OK; }
        if (rc != PM_ERR_UNSUPPORTED) { return rc; }
    }
    int it = 0;
    if (mg->opt.relax == 0 && mg->opt.fused && mg->dim == 3 && tmp) {
        // one-pass sweeps alternating phi -> tmp -> phi; an odd count ends with
        // the two-colour form below
        for (; it + 1 < n; it += 2) {
            int rc = PM_ERR_UNSUPPORTED;
            if (in_mode != 0) {
                rc = level_gsrb_sweep_in(L.layout, dtype, phi, mg->g1, tmp, mg->g0, rhs, mg->g0, L.dx, in_mode,
                                         in_mode == 2 ? mg->lv[l + 1].layout : nullptr, cphi, mg->g1, s);
                if (rc != PM_OK && rc != PM_ERR_UNSUPPORTED) { return rc; }
                if (rc == PM_OK) {
                    in_mode = 0;
                } else {
                    MG_TRY(prepare());
                }
            }
            if (rc != PM_OK) {
                rc = pm_level_gsrb_sweep(L.layout, dtype, phi, mg->g1, tmp, mg->g0, rhs, mg->g0, L.dx, s);
                if (rc == PM_ERR_UNSUPPORTED) { break; }
                if (rc != PM_OK) { return rc; }
            }
            if (restricted && it + 2 == n) {
                rc = level_sweep_restrict(L.layout, mg->lv[l + 1].layout, dtype, tmp, mg->g0, phi, mg->g1, rhs, mg->g0,
                                          L.dx, crhs, mg->g0, s);
                if (rc != PM_OK && rc != PM_ERR_UNSUPPORTED) { return rc; }
                *restricted = rc == PM_OK;
            }
            if (acc64 && it + 2 == n) {
                MG_TRY(level_gsrb_sweep_in(L.layout, dtype, tmp, mg->g0, phi, mg->g1, rhs, mg->g0, L.dx, 0, nullptr,
                                           nullptr, nullptr, s, acc64));
                *folded = true;
            } else if (!(restricted && *restricted)) {
                MG_TRY(pm_level_gsrb_sweep(L.layout, dtype, tmp, mg->g0, phi, mg->g1, rhs, mg->g0, L.dx, s));
            }
        }
    }
    MG_TRY(prepare());
    for (; it < n; it++) {
        if (mg->opt.relax == 0) {
            for (int color = 0; color < 2; color++) {
                MG_TRY(pm_level_exchange(L.layout, dtype, phi, mg->g1, 1, s));
                MG_TRY(pm_level_gsrb(L.layout, dtype, phi, mg->g1, rhs, mg->g0, L.dx, color, s));
            }
        } else {
            MG_TRY(pm_level_exchange(L.layout, dtype, phi, mg->g1, 1, s));
            MG_TRY(pm_level_jacobi(L.layout, dtype, phi, mg->g1, rhs, mg->g0, tmp, L.dx, s));
        }
    }
    return PM_OK;
}

// vCycle (LevelMultigrid.cpp:142-160) on level l with (phi, rhs) of dtype;
// zero_in: phi is the zero initial guess of a correction, not yet written;
// acc64 / folded: see relax (level l's last sweep).
static int vcycle(pm_mg* mg, int l, int dtype, void* phi, const void* rhs, hipStream_t s, bool zero_in = false,
                  double* acc64 = nullptr, bool* folded = nullptr)
{
    auto& L = mg->lv[l];
    const bool f32 = dtype == PM_F32;
    void* tmp = f32 ? L.tmp32 : L.tmp;
    int pre = mg->opt.pre, post = mg->opt.post, bottom = mg->opt.bottom;
    if (mg->opt.relax == 1) { pre = post = 2 * mg->dim; bottom = 10; }
    const int zin = zero_in ? 1 : 0;
    if (l < mg->nlev) {
        auto& C = mg->lv[l + 1];
        void* cphi = f32 ? C.phi32 : C.phi;
        void* crhs = f32 ? C.rhs32 : C.rhs;
        // coarseResidual (51-73): exchange, residual, AvgDown, copy to the coarse
        // layout; where the level qualifies it is computed by the last
        // pre-smoothing sweep, else the halo is read from the neighbouring boxes
        // instead of an exchange
        bool restricted = false;
        MG_TRY(relax(mg, l, dtype, phi, rhs, tmp, pre, s, zin, nullptr, nullptr, nullptr, crhs, &restricted));
        const int rc = restricted ? PM_OK
                                  : level_restrict_nb(L.layout, C.layout, dtype, phi, mg->g1, rhs, mg->g0, L.dx, crhs,
                                                      mg->g0, s);
        if (rc == PM_ERR_UNSUPPORTED) {
            MG_TRY(pm_level_exchange(L.layout, dtype, phi, mg->g1, 1, s));
            MG_TRY(pm_level_restrict_residual(L.layout, C.layout, dtype, phi, mg->g1, rhs, mg->g0, L.dx, crhs, mg->g0,
                                              s));
        } else if (rc != PM_OK) {
            return rc;
        }
        MG_TRY(vcycle(mg, l + 1, dtype, cphi, crhs, s, true));  // m_delta.setToZero() folded into its first sweep
        // fineInterp (125-140): phi += I(delta), folded into the first post sweep for Constant
        if (mg->opt.interp == 0) {
            MG_TRY(relax(mg, l, dtype, phi, rhs, tmp, post, s, 2, cphi, acc64, folded));
        } else {
            MG_TRY(pm_level_prolong_add(L.layout, C.layout, dtype, cphi, mg->g1, phi, mg->g1, mg->opt.interp, s));
            MG_TRY(relax(mg, l, dtype, phi, rhs, tmp, post, s, 0, nullptr, acc64, folded));
        }
    } else {
        MG_TRY(relax(mg, l, dtype, phi, rhs, tmp, bottom, s, zin, nullptr, acc64, folded));
    }
    return PM_OK;
}

static int read_norm(pm_mg* mg, double* out, hipStream_t s)
{
    PM_CHECK_HIP(hipMemcpyAsync(out, mg->d_norm, sizeof(double), hipMemcpyDeviceToHost, s));
    PM_CHECK_HIP(hipStreamSynchronize(s));
    return PM_OK;
}

// resnorm (LevelMultigrid.cpp:75-97) of level 0; the exchange is skipped
// where the norm kernel reads the neighbouring boxes itself (pm_mg_solve
// exchanges phi once at its end, so callers see the ghosts Proto leaves)
static int resnorm(pm_mg* mg, void* phi, const void* rhs, double* out, hipStream_t s)
{
    auto& L = mg->lv[0];
    const int rc = level_resnorm_nb(L.layout, PM_F64, phi, mg->g1, rhs, mg->g0, L.dx, mg->d_norm, 0, s);
    if (rc == PM_ERR_UNSUPPORTED) {
        MG_TRY(pm_level_exchange(L.layout, PM_F64, phi, mg->g1, 1, s));
        MG_TRY(pm_level_resnorm(L.layout, PM_F64, phi, mg->g1, rhs, mg->g0, L.dx, mg->d_norm, 0, s));
    } else if (rc != PM_OK) {
        return rc;
    }
    return read_norm(mg, out, s);
}

// FP64 residual of level 0, rounded into the FP32 right-hand side, and its norm
static int residual32(pm_mg* mg, void* phi, const void* rhs, hipStream_t s)
{
    auto& L = mg->lv[0];
    const int rc = level_residual_nb(L.layout, PM_F64, phi, mg->g1, rhs, mg->g0, L.dx, PM_F32, L.rhs32, mg->g0,
                                     mg->d_norm, s);
    if (rc == PM_ERR_UNSUPPORTED) {
        MG_TRY(pm_level_exchange(L.layout, PM_F64, phi, mg->g1, 1, s));
        MG_TRY(pm_level_residual(L.layout, PM_F64, phi, mg->g1, rhs, mg->g0, L.dx, PM_F32, L.rhs32, mg->g0,
                                 mg->d_norm, s));
    } else if (rc != PM_OK) {
        return rc;
    }
    return PM_OK;
}

// One step of the FP32-preconditioned refinement after the residual r32 of
// level 0 is in place: e = 0, `inner` FP32 V-cycles on A e = r32 (each from
// the previous one's e), phi += e.
static int ir_correct(pm_mg* mg, void* phi, hipStream_t s)
{
    auto& L = mg->lv[0];
    const int inner = mg->opt.inner > 1 ? mg->opt.inner : 1;
    // e = 0 folded into the first sweep, phi += e into the last where it can be
    bool folded = false;
    for (int c = 0; c < inner; c++) {
        const bool last = c == inner - 1;
        MG_TRY(vcycle(mg, 0, PM_F32, L.phi32, L.rhs32, s, c == 0, last ? (double*)phi : nullptr,
                      last ? &folded : nullptr));
    }
    if (!folded) { MG_TRY(pm_level_add_correction(L.layout, PM_F64, phi, mg->g1, PM_F32, L.phi32, mg->g1, s)); }
    return PM_OK;
}

static int ir_residual(pm_mg* mg, void* phi, const void* rhs, double* out, hipStream_t s)
{
    MG_TRY(residual32(mg, phi, rhs, s));
    return read_norm(mg, out, s);
}

// The work of one cycle (no norm read-back), captured into a graph when asked.
static int cycle_body(pm_mg* mg, void* phi, const void* rhs, hipStream_t s)
{
    if (mg->opt.precision == 0) { return vcycle(mg, 0, PM_F64, phi, rhs, s); }
    MG_TRY(ir_correct(mg, phi, s));
    return residual32(mg, phi, rhs, s);
}

static int run_cycle(pm_mg* mg, void* phi, const void* rhs, hipStream_t s)
{
    if (!mg->opt.use_graph) { return cycle_body(mg, phi, rhs, s); }
    if (!mg->gexec || mg->g_phi != phi || mg->g_rhs != rhs || mg->g_stream != s || mg->g_prec != mg->opt.precision) {
        if (mg->gexec) { hipGraphExecDestroy(mg->gexec); mg->gexec = nullptr; }
        if (mg->graph) { hipGraphDestroy(mg->graph); mg->graph = nullptr; }
        // the residual kernels' per-stream scratch must exist before capture
        void* ws = nullptr;
        MG_TRY(workspace(s, sizeof(double) * 64 * 4096, &ws));
        PM_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = cycle_body(mg, phi, rhs, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != PM_OK) {
            if (g) { hipGraphDestroy(g); }
            return rc;
        }
        PM_CHECK_HIP(e);
        mg->graph = g;
        PM_CHECK_HIP(hipGraphInstantiate(&mg->gexec, g, nullptr, nullptr, 0));
        mg->g_phi = phi;
        mg->g_rhs = rhs;
        mg->g_stream = s;
        mg->g_prec = mg->opt.precision;
    }
    PM_CHECK_HIP(hipGraphLaunch(mg->gexec, s));
    return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {

void pm_mg_default_options(pm_mg_options* o)
{
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->relax = 0;
    o->pre = 2;
    o->post = 2;
    o->bottom = 10;
    o->interp = 0;
    o->precision = 0;
    o->max_iter = 30;
    o->tol = 1e-10;
    o->use_graph = 0;
    o->fused = 1;
    o->inner = 1;
}

int pm_mg_create(pm_mg** out, const pm_layout* fine, int nlevels, double dx, const pm_mg_options* opt)
{
    PM_REQUIRE(out && fine && opt, "pm_mg_create: null argument");
    *out = nullptr;
    PM_REQUIRE(nlevels >= 0 && nlevels < 24, "pm_mg_create: bad level count %d", nlevels);
    PM_REQUIRE(opt->relax == 0 || opt->relax == 1, "pm_mg_create: relax must be 0 (GSRB) or 1 (Jacobi)");
    PM_REQUIRE(opt->precision == 0 || opt->precision == 1, "pm_mg_create: precision must be 0 or 1");
    PM_REQUIRE(opt->interp == 0 || opt->interp == 1, "pm_mg_create: interp must be 0 or 1");
    PM_REQUIRE(opt->inner >= 0 && opt->inner <= 4, "pm_mg_create: inner must be 0..4");
    PM_REQUIRE(opt->pre >= 0 && opt->post >= 0 && opt->bottom >= 0 && opt->max_iter >= 0, "pm_mg_create: bad counts");
    PM_REQUIRE(dx > 0, "pm_mg_create: dx must be positive");
    int dim, domain[3], bs[3], periodic[3], nboxes;
    MG_TRY(layout_shape(fine, &dim, domain, bs, periodic, &nboxes));
    for (int k = 0; k < dim; k++) {
        PM_REQUIRE(periodic[k], "pm_mg_create: the solver needs a periodic domain");
        PM_REQUIRE(domain[k] % (1 << nlevels) == 0, "pm_mg_create: domain not coarsenable %d times", nlevels);
    }
    pm_mg* mg = new pm_mg();
    mg->opt = *opt;
    mg->nlev = nlevels;
    mg->dim = dim;
    if (dim == 2) { mg->g1[2] = 0; }
    mg->lv.resize(nlevels + 1);
    int rc = PM_OK;
    // level 0: a private copy of the fine layout; coarser: LevelMultigrid::define (21-49)
    std::vector<int> lows(3 * nboxes);
    rc = layout_lows(fine, lows.data());
    if (rc == PM_OK) { rc = pm_layout_create(&mg->lv[0].layout, dim, domain, bs, periodic, nboxes, lows.data()); }
    mg->lv[0].dx = dx;
    int dom[3] = {domain[0], domain[1], domain[2]};
    int box[3] = {bs[0], bs[1], bs[2]};
    for (int l = 1; l <= nlevels && rc == PM_OK; l++) {
        bool divisible = true;
        for (int k = 0; k < dim; k++) {
            dom[k] /= 2;
            if (dom[k] % box[k] != 0) { divisible = false; }
        }
        if (!divisible) {
            for (int k = 0; k < dim; k++) { box[k] = dom[k]; }
        }
        int nb[3] = {1, 1, 1};
        for (int k = 0; k < dim; k++) { nb[k] = dom[k] / box[k]; }
        std::vector<int> cl;
        for (int z = 0; z < nb[2]; z++)
            for (int y = 0; y < nb[1]; y++)
                for (int x = 0; x < nb[0]; x++) {
                    cl.push_back(x * box[0]);
                    cl.push_back(y * box[1]);
                    cl.push_back(dim == 3 ? z * box[2] : 0);
                }
        rc = pm_layout_create(&mg->lv[l].layout, dim, dom, box, periodic, (int)(cl.size() / 3), cl.data());
        mg->lv[l].dx = mg->lv[l - 1].dx * 2;
    }
    // resident hierarchy
    for (int l = 0; l <= nlevels && rc == PM_OK; l++) {
        auto& L = mg->lv[l];
        int nbx = 0, d_, dm_[3], bs_[3], p_[3];
        layout_shape(L.layout, &d_, dm_, bs_, p_, &nbx);
        const size_t e1 = level_elems(L.layout, mg->g1) * nbx, e0 = level_elems(L.layout, mg->g0) * nbx;
        if (l > 0 && rc == PM_OK) { rc = dalloc(mg, &L.phi, 8 * e1); }
        if (l > 0 && rc == PM_OK) { rc = dalloc(mg, &L.rhs, 8 * e0); }
        const bool need_tmp = opt->relax == 1 || (opt->fused && dim == 3);
        if (need_tmp && rc == PM_OK) { rc = dalloc(mg, &L.tmp, 8 * e0); }
        if (opt->precision == 1) {
            if (rc == PM_OK) { rc = dalloc(mg, &L.phi32, 4 * e1); }
            if (rc == PM_OK) { rc = dalloc(mg, &L.rhs32, 4 * e0); }
            if (need_tmp && rc == PM_OK) { rc = dalloc(mg, &L.tmp32, 4 * e0); }
        }
    }
    if (rc == PM_OK) { rc = dalloc(mg, &mg->d_norm, 64); }
    if (rc != PM_OK) {
        mg_free(mg);
        delete mg;
        return rc;
    }
    *out = mg;
    return PM_OK;
}

int pm_mg_destroy(pm_mg* mg)
{
    if (!mg) { return PM_OK; }
    mg_free(mg);
    delete mg;
    return PM_OK;
}

int pm_mg_bytes(const pm_mg* mg, size_t* bytes)
{
    PM_REQUIRE(mg && bytes, "pm_mg_bytes: null argument");
    *bytes = mg->bytes;
    return PM_OK;
}

int pm_mg_level_layout(const pm_mg* mg, int level, const pm_layout** layout)
{
    PM_REQUIRE(mg && layout && level >= 0 && level <= mg->nlev, "pm_mg_level_layout: bad argument");
    *layout = mg->lv[level].layout;
    return PM_OK;
}

int pm_mg_set_tolerance(pm_mg* mg, double tol, int max_iter)
{
    PM_REQUIRE(mg && tol >= 0 && max_iter >= 0, "pm_mg_set_tolerance: bad argument");
    mg->opt.tol = tol;
    mg->opt.max_iter = max_iter;
    return PM_OK;
}

int pm_mg_vcycle(pm_mg* mg, void* phi, const void* rhs, pm_stream_t stream)
{
    PM_REQUIRE(mg && phi && rhs, "pm_mg_vcycle: null argument");
    return vcycle(mg, 0, PM_F64, phi, rhs, S(stream));
}

int pm_mg_resnorm(pm_mg* mg, void* phi, const void* rhs, double* out_host, pm_stream_t stream)
{
    PM_REQUIRE(mg && phi && rhs && out_host, "pm_mg_resnorm: null argument");
    MG_TRY(pm_level_exchange(mg->lv[0].layout, PM_F64, phi, mg->g1, 1, S(stream)));  // as Proto's resnorm does
    return resnorm(mg, phi, rhs, out_host, S(stream));
}

int pm_mg_solve(pm_mg* mg, void* phi, const void* rhs, double* hist, int* iters, pm_stream_t stream)
{
    PM_REQUIRE(mg && phi && rhs, "pm_mg_solve: null argument");
    const hipStream_t s = S(stream);
    double res0 = 0, r = 0;
    if (mg->opt.precision == 0) {
        MG_TRY(resnorm(mg, phi, rhs, &res0, s));
    } else {
        MG_TRY(ir_residual(mg, phi, rhs, &res0, s));
    }
    if (hist) { hist[0] = res0; }
    int it = 0;
    bool ok = res0 == 0.0;
    for (; it < mg->opt.max_iter && !ok; it++) {
        MG_TRY(run_cycle(mg, phi, rhs, s));
        if (mg->opt.precision == 0) {
            MG_TRY(resnorm(mg, phi, rhs, &r, s));
        } else {
            MG_TRY(read_norm(mg, &r, s));
        }
        if (hist) { hist[it + 1] = r; }
        if (r < mg->opt.tol * res0) {
            ok = true;
        }
    }
    if (iters) { *iters = it; }
    // the ghost cells of phi as Proto's last resnorm leaves them (exchanged)
    MG_TRY(pm_level_exchange(mg->lv[0].layout, PM_F64, phi, mg->g1, 1, s));
    if (!ok) {
        set_error("pm_mg_solve: %d cycles left the residual at %.3e of the initial one (tol %.1e)", it,
                  res0 > 0 ? r / res0 : 0.0, mg->opt.tol);
        return PM_ERR_NOT_CONVERGED;
    }
    return PM_OK;
}

}  // extern "C"
