// LevelMultigrid: the correction-scheme V-cycle of
// examples/_archived/LevelMultigrid/src/LevelMultigrid.{H,cpp} (same
// constructor and vCycle / resnorm calls) solving Laplacian(phi)/dx^2 = rhs on a
// periodic LevelBoxData, with GSRB (examples/_archived/AMRFAS_old/src/
// LaplaceOp.H:95-123) or damped point-Jacobi relaxation, optionally as an
// FP32 V-cycle inside FP64 iterative refinement.  The hierarchy is the
// resident pm_mg of the C ABI.
#pragma once

#include "Proto_Level.H"

namespace Proto {

struct MultigridOptions : pm_mg_options {
    MultigridOptions() { pm_mg_default_options(this); }
    static MultigridOptions GSRB(int pre = 2, int post = 2, int bottom = 10)
    {
        MultigridOptions o;
        o.relax = 0;
        o.pre = pre;
        o.post = post;
        o.bottom = bottom;
        return o;
    }
    static MultigridOptions Jacobi()
    {
        MultigridOptions o;
        o.relax = 1;
        return o;
    }
    MultigridOptions& mixedPrecision(bool on = true) { precision = on ? 1 : 0; return *this; }
    MultigridOptions& linearInterp(bool on = true) { interp = on ? 1 : 0; return *this; }
    MultigridOptions& graph(bool on = true) { use_graph = on ? 1 : 0; return *this; }
};

class LevelMultigrid {
public:
    // LevelMultigrid(layout, dx, level): `level` coarsenings below `layout`
    LevelMultigrid(const DisjointBoxLayout& layout, double dx, int level,
                   const MultigridOptions& opt = MultigridOptions::GSRB())
        : m_layout(layout), m_opt(opt), m_dx(dx), m_level(level)
    {
        pm_mg* mg = nullptr;
        pm_check(pm_mg_create(&mg, layout.handle(), level, dx, &m_opt), "LevelMultigrid");
        m_mg.reset(mg, [](pm_mg* p) { pm_mg_destroy(p); });
    }
    void vCycle(LevelBoxData<double>& phi, LevelBoxData<double>& rhs)
    {
        check(phi, rhs);
        pm_check(pm_mg_vcycle(m_mg.get(), phi.data(), rhs.data(), protoCurrentStream()), "LevelMultigrid::vCycle");
    }
    double resnorm(LevelBoxData<double>& phi, LevelBoxData<double>& rhs)
    {
        check(phi, rhs);
        double r = 0;
        pm_check(pm_mg_resnorm(m_mg.get(), phi.data(), rhs.data(), &r, protoCurrentStream()), "LevelMultigrid::resnorm");
        return r;
    }
    // the solve loop of LevelMultigrid/exec/main.cpp:84-103; returns the cycle
    // count, throws ProtoError(PM_ERR_NOT_CONVERGED) if tol is not reached
    int solve(LevelBoxData<double>& phi, LevelBoxData<double>& rhs, int maxIter, double tol,
              std::vector<double>* hist = nullptr)
    {
        check(phi, rhs);
        m_opt.max_iter = maxIter;
        m_opt.tol = tol;
        pm_check(pm_mg_set_tolerance(m_mg.get(), tol, maxIter), "LevelMultigrid::solve");
        std::vector<double> h(maxIter + 1, 0.0);
        int iters = 0;
        const int rc = pm_mg_solve(m_mg.get(), phi.data(), rhs.data(), h.data(), &iters, protoCurrentStream());
        h.resize(iters + 1);
        if (hist) { *hist = h; }
        pm_check(rc, "LevelMultigrid::solve");
        return iters;
    }
    size_t bytes() const
    {
        size_t b = 0;
        pm_mg_bytes(m_mg.get(), &b);
        return b;
    }

private:
    void check(const LevelBoxData<double>& phi, const LevelBoxData<double>& rhs) const
    {
        if (!phi.layout().compatible(m_layout) || !rhs.layout().compatible(m_layout) || phi.ghost() != Point::Ones() ||
            rhs.ghost() != Point::Zeros()) {
            throw ProtoError(PM_ERR_INVALID, "LevelMultigrid: phi needs ghost 1 and rhs ghost 0 on the solver's layout");
        }
    }

    DisjointBoxLayout m_layout;
    MultigridOptions m_opt;
    double m_dx;
    int m_level;
    std::shared_ptr<pm_mg> m_mg;
};

}  // namespace Proto
