// LevelMultigridSolver<OpType, T, MEM> (include/utils/Proto_LevelMultigridSolver.H,
// implem/Proto_LevelMultigridSolverImplem.H): Proto's FAS multigrid on a
// periodic level for any BoxOp operator, with Jacobi relaxation
// (lambda = 0.5 / spectralRadius), 2 DIM pre / post and 10 bottom sweeps,
// average-down restriction and constant prolongation.
//
// Two paths, the same results:
//  - generic: Proto's MGLevel recursion written over this library's LevelOp,
//    LevelBoxData, Stencil and InterpStencil, one device launch per box and
//    step, in Proto's order -- for every operator;
//  - resident: the library's own FAS solver (pm_fas, one launch per level
//    and operator, hierarchy resident on the device), taken when the user's
//    operator is the 4th-order flux-form Laplacian.  The solver recognises it
//    when it is defined: the operator's spectral radius must equal
//    2 DIM / dx^2 on every level, and the operator applied to a random state
//    on the first box must equal the resident kernel's result bit for bit.
//    Anything else (another operator, aux variables, FP32, a layout the
//    resident solver does not take) runs the generic path.
#pragma once

#include <memory>
#include <type_traits>
#include <vector>

#include "Proto_LevelOp.H"

namespace Proto {

template<template<typename, MemType> class OpType, typename T, MemType MEM = MEMTYPE_DEFAULT>
class LevelMultigridSolver {
public:
    typedef OpType<T, MEM> BOP;
    typedef LevelOp<OpType, T, PeriodicBC, MEM> LOP;
    typedef LevelBoxData<T, BOP::numState(), MEM> LevelStateData;
    typedef LevelBoxData<T, BOP::numAux(), MEM> LevelAuxData;

    LevelMultigridSolver(DisjointBoxLayout& layout, Point refRatio, int numLevels, T dx)
    {
        define(layout, refRatio, numLevels, dx);
    }
    LevelMultigridSolver(DisjointBoxLayout& layout, Point refRatio, int numLevels, Array<T, DIM> dx)
    {
        define(layout, refRatio, numLevels, dx);
    }
    void define(DisjointBoxLayout& layout, Point refRatio, int numLevels, T dx)
    {
        Array<T, DIM> d;
        d.fill(dx);
        define(layout, refRatio, numLevels, d);
    }
    // LevelMultigridSolverImplem.H:24-36
    void define(DisjointBoxLayout& layout, Point refRatio, int numLevels, Array<T, DIM> dx)
    {
        if (numLevels < 1) { throw ProtoError(PM_ERR_INVALID, "LevelMultigridSolver: numLevels must be >= 1"); }
        m_verbose = false;
        m_layout = layout;
        m_residual.define(layout, Point::Zeros());
        m_mg = std::make_shared<MGLevel>(layout, refRatio, dx, numLevels - 1);
        m_fas.reset();
        if (m_allowResident) { tryResident(layout, refRatio, numLevels, dx); }
    }

    // solve with auxiliary data: not implemented in Proto either (66-77)
    T solve(LevelStateData&, LevelStateData&, LevelAuxData&, int, T) { return 0; }

    // solve (79-132): V-cycles until resnorm < tolerance * initial resnorm or
    // maxIter cycles; returns the last resnorm
    T solve(LevelStateData& state, LevelStateData& force, int maxIter, T tolerance)
    {
        checkState(state, force);
        if (m_fas) {
            std::vector<double> h(maxIter + 1, 0.0);
            int iters = 0;
            pm_check(pm_fas_solve(m_fas.get(), state.data(), force.data(), maxIter, tolerance, h.data(), &iters,
                                  protoCurrentStream()),
                     "LevelMultigridSolver::solve");
            m_hist.assign(h.begin(), h.begin() + iters + 1);
            // m_residual holds the last residual, as Proto's last resnorm leaves it
            double r = 0;
            pm_check(pm_fas_residual(m_fas.get(), state.data(), force.data(), m_residual.data(), &r,
                                     protoCurrentStream()),
                     "LevelMultigridSolver::solve");
            return (T)h[iters];
        }
        T res0 = resnorm(m_residual, state, force);
        m_hist.assign(1, (double)res0);
        T res = res0;
        for (int ii = 0; ii < maxIter; ii++) {
            m_mg->vCycle(state, force);
            res = resnorm(m_residual, state, force);
            m_hist.push_back((double)res);
            if (res < tolerance * res0) { break; }
        }
        return res;
    }

    // resnorm (134-160): residual = L(state) + force on every box, its max norm
    T resnorm(LevelStateData& residual, LevelStateData& state, LevelStateData& force)
    {
        checkState(state, force);
        if (!residual.layout().compatible(m_layout)) {
            throw ProtoError(PM_ERR_INVALID, "LevelMultigridSolver::resnorm: residual on another layout");
        }
        if (m_fas && residual.ghost() == Point::Zeros()) {
            double r = 0;
            pm_check(pm_fas_residual(m_fas.get(), state.data(), force.data(), residual.data(), &r,
                                     protoCurrentStream()),
                     "LevelMultigridSolver::resnorm");
            return (T)r;
        }
        state.exchange();
        Reduction<T, Abs, MEM> rxn;
        for (auto it : state.layout()) {
            auto& res = residual[it];
            m_mg->op()[it](res, state[it]);
            res += force[it];
            res.reduce(rxn);
        }
        return rxn.fetch();
    }

    // one V-cycle of the finest level (MGLevel::vCycle, 408-433)
    void vCycle(LevelStateData& state, LevelStateData& force)
    {
        checkState(state, force);
        if (m_fas) {
            pm_check(pm_fas_vcycle(m_fas.get(), state.data(), force.data(), protoCurrentStream()),
                     "LevelMultigridSolver::vCycle");
            return;
        }
        m_mg->vCycle(state, force);
    }

    LevelStateData& res() { return m_residual; }
    // Proto's interpBC interpolates with an InterpStencil its define never sets
    // (the code that would set it is commented out, 178-188), so it is not offered.
    void interpBC(LevelStateData&)
    {
        throw ProtoError(PM_ERR_UNSUPPORTED, "LevelMultigridSolver::interpBC: no boundary interpolation is defined");
    }
    void setVerbose(bool flag) { m_verbose = flag; }

    // Not in Proto: residual norms of the last solve (initial first), whether
    // the resident path is in use, and a switch that keeps every later define
    // on the generic path (for comparing the two).
    const std::vector<double>& history() const { return m_hist; }
    bool resident() const { return (bool)m_fas; }
    void allowResident(bool allow)
    {
        m_allowResident = allow;
        if (!allow) { m_fas.reset(); }
    }

private:
    // local class implementing the recursion (Proto_LevelMultigridSolver.H:72-152)
    class MGLevel {
    public:
        MGLevel(DisjointBoxLayout& layout, Point refRatio, Array<T, DIM> dx, int level)
        {
            define(layout, refRatio, dx, level);
        }
        // MGLevel::define (200-244)
        void define(DisjointBoxLayout& layout, Point refRatio, Array<T, DIM> dx, int level)
        {
            m_level = level;
            m_refRatio = refRatio;
            m_numPreRelax = 2 * DIM;
            m_numPostRelax = 2 * DIM;
            m_numBottomRelax = 10;
            m_levelOp.define(layout, dx);
            m_levelOp.setFluxScale(-1);
            m_lambda = 0.5 / m_levelOp.spectralRadius();
            m_average = Stencil<T>::AvgDown(refRatio);
            m_interpolate = InterpStencil<T>::Constant(refRatio);
            m_increment = (-m_lambda) * Shift::Zeros();
            m_layout = layout;
            if (level > 0) {
                m_crseLocalLayout = layout.coarsen(refRatio);
                ProblemDomain crseDomain = layout.domain().coarsen(refRatio);
                Point boxSize = layout.boxSize();
                if (crseDomain.sizes() % boxSize == Point::Zeros()) {
                    m_crseLayout.define(crseDomain, boxSize);
                } else {
                    m_crseLayout.define(crseDomain, crseDomain.sizes());
                }
                Array<T, DIM> cdx;
                for (int d = 0; d < DIM; d++) { cdx[d] = dx[d] * refRatio[d]; }
                m_crseMG = std::make_shared<MGLevel>(m_crseLayout, refRatio, cdx, level - 1);
            }
        }
        // The level's arrays (244-259), allocated when the first cycle needs them:
        // a solver that takes the resident path never holds them.
        void allocate()
        {
            if (m_allocated) { return; }
            m_residual.define(m_layout, Point::Zeros());
            if (m_level > 0) {
                m_crseLocal.define(m_crseLocalLayout, Point::Zeros());
                m_crseForce.define(m_crseLayout, Point::Zeros());
                m_crseState.define(m_crseLayout, LOP::ghost());
                m_crseState_0.define(m_crseLayout, Point::Zeros());
                if (BOP::numAux() > 0) { m_crseAux.define(m_crseLayout, LOP::auxGhost()); }
            }
            m_allocated = true;
        }
        // coarseForce (262-290): the restricted fine residual minus the coarse operator
        void coarseForce(LevelStateData& crseForce, LevelStateData& state, LevelStateData& force,
                         LevelStateData& crseState)
        {
            state.exchange();
            crseState.exchange();
            for (auto it : m_levelOp.layout()) {
                auto& resC_i = m_crseLocal[it];
                auto res = m_levelOp[it](state[it], state.layout()[it]);
                res += force[it];
                resC_i |= m_average(res);
            }
            m_crseLocal.copyTo(crseForce);
            for (auto it : crseForce.layout()) {
                auto lphi = m_crseMG->m_levelOp[it](crseState[it], crseForce.layout()[it]);
                crseForce[it] -= lphi;
            }
        }
        // relax (311-331): Jacobi, state += -lambda (L(state) + force)
        void relax(LevelStateData& state, LevelStateData& force, int numIter)
        {
            for (int ii = 0; ii < numIter; ii++) {
                state.exchange();
                for (auto it : m_levelOp.layout()) {
                    auto& state_i = state[it];
                    auto res = m_levelOp[it](state_i, state.layout()[it]);
                    res += force[it];
                    state_i += m_increment(res);
                }
            }
        }
        // fineCorrect (333-356): fine += I(crse - crse0)
        void fineCorrect(LevelStateData& fine, LevelStateData& crse, LevelStateData& crse0)
        {
            for (auto it = crse.begin(); it.ok(); ++it) {
                auto& crse0_i = crse0[*it];
                crse0_i -= crse[*it];
                crse0_i *= -1;
            }
            crse0.copyTo(m_crseLocal);
            for (auto it = fine.begin(); it.ok(); ++it) { fine[*it] += m_interpolate(m_crseLocal[*it]); }
        }
        // vCycle (358-383)
        void vCycle(LevelStateData& state, LevelStateData& force)
        {
            allocate();
            if (m_level > 0) {
                relax(state, force, m_numPreRelax);
                averageDown(m_crseState, state, m_refRatio);
                m_crseState.copyTo(m_crseState_0);
                coarseForce(m_crseForce, state, force, m_crseState);
                m_crseMG->vCycle(m_crseState, m_crseForce);
                fineCorrect(state, m_crseState, m_crseState_0);
                relax(state, force, m_numPostRelax);
            } else {
                relax(state, force, m_numBottomRelax);
            }
        }
        LOP& op() { return m_levelOp; }
        T lambda() const { return m_lambda; }
        int level() const { return m_level; }
        MGLevel* coarser() { return m_crseMG.get(); }

    private:
        DisjointBoxLayout m_layout, m_crseLayout, m_crseLocalLayout;
        bool m_allocated = false;
        std::shared_ptr<MGLevel> m_crseMG;
        LOP m_levelOp;
        int m_level = 0;
        Point m_refRatio;
        T m_lambda = 0;
        int m_numPreRelax = 0, m_numPostRelax = 0, m_numBottomRelax = 0;
        LevelStateData m_crseState_0, m_crseState, m_crseForce, m_residual;
        LevelAuxData m_crseAux;
        LevelStateData m_crseLocal;
        Stencil<T> m_average;
        Stencil<T> m_increment;
        InterpStencil<T> m_interpolate;
    };

    void checkState(const LevelStateData& state, const LevelStateData& force) const
    {
        if (!state.layout().compatible(m_layout) || !force.layout().compatible(m_layout)) {
            throw ProtoError(PM_ERR_INVALID, "LevelMultigridSolver: state and force must be on the solver's layout");
        }
        if (state.ghost() != LOP::ghost() || force.ghost() != Point::Zeros()) {
            throw ProtoError(PM_ERR_INVALID, "LevelMultigridSolver: state needs the operator's ghost() layers, force none");
        }
    }

    // The resident path: the operator must be the one pm_fas applies.
    void tryResident(DisjointBoxLayout& layout, Point refRatio, int numLevels, const Array<T, DIM>& dx)
    {
        if (!std::is_same<T, double>::value || DIM != 3 || BOP::numState() != 1 || BOP::numAux() != 0 ||
            LOP::ghost() != Point::Ones(2) || refRatio != Point::Ones(2)) {
            return;
        }
        for (int d = 1; d < DIM; d++) {
            if (dx[d] != dx[0]) { return; }
        }
        // Jacobi weight of every level: lambda = 0.5 / (2 DIM / dx^2), as pm_fas uses it
        double h = (double)dx[0];
        for (MGLevel* L = m_mg.get(); L; L = L->coarser()) {
            const double want = 0.5 / ((2.0 * DIM) / std::pow(h, 2));
            if ((double)L->lambda() != want) { return; }
            h *= 2;
        }
        pm_fas* f = nullptr;
        if (pm_fas_create(&f, layout.handle(), numLevels, (double)dx[0]) != PM_OK) { return; }
        std::shared_ptr<pm_fas> fas(f, [](pm_fas* p) { pm_fas_destroy(p); });
        // the operator on a random state: resident kernel vs the user's BoxOp on box 0
        LevelStateData state(layout, LOP::ghost()), force(layout, Point::Zeros()), res(layout, Point::Zeros());
        {
            int g[3] = {0, 0, 0};
            for (int d = 0; d < DIM; d++) { g[d] = LOP::ghost()[d]; }
            size_t elems = 0;
            pm_check(pm_layout_box_elems(layout.handle(), g, &elems), "LevelMultigridSolver probe");
            pm_check(pm_fill_uniform(PM_F64, state.data(), elems * layout.size(), 0x5eed, 7, 1.0, 0.0,
                                     protoCurrentStream()),
                     "LevelMultigridSolver probe");
        }
        state.exchange();
        force.setVal(0);
        double r = 0;
        pm_check(pm_fas_residual(fas.get(), state.data(), force.data(), res.data(), &r, protoCurrentStream()),
                 "LevelMultigridSolver probe");
        auto user = m_mg->op()[0](state[0], layout[0]);
        const size_t n = (size_t)layout[0].size();
        std::vector<T> a(n), b(n);
        proto_memcpy<DEVICE, HOST>(res[0].data(), a.data(), sizeof(T) * n);
        proto_memcpy<DEVICE, HOST>(user.data(), b.data(), sizeof(T) * n);
        if (std::memcmp(a.data(), b.data(), sizeof(T) * n) != 0) { return; }
        m_fas = fas;
    }

    bool m_verbose = false;
    bool m_allowResident = true;
    DisjointBoxLayout m_layout;
    LevelStateData m_residual;
    std::shared_ptr<MGLevel> m_mg;
    std::shared_ptr<pm_fas> m_fas;
    std::vector<double> m_hist;
};

}  // namespace Proto
