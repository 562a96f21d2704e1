// LevelMultigridSolver<OpType, T, MEM> (include/utils/Proto_LevelMultigridSolver.H,
// implem/Proto_LevelMultigridSolverImplem.H) with OpType = BoxOp_Laplace
// (examples/_common/BoxOp_Laplace.H): Proto's library FAS solver on a periodic
// level, the resident pm_fas of the C ABI.  Same constructor and solve /
// resnorm calls; state needs BoxOp_Laplace::ghost() = 2 ghost layers.
#pragma once

#include <type_traits>

#include "Proto_Level.H"

namespace Proto {

// The operator tag of the 4th-order flux-form Laplacian (its numerics live in
// the library's kernels).
template<typename T, MemType MEM = MEMTYPE_DEFAULT>
class BoxOp_Laplace {
public:
    static Point ghost() { return Point::Ones(2); }
    static Point auxGhost() { return Point::Zeros(); }
    static constexpr int order() { return 4; }
    static double spectralRadius(double dx) { return (2.0 * DIM) / (dx * dx); }
};

template<template<typename, MemType> class OpType, typename T, MemType MEM = MEMTYPE_DEFAULT>
class LevelMultigridSolver {
    static_assert(std::is_same<OpType<T, MEM>, BoxOp_Laplace<T, MEM>>::value,
                  "LevelMultigridSolver: this library provides the BoxOp_Laplace operator");
    static_assert(std::is_same<T, double>::value, "LevelMultigridSolver: double precision");

public:
    LevelMultigridSolver(DisjointBoxLayout& layout, Point refRatio, int numLevels, T dx) : m_layout(layout)
    {
        if (refRatio != Point::Ones(2)) {
            throw ProtoError(PM_ERR_UNSUPPORTED, "LevelMultigridSolver: refinement ratio 2 only");
        }
        pm_fas* f = nullptr;
        pm_check(pm_fas_create(&f, layout.handle(), numLevels, dx), "LevelMultigridSolver");
        m_fas.reset(f, [](pm_fas* p) { pm_fas_destroy(p); });
    }
    // solve (LevelMultigridSolverImplem.H:85-132): returns the last resnorm
    T solve(LevelBoxData<T, 1, MEM>& state, LevelBoxData<T, 1, MEM>& force, int maxIter, T tolerance)
    {
        check(state, force);
        std::vector<double> h(maxIter + 1, 0.0);
        int iters = 0;
        pm_check(pm_fas_solve(m_fas.get(), state.data(), force.data(), maxIter, tolerance, h.data(), &iters,
                              protoCurrentStream()),
                 "LevelMultigridSolver::solve");
        m_hist.assign(h.begin(), h.begin() + iters + 1);
        return h[iters];
    }
    // resnorm (134-160): residual = L(state) + force, returns its max norm
    T resnorm(LevelBoxData<T, 1, MEM>& residual, LevelBoxData<T, 1, MEM>& state, LevelBoxData<T, 1, MEM>& force)
    {
        check(state, force);
        if (!residual.layout().compatible(m_layout) || residual.ghost() != Point::Zeros()) {
            throw ProtoError(PM_ERR_INVALID, "LevelMultigridSolver::resnorm: residual needs no ghosts");
        }
        double r = 0;
        pm_check(pm_fas_residual(m_fas.get(), state.data(), force.data(), residual.data(), &r, protoCurrentStream()),
                 "LevelMultigridSolver::resnorm");
        return r;
    }
    void vCycle(LevelBoxData<T, 1, MEM>& state, LevelBoxData<T, 1, MEM>& force)
    {
        check(state, force);
        pm_check(pm_fas_vcycle(m_fas.get(), state.data(), force.data(), protoCurrentStream()),
                 "LevelMultigridSolver::vCycle");
    }
    void setVerbose(bool) {}
    // residual norms of the last solve, initial first
    const std::vector<double>& history() const { return m_hist; }

private:
    void check(const LevelBoxData<T, 1, MEM>& state, const LevelBoxData<T, 1, MEM>& force) const
    {
        if (!state.layout().compatible(m_layout) || !force.layout().compatible(m_layout) ||
            state.ghost() != Point::Ones(2) || force.ghost() != Point::Zeros()) {
            throw ProtoError(PM_ERR_INVALID,
                             "LevelMultigridSolver: state needs 2 ghost layers and force none, on the solver's layout");
        }
    }

    DisjointBoxLayout m_layout;
    std::shared_ptr<pm_fas> m_fas;
    std::vector<double> m_hist;
};

}  // namespace Proto
