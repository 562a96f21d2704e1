// LevelOp<OpType, T, BCType, MEM>: a user BoxOp applied on every box of a
// level (include/base/Proto_LevelOp.H, implem/Proto_LevelOpImplem.H), and the
// level boundary conditions it carries (Proto_LevelBC.H, Proto_LevelBCLib.H:
// PeriodicBC, whose apply is the level exchange).
#pragma once

#include <memory>
#include <vector>

#include "Proto_BoxOp.H"

namespace Proto {

template<typename T, unsigned int C, MemType MEM, Centering CTR>
class LevelBC {
public:
    LevelBC() {}
    explicit LevelBC(const DisjointBoxLayout& layout) { define(layout); }
    virtual ~LevelBC() {}
    virtual void define(const DisjointBoxLayout& layout) { m_layout = layout; }
    void setTime(T t) { m_time = t; }
    T time() const { return m_time; }
    const DisjointBoxLayout& layout() const { return m_layout; }
    virtual void apply(LevelBoxData<T, C, MEM, CTR>&) const {}

protected:
    DisjointBoxLayout m_layout;
    T m_time = 0;
};

// Proto_LevelBCLib.H:12-18: fill the ghosts from the periodic images and
// neighbouring boxes
template<typename T, unsigned int C, MemType MEM, Centering CTR>
class PeriodicBC : public LevelBC<T, C, MEM, CTR> {
public:
    using LevelBC<T, C, MEM, CTR>::LevelBC;
    void apply(LevelBoxData<T, C, MEM, CTR>& data) const override { data.exchange(); }
};

template<template<typename, MemType> class OpType, typename T,
         template<typename, unsigned int, MemType, Centering> class BCType = PeriodicBC, MemType MEM = MEMTYPE_DEFAULT>
class LevelOp {
public:
    typedef OpType<T, MEM> OP;
    typedef BCType<T, OP::numState(), MEM, PR_CELL> BC;
    typedef BoxData<T, OP::numState(), MEM> StateData;
    typedef BoxData<T, OP::numAux(), MEM> AuxData;
    typedef LevelBoxData<T, OP::numState(), MEM, PR_CELL> LevelStateData;
    typedef LevelBoxData<T, OP::numAux(), MEM, PR_CELL> LevelAuxData;

    static constexpr unsigned int numState() { return OP::numState(); }
    static constexpr unsigned int numAux() { return OP::numAux(); }
    static Point ghost() { return OP::ghost(); }
    static Point auxGhost() { return auxGhostOf<OP>(0); }
    static int order() { return OP::order(); }

    LevelOp() {}
    LevelOp(DisjointBoxLayout& layout, T dx) { define(layout, dx); }
    LevelOp(DisjointBoxLayout& layout, Array<T, DIM> dx) { define(layout, dx); }
    LevelOp(LevelOp&&) = default;
    LevelOp& operator=(LevelOp&&) = default;

    void define(DisjointBoxLayout& layout, T dx)
    {
        Array<T, DIM> d;
        d.fill(dx);
        define(layout, d);
    }
    // one operator per box, each constructed on (layout, index, dx)
    // (Proto_LevelOpImplem.H:36-56; the constructor's define calls init())
    void define(DisjointBoxLayout& layout, Array<T, DIM> dx)
    {
        m_layout = layout;
        m_dx = dx;
        m_bc.define(layout);
        m_ops.clear();
        for (auto it : layout) { m_ops.push_back(std::make_shared<OP>(layout, it, dx)); }
    }

    T spectralRadius() const { return (*m_ops[0]).spectralRadius(); }

    // Proto_LevelOpImplem.H:58-83
    void operator()(LevelStateData& out, const LevelStateData& state, const LevelAuxData& aux, T scale = 1.0) const
    {
        require(out.layout().compatible(m_layout) && state.layout().compatible(m_layout) &&
                aux.layout().compatible(m_layout));
        for (auto it : m_layout) { (*m_ops[it])(out[it], state[it], aux[it], scale); }
    }
    // Proto_LevelOpImplem.H:85-109
    void operator()(LevelStateData& out, const LevelStateData& state, T scale = 1.0) const
    {
        require(out.layout().compatible(m_layout) && state.layout().compatible(m_layout));
        for (auto it : m_layout) {
            auto& out_i = out[it];
            Array<StateData, DIM> fluxes;
            for (int dir = 0; dir < DIM; dir++) { fluxes[dir].define(out_i.box().grow(dir, Side::Hi, 1)); }
            (*m_ops[it])(out_i, fluxes, state[it], scale);
        }
    }
    void applyBC(LevelStateData& state) const { m_bc.apply(state); }

    void setDiagScale(T v)
    {
        m_diagScale = v;
        for (auto it : m_layout) { (*m_ops[it]).setDiagScale(v); }
    }
    void setFluxScale(T v)
    {
        m_fluxScale = v;
        for (auto it : m_layout) { (*m_ops[it]).setFluxScale(v); }
    }
    void setTime(T t)
    {
        m_time = t;
        m_bc.setTime(t);
        for (auto it : m_layout) { (*m_ops[it]).setTime(t); }
    }
    void setRKStage(unsigned int s)
    {
        m_rkStage = s;
        for (auto it : m_layout) { (*m_ops[it]).setRKStage(s); }
    }
    T diagScale() const { return m_diagScale; }
    T fluxScale() const { return m_fluxScale; }
    T time() const { return m_time; }
    unsigned int RKStage() const { return m_rkStage; }
    Array<T, DIM> dx() const { return m_dx; }
    BC& bc() { return m_bc; }
    const DisjointBoxLayout& layout() const { return m_layout; }
    const OP& operator[](const LevelIndex& i) const { return *m_ops[i]; }
    OP& operator[](const LevelIndex& i) { return *m_ops[i]; }

private:
    template<typename O> static auto auxGhostOf(int) -> decltype(O::auxGhost())
    {
        return numAux() > 0 ? O::auxGhost() : Point::Zeros();
    }
    template<typename O> static Point auxGhostOf(long) { return Point::Zeros(); }
    static void require(bool ok)
    {
        if (!ok) { throw ProtoError(PM_ERR_INVALID, "LevelOp::operator(): data on an incompatible layout"); }
    }

    T m_diagScale = 0;
    T m_fluxScale = 1;
    T m_time = 0;
    unsigned int m_rkStage = 0;
    Array<T, DIM> m_dx;
    DisjointBoxLayout m_layout;
    std::vector<std::shared_ptr<OP>> m_ops;
    BC m_bc;
};

}  // namespace Proto
