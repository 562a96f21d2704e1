// BoxOp<T, C_STATE, C_AUX, MEM>: Proto's box-scope operator
// (include/base/Proto_BoxOp.H, implem/Proto_BoxOpImplem.H)
//
//   L(phi, rho) = d * D(phi, rho) + f * [Div F(phi, rho) + S(phi, rho)]
//
// with user-defined flux / source / diag and the scalings d (setDiagScale)
// and f (setFluxScale).  A user operator derives from BoxOp and overrides the
// pieces it has, exactly as with Proto (examples/_common/BoxOp_Laplace.H
// compiles unchanged).  The composition runs on the device: every step is a
// pm_stencil_apply or a pm_boxdata_* launch, in Proto's order (output zeroed,
// flux divergence accumulated direction by direction, then source, flux
// scale, diagonal, overall scale), so the results are Proto's bits.
#pragma once

#include "Proto_Level.H"

namespace Proto {

using LevelIndex = DataIndex;

template<typename T, unsigned int C_STATE, unsigned int C_AUX, MemType MEM = MEMTYPE_DEFAULT>
class BoxOp {
public:
    typedef BoxData<T, C_STATE, MEM> StateData;
    typedef BoxData<T, C_AUX, MEM> AuxData;

    static constexpr unsigned int numState() { return C_STATE; }
    static constexpr unsigned int numAux() { return C_AUX; }

    BoxOp() {}
    BoxOp(const DisjointBoxLayout& layout, const LevelIndex& index, T dx) { define(layout, index, dx); }
    BoxOp(const DisjointBoxLayout& layout, const LevelIndex& index, Array<T, DIM> dx) { define(layout, index, dx); }
    BoxOp(BoxOp&&) = default;
    BoxOp& operator=(BoxOp&&) = default;
    virtual ~BoxOp() {}

    void define(const DisjointBoxLayout& layout, const LevelIndex& index, T dx)
    {
        Array<T, DIM> d;
        d.fill(dx);
        define(layout, index, d);
    }
    // BoxOp::define (Proto_BoxOpImplem.H:46-66): d = 0, f = 1, then init()
    void define(const DisjointBoxLayout& layout, const LevelIndex& index, Array<T, DIM> dx)
    {
        m_scaleDiag = 0;
        m_scaleFlux = 1;
        m_dx = dx;
        m_layout = layout;
        m_index = index;
        m_time = 0;
        m_definedFlux = true;
        m_definedSrce = true;
        init();
    }

    virtual void init() {}
    virtual T spectralRadius() const { return 1; }

    // ------------------------------------------------------------- application
    // with aux (Proto_BoxOpImplem.H:68-128): the first direction's divergence
    // replaces, the others accumulate
    virtual void operator()(StateData& out, Array<StateData, DIM>& fluxes, const StateData& state, const AuxData& aux,
                            T scale = 1.0) const
    {
        out.setVal(0);
        if (m_scaleFlux != 0) {
            StateData divFlux(out.box());
            for (int dir = 0; dir < DIM; dir++) {
                const Stencil<T> DIV = divergence(dir);
                auto& flux_d = fluxes[dir];
                flux(flux_d, state, aux, dir);
                if (!m_definedFlux) { break; }
                if (dir == 0) {
                    divFlux |= DIV(flux_d, (T)(1.0 / m_dx[dir]));
                } else {
                    divFlux += DIV(flux_d, (T)(1.0 / m_dx[dir]));
                }
            }
            if (m_definedFlux) { out += divFlux; }
            StateData sourceData(out.box());
            source(sourceData, state, aux);
            if (m_definedSrce) { out += sourceData; }
            if (m_definedFlux || m_definedSrce) { out *= m_scaleFlux; }
        }
        if (m_scaleDiag != 0) {
            StateData diagData(out.box());
            diag(diagData, state, aux);
            diagData *= m_scaleDiag;
            out += diagData;
        }
        if (scale != (T)1.0) { out *= scale; }
    }
    // without aux (130-176): divergence accumulated into a zeroed array
    virtual void operator()(StateData& out, Array<StateData, DIM>& fluxes, const StateData& state,
                            T scale = 1.0) const
    {
        out.setVal(0);
        if (m_scaleFlux != 0) {
            StateData divFlux(out.box());
            divFlux.setVal(0);
            for (int dir = 0; dir < DIM; dir++) {
                const Stencil<T> DIV = divergence(dir);
                auto& flux_d = fluxes[dir];
                flux(flux_d, state, dir);
                if (!m_definedFlux) { break; }
                divFlux += DIV(flux_d, (T)(1.0 / m_dx[dir]));
            }
            if (m_definedFlux) { out += divFlux; }
            StateData sourceData(out.box());
            source(sourceData, state);
            if (m_definedSrce) { out += sourceData; }
            if (m_definedFlux || m_definedSrce) { out *= m_scaleFlux; }
        }
        if (m_scaleDiag != 0) {
            StateData diagData(out.box());
            diag(diagData, state);
            diagData *= m_scaleDiag;
            out += diagData;
        }
        if (scale != (T)1.0) { out *= scale; }
    }
    // fluxes on the output box grown by one face in each direction (178-210)
    virtual void operator()(StateData& out, const StateData& state, const AuxData& aux, T scale = 1.0) const
    {
        Array<StateData, DIM> fluxes;
        for (int dir = 0; dir < DIM; dir++) { fluxes[dir].define(out.box().grow(dir, Side::Hi, 1)); }
        (*this)(out, fluxes, state, aux, scale);
    }
    virtual void operator()(StateData& out, const StateData& state, T scale = 1.0) const
    {
        Array<StateData, DIM> fluxes;
        for (int dir = 0; dir < DIM; dir++) { fluxes[dir].define(out.box().grow(dir, Side::Hi, 1)); }
        (*this)(out, fluxes, state, scale);
    }
    // out of place (212-240)
    virtual StateData operator()(const StateData& state, const AuxData& aux, Box range, T scale = 1.0) const
    {
        StateData out(range);
        (*this)(out, state, aux, scale);
        return out;
    }
    virtual StateData operator()(const StateData& state, Box range, T scale = 1.0) const
    {
        StateData out(range);
        (*this)(out, state, scale);
        return out;
    }

    // ------------------------------------------------------------- boundary conditions
    // the user's condition on every non-periodic face of the domain (242-256)
    void applyBC(Array<StateData, DIM>& fluxes, const StateData& state) const
    {
        for (int d = 0; d < DIM; d++) {
            if (layout().domain().isPeriodic(d)) { continue; }
            this->applyBC(fluxes, state, Face(d, Side::Lo));
            this->applyBC(fluxes, state, Face(d, Side::Hi));
        }
    }
    virtual void applyBC(Array<StateData, DIM>&, const StateData&, Face) const {}

    // ------------------------------------------------------------- user pieces
    // The defaults mark the piece as absent (flux, source) or return the state
    // (diag), as Proto_BoxOpImplem.H:259-345 does.
    virtual void flux(StateData&, const StateData&, const AuxData&, int) const { m_definedFlux = false; }
    virtual void flux(StateData&, const StateData&, int) const { m_definedFlux = false; }
    virtual void source(StateData&, const StateData&, const AuxData&) const { m_definedSrce = false; }
    virtual void source(StateData&, const StateData&) const { m_definedSrce = false; }
    virtual void diag(StateData& d, const StateData& state, const AuxData&) const { state.copyTo(d); }
    virtual void diag(StateData& d, const StateData& state) const { state.copyTo(d); }

    // ------------------------------------------------------------- parameters
    void setDiagScale(T v) { m_scaleDiag = v; }
    void setFluxScale(T v) { m_scaleFlux = v; }
    void setTime(T t) { m_time = t; }
    void setRKStage(unsigned int s) { m_RKStage = s; }
    T diagScale() const { return m_scaleDiag; }
    T fluxScale() const { return m_scaleFlux; }
    unsigned int RKStage() const { return m_RKStage; }
    T time() const { return m_time; }
    Array<T, DIM> dx() const { return m_dx; }
    T dxMin() const
    {
        T m = m_dx[0];
        for (int d = 1; d < DIM; d++) { m = std::min(m, m_dx[d]); }
        return m;
    }
    T dxMax() const
    {
        T m = m_dx[0];
        for (int d = 1; d < DIM; d++) { m = std::max(m, m_dx[d]); }
        return m;
    }
    const LevelIndex& index() const { return m_index; }
    const DisjointBoxLayout& layout() const { return m_layout; }
    Box box() const { return m_layout[m_index]; }

protected:
    T m_scaleDiag = 0;
    T m_scaleFlux = 1;
    T m_time = 0;
    unsigned int m_RKStage = 0;
    LevelIndex m_index = 0;
    DisjointBoxLayout m_layout;
    mutable bool m_definedSrce = true;
    mutable bool m_definedFlux = true;
    Array<T, DIM> m_dx;

private:
    // the face-to-cell divergence e_dir - 0 of Proto_BoxOpImplem.H:87-88
    static Stencil<T> divergence(int dir)
    {
        return ((T)1.0) * Shift::Basis(dir, 1) - ((T)1.0) * Shift::Zeros();
    }
};

}  // namespace Proto
