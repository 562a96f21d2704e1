// InterpStencil1D / InterpStencil with Constant and Linear
// (include/base/Proto_InterpStencil.H, Proto_InterpStencil1D.H,
// implem/Proto_InterpStencilImplem.H:83-112, Proto_InterpStencilDefs.H:1-36,
// Proto_InterpStencil1DImplem.H:20-113): prolongation staged one direction at
// a time, each stage a set of destination-refined stencils, then
// output (|= or +=) scale*stage.  Every stage runs through pm_stencil_apply.
#pragma once

#include <array>
#include <cstdlib>

#include "Proto_Stencil.H"

namespace Proto {

template<typename T>
class InterpStencil1D {
public:
    void define(int dir, int ratio)
    {
        m_dir = dir;
        m_r = ratio;
        m_s.assign(ratio, Stencil<T>());
    }
    int ratio() const { return m_r; }
    const Stencil<T>& get(int shift) const { return m_s[shift]; }
    // union of the stencils' spans (Proto_InterpStencil1DImplem.H:36-46)
    Box span() const
    {
        Box b;
        for (auto& S : m_s) {
            b = b & S.span().low();
            b = b & S.span().high();
        }
        return b;
    }
    // per direction the largest |offset| of the span (48-57)
    Point ghost() const
    {
        const Box b = span();
        Point v;
        for (int d = 0; d < DIM; d++) { v[d] = std::max(std::abs(b.low()[d]), std::abs(b.high()[d])); }
        return v;
    }
    void set(Stencil<T> S, int shift)
    {
        Point r = Point::Ones();
        r[m_dir] = m_r;
        S.destRatio() = r;
        S.destShift() = Point::Basis(m_dir, shift);
        m_s[shift] = S;
    }
    // operator()(src, scale): one lazy stencil per shift, output range the
    // union of their ranges (Proto_InterpStencil1DImplem.H:90-113)
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    LazyStencil<T, C, MEM, D, E> operator()(const BoxData<T, C, MEM, D, E>& src, T scale = 1) const
    {
        LazyStencil<T, C, MEM, D, E> ret;
        const Box bsrc = src.box();
        for (size_t i = 0; i < m_s.size(); i++) {
            if (i == 0) { ret.m_range = m_s[i].range(bsrc); }
            else {
                ret.m_range = ret.m_range & m_s[i].range(bsrc).low();
                ret.m_range = ret.m_range & m_s[i].range(bsrc).high();
            }
            Box b(bsrc.low() - m_s[i].span().low(), bsrc.high() - m_s[i].span().high());
            b = b.taperCoarsen(m_s[i].srcRatio());
            ret.m_stencil.push_back(&m_s[i]);
            ret.m_src.push_back(&src);
            ret.m_box.push_back(b);
            ret.m_scale.push_back(scale);
        }
        return ret;
    }

private:
    int m_dir = 0, m_r = 1;
    std::vector<Stencil<T>> m_s;
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E> class LazyInterpStencil;

template<typename T>
class InterpStencil {
public:
    InterpStencil() {}
    explicit InterpStencil(Point ratio) { define(ratio); }
    void define(Point ratio)
    {
        m_ratio = ratio;
        for (int d = 0; d < DIM; d++) { m_interp[d].define(d, ratio[d]); }
    }
    static InterpStencil Constant(Point ratio)
    {
        InterpStencil r(ratio);
        auto S = ((T)1.0) * Shift::Zeros();
        for (int d = 0; d < DIM; d++)
            for (int s = 0; s < ratio[d]; s++) { r.m_interp[d].set(S, s); }
        return r;
    }
    static InterpStencil Linear(Point ratio)
    {
        InterpStencil r(ratio);
        for (int d = 0; d < DIM; d++)
            for (int s = 0; s < ratio[d]; s++) {
                const double rr = ratio[d], ss = s;
                const double p = ss / rr, q = (rr - ss) / rr;
                auto S = ((T)q) * Shift::Zeros() + ((T)p) * Shift::Basis(d, 1);
                r.m_interp[d].set(S, s);
            }
        return r;
    }
    // Quadratic (Proto_InterpStencilDefs.H:37-75): per direction and shift n,
    // coefficients qcoef(n, r, side) at offsets -1, 0, +1
    static double qcoef(int ni, int ri, int side)
    {
        const double n = ni, r = ri;
        double c = 0;
        if (side == -1) { c = 1.0 + 3.0 * n + 3.0 * n * n - 3.0 * r - 6.0 * n * r + 2.0 * r * r; }
        else if (side == 0) { c = -2.0 - 6.0 * n - 6.0 * n * n + 3.0 * r + 6.0 * n * r + 5.0 * r * r; }
        else if (side == 1) { c = -(-1.0 - 3.0 * n - 3.0 * n * n + r * r); }
        c /= (6.0 * r * r);
        return c;
    }
    static InterpStencil Quadratic(Point ratio)
    {
        InterpStencil r(ratio);
        for (int d = 0; d < DIM; d++)
            for (int s = 0; s < ratio[d]; s++) {
                Stencil<T> S;
                for (int ii = -1; ii <= 1; ii++) {
                    const T coef = (T)qcoef(s, ratio[d], ii);
                    S += coef * Shift::Basis(d, ii);
                }
                r.m_interp[d].set(S, s);
            }
        return r;
    }
    // FiniteVolume (Proto_InterpStencilDefs.H:77-214): 5-point tables at
    // offsets -2 .. +2 for ratio 2 (orders 2-5) and ratio 4 (orders 2-5)
    static InterpStencil FiniteVolume(Point ratio, int order)
    {
        InterpStencil r(ratio);
        for (int d = 0; d < DIM; d++) {
            const int rr = ratio[d];
            std::vector<std::array<T, 5>> c;
            if (rr == 2 && (order == 2 || order == 3)) {
                c = {{0, (T)(1.0 / 8.0), (T)1.0, (T)(-1.0 / 8.0), 0}, {0, (T)(-1.0 / 8.0), (T)1.0, (T)(1.0 / 8.0), 0}};
            } else if (rr == 2 && (order == 4 || order == 5)) {
                c = {{(T)(-3.0 / 128), (T)(22.0 / 128), (T)(128.0 / 128), (T)(-22.0 / 128), (T)(3.0 / 128)},
                     {(T)(3.0 / 128), (T)(-22.0 / 128), (T)(128.0 / 128), (T)(22.0 / 128), (T)(-3.0 / 128)}};
            } else if (rr == 4 && order == 2) {
                c = {{0, (T)(3.0 / 16.0), (T)1.0, (T)(-3.0 / 16.0), 0}, {0, (T)(1.0 / 16.0), (T)1.0, (T)(-1.0 / 16.0), 0},
                     {0, (T)(-1.0 / 16.0), (T)1.0, (T)(1.0 / 16.0), 0}, {0, (T)(-3.0 / 16.0), (T)1.0, (T)(3.0 / 16.0), 0}};
            } else if (rr == 4 && order == 3) {
                c = {{0, (T)(7.0 / 32.0), (T)(30.0 / 32.0), (T)(-5.0 / 32.0), 0},
                     {0, (T)(1.0 / 32.0), (T)(34.0 / 32.0), (T)(-3.0 / 32.0), 0},
                     {0, (T)(-3.0 / 32.0), (T)(34.0 / 32.0), (T)(1.0 / 32.0), 0},
                     {0, (T)(-5.0 / 32.0), (T)(30.0 / 32.0), (T)(7.0 / 32.0), 0}};
            } else if (rr == 4 && order == 4) {
                c = {{(T)(-35.0 / 1024), (T)(294.0 / 1024), (T)(960.0 / 1024), (T)(-230.0 / 1024), (T)(35.0 / 1024)},
                     {(T)(-13.0 / 1024), (T)(58.0 / 1024), (T)(1088.0 / 1024), (T)(-122.0 / 1024), (T)(13.0 / 1024)},
                     {(T)(13.0 / 1024), (T)(-122.0 / 1024), (T)(1088.0 / 1024), (T)(58.0 / 1024), (T)(-13.0 / 1024)},
                     {(T)(35.0 / 1024), (T)(-230.0 / 1024), (T)(960.0 / 1024), (T)(294.0 / 1024), (T)(-35.0 / 1024)}};
            } else if (rr == 4 && order == 5) {
                c = {{(T)(-77.0 / 2048), (T)(616.0 / 2048), (T)(1878.0 / 2048), (T)(-432.0 / 2048), (T)(63.0 / 2048)},
                     {(T)(-19.0 / 2048), (T)(88.0 / 2048), (T)(2218.0 / 2048), (T)(-272.0 / 2048), (T)(33.0 / 2048)},
                     {(T)(33.0 / 2048), (T)(-272.0 / 2048), (T)(2218.0 / 2048), (T)(88.0 / 2048), (T)(-19.0 / 2048)},
                     {(T)(63.0 / 2048), (T)(-432.0 / 2048), (T)(1878.0 / 2048), (T)(616.0 / 2048), (T)(-77.0 / 2048)}};
            } else {
                throw ProtoError(PM_ERR_UNSUPPORTED, "InterpStencil::FiniteVolume: ratio 2 or 4, order 2..5");
            }
            for (int s = 0; s < rr; s++) {
                Stencil<T> S;
                for (int ii = 0; ii < 5; ii++) { S += c[s][ii] * Shift::Basis(d, ii - 2); }
                r.m_interp[d].set(S, s);
            }
        }
        return r;
    }
    Box range(const Box& domain) const
    {
        Point L, H;
        for (int d = 0; d < DIM; d++)
            for (int s = 0; s < m_interp[d].ratio(); s++) {
                Box di = m_interp[d].get(s).range(domain);
                if (s == 0) { L[d] = di.low()[d]; H[d] = di.high()[d]; }
                else { L[d] = std::min(di.low()[d], L[d]); H[d] = std::max(di.high()[d], H[d]); }
            }
        return Box(L, H);
    }
    Point ratio() const { return m_ratio; }
    // the largest per-direction ghost of the 1D stencils
    // (Proto_InterpStencilImplem.H:56-68)
    Point ghost() const
    {
        Point g = m_interp[0].ghost();
        for (int j = 1; j < DIM; j++) {
            const Point h = m_interp[j].ghost();
            for (int d = 0; d < DIM; d++) { g[d] = std::max(h[d], g[d]); }
        }
        return g;
    }

    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    void apply(BoxData<T, C, MEM, D, E>& out, const BoxData<T, C, MEM, D, E>& in, Box range, bool overwrite,
               T scale = 1) const
    {
        Box rangeBox = range;
        if (rangeBox.empty()) { rangeBox = this->range(in.box()) & out.box(); }
        BoxData<T, C, MEM, D, E> stage[DIM];
        stage[0] = BoxData<T, C, MEM, D, E>(m_interp[0](in));
        for (int d = 0; d < DIM - 1; d++) { stage[d + 1] = BoxData<T, C, MEM, D, E>(m_interp[d + 1](stage[d])); }
        Stencil<T> S = scale * Shift::Zeros();
        if (overwrite) { out |= S(stage[DIM - 1], rangeBox); }
        else { out += S(stage[DIM - 1], rangeBox); }
    }
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    LazyInterpStencil<T, C, MEM, D, E> operator()(const BoxData<T, C, MEM, D, E>& in, T scale = 1) const
    {
        return LazyInterpStencil<T, C, MEM, D, E>(*this, in, Box(), scale);
    }
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    LazyInterpStencil<T, C, MEM, D, E> operator()(const BoxData<T, C, MEM, D, E>& in, Box range, T scale = 1) const
    {
        return LazyInterpStencil<T, C, MEM, D, E>(*this, in, range, scale);
    }

private:
    Point m_ratio;
    InterpStencil1D<T> m_interp[DIM];
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
class LazyInterpStencil {
public:
    LazyInterpStencil(const InterpStencil<T>& op, const BoxData<T, C, MEM, D, E>& src, Box range, T scale)
        : m_op(&op), m_src(&src), m_range(range), m_scale(scale)
    {
    }
    void apply(BoxData<T, C, MEM, D, E>& dst, bool overwrite) const { m_op->apply(dst, *m_src, m_range, overwrite, m_scale); }

private:
    const InterpStencil<T>* m_op;
    const BoxData<T, C, MEM, D, E>* m_src;
    Box m_range;
    T m_scale;
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
BoxData<T, C, MEM, D, E>& operator|=(BoxData<T, C, MEM, D, E>& dst, const LazyInterpStencil<T, C, MEM, D, E>&& op)
{
    op.apply(dst, true);
    return dst;
}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
BoxData<T, C, MEM, D, E>& operator+=(BoxData<T, C, MEM, D, E>& dst, const LazyInterpStencil<T, C, MEM, D, E>&& op)
{
    op.apply(dst, false);
    return dst;
}

}  // namespace Proto
