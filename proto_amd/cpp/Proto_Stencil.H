// Shift, Stencil<T>, LazyStencil and the stencil library
// (include/base/Proto_Stencil.H, implem/Proto_StencilImplem.H,
// implem/Proto_StencilDefs.H): same construction rules (addCoef merges
// repeated offsets, first-insertion order), same box inference, same
// |= / += semantics.  Application runs on the device through
// pm_stencil_apply, bit-identical to Proto's hostApply.
#pragma once

#include <cstring>
#include <limits>
#include <vector>

#include "Proto_BoxData.H"

namespace Proto {

class Shift {
public:
    Shift() {}
    explicit Shift(const Point& p) : m_shift(p) {}
    template<typename... Vals, typename = typename std::enable_if<(sizeof...(Vals) == DIM) && (DIM > 1)>::type>
    explicit Shift(Vals... v) : m_shift(Point(v...)) {}
    static Shift Zeros() { return Shift(Point::Zeros()); }
    static Shift Zeroes() { return Zeros(); }
    static Shift Ones(int s = 1) { return Shift(Point::Ones(s)); }
    static Shift Basis(int dir, int s = 1) { return Shift(Point::Basis(dir, s)); }
    const Point& point() const { return m_shift; }
    Shift operator*(const Shift& o) const { return Shift(m_shift + o.m_shift); }

private:
    Point m_shift;
};

template<typename T> class Stencil;
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1>
class LazyStencil;

template<typename T>
class Stencil {
public:
    Stencil() : m_destRefratio(Point::Ones()), m_destShift(Point::Zeros()), m_srcRefratio(Point::Ones()) {}
    Stencil(Shift s, T coef, Point destRatio = Point::Ones(), Point destShift = Point::Zeros(),
            Point srcRatio = Point::Ones())
        : m_destRefratio(destRatio), m_destShift(destShift), m_srcRefratio(srcRatio)
    {
        addCoef(coef, s.point());
    }
    Stencil(const Stencil& o)
        : m_coefs(o.m_coefs), m_offsets(o.m_offsets), m_destRefratio(o.m_destRefratio), m_destShift(o.m_destShift),
          m_srcRefratio(o.m_srcRefratio), m_span(o.m_span)
    {
    }
    Stencil& operator=(const Stencil& o)
    {
        if (this != &o) {
            release();
            m_coefs = o.m_coefs;
            m_offsets = o.m_offsets;
            m_destRefratio = o.m_destRefratio;
            m_destShift = o.m_destShift;
            m_srcRefratio = o.m_srcRefratio;
            m_span = o.m_span;
        }
        return *this;
    }
    ~Stencil() { release(); }

    // -------------------------------------------------------- construction
    // Stencil::addCoef (Proto_StencilImplem.H:237-276)
    void addCoef(T c, const Point& off)
    {
        release();
        for (size_t j = 0; j < m_offsets.size(); j++) {
            if (m_offsets[j] == off) {
                m_coefs[j] += c;
                return;
            }
        }
        m_coefs.push_back(c);
        m_offsets.push_back(off);
        m_span &= off;
    }
    Stencil operator+(const Stencil& o) const
    {
        Stencil r = *this;
        r += o;
        return r;
    }
    Stencil& operator+=(const Stencil& o)
    {
        for (size_t j = 0; j < o.m_coefs.size(); j++) { addCoef(o.m_coefs[j], o.m_offsets[j]); }
        return *this;
    }
    Stencil operator-(const Stencil& o) const
    {
        Stencil r = *this;
        r -= o;
        return r;
    }
    Stencil& operator-=(const Stencil& o)
    {
        for (size_t j = 0; j < o.m_coefs.size(); j++) { addCoef(-o.m_coefs[j], o.m_offsets[j]); }
        return *this;
    }
    // composition (Proto_StencilImplem.H:64-82)
    Stencil operator*(const Stencil& o) const
    {
        Stencil r;
        r.m_destRefratio = m_destRefratio;
        r.m_srcRefratio = m_srcRefratio;
        r.m_destShift = m_destShift;
        for (size_t i = 0; i < m_coefs.size(); i++)
            for (size_t j = 0; j < o.m_coefs.size(); j++) {
                r += Stencil(Shift(m_offsets[i] + o.m_offsets[j]), m_coefs[i] * o.m_coefs[j]);
            }
        return r;
    }
    Stencil operator*(T a) const
    {
        Stencil r = *this;
        for (auto& c : r.m_coefs) { c *= a; }
        return r;
    }
    Stencil& operator*=(const Stencil& o)
    {
        *this = (*this) * o;
        return *this;
    }
    Stencil& operator*=(T a)
    {
        release();
        for (auto& c : m_coefs) { c *= a; }
        return *this;
    }
    bool operator==(const Stencil& o) const
    {
        if (m_coefs.size() != o.m_coefs.size()) { return false; }
        for (size_t j = 0; j < m_coefs.size(); j++) {
            bool found = false;
            for (size_t k = 0; k < o.m_coefs.size(); k++) {
                if (m_offsets[j] == o.m_offsets[k] && m_coefs[j] == o.m_coefs[k]) { found = true; }
            }
            if (!found) { return false; }
        }
        return m_srcRefratio == o.m_srcRefratio && m_destRefratio == o.m_destRefratio && m_destShift == o.m_destShift;
    }
    void invert(int dir)
    {
        release();
        Box nspan;
        for (auto& o : m_offsets) { o[dir] = -o[dir]; nspan &= o; }
        m_span = nspan;
    }
    void transpose(int a, int b)
    {
        release();
        Box nspan;
        for (auto& o : m_offsets) { std::swap(o[a], o[b]); nspan &= o; }
        m_span = nspan;
    }

    // -------------------------------------------------------- accessors
    const std::vector<T>& coefs() const { return m_coefs; }
    const std::vector<Point>& offsets() const { return m_offsets; }
    unsigned long long size() const { return m_coefs.size(); }
    const Box& span() const { return m_span; }
    Point& srcRatio() { release(); return m_srcRefratio; }
    Point& destRatio() { release(); return m_destRefratio; }
    Point& destShift() { release(); return m_destShift; }
    const Point& srcRatio() const { return m_srcRefratio; }
    const Point& destRatio() const { return m_destRefratio; }
    const Point& destShift() const { return m_destShift; }
    Point ghost() const
    {
        Point g;
        for (int d = 0; d < DIM; d++) { g[d] = std::max(std::abs(m_span.low()[d]), std::abs(m_span.high()[d])); }
        return g;
    }
    T diagonal() const
    {
        T r = 0;
        for (size_t j = 0; j < m_offsets.size(); j++) {
            if (m_offsets[j] == Point::Zeros()) { r += m_coefs[j]; }
        }
        return r;
    }
    unsigned long long numFlops(const Box& b) const { return b.size() * (2 * size() + 1); }

    // -------------------------------------------------------- box inference
    Box indexRange(const Box& domain) const
    {
        if (domain.empty()) { return Box(); }
        return Box(domain.low() - m_span.low(), domain.high() - m_span.high()).taperCoarsen(m_srcRefratio);
    }
    Box indexDomain(const Box& range) const
    {
        if (range.empty()) { return Box(); }
        return range.shift(-m_destShift).taperCoarsen(m_destRefratio);
    }
    Box range(const Box& domain) const
    {
        if (domain.empty()) { return Box(); }
        Box r = indexRange(domain);
        r = Box(r.low() * m_destRefratio, r.high() * m_destRefratio);
        return r.shift(m_destShift);
    }
    Box domain(const Box& range) const
    {
        if (range.empty()) { return Box(); }
        Box d = indexDomain(range);
        d = Box(d.low() * m_srcRefratio, d.high() * m_srcRefratio);
        return Box(d.low() + m_span.low(), d.high() + m_span.high());
    }

    // -------------------------------------------------------- application
    // Stencil::apply (Proto_StencilImplem.H:412-431)
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    void apply(const BoxData<T, C, MEM, D, E>& src, BoxData<T, C, MEM, D, E>& dst, const Box& box, bool initToZero,
               T scale = 1) const
    {
        static_assert(MEM == DEVICE, "Stencil application needs DEVICE data");
        Box b = box;
        if (b.empty()) { b = indexDomain(dst.box()) & indexRange(src.box()); }
        if (b.empty() || m_coefs.empty()) { return; }
        closeForDevice();
        pm_box sb = src.box().pm(), db = dst.box().pm(), ib = b.pm();
        pm_check(pm_stencil_apply(m_dev, src.data(), &sb, dst.data(), &db, &ib, (int)(C * D * E), initToZero ? 1 : 0,
                                  (double)scale, protoCurrentStream()),
                 "Stencil::apply");
    }
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    LazyStencil<T, C, MEM, D, E> operator()(const BoxData<T, C, MEM, D, E>& src, T scale = 1) const
    {
        return LazyStencil<T, C, MEM, D, E>(this, &src, Box(), scale);
    }
    template<unsigned int C, MemType MEM, unsigned int D, unsigned int E>
    LazyStencil<T, C, MEM, D, E> operator()(const BoxData<T, C, MEM, D, E>& src, Box range, T scale = 1) const
    {
        return LazyStencil<T, C, MEM, D, E>(this, &src, range, scale);
    }

    // closeForDevice (Proto_StencilImplem.H:539-556): upload once per definition
    void closeForDevice() const
    {
        if (m_dev) { return; }
        std::vector<int> offs(3 * m_offsets.size(), 0);
        for (size_t j = 0; j < m_offsets.size(); j++)
            for (int d = 0; d < DIM; d++) { offs[3 * j + d] = m_offsets[j][d]; }
        int sr[3] = {1, 1, 1}, dr[3] = {1, 1, 1}, ds[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { sr[d] = m_srcRefratio[d]; dr[d] = m_destRefratio[d]; ds[d] = m_destShift[d]; }
        pm_check(pm_stencil_create(&m_dev, pm_type<T>::value, DIM, (int)m_coefs.size(), m_coefs.data(), offs.data(), sr,
                                   dr, ds),
                 "Stencil::closeForDevice");
    }
    int kernelKind() const
    {
        closeForDevice();
        int k = 0;
        pm_check(pm_stencil_kind(m_dev, &k), "Stencil::kernelKind");
        return k;
    }

    // -------------------------------------------------------- library (Proto_StencilDefs.H)
    static Stencil Laplacian()
    {
        Stencil r = ((T)(-2 * DIM)) * Shift::Zeros();
        for (int d = 0; d < DIM; d++) {
            r += ((T)1) * Shift::Basis(d, 1);
            r += ((T)1) * Shift::Basis(d, -1);
        }
        return r;
    }
#if DIM == 2
    static Stencil Laplacian_9()
    {
        Stencil r = ((T)(-20. / 6.)) * Shift::Zeros();
        Box K = Box::Cube(3).shift(Point::Ones(-1));
        for (auto it = K.begin(); it != K.end(); ++it) {
            const int s = (*it).sumAbs();
            if (s == 1) { r += ((T)(2.0 / 3.0)) * Shift(*it); }
            else if (s == 2) { r += ((T)(1.0 / 6.0)) * Shift(*it); }
        }
        return r;
    }
#endif
#if DIM == 3
    static Stencil Laplacian_19()
    {
        Stencil r = ((T)(-24.0)) * Shift::Zeros();
        Box K = Box::Cube(3).shift(Point::Ones(-1));
        for (auto it = K.begin(); it != K.end(); ++it) {
            const int s = (*it).sumAbs();
            if (s == 1) { r += ((T)2.0) * Shift(*it); }
            else if (s == 2) { r += ((T)1.0) * Shift(*it); }
        }
        r *= (T)(1.0 / 6.0);
        return r;
    }
    static Stencil Laplacian_27()
    {
        Stencil r = ((T)(-64.0 / 15.0)) * Shift::Zeros();
        Box K = Box::Cube(3).shift(Point::Ones(-1));
        for (auto it = K.begin(); it != K.end(); ++it) {
            const int s = (*it).sumAbs();
            if (s == 1) { r += ((T)(7.0 / 15.0)) * Shift(*it); }
            else if (s == 2) { r += ((T)(1.0 / 10.0)) * Shift(*it); }
            else if (s == 3) { r += ((T)(1.0 / 30.0)) * Shift(*it); }
        }
        return r;
    }
#endif
    static Stencil AvgDown(int r) { return AvgDown(Point::Ones(r)); }
    static Stencil AvgDown(Point ratio)
    {
        Stencil r;
        Box K(ratio);
        T c = 1.0 / K.size();
        for (auto it = K.begin(); it != K.end(); ++it) { r += c * Shift(*it); }
        r.m_srcRefratio = ratio;
        return r;
    }
    static Stencil Sum(int r) { return Sum(Point::Ones(r)); }
    static Stencil Sum(Point ratio)
    {
        Stencil r;
        Box K(ratio);
        for (auto it = K.begin(); it != K.end(); ++it) { r += ((T)1.0) * Shift(*it); }
        r.m_srcRefratio = ratio;
        return r;
    }
    static Stencil Identity() { return ((T)1.0) * Shift::Zeros(); }
    // second-order Laplacian in the directions other than dir (Proto_StencilDefs.H:152-182)
    static Stencil LaplacianFace(int dir, int order = 2)
    {
        if (order != 2) { throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::LaplacianFace: order 2 only"); }
        Stencil r = ((T)(-2 * (DIM - 1))) * Shift(Point::Zeros());
        for (int d = 0; d < DIM; d++) {
            if (d == dir) { continue; }
            r += ((T)1) * Shift(Point::Basis(d));
            r += ((T)1) * Shift(Point::Basis(d) * (-1));
        }
        return r;
    }
    // cell averages to the low face of direction dir, orders 2, 4, 5; Side::Hi
    // mirrors it (Proto_StencilDefs.H:184-229); terms in Proto's order
    static Stencil CellToFace(int dir, Side::LoHiSide side = Side::Lo, int order = 4)
    {
        Stencil r;
        if (order == 2) {
            r = ((T).5) * Shift::Basis(dir, -1) + ((T).5) * Shift::Zeros();
        } else if (order == 4) {
            T c[4] = {-1, 7, 7, -1};
            r = c[0] * Shift::Basis(dir, -2) + c[3] * Shift::Basis(dir, 1) + c[1] * Shift::Basis(dir, -1) +
                c[2] * Shift::Zeros();
            r *= (T)(1.0 / 12.0);
        } else if (order == 5) {
            T c[5] = {2, -13, 47, 27, -3};
            r = c[0] * Shift::Basis(dir, -3) + c[4] * Shift::Basis(dir, 1) + c[1] * Shift::Basis(dir, -2) +
                c[3] * Shift::Zeros() + c[2] * Shift::Basis(dir, -1);
            r *= (T)(1 / 60.0);
        } else {
            throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::CellToFace: order must be 2, 4 or 5");
        }
        if (side == Side::Hi) { r.invert(dir); }
        return r;
    }
    // Proto_StencilDefs.H:270-300
    static Stencil CellToFaceL(int dir, int order = 4) { return CellToFace(dir, Side::Lo, order); }
    static Stencil CellToFaceH(int dir, int order = 4)
    {
        if (order % 2 == 0) { return CellToFace(dir, Side::Lo, order); }
        Stencil S = CellToFace(dir, Side::Hi, order);
        return S * (((T)1.0) * Shift(Point::Basis(dir, -1)));
    }
    // AvgDown over the cells of one face of the refinement kernel (356-382)
    static Stencil AvgDownFace(int dir, Side::LoHiSide side, int ratio)
    {
        return AvgDownFace(dir, side, Point::Ones(ratio));
    }
    static Stencil AvgDownFace(int dir, Side::LoHiSide side, Point ratio)
    {
        Stencil r;
        Box K(ratio);
        K = K.face(dir, side);
        T c = 1.0 / K.size();
        for (auto it = K.begin(); it != K.end(); ++it) { r += c * Shift(*it); }
        r.m_srcRefratio = ratio;
        return r;
    }
    // n-th derivative in direction dir to the given (even) order, from powers
    // of the second difference (Proto_StencilDefs.H:392-457)
    static Stencil Derivative(int n, int dir, int order = 2)
    {
        constexpr int MAXORDH = 6;
        if (dir < 0 || dir >= DIM || n < 1 || n > 2 * MAXORDH || order < 1 || order > 2 * MAXORDH ||
            2 * MAXORDH + 2 - n - order < 0) {
            throw ProtoError(PM_ERR_INVALID, "Stencil::Derivative: derivative, direction or order out of range");
        }
        if (order % 2 != 0) { order++; }
        const Point e = Point::Basis(dir);
        Stencil factor = ((T)(1.0)) * Shift(e) + ((T)(-2.0)) * Shift(Point::Zeros()) + ((T)(1.0)) * Shift(e * (-1));
        Stencil deriv;
        Stencil temps[MAXORDH];
        if (n % 2 == 0) {
            // coefficients of the even series in powers of the second difference
            static const T ev[MAXORDH][MAXORDH] = {{1., -1. / 12, 1. / 90, -1. / 560, 1. / 3150, -1. / 16632},
                                                   {0, 1., -1. / 6, 7. / 240, -41. / 7560, 479. / 453600},
                                                   {0, 0, 1., -1. / 4, 13. / 240, -139. / 12096},
                                                   {0, 0, 0, 1., -1. / 3, 31. / 360},
                                                   {0, 0, 0, 0, 1., -5. / 12},
                                                   {0, 0, 0, 0, 0, 1.}};
            const int p = (n - 2) / 2;
            temps[0] = factor;
            for (int k = 1; k < MAXORDH; k++) { temps[k] = temps[k - 1] * factor; }
            for (int k = 0; k < (n + order - 2) / 2; k++) { deriv += ev[p][k] * temps[k]; }
        } else {
            static const T od[MAXORDH][MAXORDH] = {{1., -1. / 6, 1. / 30, -1. / 140, 1. / 630, -1. / 2772},
                                                   {0, 1., -1. / 4, 7. / 120, -41. / 3024, 479. / 151200},
                                                   {0, 0, 1., -1. / 3, 13. / 144, -139. / 6048},
                                                   {0, 0, 0, 1., -5. / 12, 31. / 240},
                                                   {0, 0, 0, 0, 1., -1. / 2},
                                                   {0, 0, 0, 0, 0, 1.}};
            const int p = (n - 1) / 2;
            temps[0] = ((T)(0.5)) * Shift(e) + ((T)(-0.5)) * Shift(e * (-1));
            for (int k = 1; k < MAXORDH; k++) { temps[k] = temps[k - 1] * factor; }
            for (int k = 0; k < (n + 1 + order - 2) / 2; k++) { deriv += od[p][k] * temps[k]; }
        }
        return deriv;
    }
    // face averages to cell averages in direction dir, orders 2 and 4 (459-477)
    static Stencil faceToCell(int dir, int order = 4)
    {
        Stencil av2 = ((T).5) * Shift::Zeros() + ((T).5) * Shift::Basis(dir, 1);
        if (order == 2) { return av2; }
        if (order != 4) { throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::faceToCell: order must be 2 or 4"); }
        return (((T)1.0) * Shift::Zeros() + ((T)(-1. / 8.)) * (((T)(-2.0)) * Shift::Zeros() +
                                                               ((T)1.0) * Shift::Basis(dir, 1) +
                                                               ((T)1.0) * Shift::Basis(dir, -1))) *
               av2;
    }
    // corner values to face averages normal to dir, orders 2 and 4 (478-502)
    static Stencil cornersToFaces(int dir, int order = 4)
    {
        Stencil av2 = ((T)1.0) * Shift::Zeros();
        for (int d = 0; d < DIM; d++) {
            if (d != dir) { av2 = av2 * (((T).5) * Shift::Zeros() + ((T).5) * Shift::Basis(d, 1)); }
        }
        if (order == 2) { return av2; }
        if (order != 4) { throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::cornersToFaces: order must be 2 or 4"); }
        return (((T)1.0) * Shift::Zeros() + ((T)(-1. / 12.)) * LaplacianFace(dir, 2)) * av2;
    }
    // corner values to cell averages, orders 2 and 4 (503-527)
    static Stencil CornersToCells(int order = 4)
    {
        Stencil av2 = ((T)1.0) * Shift::Zeros();
        for (int d = 0; d < DIM; d++) { av2 = av2 * (((T).5) * Shift::Zeros() + ((T).5) * Shift::Basis(d, 1)); }
        if (order == 2) { return av2; }
        if (order != 4) { throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::CornersToCells: order must be 2 or 4"); }
        return (((T)1.0) * Shift::Zeros() + ((T)(-1. / 12.)) * Laplacian()) * av2;
    }
    static Stencil FluxDivergence(int dir)
    {
        return ((T)(-1)) * Shift(Point::Zeros()) + ((T)(1.0)) * Shift(Point::Basis(dir));
    }
    static Stencil DiffCellToFace(int dir, Side::LoHiSide side = Side::Lo, int order = 4)
    {
        Stencil r;
        if (order == 4) {
            T c[4] = {1, -15, 15, -1};
            r = c[0] * Shift::Basis(dir, -2) + c[1] * Shift::Basis(dir, -1) + c[2] * Shift::Zeros() +
                c[3] * Shift::Basis(dir, 1);
            r *= (T)(1.0 / 12.0);
        } else if (order == 2) {
            r = ((T)(-1.0)) * Shift::Basis(dir, -1) + ((T)1.0) * Shift::Zeros();
        } else {
            throw ProtoError(PM_ERR_UNSUPPORTED, "Stencil::DiffCellToFace: order must be 2 or 4");
        }
        if (side == Side::Hi) { r.invert(dir); }
        return r;
    }

private:
    void release() const
    {
        if (m_dev) {
            pm_stencil_destroy(m_dev);
            m_dev = nullptr;
        }
    }

    std::vector<T> m_coefs;
    std::vector<Point> m_offsets;
    Point m_destRefratio, m_destShift, m_srcRefratio;
    Box m_span;
    mutable pm_stencil* m_dev = nullptr;
};

// coef * Shift -> Stencil (Proto_Stencil.H:46-164)
template<typename T, typename std::enable_if<std::is_floating_point<T>::value, int>::type = 0>
inline Stencil<T> operator*(T c, const Shift& s) { return Stencil<T>(s, c); }
template<typename T>
inline Stencil<T> operator*(T c, const Stencil<T>& s) { return s * c; }

// LazyStencil (Proto_Stencil.H:181-199, Proto_StencilImplem.H:492-537)
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
class LazyStencil {
public:
    LazyStencil() {}
    LazyStencil(const Stencil<T>* st, const BoxData<T, C, MEM, D, E>* src, Box range, T scale)
    {
        m_range = range;
        m_stencil.push_back(st);
        m_src.push_back(src);
        m_box.push_back(st->indexDomain(range));
        m_scale.push_back(scale);
    }
    unsigned long long size() const { return m_stencil.size(); }
    void apply(BoxData<T, C, MEM, D, E>& dst, bool overwrite)
    {
        for (size_t i = 0; i < m_src.size(); i++) {
            m_stencil[i]->apply(*m_src[i], dst, m_box[i], overwrite, m_scale[i]);
        }
    }
    Box inferredRange() const
    {
        if (size() == 0) { return Box(); }
        Box range;
        for (size_t i = 0; i < size(); i++) {
            Box ri = m_stencil[0]->range(m_src[0]->box());
            if (i == 0) { range = ri; } else { range = range & ri.low(); range = range & ri.high(); }
        }
        return range;
    }

    std::vector<const Stencil<T>*> m_stencil;
    std::vector<const BoxData<T, C, MEM, D, E>*> m_src;
    std::vector<Box> m_box;
    std::vector<T> m_scale;
    Box m_range;
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
BoxData<T, C, MEM, D, E>& operator|=(BoxData<T, C, MEM, D, E>& dst, LazyStencil<T, C, MEM, D, E>&& op)
{
    for (auto* s : op.m_src) {
        if (s->isAlias(dst)) { throw ProtoError(PM_ERR_INVALID, "operator|=: stencil source and destination alias"); }
    }
    if (op.size() > 0) { op.apply(dst, true); }
    return dst;
}

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
BoxData<T, C, MEM, D, E>& operator+=(BoxData<T, C, MEM, D, E>& dst, LazyStencil<T, C, MEM, D, E>&& op)
{
    for (auto* s : op.m_src) {
        if (s->isAlias(dst)) { throw ProtoError(PM_ERR_INVALID, "operator+=: stencil source and destination alias"); }
    }
    if (op.size() > 0) { op.apply(dst, false); }
    return dst;
}

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
BoxData<T, C, MEM, D, E>::BoxData(LazyStencil<T, C, MEM, D, E>&& op)
{
    Box range = op.m_range;
    if (range.empty()) { range = op.inferredRange(); }
    define(range);
    op.apply(*this, true);
}

}  // namespace Proto
