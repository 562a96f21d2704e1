// Point and Box: Proto's integer index algebra (include/base/Proto_Point.H,
// Proto_Box.H, implem/Proto_BoxImplem.H) -- same names and semantics, written
// for this library.  Point is usable in device functors (forall); Box is host
// only, kernels receive plain pm_box structs.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <initializer_list>
#include <iostream>
#include <vector>

#ifndef DIM
#define DIM 3
#endif

#include "protomi.h"

#ifdef __HIPCC__
#define PR_HD __host__ __device__
#else
#define PR_HD
#endif

namespace Proto {

class Point {
public:
    int m_tuple[DIM];

    PR_HD Point() { for (int d = 0; d < DIM; d++) { m_tuple[d] = 0; } }
    PR_HD Point(const int (&t)[DIM]) { for (int d = 0; d < DIM; d++) { m_tuple[d] = t[d]; } }
    PR_HD Point(std::initializer_list<int> v)
    {
        int d = 0;
        for (int x : v) { if (d < DIM) { m_tuple[d++] = x; } }
        for (; d < DIM; d++) { m_tuple[d] = 0; }
    }
    template<typename... Vals, typename = typename std::enable_if<(sizeof...(Vals) == DIM) && (DIM > 1)>::type>
    PR_HD explicit Point(Vals... vals) : Point({static_cast<int>(vals)...}) {}

    PR_HD static Point Zeros() { return Point(); }
    PR_HD static Point Zeroes() { return Point(); }
    PR_HD static Point Ones(int s = 1) { Point p; for (int d = 0; d < DIM; d++) { p.m_tuple[d] = s; } return p; }
    PR_HD static Point Basis(int dir, int s = 1) { Point p; p.m_tuple[dir] = s; return p; }

    PR_HD int& operator[](int d) { return m_tuple[d]; }
    PR_HD const int& operator[](int d) const { return m_tuple[d]; }

#define PR_POINT_OP(op)                                                                         \
    PR_HD Point operator op(const Point& b) const { Point r; for (int d = 0; d < DIM; d++) { r.m_tuple[d] = m_tuple[d] op b.m_tuple[d]; } return r; } \
    PR_HD Point operator op(int s) const { Point r; for (int d = 0; d < DIM; d++) { r.m_tuple[d] = m_tuple[d] op s; } return r; }
    PR_POINT_OP(+)
    PR_POINT_OP(-)
    PR_POINT_OP(*)
    PR_POINT_OP(/)
    PR_POINT_OP(%)
#undef PR_POINT_OP
    PR_HD Point operator-() const { return (*this) * -1; }
    PR_HD Point& operator+=(const Point& b) { *this = *this + b; return *this; }
    PR_HD Point& operator-=(const Point& b) { *this = *this - b; return *this; }
    PR_HD bool operator==(const Point& b) const
    {
        for (int d = 0; d < DIM; d++) { if (m_tuple[d] != b.m_tuple[d]) { return false; } }
        return true;
    }
    PR_HD bool operator!=(const Point& b) const { return !(*this == b); }
    PR_HD bool operator<(const Point& b) const
    {
        for (int d = DIM - 1; d >= 0; d--) {
            if (m_tuple[d] != b.m_tuple[d]) { return m_tuple[d] < b.m_tuple[d]; }
        }
        return false;
    }
    PR_HD int sum() const { int s = 0; for (int d = 0; d < DIM; d++) { s += m_tuple[d]; } return s; }
    PR_HD int sumAbs() const { int s = 0; for (int d = 0; d < DIM; d++) { s += m_tuple[d] < 0 ? -m_tuple[d] : m_tuple[d]; } return s; }
    PR_HD int min() const { int s = m_tuple[0]; for (int d = 1; d < DIM; d++) { s = s < m_tuple[d] ? s : m_tuple[d]; } return s; }
    PR_HD int max() const { int s = m_tuple[0]; for (int d = 1; d < DIM; d++) { s = s > m_tuple[d] ? s : m_tuple[d]; } return s; }
    PR_HD bool isotropic() const { for (int d = 1; d < DIM; d++) { if (m_tuple[d] != m_tuple[0]) { return false; } } return true; }
};

inline std::ostream& operator<<(std::ostream& os, const Point& p)
{
    os << "(";
    for (int d = 0; d < DIM; d++) { os << p[d] << (d + 1 < DIM ? "," : ")"); }
    return os;
}

// Side of a cell face (Proto_Face.H: class Side with enum LoHiSide).
class Side {
public:
    enum LoHiSide { Invalid = -1, Lo = 0, Hi, NUMSIDES };
};
inline Side::LoHiSide flip(const Side::LoHiSide& s) { return s == Side::Lo ? Side::Hi : Side::Lo; }
inline int sign(const Side::LoHiSide& s) { return s == Side::Lo ? -1 : 1; }

// Data centering (Proto_Centering.H): cells, the faces normal to direction d, nodes.
enum Centering {
    PR_CELL = -1, PR_FACE_0 = 0, PR_FACE_1, PR_FACE_2, PR_FACE_3, PR_FACE_4, PR_FACE_5, PR_FACE_6,
    PR_NODE = 7, PR_FACE, PR_EDGE
};

// A face of a box: its normal direction and side (Proto_Face.H).
struct Face {
    int dir = 0;
    Side::LoHiSide side = Side::Lo;
    Face() {}
    Face(int d, Side::LoHiSide s) : dir(d), side(s) {}
    bool operator==(const Face& o) const { return dir == o.dir && side == o.side; }
};

// Fixed-size array with element-wise arithmetic (Proto_Array.H): grid
// spacings, per-direction fluxes.  Elements are value-initialised.
template<typename T, size_t N>
class Array {
public:
    Array() : m_data{} {}
    Array(T v) { fill(v); }
    // the first N values of the list (Proto_ArrayImplem.H:31-40)
    Array(std::initializer_list<T> list) : m_data{}
    {
        size_t i = 0;
        for (const T& x : list) {
            if (i >= N) { break; }
            m_data[i++] = x;
        }
    }
    Array(const Point& p) : m_data{}
    {
        for (int d = 0; d < (int)N && d < DIM; d++) { m_data[d] = (T)p[d]; }
    }
    Array(const std::array<T, N>& a) : m_data(a) {}
    static Array Zeros() { return Array(T(0)); }
    static Array Ones(T scale = 1) { return Array(scale); }

    void fill(T v) { for (auto& x : m_data) { x = v; } }
    T* data() { return m_data.data(); }
    const T* data() const { return m_data.data(); }
    T& operator[](size_t i) { return m_data[i]; }
    const T& operator[](size_t i) const { return m_data[i]; }
    static constexpr size_t size() { return N; }
    bool operator==(const Array& o) const { return m_data == o.m_data; }
    bool operator!=(const Array& o) const { return !(*this == o); }

#define PR_ARRAY_OP(op)                                                                                \
    Array operator op(const Array& o) const { Array r; for (size_t i = 0; i < N; i++) { r[i] = m_data[i] op o[i]; } return r; } \
    Array operator op(T v) const { Array r; for (size_t i = 0; i < N; i++) { r[i] = m_data[i] op v; } return r; }             \
    void operator op##=(const Array& o) { for (size_t i = 0; i < N; i++) { m_data[i] = m_data[i] op o[i]; } }                  \
    void operator op##=(T v) { for (size_t i = 0; i < N; i++) { m_data[i] = m_data[i] op v; } }
    PR_ARRAY_OP(+)
    PR_ARRAY_OP(-)
    PR_ARRAY_OP(*)
    PR_ARRAY_OP(/)
#undef PR_ARRAY_OP
    Array operator-() const { Array r; for (size_t i = 0; i < N; i++) { r[i] = -m_data[i]; } return r; }

    T max() const { T r = m_data[0]; for (size_t i = 1; i < N; i++) { r = std::max(r, m_data[i]); } return r; }
    T min() const { T r = m_data[0]; for (size_t i = 1; i < N; i++) { r = std::min(r, m_data[i]); } return r; }
    T absMax() const { T r = 0; for (size_t i = 0; i < N; i++) { r = std::max(r, (T)std::abs(m_data[i])); } return r; }
    T sum() const { T r = 0; for (size_t i = 0; i < N; i++) { r += m_data[i]; } return r; }
    T product() const { T r = 1; for (size_t i = 0; i < N; i++) { r *= m_data[i]; } return r; }
    T dot(const Array& o) const { T r = 0; for (size_t i = 0; i < N; i++) { r += m_data[i] * o[i]; } return r; }
    T norm() const { return std::sqrt(dot(*this)); }
    T* begin() { return m_data.data(); }
    T* end() { return m_data.data() + N; }
    const T* begin() const { return m_data.data(); }
    const T* end() const { return m_data.data() + N; }

private:
    std::array<T, N> m_data;
};

class Box;

class BoxIterator {
public:
    BoxIterator(const Box& b, long pos);
    Point operator*() const;
    BoxIterator& operator++() { ++m_pos; return *this; }
    bool operator!=(const BoxIterator& o) const { return m_pos != o.m_pos; }
    bool operator==(const BoxIterator& o) const { return m_pos == o.m_pos; }
    bool ok() const;
    void begin() { m_pos = 0; }
    long pos() const { return m_pos; }

private:
    const Box* m_box;
    long m_pos;
};

class Box {
public:
    Box() : m_low(Point::Zeros()), m_high(Point::Ones(-1)) {}
    Box(const Point& lo, const Point& hi) : m_low(lo), m_high(hi) {}
    explicit Box(const Point& size) : m_low(Point::Zeros()), m_high(size - Point::Ones()) {}

    static Box Cube(int n) { return Box(Point::Zeros(), Point::Ones(n - 1)); }
    static Box Kernel(int r) { return Box::Cube(2 * r + 1).shift(Point::Ones(-r)); }

    const Point& low() const { return m_low; }
    const Point& high() const { return m_high; }
    bool empty() const { for (int d = 0; d < DIM; d++) { if (m_high[d] < m_low[d]) { return true; } } return false; }
    unsigned long long size() const
    {
        if (empty()) { return 0; }
        unsigned long long s = 1;
        for (int d = 0; d < DIM; d++) { s *= (unsigned long long)(m_high[d] - m_low[d] + 1); }
        return s;
    }
    int size(int d) const { return std::max(0, m_high[d] - m_low[d] + 1); }
    Point sizes() const { Point p; for (int d = 0; d < DIM; d++) { p[d] = size(d); } return p; }

    bool containsPoint(const Point& p) const
    {
        for (int d = 0; d < DIM; d++) { if (p[d] < m_low[d] || p[d] > m_high[d]) { return false; } }
        return true;
    }
    bool contains(const Point& p) const { return containsPoint(p); }
    bool containsBox(const Box& b) const { return b.empty() || (containsPoint(b.low()) && containsPoint(b.high())); }

    Box shift(const Point& s) const { return Box(m_low + s, m_high + s); }
    Box shift(int dir, int s) const { return shift(Point::Basis(dir, s)); }
    Box grow(int g) const { return grow(Point::Ones(g)); }
    Box grow(const Point& g) const { return Box(m_low - g, m_high + g); }
    Box grow(int dir, int g) const { return grow(Point::Basis(dir, g)); }
    // one side only (Proto_BoxImplem.H:409-413).  The direction is
    // Point::Basis(dir, side) as Proto forms it, so Side::Lo (0) gives the zero
    // vector and the box comes back unchanged; Side::Hi grows the high end.
    Box grow(int dir, Side::LoHiSide side, int g) const { return extrude(Point::Basis(dir, (int)side), g); }
    // face / node boxes of a cell box (Proto_BoxImplem.H:415-433)
    Box grow(Centering ctr) const
    {
        if (ctr == PR_CELL) { return *this; }
        if (ctr == PR_NODE) { return extrude(Point::Ones(), 1); }
        if ((int)ctr >= 0 && (int)ctr < DIM) { return extrude(Point::Basis((int)ctr, 1)); }
        throw std::invalid_argument("Box::grow(Centering): undefined for FACE / EDGE centering");
    }
    // every direction of dir that is positive grows the high end, negative the low end
    Box extrude(const Point& dir, int dist = 1) const
    {
        Point lo = m_low, hi = m_high;
        for (int d = 0; d < DIM; d++) {
            if (dir[d] > 0) { hi[d] += dist; } else if (dir[d] < 0) { lo[d] -= dist; }
        }
        return Box(lo, hi);
    }

    // intersection
    Box operator&(const Box& b) const
    {
        Point lo, hi;
        for (int d = 0; d < DIM; d++) {
            lo[d] = std::max(m_low[d], b.m_low[d]);
            hi[d] = std::min(m_high[d], b.m_high[d]);
        }
        return Box(lo, hi);
    }
    Box& operator&=(const Box& b) { *this = *this & b; return *this; }
    // smallest box containing *this and p (Proto_BoxImplem.H:297-318)
    Box operator&(const Point& p) const
    {
        if (containsPoint(p)) { return *this; }
        if (empty()) { return Box(p, p); }
        Point lo = m_low, hi = m_high;
        for (int d = 0; d < DIM; d++) { lo[d] = std::min(p[d], lo[d]); hi[d] = std::max(p[d], hi[d]); }
        return Box(lo, hi);
    }
    Box& operator&=(const Point& p) { *this = *this & p; return *this; }
    bool operator==(const Box& b) const
    {
        if (empty() && b.empty()) { return true; }
        return m_low == b.m_low && m_high == b.m_high;
    }
    bool operator!=(const Box& b) const { return !(*this == b); }

    // floor division of both corners (Proto_BoxImplem.H:506-553)
    Box coarsen(int r) const { return coarsen(Point::Ones(r)); }
    Box coarsen(const Point& r) const
    {
        Point lo, hi;
        for (int d = 0; d < DIM; d++) {
            lo[d] = floordiv(m_low[d], r[d]);
            hi[d] = floordiv(m_high[d], r[d]);
        }
        return Box(lo, hi);
    }
    Box refine(int r) const { return refine(Point::Ones(r)); }
    Box refine(const Point& r) const { return Box(m_low * r, (m_high + Point::Ones()) * r - Point::Ones()); }
    // high by truncating division, low by ceiling (Proto_BoxImplem.H:555-564)
    Box taperCoarsen(const Point& r) const
    {
        Point lo, hi = m_high / r;
        for (int d = 0; d < DIM; d++) { lo[d] = (int)std::ceil(m_low[d] / (double)r[d]); }
        return Box(lo, hi);
    }
    bool coarsenable(const Point& r) const
    {
        return coarsen(r).refine(r) == *this;
    }
    bool coarsenable(int r) const { return coarsenable(Point::Ones(r)); }
    Box flatten(int dir, bool up = false) const
    {
        Point lo = m_low, hi = m_high;
        if (up) { lo[dir] = hi[dir]; } else { hi[dir] = lo[dir]; }
        return Box(lo, hi);
    }
    Box extrude(int dir, int d = 1) const
    {
        Point lo = m_low, hi = m_high;
        if (d > 0) { hi[dir] += d; } else { lo[dir] += d; }
        return Box(lo, hi);
    }
    // the region of thickness dist just outside the box in the directions of
    // dir (Proto_BoxImplem.H:685-722); dist < 0: inside, as edge
    Box adjacent(const Point& dir, int dist) const
    {
        if (dist == 0) { return Box(); }
        if (dist < 0) { return edge(dir, -dist); }
        Point lo = m_low, hi = m_high;
        for (int k = 0; k < DIM; k++) {
            if (dir[k] > 0) { lo[k] = m_high[k] + 1; hi[k] = m_high[k] + dist; }
            else if (dir[k] < 0) { hi[k] = m_low[k] - 1; lo[k] = m_low[k] - dist; }
        }
        return Box(lo, hi);
    }
    Box adjacent(const Point& dir) const
    {
        Point lo = m_low, hi = m_high;
        for (int k = 0; k < DIM; k++) {
            if (dir[k] > 0) { lo[k] = m_high[k] + 1; hi[k] = m_high[k] + dir[k]; }
            else if (dir[k] < 0) { hi[k] = m_low[k] - 1; lo[k] = m_low[k] + dir[k]; }
        }
        return Box(lo, hi);
    }
    // the layer of cells inside the box along the directions of dir, dir[k]
    // cells thick (Proto_BoxImplem.H:608-640)
    Box edge(const Point& dir) const
    {
        Point lo = m_low, hi = m_high;
        for (int k = 0; k < DIM; k++) {
            if (dir[k] > 0) { lo[k] = m_high[k] - dir[k] + 1; }
            if (dir[k] < 0) { hi[k] = m_low[k] - dir[k] - 1; }
        }
        return Box(lo, hi);
    }
    Box edge(const Point& dir, int dist) const
    {
        if (dist == 0) { return Box(); }
        if (dist < 0) { return adjacent(dir, -dist); }
        Point d;
        for (int k = 0; k < DIM; k++) { d[k] = dir[k] > 0 ? dist : (dir[k] < 0 ? -dist : 0); }
        return edge(d);
    }
    // Point::Basis(dir, side) as Proto forms it (Side::Lo gives the zero vector, 636-640)
    Box edge(int dir, Side::LoHiSide side, int dist) const { return edge(Point::Basis(dir, (int)side), dist); }
    // the cells of the low or high face in direction dir, at most dist thick (642-658)
    Box face(int dir, Side::LoHiSide side, int dist = 1) const
    {
        const int d = std::min(dist, size(dir));
        if (side == Side::Lo) { return edge(Point::Basis(dir, -1), d); }
        if (side == Side::Hi) { return edge(Point::Basis(dir, 1), d); }
        return Box();
    }

    // Box::sindex / index and its inverse, x fastest (Proto_BoxImplem.H:76-86, 208-222)
    unsigned int index(const Point& p) const
    {
        unsigned int f = 1, li = p[0] - m_low[0];
        for (int d = 1; d < DIM; d++) { f *= (m_high[d - 1] - m_low[d - 1] + 1); li += (p[d] - m_low[d]) * f; }
        return li;
    }
    Point operator[](unsigned long long i) const
    {
        Point p;
        for (int d = 0; d < DIM; d++) {
            const long n = m_high[d] - m_low[d] + 1;
            p[d] = (int)(i % n) + m_low[d];
            i /= n;
        }
        return p;
    }

    BoxIterator begin() const { return BoxIterator(*this, 0); }
    BoxIterator end() const { return BoxIterator(*this, (long)size()); }

    pm_box pm() const
    {
        pm_box b;
        for (int d = 0; d < 3; d++) {
            b.lo[d] = d < DIM ? m_low[d] : 0;
            b.hi[d] = d < DIM ? m_high[d] : 0;
        }
        return b;
    }

    static int floordiv(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

private:
    Point m_low, m_high;
};

inline BoxIterator::BoxIterator(const Box& b, long pos) : m_box(&b), m_pos(pos) {}
inline Point BoxIterator::operator*() const { return (*m_box)[m_pos]; }
inline bool BoxIterator::ok() const { return m_pos < (long)m_box->size(); }

inline std::ostream& operator<<(std::ostream& os, const Box& b)
{
    return os << "[" << b.low() << ", " << b.high() << "]";
}

}  // namespace Proto
