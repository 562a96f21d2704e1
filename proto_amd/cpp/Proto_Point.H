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
