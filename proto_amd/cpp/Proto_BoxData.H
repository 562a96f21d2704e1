// MemType, proto_malloc/free/memcpy and BoxData<T,C,MEM,D,E>: Proto's patch
// array (include/base/Proto_BoxData.H, Proto_Memory.H, Proto_MemType.H) with
// device storage served by the protomi C ABI.  Layout is Proto's: component
// c of box B at data() + c*|B|, x fastest.  HOST BoxData holds storage for
// transfers and element access; compute is done on DEVICE BoxData.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "Proto_Point.H"

namespace Proto {

enum MemType { HOST = 0, DEVICE = 1, BOTH = 2 };
#ifndef MEMTYPE_DEFAULT
#define MEMTYPE_DEFAULT DEVICE
#endif

// Errors from the C ABI become exceptions on the C++ side (Proto aborts in
// the same places; an exception lets a server or test recover).
struct ProtoError : std::runtime_error {
    int code;
    ProtoError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void pm_check(int rc, const char* what)
{
    if (rc != PM_OK) {
        throw ProtoError(rc, std::string(what) + ": " + pm_last_error());
    }
}

// The stream all Proto calls of this thread use (Proto's
// DisjointBoxLayout::getCurrentStream, Proto_DisjointBoxLayout.H:467-476).
inline pm_stream_t& protoCurrentStream()
{
    static thread_local pm_stream_t s = nullptr;
    return s;
}

template<MemType MEM> inline void* proto_malloc(size_t bytes)
{
    if (MEM == HOST) { return bytes ? ::operator new(bytes) : nullptr; }
    void* p = nullptr;
    pm_check(pm_malloc(&p, bytes), "proto_malloc<DEVICE>");
    return p;
}
template<MemType MEM> inline void proto_free(void* p)
{
    if (MEM == HOST) { ::operator delete(p); return; }
    pm_free(p);
}
template<MemType SRC, MemType DST> inline void proto_memcpy(const void* src, void* dst, size_t bytes)
{
    if (SRC == HOST && DST == HOST) { std::memcpy(dst, src, bytes); return; }
    const int kind = (SRC == HOST) ? 0 : (DST == HOST ? 1 : 2);
    pm_check(pm_memcpy(dst, src, bytes, kind, protoCurrentStream()), "proto_memcpy");
    pm_check(pm_stream_sync(protoCurrentStream()), "proto_memcpy");
}
inline void protoThreadSynchronize() { pm_check(pm_stream_sync(protoCurrentStream()), "protoThreadSynchronize"); }

template<typename T> struct pm_type;
template<> struct pm_type<double> { static constexpr int value = PM_F64; };
template<> struct pm_type<float> { static constexpr int value = PM_F32; };

enum Operation { Max, Min, Abs, Sum, SumAbs, SumSquare };
template<typename T, Operation OP, MemType MEM = MEMTYPE_DEFAULT> class Reduction;
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E> class LazyStencil;

template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1>
class BoxData {
public:
    static constexpr unsigned int NCOMP = C * D * E;

    BoxData() : m_rawPtr(nullptr) {}
    explicit BoxData(const Box& b) { define(b); }
    BoxData(const Box& b, T init) { define(b); setVal(init); }
    // non-owning alias of existing memory (Proto_BoxDataImplem.H:88-111)
    BoxData(const T* ptr, const Box& b, int ncomp = NCOMP)
    {
        (void)ncomp;
        m_box = b;
        m_rawPtr = const_cast<T*>(ptr);
        m_data = std::shared_ptr<T>(m_rawPtr, [](T*) {});
        m_alias = true;
    }
    BoxData(BoxData&& o) noexcept : m_box(o.m_box), m_data(std::move(o.m_data)), m_rawPtr(o.m_rawPtr), m_alias(o.m_alias) {}
    BoxData& operator=(BoxData&& o) noexcept
    {
        m_box = o.m_box;
        m_data = std::move(o.m_data);
        m_rawPtr = o.m_rawPtr;
        m_alias = o.m_alias;
        return *this;
    }
    BoxData(const BoxData&) = delete;
    BoxData& operator=(const BoxData&) = delete;
    // output-allocating stencil form (Proto_BoxDataImplem.H:113-127)
    BoxData(LazyStencil<T, C, MEM, D, E>&& op);

    void define(const Box& b)
    {
        m_box = b;
        const size_t bytes = sizeof(T) * (size_t)b.size() * NCOMP;
        m_rawPtr = (T*)proto_malloc<MEM>(bytes);
        m_data = std::shared_ptr<T>(m_rawPtr, [](T* p) { proto_free<MEM>(p); });
        m_alias = false;
    }

    const Box& box() const { return m_box; }
    size_t size() const { return (size_t)m_box.size() * NCOMP; }
    size_t linearSize() const { return size(); }
    T* data() { return m_rawPtr; }
    const T* data() const { return m_rawPtr; }
    T* data(unsigned int c, unsigned int d = 0, unsigned int e = 0) { return m_rawPtr + compOffset(c, d, e); }
    const T* data(unsigned int c, unsigned int d = 0, unsigned int e = 0) const { return m_rawPtr + compOffset(c, d, e); }
    size_t compOffset(unsigned int c, unsigned int d, unsigned int e) const
    {
        return ((size_t)e * D * C + (size_t)d * C + c) * (size_t)m_box.size();
    }
    bool isAlias(const BoxData& o) const { return m_rawPtr != nullptr && m_rawPtr == o.m_rawPtr; }
    bool defined() const { return m_rawPtr != nullptr; }

    // host element access (HOST storage only)
    T& operator()(const Point& p, unsigned int c = 0, unsigned int d = 0, unsigned int e = 0)
    {
        static_assert(MEM == HOST, "element access needs HOST storage");
        return m_rawPtr[compOffset(c, d, e) + m_box.index(p)];
    }
    const T& operator()(const Point& p, unsigned int c = 0, unsigned int d = 0, unsigned int e = 0) const
    {
        static_assert(MEM == HOST, "element access needs HOST storage");
        return m_rawPtr[compOffset(c, d, e) + m_box.index(p)];
    }

    // ------------------------------------------------------------- fills
    void setVal(T v) { setVal(v, m_box); }
    void setVal(T v, const Box& region)
    {
        if (MEM == HOST) {
            for (unsigned c = 0; c < NCOMP; c++)
                for (auto it = region.begin(); it != region.end(); ++it) { m_rawPtr[c * m_box.size() + m_box.index(*it)] = v; }
            return;
        }
        pm_box b = m_box.pm(), r = (region & m_box).pm();
        pm_check(pm_boxdata_scalar(4, pm_type<T>::value, m_rawPtr, &b, &r, (double)v, NCOMP, protoCurrentStream()),
                 "BoxData::setVal");
    }
    void setToZero() { setVal(T(0)); }

    // ------------------------------------------------------------- copies
    // copyTo over the intersection of the boxes (Proto_BoxDataImplem.H:490-640)
    template<MemType DST>
    void copyTo(BoxData<T, C, DST, D, E>& dst) const { copyTo(dst, m_box & dst.box(), Point::Zeros()); }
    template<MemType DST>
    void copyTo(BoxData<T, C, DST, D, E>& dst, const Box& region, const Point& shift = Point::Zeros()) const
    {
        Box r = region & m_box;
        if (r.empty()) { return; }
        if (MEM == DEVICE && DST == DEVICE) {
            pm_box sb = m_box.pm(), db = dst.box().pm(), rb = r.pm();
            int sh[3] = {0, 0, 0};
            for (int d = 0; d < DIM; d++) { sh[d] = shift[d]; }
            pm_check(pm_boxdata_copy(pm_type<T>::value, m_rawPtr, &sb, dst.data(), &db, &rb, sh, NCOMP,
                                     protoCurrentStream()),
                     "BoxData::copyTo");
            return;
        }
        if (r == m_box && r.shift(shift) == dst.box()) {
            proto_memcpy<MEM, DST>(m_rawPtr, dst.data(), sizeof(T) * size());
            return;
        }
        throw ProtoError(PM_ERR_UNSUPPORTED, "BoxData::copyTo between memory types needs whole, equal boxes");
    }

    // ------------------------------------------------------------- arithmetic
    BoxData& operator+=(const BoxData& o) { return binop(0, o); }
    BoxData& operator-=(const BoxData& o) { return binop(1, o); }
    BoxData& operator*=(const BoxData& o) { return binop(2, o); }
    BoxData& operator/=(const BoxData& o) { return binop(3, o); }
    BoxData& operator+=(T v) { return scalar(0, v); }
    BoxData& operator-=(T v) { return scalar(1, v); }
    BoxData& operator*=(T v) { return scalar(2, v); }
    BoxData& operator/=(T v) { return scalar(3, v); }
    // out += in * scale (Proto_BoxDataImplem.H:1219-1234)
    void increment(const BoxData& in, T scale = 1)
    {
        static_assert(MEM == DEVICE, "compute needs DEVICE storage");
        Box r = m_box & in.box();
        if (r.empty()) { return; }
        pm_box a = m_box.pm(), b = in.box().pm(), rb = r.pm();
        pm_check(pm_boxdata_increment(pm_type<T>::value, m_rawPtr, &a, in.data(), &b, &rb, (double)scale, NCOMP,
                                      protoCurrentStream()),
                 "BoxData::increment");
    }

    // ------------------------------------------------------------- reductions
    // over the whole buffer, ghosts included, as Proto's BoxData::reduce (1375-1382)
    template<Operation OP> void reduce(Reduction<T, OP, MEM>& rxn) const;
    T absMax() const;
    T sum() const;
    T max() const;
    T min() const;

private:
    BoxData& binop(int op, const BoxData& o)
    {
        static_assert(MEM == DEVICE, "compute needs DEVICE storage");
        Box r = m_box & o.box();
        if (r.empty()) { return *this; }
        pm_box a = m_box.pm(), b = o.box().pm(), rb = r.pm();
        pm_check(pm_boxdata_binop(op, pm_type<T>::value, m_rawPtr, &a, o.data(), &b, &rb, NCOMP, protoCurrentStream()),
                 "BoxData operator");
        return *this;
    }
    BoxData& scalar(int op, T v)
    {
        static_assert(MEM == DEVICE, "compute needs DEVICE storage");
        pm_box a = m_box.pm();
        pm_check(pm_boxdata_scalar(op, pm_type<T>::value, m_rawPtr, &a, &a, (double)v, NCOMP, protoCurrentStream()),
                 "BoxData operator");
        return *this;
    }

    Box m_box;
    std::shared_ptr<T> m_data;
    T* m_rawPtr = nullptr;
    bool m_alias = false;
};

// ---------------------------------------------------------------- Reduction
// Reduction<T,OP,MEM> (include/base/Proto_Reduction.H:29-110,
// implem/Proto_ReductionImplem.H): a device scalar folded by every reduce()
// until reset(); fetchLocal() copies it to the host.
template<typename T, Operation OP, MemType MEM>
class Reduction {
    static_assert(MEM == DEVICE, "Reduction on this library runs on the device");

public:
    Reduction(bool dynamic = false)
    {
        (void)dynamic;
        m_dev = (T*)proto_malloc<DEVICE>(sizeof(T));
        reset();
    }
    ~Reduction() { proto_free<DEVICE>(m_dev); }
    Reduction(const Reduction&) = delete;
    Reduction& operator=(const Reduction&) = delete;

    static T init()
    {
        switch (OP) {
            case Max: return std::numeric_limits<T>::min();
            case Min: return std::numeric_limits<T>::max();
            default: return T(0);
        }
    }
    void reset() { pm_check(pm_reduce_init((int)OP, pm_type<T>::value, m_dev, protoCurrentStream()), "Reduction::reset"); }
    void reduce(const T* data, size_t n)
    {
        pm_check(pm_reduce((int)OP, pm_type<T>::value, data, n, m_dev, 1, protoCurrentStream()), "Reduction::reduce");
    }
    void reduceBox(const T* data, const Box& box, const Box& region, int comp = 0)
    {
        pm_box b = box.pm(), r = region.pm();
        pm_check(pm_reduce_box((int)OP, pm_type<T>::value, data, &b, &r, comp, m_dev, 1, protoCurrentStream()),
                 "Reduction::reduceBox");
    }
    T fetchLocal()
    {
        T v;
        proto_memcpy<DEVICE, HOST>(m_dev, &v, sizeof(T));
        return v;
    }
    T fetch() { return fetchLocal(); }
    T* deviceValue() { return m_dev; }

private:
    T* m_dev = nullptr;
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
template<Operation OP>
void BoxData<T, C, MEM, D, E>::reduce(Reduction<T, OP, MEM>& rxn) const
{
    rxn.reduce(m_rawPtr, size());
}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
T BoxData<T, C, MEM, D, E>::absMax() const
{
    Reduction<T, Abs, MEM> r(true);
    reduce(r);
    return r.fetchLocal();
}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
T BoxData<T, C, MEM, D, E>::sum() const
{
    Reduction<T, Sum, MEM> r(true);
    reduce(r);
    return r.fetchLocal();
}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
T BoxData<T, C, MEM, D, E>::max() const
{
    Reduction<T, Max, MEM> r(true);
    reduce(r);
    return r.fetchLocal();
}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
T BoxData<T, C, MEM, D, E>::min() const
{
    Reduction<T, Min, MEM> r(true);
    reduce(r);
    return r.fetchLocal();
}

}  // namespace Proto
