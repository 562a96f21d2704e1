// ProblemDomain, DisjointBoxLayout, LevelBoxData and averageDown
// (include/base/Proto_ProblemDomain.H, Proto_DisjointBoxLayout.H,
// Proto_LevelBoxData.H, implem/Proto_LevelBoxDataImplem.H).  Boxes iterate in
// Proto's order (Morton order of the patch points, Proto_BoxPartitionImplem.H:
// 51-53 with Proto_Morton.H); one device allocation holds every box of a
// LevelBoxData, and exchange / copyTo / setVal run as one launch per level.
#pragma once

#include <array>
#include <map>
#include <memory>

#include "Proto_InterpStencil.H"

namespace Proto {

class ProblemDomain {
public:
    ProblemDomain() {}
    ProblemDomain(const Box& b, bool periodic) : m_box(b) { m_periodic.fill(periodic); }
    ProblemDomain(const Box& b, const std::array<bool, DIM>& periodic) : m_box(b), m_periodic(periodic) {}
    const Box& box() const { return m_box; }
    Point sizes() const { return m_box.sizes(); }
    bool isPeriodic(int d) const { return m_periodic[d]; }
    const std::array<bool, DIM>& periodicity() const { return m_periodic; }
    ProblemDomain coarsen(const Point& r) const { return ProblemDomain(m_box.coarsen(r), m_periodic); }
    ProblemDomain refine(const Point& r) const { return ProblemDomain(m_box.refine(r), m_periodic); }
    bool operator==(const ProblemDomain& o) const { return m_box == o.m_box && m_periodic == o.m_periodic; }

private:
    Box m_box;
    std::array<bool, DIM> m_periodic{};
};

// Morton index: bit i of coordinate d goes to bit DIM*i + d (Proto_MortonImplem.H).
inline uint64_t mortonIndex(const Point& p)
{
    uint64_t r = 0;
    for (int i = 0; i < 21; i++)
        for (int d = 0; d < DIM; d++) { r |= (uint64_t)((p[d] >> i) & 1) << (DIM * i + d); }
    return r;
}

using DataIndex = int;

class DisjointBoxLayout {
public:
    DisjointBoxLayout() {}
    DisjointBoxLayout(const ProblemDomain& pd, const Point& boxSize) { define(pd, boxSize); }
    DisjointBoxLayout(const ProblemDomain& pd, const std::vector<Point>& patches, const Point& boxSize)
    {
        define(pd, patches, boxSize);
    }
    void define(const ProblemDomain& pd, const Point& boxSize)
    {
        Box pb = pd.box().coarsen(boxSize);
        std::vector<Point> patches;
        for (auto it = pb.begin(); it != pb.end(); ++it) { patches.push_back(*it); }
        define(pd, patches, boxSize);
    }
    void define(const ProblemDomain& pd, std::vector<Point> patches, const Point& boxSize)
    {
        if (!pd.box().coarsenable(boxSize)) {
            throw ProtoError(PM_ERR_INVALID, "DisjointBoxLayout: domain cannot be tiled by the box size");
        }
        std::stable_sort(patches.begin(), patches.end(),
                         [](const Point& a, const Point& b) { return mortonIndex(a) < mortonIndex(b); });
        m_impl = std::make_shared<Impl>();
        m_impl->domain = pd;
        m_impl->boxSize = boxSize;
        m_impl->patches = patches;
    }
    int size() const { return m_impl ? (int)m_impl->patches.size() : 0; }
    Box operator[](DataIndex i) const
    {
        const Point lo = m_impl->patches[i] * m_impl->boxSize;
        return Box(lo, lo + m_impl->boxSize - Point::Ones());
    }
    Point point(DataIndex i) const { return m_impl->patches[i]; }
    const Point& boxSize() const { return m_impl->boxSize; }
    const ProblemDomain& domain() const { return m_impl->domain; }
    bool coarsenable(const Point& r) const { return (m_impl->boxSize % r) == Point::Zeros(); }
    DisjointBoxLayout coarsen(const Point& r) const
    {
        DisjointBoxLayout c;
        c.m_impl = std::make_shared<Impl>();
        c.m_impl->domain = m_impl->domain.coarsen(r);
        c.m_impl->boxSize = m_impl->boxSize / r;
        c.m_impl->patches = m_impl->patches;
        return c;
    }
    // layout index of a patch point or of its periodic image, -1 if none
    // (DisjointBoxLayout::find, Proto_DisjointBoxLayoutImplem.H:262-275 with
    // ProblemDomain::image of the patch domain)
    int find(const Point& pt) const
    {
        const Box pb = m_impl->domain.box().coarsen(m_impl->boxSize);
        Point q = pt;
        for (int d = 0; d < DIM; d++) {
            const int n = pb.size(d);
            if (q[d] < pb.low()[d] || q[d] > pb.high()[d]) {
                if (!m_impl->domain.isPeriodic(d)) { return -1; }
                q[d] = pb.low()[d] + (((q[d] - pb.low()[d]) % n) + n) % n;
            }
        }
        for (int i = 0; i < size(); i++) {
            if (m_impl->patches[i] == q) { return i; }
        }
        return -1;
    }
    bool contains(const Point& pt) const { return find(pt) >= 0; }
    // a patch with a missing neighbour patch among its 3^DIM - 1 neighbours
    // (onLevelBoundary, Proto_DisjointBoxLayoutImplem.H:316-330)
    bool onLevelBoundary(const Point& pt) const
    {
        const Box K = Box::Kernel(1).shift(pt);
        for (auto it = K.begin(); it != K.end(); ++it) {
            if (*it == pt) { continue; }
            if (find(*it) < 0) { return true; }
        }
        return false;
    }
    bool operator==(const DisjointBoxLayout& o) const { return m_impl == o.m_impl; }
    bool compatible(const DisjointBoxLayout& o) const
    {
        return m_impl && o.m_impl && m_impl->domain == o.m_impl->domain && m_impl->boxSize == o.m_impl->boxSize &&
               m_impl->patches == o.m_impl->patches;
    }

    struct iterator {
        int i;
        int n;
        DataIndex operator*() const { return i; }
        iterator& operator++() { ++i; return *this; }
        bool operator!=(const iterator& o) const { return i != o.i; }
        bool ok() const { return i < n; }
    };
    iterator begin() const { return {0, size()}; }
    iterator end() const { return {size(), size()}; }

    // the device layout handle (created on first use)
    pm_layout* handle() const
    {
        if (!m_impl->handle) {
            int dom[3] = {1, 1, 1}, bs[3] = {1, 1, 1}, per[3] = {0, 0, 0};
            for (int d = 0; d < DIM; d++) {
                dom[d] = m_impl->domain.box().size(d);
                bs[d] = m_impl->boxSize[d];
                per[d] = m_impl->domain.isPeriodic(d);
            }
            std::vector<int> lows(3 * size(), 0);
            for (int b = 0; b < size(); b++)
                for (int d = 0; d < DIM; d++) { lows[3 * b + d] = (*this)[b].low()[d]; }
            pm_layout* h = nullptr;
            pm_check(pm_layout_create(&h, DIM, dom, bs, per, size(), lows.data()), "DisjointBoxLayout");
            m_impl->handle.reset(h, [](pm_layout* p) { pm_layout_destroy(p); });
        }
        return m_impl->handle.get();
    }

private:
    struct Impl {
        ProblemDomain domain;
        Point boxSize;
        std::vector<Point> patches;
        std::shared_ptr<pm_layout> handle;
    };
    std::shared_ptr<Impl> m_impl;
};

template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, Centering CTR = PR_CELL>
class LevelBoxData {
    static_assert(MEM == DEVICE, "LevelBoxData on this library lives on the device");
    static_assert(CTR == PR_CELL, "LevelBoxData on this library holds cell-centred data");

public:
    LevelBoxData() {}
    LevelBoxData(const DisjointBoxLayout& layout, const Point& ghost) { define(layout, ghost); }
    LevelBoxData(const LevelBoxData&) = delete;
    LevelBoxData& operator=(const LevelBoxData&) = delete;

    void define(const DisjointBoxLayout& layout, const Point& ghost)
    {
        m_layout = layout;
        m_ghost = ghost;
        const size_t per = (size_t)layout[0].grow(ghost).size() * C;
        const size_t bytes = sizeof(T) * per * layout.size();
        T* p = (T*)proto_malloc<DEVICE>(bytes);
        m_data = std::shared_ptr<T>(p, [](T* q) { proto_free<DEVICE>(q); });
        m_boxes.clear();
        for (int b = 0; b < layout.size(); b++) { m_boxes.emplace_back(p + per * b, layout[b].grow(ghost), C); }
    }
    BoxData<T, C, MEM>& operator[](DataIndex i) { return m_boxes[i]; }
    const BoxData<T, C, MEM>& operator[](DataIndex i) const { return m_boxes[i]; }
    const DisjointBoxLayout& layout() const { return m_layout; }
    const Point& ghost() const { return m_ghost; }
    Box patchBox(DataIndex i) const { return m_layout[i]; }
    T* data() { return m_data.get(); }
    const T* data() const { return m_data.get(); }
    DisjointBoxLayout::iterator begin() const { return m_layout.begin(); }
    DisjointBoxLayout::iterator end() const { return m_layout.end(); }

    // LevelBoxData::exchange (Proto_LevelBoxDataImplem.H:313-341)
    void exchange()
    {
        if (m_ghost == Point::Zeros()) { return; }
        int g[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { g[d] = m_ghost[d]; }
        pm_check(pm_level_exchange(m_layout.handle(), pm_type<T>::value, data(), g, C, protoCurrentStream()),
                 "LevelBoxData::exchange");
    }
    void setVal(T v)
    {
        int g[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { g[d] = m_ghost[d]; }
        pm_check(pm_level_setval(m_layout.handle(), pm_type<T>::value, data(), g, C, (double)v, protoCurrentStream()),
                 "LevelBoxData::setVal");
    }
    void setToZero() { setVal(T(0)); }
    // copyTo another layout of the same domain (343-386): compatible layouts
    // copy box by box over the common ghosted region and exchange the
    // destination (copyToSimple); others fill every ghosted destination cell
    // that a source patch (or its periodic image) covers (LevelCopier)
    void copyTo(LevelBoxData& dst) const
    {
        int gs[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { gs[d] = m_ghost[d]; gd[d] = dst.m_ghost[d]; }
        pm_check(pm_level_copy_to(m_layout.handle(), data(), gs, dst.m_layout.handle(), dst.data(), gd,
                                  pm_type<T>::value, C, protoCurrentStream()),
                 "LevelBoxData::copyTo");
    }
    // LevelBoxData::reduce / absMax / sum / integrate over valid cells (595-678)
    template<Operation OP> T reduce(unsigned int comp = 0) const
    {
        Reduction<T, OP, MEM> rxn;
        for (auto it : m_layout) { rxn.reduceBox(m_boxes[it].data(), m_boxes[it].box(), m_layout[it], comp); }
        return rxn.fetch();
    }
    T absMax(unsigned int comp = 0) const { return reduce<Abs>(comp); }
    T sum(unsigned int comp = 0) const { return reduce<Sum>(comp); }
    T max(unsigned int comp = 0) const { return reduce<Max>(comp); }
    T min(unsigned int comp = 0) const { return reduce<Min>(comp); }
    T integrate(T dx, unsigned int comp = 0) const { return reduce<Sum>(comp) * std::pow(dx, DIM); }
    // linearSize / linearOut / linearIn (Proto_LevelBoxDataImplem.H:412-507): the
    // buffer is device memory (Proto's MemType rule), box after box in layout
    // order, component after component over each ghosted box
    size_t linearSize() const
    {
        size_t b = 0;
        int g[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { g[d] = m_ghost[d]; }
        pm_check(pm_level_linear_size(m_layout.handle(), pm_type<T>::value, g, C, &b), "LevelBoxData::linearSize");
        return b;
    }
    void linearOut(void* buffer) const
    {
        int g[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { g[d] = m_ghost[d]; }
        pm_check(pm_level_linear_out(m_layout.handle(), pm_type<T>::value, data(), g, C, buffer, 0,
                                     protoCurrentStream()),
                 "LevelBoxData::linearOut");
    }
    void linearIn(const void* buffer)
    {
        int g[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) { g[d] = m_ghost[d]; }
        pm_check(pm_level_linear_in(m_layout.handle(), pm_type<T>::value, data(), g, C, buffer, 0,
                                    protoCurrentStream()),
                 "LevelBoxData::linearIn");
    }
    void increment(const LevelBoxData& o, T scale = 1)
    {
        for (auto it : m_layout) { m_boxes[it].increment(o[it], scale); }
    }
    // initialize (Proto_LevelBoxDataImplem.H:176-187): a pointwise functor run
    // with forallInPlace_p on every ghosted box (needs hipcc); the extra
    // arguments are passed to every box as they are
    template<typename Func, typename... Srcs>
    void initialize(const Func& f, Srcs&&... srcs)
    {
        for (auto it : m_layout) { forallInPlace_p(f, m_boxes[it], std::forward<Srcs>(srcs)...); }
    }

private:
    DisjointBoxLayout m_layout;
    Point m_ghost;
    std::shared_ptr<T> m_data;
    std::vector<BoxData<T, C, MEM>> m_boxes;
};

// averageDown (Proto_LevelBoxDataImplem.H:992-1038): exchange the fine level,
// AvgDown each box onto the coarsened fine layout, copy into the coarse layout.
template<typename T, unsigned int C, MemType MEM>
void averageDown(LevelBoxData<T, C, MEM>& crse, LevelBoxData<T, C, MEM>& fine, Point ratio)
{
    LevelBoxData<T, C, MEM> temp(fine.layout().coarsen(ratio), Point::Zeros());
    fine.exchange();
    auto AVG = Stencil<T>::AvgDown(ratio);
    for (auto it : fine.layout()) { temp[it] |= AVG(fine[it]); }
    temp.copyTo(crse);
}
template<typename T, unsigned int C, MemType MEM>
void averageDown(LevelBoxData<T, C, MEM>& crse, LevelBoxData<T, C, MEM>& fine, int ratio)
{
    averageDown(crse, fine, Point::Ones(ratio));
}

// interpBoundaries (Proto_LevelBoxDataImplem.H:953-990): coarse data copied
// onto the coarsened fine layout (with the ghosts the interpolation needs),
// then on every fine patch at the level's boundary the InterpStencil output
// overwrites the ghosted box and the valid cells are put back; finally the
// fine level is exchanged, so ghosts facing other fine patches hold fine data.
template<typename T, unsigned int C, MemType MEM>
void interpBoundaries(LevelBoxData<T, C, MEM>& crse, LevelBoxData<T, C, MEM>& fine,
                      LevelBoxData<T, C, MEM>& crseFine, InterpStencil<T>& interp)
{
    crse.exchange();
    crse.copyTo(crseFine);
    const DisjointBoxLayout& fl = fine.layout();
    for (auto it : fl) {
        if (!fl.onLevelBoundary(fl.point(it))) { continue; }
        BoxData<T, C, MEM>& fine_i = fine[it];
        BoxData<T, C, MEM>& crse_i = crseFine[it];
        BoxData<T, C, MEM> fine_0(fl[it]);
        fine_i.copyTo(fine_0);  // the valid cells
        fine_i |= interp(crse_i);
        fine_0.copyTo(fine_i);
    }
    fine.exchange();
}
template<typename T, unsigned int C, MemType MEM>
void interpBoundaries(LevelBoxData<T, C, MEM>& crse, LevelBoxData<T, C, MEM>& fine, InterpStencil<T>& interp)
{
    const Point r = interp.ratio();
    LevelBoxData<T, C, MEM> temp(fine.layout().coarsen(r), fine.ghost() / r + Point::Ones() + interp.ghost());
    interpBoundaries(crse, fine, temp, interp);
}

}  // namespace Proto
