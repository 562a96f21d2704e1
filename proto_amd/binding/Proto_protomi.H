// Proto-side binding of the protomi C ABI (include/protomi.h): what Proto's
// own MemType::DEVICE code paths call in place of their kernels.  It is written
// against Proto's public class interfaces only (Stencil::coefs / offsets /
// srcRatio / destRatio / destShift, BoxData::data / box, LevelBoxData,
// DisjointBoxLayout), so a maintainer includes it from Proto's implem headers
// and each DEVICE body becomes one call:
//
//   Stencil<T>::protoApply<DEVICE>      Proto_StencilImplem.H:843-900
//       -> protomi::apply(m_pm, a_src, a_dst, a_box, a_initToZero, a_scale, stream)
//          with m_pm a protomi::StencilHandle<T> built once (replaces
//          closeForDevice, Proto_StencilImplem.H:539-556)
//   Reduction<T,OP,DEVICE>::reduce       Proto_ReductionImplem.H:157-192
//       -> protomi::reduce<T,OP>(a_data, a_size, m_deviTotal, stream)
//   Reduction<T,OP,DEVICE>::reset        Proto_ReductionImplem.H:104-118
//       -> protomi::reset<T,OP>(m_deviTotal, stream)
//   LevelBoxData::exchange (one device) Proto_LevelBoxDataImplem.H:313-341
//       -> protomi::exchange(layoutHandle, *this, stream)
//   LevelMultigridSolver<BoxOp_Laplace> Proto_LevelMultigridSolverImplem.H:85-160
//       -> protomi::FasSolver (pm_fas_*)
//
// Errors: every C-ABI status other than PM_OK becomes MayDay-style abort in
// Proto; here the functions return the status and the caller decides
// (pm_last_error() holds the message).  Nothing in this header allocates
// device memory except the handles' own tables.
#pragma once

#include <cstring>
#include <type_traits>
#include <vector>

#include "protomi.h"

namespace Proto {
namespace protomi {

inline pm_box toPm(const Box& b)
{
    pm_box r;
    for (int d = 0; d < 3; d++) {
        r.lo[d] = d < DIM ? b.low()[d] : 0;
        r.hi[d] = d < DIM ? b.high()[d] : 0;
    }
    return r;
}

template<typename T> constexpr int dtype() { return std::is_same<T, double>::value ? PM_F64 : PM_F32; }

// Proto's Operation (Proto_Reduction.H:15) has the C ABI's order.
template<Operation OP> constexpr int redOp() { return (int)OP; }

// The device copy of a Stencil's coefficients, offsets, ratios and shift.
template<typename T>
class StencilHandle {
public:
    StencilHandle() {}
    explicit StencilHandle(const Stencil<T>& S) { define(S); }
    StencilHandle(const StencilHandle&) = delete;
    StencilHandle& operator=(const StencilHandle&) = delete;
    ~StencilHandle() { release(); }
    int define(const Stencil<T>& S)
    {
        release();
        const std::vector<T>& c = S.coefs();
        const std::vector<Point>& o = S.offsets();
        std::vector<int> offs(3 * c.size(), 0);
        for (size_t j = 0; j < c.size(); j++)
            for (int d = 0; d < DIM; d++) { offs[3 * j + d] = o[j][d]; }
        int sr[3] = {1, 1, 1}, dr[3] = {1, 1, 1}, ds[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) {
            sr[d] = S.srcRatio()[d];
            dr[d] = S.destRatio()[d];
            ds[d] = S.destShift()[d];
        }
        return pm_stencil_create(&m_h, dtype<T>(), DIM, (int)c.size(), c.data(), offs.data(), sr, dr, ds);
    }
    pm_stencil* get() const { return m_h; }

private:
    void release()
    {
        if (m_h) { pm_stencil_destroy(m_h); }
        m_h = nullptr;
    }
    pm_stencil* m_h = nullptr;
};

// Stencil<T>::protoApply on DEVICE data: dst(box) (|= or +=) S(src, scale),
// every component, Proto's hostApply arithmetic bit for bit.
template<typename T, unsigned int C, unsigned int D, unsigned int E>
int apply(const StencilHandle<T>& S, const BoxData<T, C, DEVICE, D, E>& src, BoxData<T, C, DEVICE, D, E>& dst,
          const Box& box, bool initToZero, T scale, pm_stream_t stream)
{
    if (box.size() == 0) { return PM_OK; }
    const pm_box sb = toPm(src.box()), db = toPm(dst.box()), ib = toPm(box);
    return pm_stencil_apply(S.get(), src.data(), &sb, dst.data(), &db, &ib, (int)(C * D * E), initToZero ? 1 : 0,
                            (double)scale, stream);
}

// Reduction<T,OP,DEVICE>::reduce / reset on its device total.
template<typename T, Operation OP>
int reduce(const T* data, size_t n, T* devTotal, pm_stream_t stream)
{
    return pm_reduce(redOp<OP>(), dtype<T>(), data, n, devTotal, 1, stream);
}
template<typename T, Operation OP>
int reset(T* devTotal, pm_stream_t stream)
{
    return pm_reduce_init(redOp<OP>(), dtype<T>(), devTotal, stream);
}

// The device table of a DisjointBoxLayout: its domain, box size, periodicity
// and the low corners of its boxes in Proto's iteration order.
class LayoutHandle {
public:
    LayoutHandle() {}
    explicit LayoutHandle(const DisjointBoxLayout& L) { define(L); }
    LayoutHandle(const LayoutHandle&) = delete;
    LayoutHandle& operator=(const LayoutHandle&) = delete;
    ~LayoutHandle() { release(); }
    int define(const DisjointBoxLayout& L)
    {
        release();
        int dom[3] = {1, 1, 1}, bs[3] = {1, 1, 1}, per[3] = {0, 0, 0};
        const ProblemDomain pd = L.domain();
        for (int d = 0; d < DIM; d++) {
            dom[d] = pd.box().size(d);
            bs[d] = L.boxSize()[d];
            per[d] = pd.isPeriodic(d) ? 1 : 0;
        }
        std::vector<int> lows;
        for (auto iter : L) {
            const Box b = L[iter];
            for (int d = 0; d < 3; d++) { lows.push_back(d < DIM ? b.low()[d] : 0); }
        }
        return pm_layout_create(&m_h, DIM, dom, bs, per, (int)(lows.size() / 3), lows.data());
    }
    pm_layout* get() const { return m_h; }

private:
    void release()
    {
        if (m_h) { pm_layout_destroy(m_h); }
        m_h = nullptr;
    }
    pm_layout* m_h = nullptr;
};

// Base pointer of level data stored as the C ABI expects it: box b, component
// c at (b * C + c) * |ghosted box|.  Proto allocates one BoxData per box; a
// maintainer allocates the level once and aliases each box into it (raw
// pointer BoxData, Proto_BoxDataImplem.H:99-111).  nullptr when the boxes are
// not laid out that way.
template<typename T, unsigned int C>
T* contiguousBase(LevelBoxData<T, C, DEVICE>& L)
{
    T* base = nullptr;
    size_t per = 0, b = 0;
    for (auto iter : L.layout()) {
        BoxData<T, C, DEVICE>& B = L[iter];
        if (b == 0) {
            base = B.data();
            per = (size_t)B.box().size() * C;
        } else if (B.data() != base + b * per) {
            return nullptr;
        }
        b++;
    }
    return base;
}

// LevelBoxData::exchange on one device: every ghost cell from its neighbour
// box or periodic image.
template<typename T, unsigned int C>
int exchange(const LayoutHandle& layout, LevelBoxData<T, C, DEVICE>& L, pm_stream_t stream)
{
    T* base = contiguousBase(L);
    if (!base) { return PM_ERR_UNSUPPORTED; }
    int g[3] = {0, 0, 0};
    for (int d = 0; d < DIM; d++) { g[d] = L.ghost()[d]; }
    return pm_level_exchange(layout.get(), dtype<T>(), base, g, (int)C, stream);
}

// LevelMultigridSolver<BoxOp_Laplace, double, DEVICE>: the resident FAS solver.
class FasSolver {
public:
    FasSolver() {}
    FasSolver(const FasSolver&) = delete;
    FasSolver& operator=(const FasSolver&) = delete;
    ~FasSolver()
    {
        if (m_h) { pm_fas_destroy(m_h); }
    }
    int define(const LayoutHandle& layout, int numLevels, double dx)
    {
        return pm_fas_create(&m_h, layout.get(), numLevels, dx);
    }
    // solve (85-132): returns the status; hist (maxIter + 1 values) and iters as Proto's loop leaves them
    int solve(LevelBoxData<double, 1, DEVICE>& state, LevelBoxData<double, 1, DEVICE>& force, int maxIter,
              double tol, double* hist, int* iters, pm_stream_t stream)
    {
        double* s = contiguousBase(state);
        double* f = contiguousBase(force);
        if (!s || !f) { return PM_ERR_UNSUPPORTED; }
        return pm_fas_solve(m_h, s, f, maxIter, tol, hist, iters, stream);
    }
    // resnorm (134-160)
    int resnorm(LevelBoxData<double, 1, DEVICE>& residual, LevelBoxData<double, 1, DEVICE>& state,
                LevelBoxData<double, 1, DEVICE>& force, double* out, pm_stream_t stream)
    {
        double* r = contiguousBase(residual);
        double* s = contiguousBase(state);
        double* f = contiguousBase(force);
        if (!r || !s || !f) { return PM_ERR_UNSUPPORTED; }
        return pm_fas_residual(m_h, s, f, r, out, stream);
    }

private:
    pm_fas* m_h = nullptr;
};

}  // namespace protomi
}  // namespace Proto
