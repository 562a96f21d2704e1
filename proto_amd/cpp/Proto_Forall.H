// forall / forallInPlace / forall_p / forallInPlace_p and Var
// (include/base/Proto_Forall.H, implem/Proto_ForallImplem.H,
// Proto_BoxData.H:115-292).  A user functor cannot cross a C ABI, so this is
// a header template: its kernel is instantiated in the user's translation
// unit, which must be compiled by hipcc for gfx950 (-ffp-contract=off keeps
// the functor's arithmetic rounded the way Proto's CPU build rounds it).
// The iteration box is the intersection of the BoxData arguments' boxes
// (forall_parse_args, Proto_ForallImplem.H:57-106).
#pragma once

#include <tuple>
#include <type_traits>
#include <utility>

#include "Proto_Stencil.H"

#ifndef __HIPCC__
#error "Proto_Forall.H must be compiled with hipcc (device lambdas)"
#endif
#include <hip/hip_runtime.h>

#define PROTO_KERNEL_START __host__ __device__
#define PROTO_KERNEL_END(local, name) auto name = [] __host__ __device__ (auto&&... a) { return local(a...); };
#define PROTO_LAMBDA __host__ __device__

namespace Proto {

// A point-wise view: v(c, d, e) is component (c, d, e) at the current cell.
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1>
struct Var {
    T* m_ptr;
    long m_stride;  // elements between components (|box|)
    __host__ __device__ T& operator()(unsigned int c, unsigned int d = 0, unsigned int e = 0) const
    {
        return m_ptr[((long)e * D * C + (long)d * C + c) * m_stride];
    }
};

namespace forall_detail {

struct Pt {
    int v[3];
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
struct VarDesc {
    T* base;
    pm_box box;
    __device__ Var<T, C, MEM, D, E> at(int x, int y, int z) const
    {
        const long nx = box.hi[0] - box.lo[0] + 1, ny = box.hi[1] - box.lo[1] + 1, nz = box.hi[2] - box.lo[2] + 1;
        const long i = (long)(x - box.lo[0]) + nx * ((long)(y - box.lo[1]) + ny * (long)(z - box.lo[2]));
        return Var<T, C, MEM, D, E>{base + i, nx * ny * nz};
    }
};

template<typename V>
struct ScalarDesc {
    V v;
    __device__ V at(int, int, int) const { return v; }
};

template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
VarDesc<T, C, MEM, D, E> desc(BoxData<T, C, MEM, D, E>& b) { return VarDesc<T, C, MEM, D, E>{b.data(), b.box().pm()}; }
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
VarDesc<T, C, MEM, D, E> desc(const BoxData<T, C, MEM, D, E>& b)
{
    return VarDesc<T, C, MEM, D, E>{const_cast<T*>(b.data()), b.box().pm()};
}
template<typename V>
ScalarDesc<typename std::decay<V>::type> desc(V&& v) { return ScalarDesc<typename std::decay<V>::type>{v}; }

template<typename A> void clip(Box& b, const A&) {}
template<typename T, unsigned int C, MemType MEM, unsigned int D, unsigned int E>
void clip(Box& b, const BoxData<T, C, MEM, D, E>& a) { b = b.empty() ? a.box() : (b & a.box()); }
template<typename... A> Box intersect(const A&... a)
{
    Box b;
    int dummy[] = {0, (clip(b, a), 0)...};
    (void)dummy;
    return b;
}

template<bool WITHPOINT, typename F, typename... Ds>
__global__ __launch_bounds__(256) void k_forall(F f, pm_box box, Ds... ds)
{
    const long nx = box.hi[0] - box.lo[0] + 1, ny = box.hi[1] - box.lo[1] + 1;
    const long n = nx * ny * (box.hi[2] - box.lo[2] + 1);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = box.lo[0] + (int)(t % nx);
        const long r = t / nx;
        const int y = box.lo[1] + (int)(r % ny);
        const int z = box.lo[2] + (int)(r / ny);
        auto args = std::make_tuple(ds.at(x, y, z)...);
        if constexpr (WITHPOINT) {
            Point p;
            const int c[3] = {x, y, z};
            for (int d = 0; d < DIM; d++) { p.m_tuple[d] = c[d]; }
            std::apply([&](auto&... a) { f(p, a...); }, args);
        } else {
            std::apply([&](auto&... a) { f(a...); }, args);
        }
    }
}

template<bool WITHPOINT, typename F, typename... Args>
void launch(const Box& b, F f, Args&&... args)
{
    if (b.empty()) { return; }
    const long n = (long)b.size();
    long blocks = (n + 255) / 256;
    if (blocks > (1L << 20)) { blocks = 1L << 20; }
    hipLaunchKernelGGL((k_forall<WITHPOINT, F, decltype(desc(std::forward<Args>(args)))...>), dim3((unsigned)blocks),
                       dim3(256), 0, (hipStream_t)protoCurrentStream(), f, b.pm(), desc(std::forward<Args>(args))...);
    if (hipGetLastError() != hipSuccess) { throw ProtoError(PM_ERR_HIP, "forall: kernel launch failed"); }
}

// flop accounting of the forallOp family (Proto feeds it to its timers)
inline unsigned long long& flopCounter()
{
    static unsigned long long n = 0;
    return n;
}
inline void countFlops(unsigned long long perPoint, const Box& b)
{
    flopCounter() += perPoint * (unsigned long long)b.size();
}

}  // namespace forall_detail

// total flops recorded by forallOp / forallInPlaceOp calls so far
inline unsigned long long forallFlops() { return forall_detail::flopCounter(); }

// forallInPlace(F, args...) (Proto_ForallImplem.H:257-268)
template<typename F, typename... Args>
void forallInPlace(const F& f, Args&&... args)
{
    forall_detail::launch<false>(forall_detail::intersect(args...), f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlace(const F& f, Box b, Args&&... args)
{
    forall_detail::launch<false>(b, f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlace_p(const F& f, Args&&... args)
{
    forall_detail::launch<true>(forall_detail::intersect(args...), f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlace_p(const F& f, Box b, Args&&... args)
{
    forall_detail::launch<true>(b, f, std::forward<Args>(args)...);
}

// forall<T,C>(F, args...) returns a new BoxData on the intersection (195-213)
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forall(const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<false>(b, f, out, std::forward<Args>(args)...);
    return out;
}
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forall_p(const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<true>(b, f, out, std::forward<Args>(args)...);
    return out;
}

template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forall(const F& f, Box b, Args&&... args)
{
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<false>(b, f, out, std::forward<Args>(args)...);
    return out;
}
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forall_p(const F& f, Box b, Args&&... args)
{
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<true>(b, f, out, std::forward<Args>(args)...);
    return out;
}

// forallOp / forallOp_p / forallInPlaceOp / forallInPlaceOp_p
// (Proto_ForallImplem.H:178-400): forall plus a flop count per point and a
// timer name (the count is accumulated; the name is not used).
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forallOp(unsigned long long flops, const char*, const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    forall_detail::countFlops(flops, b);
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<false>(b, f, out, std::forward<Args>(args)...);
    return out;
}
template<typename T, unsigned int C = 1, MemType MEM = MEMTYPE_DEFAULT, unsigned int D = 1, unsigned int E = 1,
         typename F, typename... Args>
BoxData<T, C, MEM, D, E> forallOp_p(unsigned long long flops, const char*, const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    forall_detail::countFlops(flops, b);
    BoxData<T, C, MEM, D, E> out(b);
    forall_detail::launch<true>(b, f, out, std::forward<Args>(args)...);
    return out;
}
template<typename F, typename... Args>
void forallInPlaceOp(unsigned long long flops, const char*, const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    forall_detail::countFlops(flops, b);
    forall_detail::launch<false>(b, f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlaceOp(unsigned long long flops, const char*, const F& f, Box b, Args&&... args)
{
    forall_detail::countFlops(flops, b);
    forall_detail::launch<false>(b, f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlaceOp_p(unsigned long long flops, const char*, const F& f, Args&&... args)
{
    Box b = forall_detail::intersect(args...);
    forall_detail::countFlops(flops, b);
    forall_detail::launch<true>(b, f, std::forward<Args>(args)...);
}
template<typename F, typename... Args>
void forallInPlaceOp_p(unsigned long long flops, const char*, const F& f, Box b, Args&&... args)
{
    forall_detail::countFlops(flops, b);
    forall_detail::launch<true>(b, f, std::forward<Args>(args)...);
}

}  // namespace Proto
