// Internal helpers shared by the protomi HIP sources (not part of the C ABI).
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "protomi.h"

namespace pm {

// Records the message of the last failure on this host thread.
void set_error(const char* fmt, ...);

// Grow-only device scratch private to one stream (per-block partials of the
// reductions).  Kernels on different streams never share it; growing it
// synchronises that stream first.  A buffer handed out during a graph capture
// is kept (not freed) when the scratch grows later, until the stream is
// destroyed.  Under capture it cannot grow: PM_ERR_UNSUPPORTED.
int workspace(hipStream_t stream, size_t bytes, void** out);

// Shape of a layout (pm_level.hip): dim, domain, box size, periodic flags, box count.
int layout_shape(const pm_layout* L, int* dim, int* domain, int* bs, int* periodic, int* nboxes);
// Box low corners in layout order (3 ints per box).
int layout_lows(const pm_layout* L, int* lows);
// pm_level_gsrb_sweep with a choice of input (pm_level.hip): in_mode 0 reads
// phi_in; 1 sweeps a zero initial guess without reading it (phi_in unused);
// 2 sweeps phi_in + the constant prolongation of `coarse` (layout C, levels
// box-aligned), the cells pm_level_prolong_add(interp 0) would have left in
// phi_in, which is not written.  acc (FP32 sweeps, in_mode 0): phi_out is not
// written; the result is added to the FP64 array acc (layout and ghosts of
// phi_out) as pm_level_add_correction adds it.  PM_ERR_UNSUPPORTED where the
// one-pass sweep does not apply.
int level_gsrb_sweep_in(const pm_layout* L, int dtype, const void* phi_in, const int in_ghost[3], void* phi_out,
                        const int out_ghost[3], const void* rhs, const int rhs_ghost[3], double dx, int in_mode,
                        const pm_layout* C, const void* coarse, const int coarse_ghost[3], hipStream_t stream,
                        double* acc = nullptr);
// nsweeps GSRB sweeps (pm_level_gsrb's colouring and arithmetic, bit for bit)
// of a one-box periodic 3-D level of at most 4096 cells in one launch, in
// place; in_mode as level_gsrb_sweep_in (1 zero guess, 2 + constant
// prolongation of the one-box level C).  PM_ERR_UNSUPPORTED otherwise
// (pm_level.hip).
int level_gsrb_small(const pm_layout* L, int dtype, void* phi, const int phi_ghost[3], const void* rhs,
                     const int rhs_ghost[3], double dx, int nsweeps, int in_mode, const pm_layout* C,
                     const void* coarse, const int coarse_ghost[3], hipStream_t stream);
// Every fine box's coarsened region lies in one coarse box (both layouts
// cover their domains, even fine box sizes, coarse box size a multiple of half
// the fine one) (pm_level.hip).
bool levels_box_aligned(const pm_layout* F, const pm_layout* C);
// pm_level_resnorm, pm_level_residual and pm_level_restrict_residual that
// read the halo from the neighbouring boxes' valid cells instead of the ghost
// layer, so no exchange has to precede them (pm_level.hip).  Same results as
// the exchange followed by the C-ABI call; PM_ERR_UNSUPPORTED where the level
// does not qualify (32..64-wide boxes, periodic, fully covered; restriction:
// box-aligned levels with even box sizes).
int level_resnorm_nb(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                     const int rhs_ghost[3], double dx, void* out_dev, int combine, hipStream_t stream);
int level_residual_nb(const pm_layout* L, int dtype, const void* phi, const int phi_ghost[3], const void* rhs,
                      const int rhs_ghost[3], double dx, int out_dtype, void* out, const int out_ghost[3],
                      void* norm_out, hipStream_t stream);
int level_restrict_nb(const pm_layout* F, const pm_layout* C, int dtype, const void* phi, const int phi_ghost[3],
                      const void* rhs, const int rhs_ghost[3], double dx, void* resc, const int resc_ghost[3],
                      hipStream_t stream);
// The multigrid driver's second pre-smoothing sweep (scratch, ghost 0 -> phi,
// ghost 1) with the residual restriction of level_restrict_nb in the same pass;
// PM_ERR_UNSUPPORTED where the level does not qualify (pm_level.hip).
int level_sweep_restrict(const pm_layout* F, const pm_layout* C, int dtype, const void* phi_in,
                         const int in_ghost[3], void* phi_out, const int out_ghost[3], const void* rhs,
                         const int rhs_ghost[3], double dx, void* resc, const int resc_ghost[3], hipStream_t stream);
// Kernel family (0 generic, 7, 27), dtype and coefficients of a stencil (pm_stencil.hip).
int stencil_info(const pm_stencil* st, int* kind, int* dtype, std::vector<double>* coefs);

#define PM_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::pm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                  \
            return PM_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define PM_REQUIRE(cond, ...)                   \
    do {                                        \
        if (!(cond)) {                          \
            ::pm::set_error(__VA_ARGS__);       \
            return PM_ERR_INVALID;              \
        }                                       \
    } while (0)

// Launch-error check after a kernel launch.
#define PM_CHECK_LAUNCH() PM_CHECK_HIP(hipGetLastError())

inline hipStream_t S(pm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Inclusive box with x fastest, as Proto's Box (Proto_BoxImplem.H:76-86).
struct Box3 {
    int lo[3];
    int hi[3];
};

inline Box3 box3(const pm_box* b)
{
    Box3 r;
    for (int d = 0; d < 3; d++) { r.lo[d] = b->lo[d]; r.hi[d] = b->hi[d]; }
    return r;
}

__host__ __device__ inline long box_nx(const Box3& b) { return (long)b.hi[0] - b.lo[0] + 1; }
__host__ __device__ inline long box_ny(const Box3& b) { return (long)b.hi[1] - b.lo[1] + 1; }
__host__ __device__ inline long box_nz(const Box3& b) { return (long)b.hi[2] - b.lo[2] + 1; }
__host__ __device__ inline long box_size(const Box3& b)
{
    if (b.hi[0] < b.lo[0] || b.hi[1] < b.lo[1] || b.hi[2] < b.lo[2]) { return 0; }
    return box_nx(b) * box_ny(b) * box_nz(b);
}
__host__ __device__ inline bool box_empty(const Box3& b) { return box_size(b) == 0; }
__host__ __device__ inline bool box_contains(const Box3& b, int x, int y, int z)
{
    return x >= b.lo[0] && x <= b.hi[0] && y >= b.lo[1] && y <= b.hi[1] && z >= b.lo[2] && z <= b.hi[2];
}
__host__ __device__ inline bool box_contains_box(const Box3& outer, const Box3& inner)
{
    return box_empty(inner) ||
           (box_contains(outer, inner.lo[0], inner.lo[1], inner.lo[2]) &&
            box_contains(outer, inner.hi[0], inner.hi[1], inner.hi[2]));
}
// Linear index of (x,y,z) in box b (Box::sindex).
__host__ __device__ inline long box_index(const Box3& b, int x, int y, int z)
{
    return (long)(x - b.lo[0]) + box_nx(b) * ((long)(y - b.lo[1]) + box_ny(b) * (long)(z - b.lo[2]));
}

inline bool box_valid(const pm_box* b)
{
    return b != nullptr;
}

// Offsets of Stencil<T>::Laplacian() in its coefficient order
// (Proto_StencilDefs.H:77-87): centre, +x, -x, +y, -y, +z, -z.
__host__ __device__ constexpr int lap7_off(int j, int d)
{
    constexpr int t[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    return t[j][d];
}
// Offsets of Stencil<T>::Laplacian_27() (Proto_StencilDefs.H:130-149): the
// centre, then Box::Cube(3) shifted by -1 in x-fastest order without it.
__host__ __device__ constexpr int lap27_off(int j, int d)
{
    int k = (j == 0) ? 13 : ((j - 1) >= 13 ? j : j - 1);
    int v = (d == 0) ? (k % 3) : (d == 1 ? (k / 3) % 3 : k / 9);
    return v - 1;
}
// SHAPE 7 / 27: the two Laplacians; SHAPE 1: one tap at the origin
// (Stencil<T>::Identity(), Proto_Stencil.H:620, and its multiples)
template<int SHAPE> __host__ __device__ constexpr int shape_off(int j, int d)
{
    return SHAPE == 7 ? lap7_off(j, d) : (SHAPE == 27 ? lap27_off(j, d) : 0);
}

// Compile-time loop: f(integral_constant<int, I>) for I = B .. E - 1.
template<int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// x neighbours across a wave (row-register sweeps): lane l takes lane l - 1
// (wave_shr:1) or l + 1 (wave_shl:1); the lane with no source (0 or 63) keeps
// `edge`.
constexpr int kShr1 = 0x138, kShl1 = 0x130;
template<int CTRL>
__device__ __forceinline__ float lane_shift(float edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template<int CTRL>
__device__ __forceinline__ double lane_shift(double edge, double v)
{
    const long long e = __double_as_longlong(edge), x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)e, (int)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// XCD-aware block remap: MI355X dispatches workgroup i to XCD (i mod 8); this
// hands each XCD a contiguous run of logical tiles so that tiles sharing halo
// rows share that XCD's L2.
__device__ inline unsigned xcd_remap(unsigned b, unsigned n)
{
    const unsigned full = (n / 8u) * 8u;
    if (b >= full) { return b; }
    return (b % 8u) * (n / 8u) + b / 8u;
}

template<typename T> struct dtype_of;
template<> struct dtype_of<float> { static constexpr int value = PM_F32; };
template<> struct dtype_of<double> { static constexpr int value = PM_F64; };

inline size_t dtype_size(int dt) { return dt == PM_F32 ? 4 : 8; }

// ---------------------------------------------------------------- reductions
// Reduction operators of Proto_Reduction.H:15 with the init/update rules of
// Proto_ReductionImplem.H:64-102.
template<typename T> __host__ __device__ inline T red_init(int op)
{
    switch (op) {
        case PM_RED_MAX: return sizeof(T) == 8 ? (T)2.2250738585072014e-308 : (T)1.17549435e-38f;
        case PM_RED_MIN: return sizeof(T) == 8 ? (T)1.7976931348623157e308 : (T)3.40282347e38f;
        default: return (T)0;
    }
}

template<typename T> __device__ inline T red_update(int op, T v, T x)
{
    switch (op) {
        case PM_RED_MAX: return (v < x) ? x : v;
        case PM_RED_MIN: return (x < v) ? x : v;
        case PM_RED_ABS: { T a = (x > 0) ? x : -x; return (v < a) ? a : v; }
        case PM_RED_SUM: return v + x;
        case PM_RED_SUMABS: return v + ((x > 0) ? x : -x);
        case PM_RED_SUMSQ: return v + x * x;
        default: return v;
    }
}

// Combining two partial values of the same operator.
template<typename T> __device__ inline T red_combine(int op, T a, T b)
{
    switch (op) {
        case PM_RED_MAX: return (a < b) ? b : a;
        case PM_RED_MIN: return (b < a) ? b : a;
        case PM_RED_ABS: return (a < b) ? b : a;
        default: return a + b;
    }
}

// Wave64 reduction with a fixed butterfly: the result is the same on every
// run for the same inputs (the combination tree depends only on lane ids).
template<typename T> __device__ inline T wave_reduce(int op, T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = red_combine(op, v, o);
    }
    return v;
}

// Block reduction for up to 16 waves; every thread returns the block value.
template<typename T> __device__ inline T block_reduce(int op, T v, T* lds /* >= 16 */)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    v = wave_reduce(op, v);
    __syncthreads();
    if (lane == 0) { lds[w] = v; }
    __syncthreads();
    T r = lds[0];
    for (int i = 1; i < nw; i++) { r = red_combine(op, r, lds[i]); }
    return r;
}

}  // namespace pm
