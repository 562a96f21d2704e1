// Reduction<T,OP,DEVICE> (Proto_ReductionImplem.H:157-192) for gfx950:
// grid-stride local folds, a fixed wave64 butterfly and an LDS tree per block,
// then one block folds the per-block partials in a fixed order.  The grid
// depends only on n, so results are reproducible run to run.
#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

template<typename T>
__global__ __launch_bounds__(256) void k_reduce1(int op, const T* __restrict__ data, size_t n,
                                                 T* __restrict__ partials)
{
    __shared__ T red[16];
    T v = red_init<T>(op);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        v = red_update(op, v, data[i]);
    }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { partials[blockIdx.x] = v; }
}

// Over the cells of `region` inside an array on `box` (component comp).
template<typename T>
__global__ __launch_bounds__(256) void k_reduce1_box(int op, const T* __restrict__ data, Box3 box, Box3 region,
                                                     T* __restrict__ partials)
{
    __shared__ T red[16];
    T v = red_init<T>(op);
    const long nx = box_nx(region), ny = box_ny(region);
    const long n = box_size(region);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = region.lo[0] + (int)(t % nx);
        const long r = t / nx;
        const int y = region.lo[1] + (int)(r % ny);
        const int z = region.lo[2] + (int)(r / ny);
        v = red_update(op, v, data[box_index(box, x, y, z)]);
    }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { partials[blockIdx.x] = v; }
}

template<typename T>
__global__ __launch_bounds__(1024) void k_reduce2(int op, const T* __restrict__ partials, int nb, T* out, int combine)
{
    __shared__ T red[16];
    T v = red_init<T>(op);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) { v = red_combine(op, v, partials[i]); }
    v = block_reduce(op, v, red);
    if (threadIdx.x == 0) { *out = combine ? red_combine(op, *out, v) : v; }
}

template<typename T>
__global__ void k_reduce_init(int op, T* out) { *out = red_init<T>(op); }

static int reduce_blocks(long n) { return (int)std::max(1L, std::min((n + 1023) / 1024, 1024L)); }

template<typename T>
static int reduce_t(int op, const T* data, size_t n, T* out, int combine, hipStream_t s)
{
    const int nb = reduce_blocks((long)n);
    void* ws = nullptr;
    int rc = workspace(s, sizeof(T) * nb, &ws);
    if (rc) { return rc; }
    hipLaunchKernelGGL(k_reduce1<T>, dim3(nb), dim3(256), 0, s, op, data, n, (T*)ws);
    PM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce2<T>, dim3(1), dim3(1024), 0, s, op, (const T*)ws, nb, out, combine);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

template<typename T>
static int reduce_box_t(int op, const T* data, const Box3& box, const Box3& region, T* out, int combine, hipStream_t s)
{
    const int nb = reduce_blocks(box_size(region));
    void* ws = nullptr;
    int rc = workspace(s, sizeof(T) * nb, &ws);
    if (rc) { return rc; }
    hipLaunchKernelGGL(k_reduce1_box<T>, dim3(nb), dim3(256), 0, s, op, data, box, region, (T*)ws);
    PM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce2<T>, dim3(1), dim3(1024), 0, s, op, (const T*)ws, nb, out, combine);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {

int pm_reduce(int op, int dtype, const void* data, size_t n, void* out_dev, int combine, pm_stream_t stream)
{
    PM_REQUIRE(op >= PM_RED_MAX && op <= PM_RED_SUMSQ, "pm_reduce: bad op %d", op);
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_reduce: bad dtype %d", dtype);
    PM_REQUIRE(out_dev && (data || n == 0), "pm_reduce: null pointer");
    if (dtype == PM_F64) { return reduce_t<double>(op, (const double*)data, n, (double*)out_dev, combine, S(stream)); }
    return reduce_t<float>(op, (const float*)data, n, (float*)out_dev, combine, S(stream));
}

int pm_reduce_box(int op, int dtype, const void* data, const pm_box* box, const pm_box* region, int comp,
                  void* out_dev, int combine, pm_stream_t stream)
{
    PM_REQUIRE(op >= PM_RED_MAX && op <= PM_RED_SUMSQ, "pm_reduce_box: bad op %d", op);
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_reduce_box: bad dtype %d", dtype);
    PM_REQUIRE(data && box && region && out_dev && comp >= 0, "pm_reduce_box: bad argument");
    const Box3 b = box3(box), r = box3(region);
    PM_REQUIRE(box_contains_box(b, r), "pm_reduce_box: region outside the array's box");
    const long off = (long)comp * box_size(b);
    if (dtype == PM_F64) {
        return reduce_box_t<double>(op, (const double*)data + off, b, r, (double*)out_dev, combine, S(stream));
    }
    return reduce_box_t<float>(op, (const float*)data + off, b, r, (float*)out_dev, combine, S(stream));
}

int pm_reduce_init(int op, int dtype, void* out_dev, pm_stream_t stream)
{
    PM_REQUIRE(op >= PM_RED_MAX && op <= PM_RED_SUMSQ, "pm_reduce_init: bad op %d", op);
    PM_REQUIRE(out_dev, "pm_reduce_init: null out");
    if (dtype == PM_F64) { hipLaunchKernelGGL(k_reduce_init<double>, dim3(1), dim3(1), 0, S(stream), op, (double*)out_dev); }
    else { hipLaunchKernelGGL(k_reduce_init<float>, dim3(1), dim3(1), 0, S(stream), op, (float*)out_dev); }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // extern "C"
