// BoxData pointwise operations on a region: copyTo with shift
// (Proto_BoxDataImplem.H:490-640), operatorT (937-983), scalar ops and setVal
// (430-486, 1260-1278), increment (1219-1234).  One thread per region cell,
// x fastest, so consecutive lanes touch consecutive addresses.
#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

enum { OP_ADD = 0, OP_SUB = 1, OP_MUL = 2, OP_DIV = 3, OP_SET = 4, OP_COPY = 5, OP_INCR = 6 };

template<typename T>
struct RegionArgs {
    T* dst;
    const T* src;
    Box3 dbox, sbox, region;  // region in source coordinates (dst = region + shift)
    int shift[3];
    int ncomp, op;
    T value;
};

template<typename T>
__global__ __launch_bounds__(256) void k_region(RegionArgs<T> a)
{
    const long nx = box_nx(a.region), ny = box_ny(a.region);
    const long n = box_size(a.region);
    const long dsz = box_size(a.dbox), ssz = a.src ? box_size(a.sbox) : 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = a.region.lo[0] + (int)(t % nx);
        const long r = t / nx;
        const int y = a.region.lo[1] + (int)(r % ny);
        const int z = a.region.lo[2] + (int)(r / ny);
        const long di = box_index(a.dbox, x + a.shift[0], y + a.shift[1], z + a.shift[2]);
        const long si = a.src ? box_index(a.sbox, x, y, z) : 0;
        for (int c = 0; c < a.ncomp; c++) {
            T* d = a.dst + di + c * dsz;
            switch (a.op) {
                case OP_COPY: *d = a.src[si + c * ssz]; break;
                case OP_ADD: *d = *d + (a.src ? a.src[si + c * ssz] : a.value); break;
                case OP_SUB: *d = *d - (a.src ? a.src[si + c * ssz] : a.value); break;
                case OP_MUL: *d = *d * (a.src ? a.src[si + c * ssz] : a.value); break;
                case OP_DIV: *d = *d / (a.src ? a.src[si + c * ssz] : a.value); break;
                case OP_SET: *d = a.value; break;
                case OP_INCR: *d = *d + a.src[si + c * ssz] * a.value; break;
            }
        }
    }
}

template<typename T>
static int region_t(int op, void* dst, const Box3& dbox, const void* src, const Box3& sbox, const Box3& region,
                    const int* shift, int ncomp, double value, hipStream_t s)
{
    RegionArgs<T> a;
    a.dst = (T*)dst;
    a.src = (const T*)src;
    a.dbox = dbox;
    a.sbox = sbox;
    a.region = region;
    for (int d = 0; d < 3; d++) { a.shift[d] = shift ? shift[d] : 0; }
    a.ncomp = ncomp;
    a.op = op;
    a.value = (T)value;
    const long n = box_size(region);
    if (n == 0) { return PM_OK; }
    const long blocks = std::min((n + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(k_region<T>, dim3((unsigned)blocks), dim3(256), 0, s, a);
    PM_CHECK_LAUNCH();
    return PM_OK;
}

static int region_op(int op, int dtype, void* dst, const pm_box* dst_box, const void* src, const pm_box* src_box,
                     const pm_box* region, const int* shift, int ncomp, double value, pm_stream_t stream)
{
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "bad dtype %d", dtype);
    PM_REQUIRE(dst && dst_box && region && ncomp >= 1, "null argument or ncomp < 1");
    const Box3 db = box3(dst_box), rg = box3(region);
    Box3 sb = rg;
    if (box_empty(rg)) { return PM_OK; }
    if (src) {
        PM_REQUIRE(src_box, "null source box");
        sb = box3(src_box);
        PM_REQUIRE(box_contains_box(sb, rg), "region outside the source box");
    }
    Box3 shifted = rg;
    for (int d = 0; d < 3; d++) {
        const int sh = shift ? shift[d] : 0;
        shifted.lo[d] += sh;
        shifted.hi[d] += sh;
    }
    PM_REQUIRE(box_contains_box(db, shifted), "region (shifted) outside the destination box");
    if (dtype == PM_F64) { return region_t<double>(op, dst, db, src, sb, rg, shift, ncomp, value, S(stream)); }
    return region_t<float>(op, dst, db, src, sb, rg, shift, ncomp, value, S(stream));
}

}  // namespace pm

using namespace pm;

extern "C" {

int pm_boxdata_copy(int dtype, const void* src, const pm_box* src_box, void* dst, const pm_box* dst_box,
                    const pm_box* region, const int shift[3], int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(src, "pm_boxdata_copy: null source");
    return region_op(OP_COPY, dtype, dst, dst_box, src, src_box, region, shift, ncomp, 0.0, stream);
}

int pm_boxdata_binop(int op, int dtype, void* dst, const pm_box* dst_box, const void* src, const pm_box* src_box,
                     const pm_box* region, int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(op >= OP_ADD && op <= OP_DIV, "pm_boxdata_binop: bad op %d", op);
    PM_REQUIRE(src, "pm_boxdata_binop: null source");
    return region_op(op, dtype, dst, dst_box, src, src_box, region, nullptr, ncomp, 0.0, stream);
}

int pm_boxdata_scalar(int op, int dtype, void* dst, const pm_box* dst_box, const pm_box* region, double value,
                      int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(op >= OP_ADD && op <= OP_SET, "pm_boxdata_scalar: bad op %d", op);
    return region_op(op, dtype, dst, dst_box, nullptr, nullptr, region, nullptr, ncomp, value, stream);
}

int pm_boxdata_increment(int dtype, void* dst, const pm_box* dst_box, const void* src, const pm_box* src_box,
                         const pm_box* region, double scale, int ncomp, pm_stream_t stream)
{
    PM_REQUIRE(src, "pm_boxdata_increment: null source");
    return region_op(OP_INCR, dtype, dst, dst_box, src, src_box, region, nullptr, ncomp, scale, stream);
}

}  // extern "C"
