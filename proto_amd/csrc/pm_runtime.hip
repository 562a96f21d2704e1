// Runtime entry points: errors, device memory, streams, seeded inputs.
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <algorithm>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// Per-stream scratch.  A buffer handed out while the stream was capturing a
// graph may be referenced by that graph for as long as it lives, which this
// library cannot see (pm_graph_*, the multigrid driver's cycle graph, or a
// caller's own capture).  Such a buffer is never freed on growth: it moves to
// the stream's retired list and is released only by pm_stream_destroy.
// Buffers never seen by a capture are freed on growth after the stream
// drains, as before.
struct WsEntry {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool captured = false;       // handed out while the stream was capturing
    std::vector<void*> retired;  // superseded buffers a graph may still use
};
static std::mutex g_ws_mu;
static std::map<hipStream_t, WsEntry>* g_ws = nullptr;

static bool stream_capturing(hipStream_t stream)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // unknown: treat the buffer as captured
    }
    return st != hipStreamCaptureStatusNone;
}

int workspace(hipStream_t stream, size_t bytes, void** out)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (!g_ws) { g_ws = new std::map<hipStream_t, WsEntry>(); }
    WsEntry& w = (*g_ws)[stream];
    const bool capturing = stream_capturing(stream);
    if (bytes > w.bytes) {
        if (capturing) {
            // neither a stream drain nor an allocation may be recorded into a graph
            set_error("workspace: %zu bytes of per-stream scratch needed under graph capture, %zu available; "
                      "run the call once on the stream before capturing it", bytes, w.bytes);
            return PM_ERR_UNSUPPORTED;
        }
        if (w.ptr) {
            if (w.captured) {
                w.retired.push_back(w.ptr);
            } else {
                PM_CHECK_HIP(hipStreamSynchronize(stream));
                PM_CHECK_HIP(hipFree(w.ptr));
            }
            w.ptr = nullptr;
            w.bytes = 0;
            w.captured = false;
        }
        size_t nb = bytes < ((size_t)1 << 16) ? ((size_t)1 << 16) : bytes;
        hipError_t e = hipMalloc(&w.ptr, nb);
        if (e != hipSuccess) {
            set_error("workspace hipMalloc(%zu): %s", nb, hipGetErrorString(e));
            w.ptr = nullptr;
            return PM_ERR_NOMEM;
        }
        w.bytes = nb;
    }
    if (capturing) { w.captured = true; }
    *out = w.ptr;
    return PM_OK;
}

// Scratch of a stream: current size and number of retired buffers kept for graphs.
static int workspace_info(hipStream_t stream, size_t* bytes, int* retired)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    *bytes = 0;
    *retired = 0;
    if (!g_ws) { return PM_OK; }
    auto it = g_ws->find(stream);
    if (it == g_ws->end()) { return PM_OK; }
    *bytes = it->second.bytes;
    *retired = (int)it->second.retired.size();
    return PM_OK;
}

// Drops the scratch of a stream that is about to be destroyed (current and
// retired buffers: graphs that used them must not be launched afterwards).
static void workspace_release(hipStream_t stream)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (!g_ws) { return; }
    auto it = g_ws->find(stream);
    if (it == g_ws->end()) { return; }
    (void)hipStreamSynchronize(stream);
    if (it->second.ptr) { (void)hipFree(it->second.ptr); }
    for (void* p : it->second.retired) { (void)hipFree(p); }
    g_ws->erase(it);
}

// SplitMix64 (Steele, Lea, Flood 2014) in counter form: z_i = mix(s0 + (i+1) G),
// s0 = seed + stream * 0x632BE59BD9B4E019.  u = ((z >> 11) 2^-53) 2 - 1.
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template<typename T>
__global__ void k_fill_uniform(T* __restrict__ out, size_t n, uint64_t s0, double a, double b, int affine)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t z = mix64(s0 + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        double u = ((double)(z >> 11) * 0x1.0p-53) * 2.0 - 1.0;
        if (affine) { u = a * u + b; }
        out[i] = (T)u;
    }
}

template<typename T>
__global__ void k_fill_separable(T* __restrict__ out, Box3 box, Box3 region, int px, int py, int pz,
                                 const double* __restrict__ tx, const double* __restrict__ ty,
                                 const double* __restrict__ tz, double amp, uint64_t s0)
{
    const long nx = box_nx(region), ny = box_ny(region);
    const long n = box_size(region);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int x = region.lo[0] + (int)(t % nx);
        const long r = t / nx;
        const int y = region.lo[1] + (int)(r % ny);
        const int z = region.lo[2] + (int)(r / ny);
        const int wx = ((x % px) + px) % px, wy = ((y % py) + py) % py, wz = ((z % pz) + pz) % pz;
        double v = (tx[wx] * ty[wy]) * tz[wz];
        if (amp != 0.0) {
            const uint64_t i = (uint64_t)wx + (uint64_t)px * ((uint64_t)wy + (uint64_t)py * (uint64_t)wz);
            const uint64_t zz = mix64(s0 + (i + 1) * 0x9E3779B97F4A7C15ULL);
            const double u = ((double)(zz >> 11) * 0x1.0p-53) * 2.0 - 1.0;
            v = v + amp * u;
        }
        out[box_index(box, x, y, z)] = (T)v;
    }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template<int NT>
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // four independent 16-byte loads in flight per thread
    for (; i + 3 * stride < n; i += 4 * stride) {
        u32x4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        if (NT) {
            __builtin_nontemporal_store(a, &dst[i]);
            __builtin_nontemporal_store(b, &dst[i + stride]);
            __builtin_nontemporal_store(c, &dst[i + 2 * stride]);
            __builtin_nontemporal_store(d, &dst[i + 3 * stride]);
        } else {
            dst[i] = a;
            dst[i + stride] = b;
            dst[i + 2 * stride] = c;
            dst[i + 3 * stride] = d;
        }
    }
    for (; i < n; i += stride) {
        if (NT) { __builtin_nontemporal_store(src[i], &dst[i]); } else { dst[i] = src[i]; }
    }
}

// Block-contiguous copy: workgroup b copies 64 KB chunks b, b + grid, ...;
// each thread has 16 sixteen-byte loads in flight, then 16 stores.  NT: the
// loads and stores are non-temporal.  A probe of the 1:1 read:write mix with
// DRAM-page-friendly streams, beside the grid-stride copy.
template<int NT>
__global__ __launch_bounds__(256) void k_copy_chunk(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n)
{
    constexpr int CH = 4096, U = CH / 256;
    const size_t nch = n / CH;
    for (size_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const size_t base = c * CH + threadIdx.x;
        u32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) { v[u] = NT ? __builtin_nontemporal_load(&src[base + 256 * u]) : src[base + 256 * u]; }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (NT) {
                __builtin_nontemporal_store(v[u], &dst[base + 256 * u]);
            } else {
                dst[base + 256 * u] = v[u];
            }
        }
    }
    for (size_t i = nch * CH + (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        dst[i] = src[i];
    }
}

}  // namespace pm

using namespace pm;

extern "C" {

const char* pm_last_error(void) { return g_err; }
const char* pm_version(void) { return "protomi 0.1 (gfx950)"; }

int pm_device_count(int* count)
{
    PM_REQUIRE(count, "pm_device_count: null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return PM_OK;
}

int pm_malloc(void** ptr, size_t bytes)
{
    PM_REQUIRE(ptr, "pm_malloc: null ptr");
    *ptr = nullptr;
    if (bytes == 0) { return PM_OK; }
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *ptr = nullptr;
        return e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP;
    }
    return PM_OK;
}

int pm_free(void* ptr)
{
    if (!ptr) { return PM_OK; }
    PM_CHECK_HIP(hipFree(ptr));
    return PM_OK;
}

int pm_memcpy(void* dst, const void* src, size_t bytes, int kind, pm_stream_t stream)
{
    if (bytes == 0) { return PM_OK; }
    PM_REQUIRE(dst && src, "pm_memcpy: null pointer");
    hipMemcpyKind k;
    switch (kind) {
        case 0: k = hipMemcpyHostToDevice; break;
        case 1: k = hipMemcpyDeviceToHost; break;
        case 2: k = hipMemcpyDeviceToDevice; break;
        default: PM_REQUIRE(false, "pm_memcpy: bad kind %d", kind);
    }
    PM_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, k, S(stream)));
    return PM_OK;
}

int pm_memset_zero(void* ptr, size_t bytes, pm_stream_t stream)
{
    if (bytes == 0) { return PM_OK; }
    PM_REQUIRE(ptr, "pm_memset_zero: null pointer");
    PM_CHECK_HIP(hipMemsetAsync(ptr, 0, bytes, S(stream)));
    return PM_OK;
}

int pm_copy_device(void* dst, const void* src, size_t bytes, int nontemporal, pm_stream_t stream)
{
    if (bytes == 0) { return PM_OK; }
    PM_REQUIRE(dst && src, "pm_copy_device: null pointer");
    PM_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
               "pm_copy_device: bytes and pointers must be multiples of 16");
    const size_t n = bytes / 16;
    // 8 resident 256-thread workgroups on each of the 256 CUs
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 8);
    if (nontemporal) {
        hipLaunchKernelGGL(k_copy16<1>, dim3(blocks), dim3(256), 0, S(stream), (const u32x4*)src, (u32x4*)dst, n);
    } else {
        hipLaunchKernelGGL(k_copy16<0>, dim3(blocks), dim3(256), 0, S(stream), (const u32x4*)src, (u32x4*)dst, n);
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

// Streaming-bandwidth probes: read-only (eight 16-byte loads in flight per
// thread, folded into one word stored only when it matches a value the data
// never produces, so the loads stay live) and write-only.
__global__ __launch_bounds__(256) void k_probe_read(const u32x4* __restrict__ src, size_t n, unsigned* __restrict__ sink)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned acc = 0;
    for (; i + 7 * stride < n; i += 8 * stride) {
        u32x4 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) { v[u] = __builtin_nontemporal_load(&src[i + u * stride]); }
#pragma unroll
        for (int u = 0; u < 8; u++) { acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w; }
    }
    for (; i < n; i += stride) {
        const u32x4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9e3779b9u && threadIdx.x == 0x3ff) { sink[0] = acc; }  // never true: threadIdx.x < 256
}

__global__ __launch_bounds__(256) void k_probe_write(u32x4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned w = (unsigned)i;
        __builtin_nontemporal_store(u32x4{w, w + 1u, w + 2u, w + 3u}, &dst[i]);
    }
}

int pm_stream_probe(int kind, void* dst, const void* src, size_t bytes, int blocks_per_cu, pm_stream_t stream)
{
    PM_REQUIRE(kind >= 0 && kind <= 4,
               "pm_stream_probe: kind must be 0 (read), 1 (write), 2 (copy), 3 / 4 (chunked copy, NT / plain)");
    PM_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 64, "pm_stream_probe: blocks_per_cu outside [1, 64]");
    PM_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
               "pm_stream_probe: bytes and pointers must be multiples of 16");
    PM_REQUIRE(kind == 1 || src, "pm_stream_probe: null source");
    PM_REQUIRE(kind == 0 || dst, "pm_stream_probe: null destination");
    if (bytes == 0) { return PM_OK; }
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)256 * blocks_per_cu);
    if (kind == 0) {
        hipLaunchKernelGGL(k_probe_read, dim3(blocks), dim3(256), 0, S(stream), (const u32x4*)src, n, (unsigned*)dst);
    } else if (kind == 1) {
        hipLaunchKernelGGL(k_probe_write, dim3(blocks), dim3(256), 0, S(stream), (u32x4*)dst, n);
    } else if (kind == 2) {
        hipLaunchKernelGGL(k_copy16<1>, dim3(blocks), dim3(256), 0, S(stream), (const u32x4*)src, (u32x4*)dst, n);
    } else {
        // chunked: one 64 KB chunk per workgroup per pass
        const unsigned cb = (unsigned)std::max<size_t>(1, std::min<size_t>(n / 4096, (size_t)256 * blocks_per_cu));
        if (kind == 3) {
            hipLaunchKernelGGL(k_copy_chunk<1>, dim3(cb), dim3(256), 0, S(stream), (const u32x4*)src, (u32x4*)dst, n);
        } else {
            hipLaunchKernelGGL(k_copy_chunk<0>, dim3(cb), dim3(256), 0, S(stream), (const u32x4*)src, (u32x4*)dst, n);
        }
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

int pm_stream_create(pm_stream_t* stream)
{
    PM_REQUIRE(stream, "pm_stream_create: null out");
    hipStream_t s;
    PM_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (pm_stream_t)s;
    return PM_OK;
}

int pm_stream_destroy(pm_stream_t stream)
{
    if (!stream) { return PM_OK; }
    workspace_release(S(stream));
    PM_CHECK_HIP(hipStreamDestroy(S(stream)));
    return PM_OK;
}

int pm_stream_scratch_info(pm_stream_t stream, size_t* bytes, int* retired)
{
    PM_REQUIRE(bytes && retired, "pm_stream_scratch_info: null argument");
    return workspace_info(S(stream), bytes, retired);
}

int pm_stream_sync(pm_stream_t stream)
{
    PM_CHECK_HIP(hipStreamSynchronize(S(stream)));
    return PM_OK;
}

int pm_device_sync(void)
{
    PM_CHECK_HIP(hipDeviceSynchronize());
    return PM_OK;
}

int pm_event_create(pm_event_t* ev)
{
    PM_REQUIRE(ev, "pm_event_create: null out");
    hipEvent_t e;
    PM_CHECK_HIP(hipEventCreate(&e));
    *ev = (pm_event_t)e;
    return PM_OK;
}

int pm_event_destroy(pm_event_t ev)
{
    if (!ev) { return PM_OK; }
    PM_CHECK_HIP(hipEventDestroy((hipEvent_t)ev));
    return PM_OK;
}

int pm_event_record(pm_event_t ev, pm_stream_t stream)
{
    PM_REQUIRE(ev, "pm_event_record: null event");
    PM_CHECK_HIP(hipEventRecord((hipEvent_t)ev, S(stream)));
    return PM_OK;
}

int pm_event_sync(pm_event_t ev)
{
    PM_REQUIRE(ev, "pm_event_sync: null event");
    PM_CHECK_HIP(hipEventSynchronize((hipEvent_t)ev));
    return PM_OK;
}

int pm_event_elapsed_ms(pm_event_t start, pm_event_t stop, float* ms)
{
    PM_REQUIRE(start && stop && ms, "pm_event_elapsed_ms: null argument");
    PM_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return PM_OK;
}

struct GraphBox {
    hipGraph_t g;
    hipGraphExec_t e;
};

int pm_graph_begin(pm_stream_t stream)
{
    PM_REQUIRE(stream, "pm_graph_begin: capture needs a non-null stream");
    PM_CHECK_HIP(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal));
    return PM_OK;
}

int pm_graph_end(pm_stream_t stream, pm_graph_t* graph)
{
    PM_REQUIRE(stream && graph, "pm_graph_end: null argument");
    *graph = nullptr;
    hipGraph_t g = nullptr;
    PM_CHECK_HIP(hipStreamEndCapture(S(stream), &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        hipGraphDestroy(g);
        set_error("pm_graph_end: instantiate failed: %s", hipGetErrorString(err));
        return PM_ERR_HIP;
    }
    *graph = (pm_graph_t) new GraphBox{g, e};
    return PM_OK;
}

int pm_graph_launch(pm_graph_t graph, pm_stream_t stream)
{
    PM_REQUIRE(graph, "pm_graph_launch: null graph");
    PM_CHECK_HIP(hipGraphLaunch(((GraphBox*)graph)->e, S(stream)));
    return PM_OK;
}

int pm_graph_destroy(pm_graph_t graph)
{
    if (!graph) { return PM_OK; }
    GraphBox* b = (GraphBox*)graph;
    hipGraphExecDestroy(b->e);
    hipGraphDestroy(b->g);
    delete b;
    return PM_OK;
}

int pm_fill_separable(int dtype, void* data, const pm_box* box, const pm_box* region, const int period[3],
                      const double* tx_host, const double* ty_host, const double* tz_host, double amp,
                      uint64_t seed, uint64_t stream_id, pm_stream_t stream)
{
    PM_REQUIRE(data && box && region && period && tx_host && ty_host && tz_host, "pm_fill_separable: null argument");
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_fill_separable: bad dtype %d", dtype);
    for (int d = 0; d < 3; d++) { PM_REQUIRE(period[d] >= 1, "pm_fill_separable: period must be >= 1"); }
    const Box3 b = box3(box), r = box3(region);
    PM_REQUIRE(box_contains_box(b, r), "pm_fill_separable: region outside the box");
    if (box_empty(r)) { return PM_OK; }
    const size_t nt = (size_t)period[0] + period[1] + period[2];
    double* dt = nullptr;
    PM_CHECK_HIP(hipMalloc(&dt, sizeof(double) * nt));
    hipError_t e1 = hipMemcpy(dt, tx_host, sizeof(double) * period[0], hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(dt + period[0], ty_host, sizeof(double) * period[1], hipMemcpyHostToDevice);
    hipError_t e3 = hipMemcpy(dt + period[0] + period[1], tz_host, sizeof(double) * period[2], hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        hipFree(dt);
        set_error("pm_fill_separable: table upload failed");
        return PM_ERR_HIP;
    }
    const uint64_t s0 = seed + stream_id * 0x632BE59BD9B4E019ULL;
    size_t blocks = std::min<size_t>(((size_t)box_size(r) + 255) / 256, 65536);
    if (dtype == PM_F64) {
        hipLaunchKernelGGL(k_fill_separable<double>, dim3((unsigned)blocks), dim3(256), 0, S(stream), (double*)data,
                           b, r, period[0], period[1], period[2], dt, dt + period[0], dt + period[0] + period[1],
                           amp, s0);
    } else {
        hipLaunchKernelGGL(k_fill_separable<float>, dim3((unsigned)blocks), dim3(256), 0, S(stream), (float*)data,
                           b, r, period[0], period[1], period[2], dt, dt + period[0], dt + period[0] + period[1],
                           amp, s0);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(S(stream));
    hipFree(dt);
    if (le != hipSuccess || se != hipSuccess) {
        set_error("pm_fill_separable: %s", hipGetErrorString(le != hipSuccess ? le : se));
        return PM_ERR_HIP;
    }
    return PM_OK;
}

int pm_fill_uniform(int dtype, void* data, size_t n, uint64_t seed, uint64_t stream_id, double a,
                    double b, pm_stream_t stream)
{
    if (n == 0) { return PM_OK; }
    PM_REQUIRE(data, "pm_fill_uniform: null data");
    PM_REQUIRE(dtype == PM_F32 || dtype == PM_F64, "pm_fill_uniform: bad dtype %d", dtype);
    uint64_t s0 = seed + stream_id * 0x632BE59BD9B4E019ULL;
    int affine = !(a == 1.0 && b == 0.0);
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) { blocks = 65536; }
    if (dtype == PM_F64) {
        hipLaunchKernelGGL(k_fill_uniform<double>, dim3((unsigned)blocks), dim3(256), 0, S(stream),
                           (double*)data, n, s0, a, b, affine);
    } else {
        hipLaunchKernelGGL(k_fill_uniform<float>, dim3((unsigned)blocks), dim3(256), 0, S(stream),
                           (float*)data, n, s0, a, b, affine);
    }
    PM_CHECK_LAUNCH();
    return PM_OK;
}

}  // extern "C"
