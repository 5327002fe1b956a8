// runtime.hip -- device, memory, stream and event plumbing of the C-ABI.
#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ch_internal.h"

namespace ch {

static thread_local std::string g_error;
static thread_local float g_last_kernel_ms = -1.0f;
static bool g_timing = false;
static int32_t g_num_cus = 0;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

ch_status hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? CH_ERR_NOMEM : CH_ERR_HIP;
}

int32_t num_cus()
{
    if (g_num_cus <= 0) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipGetDeviceProperties(&p, dev) == hipSuccess)
            g_num_cus = p.multiProcessorCount;
        else
            g_num_cus = 256;
    }
    return g_num_cus;
}

void record_kernel_ms(float ms) { g_last_kernel_ms = ms; }

bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess)
        return false;
    return st == hipStreamCaptureStatusActive;
}
bool timing_enabled() { return g_timing; }

static thread_local hipEvent_t g_launch_ev0 = nullptr;
static thread_local hipEvent_t g_launch_ev1 = nullptr;

bool take_launch_events(hipEvent_t* start, hipEvent_t* stop)
{
    if (!g_launch_ev0 || !g_launch_ev1)
        return false;
    *start = g_launch_ev0;
    *stop = g_launch_ev1;
    g_launch_ev0 = g_launch_ev1 = nullptr;
    return true;
}

// Grid-stride zero fill in 16-byte units plus a byte tail.
__global__ void zero_kernel(uint8_t* p, size_t n)
{
    size_t n16 = n / 16;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    uint4* q = reinterpret_cast<uint4*>(p);
    for (size_t j = i; j < n16; j += stride)
        q[j] = make_uint4(0, 0, 0, 0);
    for (size_t j = n16 * 16 + i; j < n; j += stride)
        p[j] = 0;
}

// Elementwise conversion between FP32 and FP64 grids, dst = src or
// dst += src (the sum rounded once in the destination type).
template <typename D, typename S>
__global__ void convert_kernel(D* __restrict__ dst, const S* __restrict__ src,
                               size_t n, int accumulate)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const D v = (D)src[i];
        dst[i] = accumulate ? dst[i] + v : v;
    }
}

// Grid-stride 16-byte copy plus a byte tail (device to device).
__global__ void copy_kernel(uint8_t* __restrict__ dst,
                            const uint8_t* __restrict__ src, size_t n)
{
    size_t n16 = n / 16;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (size_t j = i; j < n16; j += stride)
        d[j] = s[j];
    for (size_t j = n16 * 16 + i; j < n; j += stride)
        dst[j] = src[j];
}

// Reduces the partial boxes (one block of 1024 threads, grid-stride, then
// across the block's waves through LDS).
__global__ __launch_bounds__(1024) void box_finalize_kernel(
    const int32_t* __restrict__ parts, int n, int32_t* __restrict__ box,
    int widen)
{
    int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t* p = parts + (size_t)i * kBoxStride;
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], p[a]);
            v[3 + a] = max(v[3 + a], p[3 + a]);
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], __shfl_xor(v[a], o, 64));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], o, 64));
        }
    __shared__ int s[16][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 6; ++a)
            s[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int r = s[0][a];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            r = a < 3 ? min(r, s[k][a]) : max(r, s[k][a]);
        if (widen)  // union with the box already there
            r = a < 3 ? min(r, box[a]) : max(r, box[a]);
        box[a] = r;
    }
}

void launch_box_finalize(const int32_t* partials, int n_parts, int32_t* box,
                         int widen, hipStream_t s)
{
    hipLaunchKernelGGL(box_finalize_kernel, dim3(1), dim3(1024), 0, s,
                       partials, n_parts, box, widen);
}

struct StreamScratch {
    hipStream_t stream;
    void* ptr[kScratchSlots];
    size_t bytes[kScratchSlots];
};
static std::vector<StreamScratch> g_scratch;
static std::mutex g_scratch_mu;

ch_status stream_scratch(hipStream_t s, int slot, size_t bytes, void** out,
                         bool zero_new)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    StreamScratch* e = nullptr;
    for (auto& x : g_scratch)
        if (x.stream == s)
            e = &x;
    if (!e) {
        StreamScratch x;
        x.stream = s;
        for (int i = 0; i < kScratchSlots; ++i) {
            x.ptr[i] = nullptr;
            x.bytes[i] = 0;
        }
        g_scratch.push_back(x);
        e = &g_scratch.back();
    }
    if (e->bytes[slot] < bytes) {
        if (capturing(s)) {
            set_error("stream scratch must grow while the stream is being "
                      "captured: run the pipeline once before capturing it");
            return CH_ERR_ARG;
        }
        if (stream_graph_holds(s) > 0) {
            set_error("stream scratch must grow, but a live graph captured "
                      "on this stream reads it: destroy the graph or use "
                      "another stream");
            return CH_ERR_ARG;
        }
        // queued kernels of this stream may still use the old buffer
        CH_HIP(hipStreamSynchronize(s));
        CH_HIP(hipFree(e->ptr[slot]));
        e->ptr[slot] = nullptr;
        e->bytes[slot] = 0;
        CH_HIP(hipMalloc(&e->ptr[slot], bytes));
        e->bytes[slot] = bytes;
        if (zero_new) {
            // cleared before any caller (of any thread) gets the pointer
            std::vector<char> zeros(bytes, 0);
            CH_HIP(hipMemcpy(e->ptr[slot], zeros.data(), bytes,
                             hipMemcpyHostToDevice));
        }
    }
    *out = e->ptr[slot];
    return CH_OK;
}

void release_stream_scratch(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    for (size_t i = 0; i < g_scratch.size(); ++i) {
        if (g_scratch[i].stream != s)
            continue;
        (void)hipStreamSynchronize(s);
        for (int k = 0; k < kScratchSlots; ++k)
            (void)hipFree(g_scratch[i].ptr[k]);
        g_scratch.erase(g_scratch.begin() + i);
        return;
    }
}

void* stream_scratch_peek(hipStream_t s, int slot)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    for (auto& x : g_scratch)
        if (x.stream == s)
            return x.ptr[slot];
    return nullptr;
}

// Graphs and what they read.  A capture records every object an entry
// point uses on the capturing stream (capture_hold: plans, beam sets,
// convolution sets); the graph made from it pins those objects and holds
// the capture stream's scratch until it is destroyed.
struct ActiveCapture {
    hipStream_t stream;
    std::vector<GraphHeld*> held;
};
static std::vector<ActiveCapture> g_active;
static std::vector<std::pair<hipStream_t, int>> g_holds;
static std::mutex g_graph_mu;

static void unpin_locked(const std::vector<GraphHeld*>& held)
{
    for (GraphHeld* h : held)
        if (--h->pins == 0 && h->doomed)
            h->release(h);
}

void capture_hold(hipStream_t s, GraphHeld* h)
{
    if (!h || !capturing(s))
        return;
    std::lock_guard<std::mutex> lock(g_graph_mu);
    for (auto& a : g_active) {
        if (a.stream != s)
            continue;
        for (GraphHeld* q : a.held)
            if (q == h)
                return;
        a.held.push_back(h);
        ++h->pins;
        return;
    }
}

void capture_pin(hipStream_t s, const ch_plan* cp)
{
    capture_hold(s, const_cast<ch_plan*>(cp));
}

int stream_graph_holds(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    for (auto& h : g_holds)
        if (h.first == s)
            return h.second;
    return 0;
}

int held_pins(const GraphHeld* h)
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    return h->pins;
}

int plan_pins(const ch_plan* p)
{
    return held_pins(p);
}

bool defer_free(GraphHeld* h)
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    if (h->pins == 0)
        return false;
    h->doomed = true;
    return true;
}

bool plan_defer_free(ch_plan* p)
{
    return defer_free(p);
}

// Takes the capture record of `s` (its objects stay pinned).
static std::vector<GraphHeld*> take_capture(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    std::vector<GraphHeld*> plans;
    for (size_t i = 0; i < g_active.size(); ++i)
        if (g_active[i].stream == s) {
            plans.swap(g_active[i].held);
            g_active.erase(g_active.begin() + i);
            break;
        }
    return plans;
}

// Per-stream device tables (batch pointer lists, TERMA beam parameters),
// staged through pinned memory.  Each stream owns a ring of kTableSlots
// slots; a slot's pinned half is rewritten only after the event of its
// previous copy, so the host waits only when it runs kTableSlots calls ahead
// of the stream.  The device half needs no wait: the next copy into it is
// queued on the same stream behind the kernels that read it.
constexpr int kTableSlots = 64;

struct StageRing {
    hipStream_t stream;
    char* h = nullptr;
    char* d = nullptr;
    size_t slot_bytes = 0;
    hipEvent_t copied[kTableSlots] = {};
    int next = 0;
};
static std::vector<StageRing*> g_rings;
static std::mutex g_ring_mu;

static void ring_free(StageRing* t)
{
    (void)hipFree(t->d);
    (void)hipHostFree(t->h);
    for (hipEvent_t e : t->copied)
        if (e)
            (void)hipEventDestroy(e);
    delete t;
}

ch_status stage_table(hipStream_t s, const void* h_src, size_t bytes,
                      void** d_out)
{
    if (capturing(s)) {
        set_error("a staged table would be recorded by the capture: build "
                  "a device-resident set (ch_beamset, ch_convset) instead");
        return CH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_ring_mu);
    StageRing* t = nullptr;
    for (auto* x : g_rings)
        if (x->stream == s)
            t = x;
    if (!t) {
        t = new StageRing();
        t->stream = s;
        g_rings.push_back(t);
        for (hipEvent_t& e : t->copied)
            CH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (t->slot_bytes < bytes) {
        // queued copies and kernels of this stream may use the old blocks
        CH_HIP(hipStreamSynchronize(s));
        CH_HIP(hipFree(t->d));
        CH_HIP(hipHostFree(t->h));
        t->d = nullptr;
        t->h = nullptr;
        t->slot_bytes = 0;
        const size_t cap =
            (std::max(bytes, (size_t)4096) + 255) & ~(size_t)255;
        CH_HIP(hipMalloc((void**)&t->d, cap * kTableSlots));
        CH_HIP(hipHostMalloc((void**)&t->h, cap * kTableSlots,
                             hipHostMallocDefault));
        t->slot_bytes = cap;
    }
    const int k = t->next;
    t->next = (k + 1) % kTableSlots;
    CH_HIP(hipEventSynchronize(t->copied[k]));
    char* h = t->h + (size_t)k * t->slot_bytes;
    char* d = t->d + (size_t)k * t->slot_bytes;
    memcpy(h, h_src, bytes);
    CH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    CH_HIP(hipEventRecord(t->copied[k], s));
    *d_out = d;
    return CH_OK;
}

void release_stage_tables(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_ring_mu);
    for (size_t i = 0; i < g_rings.size(); ++i) {
        StageRing* t = g_rings[i];
        if (t->stream != s)
            continue;
        (void)hipStreamSynchronize(s);
        ring_free(t);
        g_rings.erase(g_rings.begin() + i);
        return;
    }
}

}  // namespace ch

using namespace ch;

extern "C" {

const char* ch_last_error(void) { return g_error.c_str(); }

int32_t ch_version(void) { return 10000; }

ch_status ch_init(int32_t device, int32_t* n_cu, int64_t* hbm_bytes,
                  char* name_out, int32_t name_cap)
{
    int count = 0;
    CH_HIP(hipGetDeviceCount(&count));
    CH_REQUIRE(device >= 0 && device < count, "device %d out of range (%d)",
               device, count);
    CH_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    CH_HIP(hipGetDeviceProperties(&p, device));
    g_num_cus = p.multiProcessorCount;
    const ch_status st = terma_init_device();
    if (st != CH_OK)
        return st;
    if (n_cu)
        *n_cu = p.multiProcessorCount;
    if (hbm_bytes)
        *hbm_bytes = (int64_t)p.totalGlobalMem;
    if (name_out && name_cap > 0) {
        snprintf(name_out, (size_t)name_cap, "%s", p.gcnArchName);
    }
    return CH_OK;
}

ch_status ch_malloc(void** d_ptr, size_t bytes)
{
    CH_REQUIRE(d_ptr != nullptr, "ch_malloc: null output");
    *d_ptr = nullptr;
    if (bytes == 0)
        bytes = 16;
    CH_HIP(hipMalloc(d_ptr, bytes));
    return CH_OK;
}

ch_status ch_free(void* d_ptr)
{
    if (d_ptr)
        CH_HIP(hipFree(d_ptr));
    return CH_OK;
}

ch_status ch_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t f = 0, t = 0;
    CH_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return CH_OK;
}

ch_status ch_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes,
                        void* stream)
{
    if (bytes == 0)
        return CH_OK;
    CH_REQUIRE(d_dst && h_src, "ch_memcpy_h2d: null pointer");
    CH_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice,
                          (hipStream_t)stream));
    return CH_OK;
}

ch_status ch_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes,
                        void* stream)
{
    if (bytes == 0)
        return CH_OK;
    CH_REQUIRE(h_dst && d_src, "ch_memcpy_d2h: null pointer");
    CH_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost,
                          (hipStream_t)stream));
    CH_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CH_OK;
}

ch_status ch_memcpy_d2h_async(void* h_dst, const void* d_src, size_t bytes,
                              void* stream)
{
    if (bytes == 0)
        return CH_OK;
    CH_REQUIRE(h_dst && d_src, "ch_memcpy_d2h_async: null pointer");
    CH_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost,
                          (hipStream_t)stream));
    return CH_OK;
}

ch_status ch_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes,
                        void* stream)
{
    if (bytes == 0)
        return CH_OK;
    CH_REQUIRE(d_dst && d_src, "ch_memcpy_d2d: null pointer");
    CH_REQUIRE((((uintptr_t)d_dst | (uintptr_t)d_src) & 15) == 0,
               "ch_memcpy_d2d: pointers must be 16-byte aligned");
    size_t blocks = (bytes / 16 + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (uint8_t*)d_dst,
                       (const uint8_t*)d_src, bytes);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_zero(void* d_ptr, size_t bytes, void* stream)
{
    if (bytes == 0)
        return CH_OK;
    CH_REQUIRE(d_ptr, "ch_zero: null pointer");
    size_t n16 = (bytes + 15) / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (uint8_t*)d_ptr, bytes);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_convert(void* d_dst, int32_t dst_type, const void* d_src,
                     int32_t src_type, size_t n, int32_t accumulate,
                     void* stream)
{
    if (n == 0)
        return CH_OK;
    CH_REQUIRE(d_dst && d_src, "ch_convert: null pointer");
    CH_REQUIRE((dst_type == CH_F32 || dst_type == CH_F64) &&
                   (src_type == CH_F32 || src_type == CH_F64),
               "ch_convert: types must be CH_F32 or CH_F64");
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    const hipStream_t s = (hipStream_t)stream;
    const int acc = accumulate ? 1 : 0;
    const dim3 g((unsigned)blocks), b(256);
    if (dst_type == CH_F64 && src_type == CH_F32)
        hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, s,
                           (double*)d_dst, (const float*)d_src, n, acc);
    else if (dst_type == CH_F32 && src_type == CH_F64)
        hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, s,
                           (float*)d_dst, (const double*)d_src, n, acc);
    else if (dst_type == CH_F32)
        hipLaunchKernelGGL((convert_kernel<float, float>), g, b, 0, s,
                           (float*)d_dst, (const float*)d_src, n, acc);
    else
        hipLaunchKernelGGL((convert_kernel<double, double>), g, b, 0, s,
                           (double*)d_dst, (const double*)d_src, n, acc);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_host_alloc(void** h_ptr, size_t bytes)
{
    CH_REQUIRE(h_ptr, "ch_host_alloc: null output");
    CH_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 16, hipHostMallocDefault));
    return CH_OK;
}

ch_status ch_host_free(void* h_ptr)
{
    if (h_ptr)
        CH_HIP(hipHostFree(h_ptr));
    return CH_OK;
}

ch_status ch_stream_create(void** stream)
{
    CH_REQUIRE(stream, "ch_stream_create: null output");
    hipStream_t s;
    // Default priority, non-blocking with respect to the null stream.
    CH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return CH_OK;
}

ch_status ch_stream_destroy(void* stream)
{
    if (stream) {
        hipStream_t s = (hipStream_t)stream;
        CH_REQUIRE(stream_graph_holds(s) == 0,
                   "ch_stream_destroy: a live graph captured on this stream "
                   "reads its scratch; destroy the graph first");
        release_gather_workspace(s);
        release_stage_tables(s);
        release_stream_scratch(s);
        CH_HIP(hipStreamDestroy(s));
    }
    return CH_OK;
}

ch_status ch_graph_begin(void* stream)
{
    CH_REQUIRE(stream, "ch_graph_begin: capture needs a created stream");
    hipStream_t s = (hipStream_t)stream;
    {
        std::vector<GraphHeld*> stale = take_capture(s);
        std::lock_guard<std::mutex> lock(g_graph_mu);
        unpin_locked(stale);
        ActiveCapture a;
        a.stream = s;
        g_active.push_back(a);
    }
    const hipError_t e =
        hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        std::vector<GraphHeld*> plans = take_capture(s);
        std::lock_guard<std::mutex> lock(g_graph_mu);
        unpin_locked(plans);
        return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    }
    return CH_OK;
}

struct ch_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipStream_t stream;            // capture stream (its scratch is held)
    std::vector<GraphHeld*> plans;   // pinned plans and sets
};

ch_status ch_graph_end(void* stream, ch_graph** graph_out)
{
    CH_REQUIRE(stream && graph_out, "ch_graph_end: null argument");
    *graph_out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    std::vector<GraphHeld*> plans = take_capture(s);
    auto unpin = [&plans]() {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        unpin_locked(plans);
    };
    hipGraph_t g = nullptr;
    // always ends the capture, also when a call inside it failed
    hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        unpin();
        return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    if (!g) {
        unpin();
        set_error("ch_graph_end: the capture produced no graph");
        return CH_ERR_ARG;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        unpin();
        return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    ch_graph* out = new ch_graph();
    out->graph = g;
    out->exec = x;
    out->stream = s;
    out->plans.swap(plans);
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        bool found = false;
        for (auto& h : g_holds)
            if (h.first == s) {
                ++h.second;
                found = true;
            }
        if (!found)
            g_holds.emplace_back(s, 1);
    }
    *graph_out = out;
    return CH_OK;
}

ch_status ch_graph_launch(const ch_graph* graph, void* stream)
{
    CH_REQUIRE(graph, "ch_graph_launch: null graph");
    CH_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return CH_OK;
}

ch_status ch_graph_destroy(ch_graph* graph)
{
    if (!graph)
        return CH_OK;
    // a replay may still be queued: wait for it before plans are freed
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        unpin_locked(graph->plans);
        for (size_t i = 0; i < g_holds.size(); ++i)
            if (g_holds[i].first == graph->stream && --g_holds[i].second == 0)
                g_holds.erase(g_holds.begin() + i);
    }
    delete graph;
    return CH_OK;
}

ch_status ch_stream_sync(void* stream)
{
    CH_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CH_OK;
}

ch_status ch_device_sync(void)
{
    CH_HIP(hipDeviceSynchronize());
    return CH_OK;
}

ch_status ch_event_create(void** event)
{
    CH_REQUIRE(event, "ch_event_create: null output");
    hipEvent_t e;
    CH_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return CH_OK;
}

ch_status ch_event_destroy(void* event)
{
    if (event)
        CH_HIP(hipEventDestroy((hipEvent_t)event));
    return CH_OK;
}

ch_status ch_event_record(void* event, void* stream)
{
    CH_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return CH_OK;
}

ch_status ch_stream_wait(void* stream, void* event)
{
    CH_REQUIRE(event, "ch_stream_wait: null event");
    CH_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return CH_OK;
}

ch_status ch_event_sync(void* event)
{
    CH_HIP(hipEventSynchronize((hipEvent_t)event));
    return CH_OK;
}

ch_status ch_event_query(void* event, int32_t* done)
{
    hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) {
        *done = 1;
        return CH_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return CH_OK;
    }
    return hip_fail(e, "hipEventQuery", __FILE__, __LINE__);
}

ch_status ch_event_elapsed(float* ms, void* start, void* stop)
{
    CH_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return CH_OK;
}

ch_status ch_set_timing(int32_t enabled)
{
    g_timing = enabled != 0;
    return CH_OK;
}

ch_status ch_last_kernel_ms(float* ms)
{
    *ms = g_last_kernel_ms;
    return CH_OK;
}

ch_status ch_set_launch_events(void* start, void* stop)
{
    CH_REQUIRE((start == nullptr) == (stop == nullptr),
               "ch_set_launch_events: give both events or neither");
    g_launch_ev0 = (hipEvent_t)start;
    g_launch_ev1 = (hipEvent_t)stop;
    return CH_OK;
}

}  // extern "C"
