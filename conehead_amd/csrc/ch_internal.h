// ch_internal.h -- shared declarations of libconehead_amd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <vector>

#include "../../include/conehead_amd.h"

namespace ch {

// Thread-local error message behind ch_last_error().
void set_error(const char* fmt, ...);
ch_status hip_fail(hipError_t e, const char* what, const char* file, int line);
int32_t num_cus();
// Resolves the current device's address of the allocation-free TERMA
// scratch (terma.hip); ch_init calls it with the device current.
ch_status terma_init_device();

#define CH_HIP(call)                                                        \
    do {                                                                    \
        hipError_t ch_e_ = (call);                                          \
        if (ch_e_ != hipSuccess)                                            \
            return ::ch::hip_fail(ch_e_, #call, __FILE__, __LINE__);        \
    } while (0)

#define CH_REQUIRE(cond, ...)                                               \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::ch::set_error(__VA_ARGS__);                                   \
            return CH_ERR_ARG;                                              \
        }                                                                   \
    } while (0)

constexpr double kBig = 1000000000.0;  // dda_3d_c.pyx:52 big_number
constexpr int32_t kMaxIndex = 5995;    // convolve_c.pyx:144
constexpr int32_t kRowLenMin = kMaxIndex + 1;
// Zero taps appended to every cone's tap list (>= the gather's unroll).
constexpr int32_t kTapPad = 8;
constexpr int kUnroll = 8;  // taps per gather step (<= kTapPad)
// x-slab support boxes of a TERMA grid: slabs of kSlabW x planes, each with
// the bounding box of its T != 0 voxels; neighbours whose union box is
// hardly larger than the pair merge.  Record 0 holds the count, record
// 1 + s the box of slab s (min x, y, z, max x, y, z), kSlabStride ints each.
constexpr int kSlabW = 16;
constexpr int kSlabStride = 8;
inline int slab_records(int X) { return 1 + (X + kSlabW - 1) / kSlabW; }
// ints of a slab workspace: per-plane extents, then the slab records
inline size_t slab_ints(int X)
{
    return (size_t)4 * X + (size_t)kSlabStride * slab_records(X);
}

// Per-cone metadata of a plan (device and host copies).
//  taps [tap_begin, tap_begin + n_taps): source byte offset and increment.
//  ftab[f_begin[a] + q], q in [0, f_len[a]): first tap index m whose offset
//  has |o_{m,a}| >= q; the last entry (q = f_len[a]-1) equals n_taps.
//  s_*: the staged gather's chunking of this cone (set by the host after
//  the tap lists are built): s_ch taps per chunk, LDS box extents s_ext
//  (block + the largest per-axis offset change over s_ch consecutive taps)
//  and s_nstage staged elements per thread.
struct ConeMeta {
    int32_t tap_begin;
    int32_t n_taps;
    int32_t f_begin[3];
    int32_t f_len[3];
    int32_t step[3];
    int32_t s_ch;
    int32_t s_ext[3];
    int32_t s_nstage;
};

// Zero voxels on both ends of every z-row (kZPad) and every y-column
// (kYPad) of the padded TERMA copy the gathers read.  They bound how far a
// read may reach past the TERMA support along y and z: a tile's extent - 1
// plus the unrolled chunk's overrun past the union range (direct gather),
// or plus a chunk's offset change (staged gather); x needs no padding
// because the gathers read T with range-checked buffer loads.
constexpr int32_t kZPad = 80;
constexpr int32_t kYPad = 32;

// Staged gather: one 256-thread workgroup per 4 x 16 x 16 output block, four
// outputs per lane (y, y + 4, y + 8, y + 12); TERMA boxes of up to
// kSLds floats are staged in LDS with rows of kSPitch floats.
constexpr int kSBX = 4, kSBY = 16, kSBZ = 16;
constexpr int kSPitch = 48;     // row pitch: conflict-free 4 x 16 lane reads
constexpr int kSLds = 8192;     // LDS floats per workgroup (32 KiB)
constexpr int kSStage = 16;     // staged elements per thread, at most
constexpr int kSChunk = 32;     // taps per chunk, at most
constexpr int kGroup = CH_MAX_GROUP;  // beams per interleaved group
constexpr int kSMaxWY = kYPad - kSBY;       // y offset change per chunk
constexpr int kSMaxWZ = kSPitch - kSBZ;     // z offset change per chunk

// Flat job descriptor of the gather kernels.
struct ConvJob {
    const float* terma;        // padded TERMA [X][Y + 2 kYPad][Z + 2 kZPad]
    void* dose;                // float* or double*, [X][Y][Z]
    const int32_t* box;        // nullable
    const int32_t* lin;        // per tap source byte offset in padded TERMA
    const float* k;            // per tap FP32 increment
    const int32_t* ftab;
    const int32_t* ftT;        // first-step tables, cone-minor (ftab_t)
    const int32_t* slabs;      // x-slab support boxes (nullable: box)
    const ConeMeta* meta;
    const int32_t* off;        // per tap offset (x, y, z)
    const int32_t* lin2;       // per tap byte offset in the cone's LDS box
    int32_t X, Y, Z;
    int32_t n_cones;
    int32_t accumulate;
    int32_t ntx, nty, ntz;     // tiles (or staged blocks) along x, y, z
    int32_t tile_begin;        // first global tile of this job
};

void record_kernel_ms(float ms);
bool timing_enabled();
// True while `s` is being captured into a graph (no allocation, no
// synchronisation may happen then).
bool capturing(hipStream_t s);

// Optional HIP-event timing of one launch (ch_set_timing): stop() waits for
// the launch and records its duration for ch_last_kernel_ms().
// The calling thread's pending launch-event pair (ch_set_launch_events):
// taken (and cleared) by the next gather or scatter launch.
bool take_launch_events(hipEvent_t* start, hipEvent_t* stop);

// Timing of one launch (or of a span of launches).  The caller's pending
// launch-event pair (ch_set_launch_events), or a pair of its own when
// ch_set_timing is on, is attached to the kernel launched through
// launch(): hipExtLaunchKernelGGL records the events at the kernel's own
// start and end, so they measure the kernel and not the dispatch gap
// before it.  A span timer (span = true) records the pair around
// everything queued between its construction and stop() instead.  stop()
// waits only when ch_set_timing is on, and then records the duration for
// ch_last_kernel_ms().  Nothing is recorded while the stream is captured.
struct LaunchTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s;
    bool on, user, span, own;
    explicit LaunchTimer(hipStream_t st, bool span_ = false)
        : s(st), on(timing_enabled() && !capturing(st)), user(false),
          span(span_), own(false)
    {
        if (!capturing(st))
            user = take_launch_events(&e0, &e1);
        if (on && !user) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            own = true;
        }
        if (span && (user || on))
            (void)hipEventRecord(e0, s);
    }
    bool armed() const { return user || on; }
    template <typename... Args, typename F = void (*)(Args...)>
    void launch(F kernel, const dim3& grid, const dim3& block,
                std::uint32_t lds, Args... args)
    {
        if (armed() && !span)
            hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0,
                                  args...);
        else
            hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
    }
    void stop()
    {
        if (span && armed())
            (void)hipEventRecord(e1, s);
        if (on) {
            (void)hipEventSynchronize(e1);
            float ms = -1.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            record_kernel_ms(ms);
        }
        if (own) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        own = false;
        on = false;
        user = false;
    }
};

// Device scratch owned by the library, one buffer per (stream, slot),
// grown on demand.  Kernels on one stream run in order, so a stream's
// buffer can be reused by its next launch without synchronisation.
enum ScratchSlot {
    kScratchBoxPartials = 0,  // per-block partial support boxes
    kScratchStatus = 1,       // TERMA geometry flag (one int32)
    kScratchSpectrum = 2,     // complex work grids of the spectral passes
    kScratchTermaList = 3,    // TERMA brick lists by cost class
    kScratchTermaCounters = 4,  // TERMA counter states (zero when allocated)
    kScratchSlots = 5
};
// zero_new: a newly allocated buffer is zero-filled (synchronously, under
// the allocator's lock) before it is handed out.
ch_status stream_scratch(hipStream_t s, int slot, size_t bytes, void** out,
                         bool zero_new = false);
// The slot's buffer if this stream has one, else nullptr (no allocation).
void* stream_scratch_peek(hipStream_t s, int slot);
// Free a stream's scratch and gather workspace (before it is destroyed).
void release_stream_scratch(hipStream_t s);
// What a live captured graph reads stays valid while the graph lives:
//  capture_hold(s, h) -- while `s` is being captured, records object h (a
//    plan, beam set or convolution set) in the capture (once) and pins it:
//    destroying a pinned object is deferred to the destruction of its last
//    graph (h->release is called then), and a plan's spectrum cannot be
//    released;
//  stream_graph_holds(s) -- live graphs captured on `s`: while there are
//    any, the stream's scratch and gather workspace must not be
//    reallocated and the stream must not be destroyed.
struct GraphHeld {
    int32_t pins = 0;        // live graphs that read the object
    bool doomed = false;     // destroyed while pinned: freed when unpinned
    void (*release)(GraphHeld*) = nullptr;
};
void capture_hold(hipStream_t s, GraphHeld* h);
void capture_pin(hipStream_t s, const ch_plan* p);
int stream_graph_holds(hipStream_t s);
int held_pins(const GraphHeld* h);
int plan_pins(const ch_plan* p);
// True when `h` is pinned: it is then marked to be freed with its last
// graph.
bool defer_free(GraphHeld* h);
bool plan_defer_free(ch_plan* p);
void plan_release(ch_plan* p);
void release_gather_workspace(hipStream_t s);
// Copies `bytes` of host data into a device table read by the kernels
// queued next on stream s (a per-stream ring of slots staged through pinned
// memory) and returns its device address.  Refused while s is captured.
ch_status stage_table(hipStream_t s, const void* h_src, size_t bytes,
                      void** d_out);
void release_stage_tables(hipStream_t s);
// Pads two grids of plan p's shape into the stream's gather workspace
// (the layout the gathers read) and returns the interior voxel (0, 0, 0)
// of each copy.
ch_status pad_pair(const ch_plan* p, const float* a, const float* b,
                   hipStream_t s, const float** pa, const float** pb);

// Support box of T != 0 from per-block partial boxes: kernels write
// kBoxStride ints per block (min x, y, z, max x, y, z) to `partials`, then
// box_finalize reduces n_parts of them into box[0..5] (lo > hi if empty).
constexpr int kBoxStride = 8;
void launch_box_finalize(const int32_t* partials, int n_parts, int32_t* box,
                         int widen, hipStream_t s);

// Reduces each thread's candidate box (lo > hi when it has none) over its
// block of 256 threads and writes the block's partial box to
// parts[blockIdx.x * kBoxStride + 0..5].  Every thread of the block must
// call it.
__device__ __forceinline__ void block_box_partial(int32_t v[6],
                                                  int32_t* parts)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], __shfl_xor(v[a], o, 64));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], o, 64));
        }
    __shared__ int32_t s[4][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a)
            s[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int32_t r = s[0][a];
#pragma unroll
        for (int k = 1; k < 4; ++k)
            r = a < 3 ? min(r, s[k][a]) : max(r, s[k][a]);
        parts[(size_t)blockIdx.x * kBoxStride + a] = r;
    }
}

constexpr int kBufferFlags = 0x00020000;  // raw 32-bit buffer resource

// Start of axis a's block in the cone-minor first-step tables of a grid
// (X, Y, Z) with n cones: blocks of (dim + 2) rows of n entries.
__host__ __device__ __forceinline__ int64_t ftab_t_base(int a, int X, int Y,
                                                        int n)
{
    return a == 0 ? 0
                  : (a == 1 ? (int64_t)(X + 2) * n
                            : (int64_t)(X + 2 + Y + 2) * n);
}

// Tap range [a, b) of one axis for coordinate v and box [lo, hi]: the taps
// whose source coordinate v - o lies in the box, read from the cone-minor
// first-step tables (ft points at the axis block, rows = the axis'
// dimension + 2, n = number of cones, c = this lane's cone).  Out-of-range
// arguments clamp to the ends of the table, which yields empty ranges and
// keeps a and b monotone in v.
__device__ __forceinline__ void vrange_t(int v, int lo, int hi, int step,
                                         const int32_t* __restrict__ ft,
                                         int rows, int n, int c, int& a,
                                         int& b)
{
    int L, U;
    if (step > 0) {
        L = v - hi;
        U = v - lo;
    } else {
        L = lo - v;
        U = hi - v;
    }
    const int last = rows - 1;
    L = L < 0 ? 0 : (L > last ? last : L);
    U = U + 1 < 0 ? 0 : (U + 1 > last ? last : U + 1);
    a = ft[(int64_t)L * n + c];
    b = ft[(int64_t)U * n + c];
}

}  // namespace ch

struct ch_plan : ch::GraphHeld {
    int32_t shape[3];
    double vsize[3];
    int32_t n_cones;
    int32_t semantics;
    int32_t tap_stride;
    int32_t f_stride;
    int64_t n_taps_total;
    int32_t max_taps;
    int32_t tile[3];
    int32_t accum_mode;
    int32_t* d_lin;
    float* d_k;
    double* d_k64;
    int32_t* d_off;
    int32_t* d_lin2;
    int32_t* d_ftab;
    // The first-step tables again, cone-minor: axis a's entry q of cone c
    // at ftab_t_base(a) + q * n_cones + c for q < dims[a] + 2, entries past
    // a cone's own table repeating its last one (so lanes that look up the
    // same q for different cones read consecutive words).
    int32_t* d_ftT;
    ch::ConeMeta* d_meta;
    std::vector<ch::ConeMeta> h_meta;
    // CH_SEM_RADIOLOGICAL plans only: the kernel rows (n_rows x row_len
    // FP64), the cones ordered by row and each row's range in that order.
    double* d_cum;
    int32_t* d_perm;
    int32_t* d_row_begin;
    int32_t n_rows;
    int32_t row_len;
    // Spectral form (ch_plan_spectrum): frame lengths, the row pitch of the
    // half spectra, K^ (Fx x Fy x pitch complex) and per-axis twiddles
    // exp(-2 pi i q / F).  Null until built.
    int32_t fft_n[3];
    int32_t fft_kp;
    // lines per workgroup of each FFT-form pass (z forward, y forward,
    // x convolution, y inverse, z inverse); 0 = the default for the length
    int32_t fft_lines[5];
    float2* d_khat;
    float2* d_tw[3];
};
// Frees a plan's spectral tables (the plan itself stays).
void spectrum_free(ch_plan* p);
